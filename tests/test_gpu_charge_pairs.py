"""GPU parity of COMB's charge-dependent pair terms (charge_exp,
charge_abs, charge_pair and their products): energy, forces, virial and their share of
the electronegativities, alone, under Ewald, and in the damped charge
relaxation.

Contract (FP64 on the device): |dE|/N <= 1e-9 eV; forces, virial and
electronegativities max|d| / max|ref| <= 1e-9; relaxed charges within
1e-9 e of the oracle's after the same steps.
"""
import math

import numpy as np
import pytest

from conftest import load_golden
from helpers import (assert_same_pairs, comb_asymmetric_terms, comb_charge_model, comb_charges,
                     rel_maxnorm)
from pysic_amd import structures as st
from pysic_amd.interactions import build_descriptors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def load(eng, sp, pos, q, cell, pd, cuts, ew=None):
    eng.create_atoms(sp, pos, charges=q)
    eng.create_cell(cell, (True, True, True))
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)))
    if ew is None:
        eng.ewald_off()
    else:
        rc, kc, K, sigma, eps = ew
        eng.set_ewald(rc, kc, [int(k) for k in K], sigma, eps)
    eng.generate_neighbor_lists(cuts)


def compare(eng, ref, n, what):
    e_ref, f_ref, w_ref, x_ref = ref
    e = eng.energy()
    f, w = eng.forces()
    x = eng.electronegativities()
    print("  %s: |dE|/N %.2e  rel|dF| %.2e  rel|dW| %.2e  rel|dchi| %.2e" % (
        what, abs(e - e_ref) / n, rel_maxnorm(f, f_ref), rel_maxnorm(w, w_ref),
        rel_maxnorm(x, x_ref)), flush=True)
    assert abs(e - e_ref) / n <= 1e-9
    assert rel_maxnorm(f, f_ref) <= 1e-9
    assert rel_maxnorm(w, w_ref) <= 1e-9
    assert rel_maxnorm(x, x_ref) <= 1e-9


def oracle_ref(oracle, eng, sp, pos, q, cell, pd, cuts, ew=None):
    nl = oracle.neighbor_list(pos, cell, [1, 1, 1], cuts)
    assert_same_pairs(eng, nl)
    ewa = None if ew is None else np.array([1, ew[0], ew[1], *ew[2], ew[3], ew[4]])
    args = (sp, pos, q, cell, nl, pd, None, ewa)
    f, w = oracle.evaluate("forces", *args)
    return (oracle.evaluate("energy", *args), f, w, oracle.evaluate("enegs", *args))


@pytest.mark.parametrize("name", ["sio2_48_comb_charge_pairs",
                                  "sio2_48_charge_terms_asymmetric"])
def test_comb_golden(eng, name):
    g = load_golden(name)
    print("case: " + name, flush=True)
    load(eng, g["species"], g["positions"], g["charges"], g["cell"], g["potdesc"], g["cutoffs"])
    compare(eng, (float(g["energy"]), g["forces"], g["stress"], g["enegs"]),
            len(g["positions"]), "golden")


def _model(s, asymmetric):
    names = sorted(set(s.symbols))
    pots = comb_charge_model() + (comb_asymmetric_terms() if asymmetric else [])
    pd, _, _ = build_descriptors(pots, {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    return sp, pd


@pytest.mark.parametrize("asymmetric", [False, True])
def test_comb_matches_oracle(eng, oracle, asymmetric):
    s = st.amorphous_sio2(512, 2.3, seed=41)
    q = comb_charges(s.symbols, seed=42)
    sp, pd = _model(s, asymmetric)
    cuts = np.full(len(s), 3.55)
    print("case: amorphous SiO2 n=%d asymmetric=%s" % (len(s), asymmetric), flush=True)
    load(eng, sp, s.positions, q, s.cell, pd, cuts)
    compare(eng, oracle_ref(oracle, eng, sp, s.positions, q, s.cell, pd, cuts), len(s), "n=1536")


def kcells(cell, kc):
    rec = 2 * math.pi * np.linalg.inv(cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
            for a in range(3)]


def _ewald_case(seed):
    s = st.amorphous_sio2(64, 2.3, seed=seed)
    q = comb_charges(s.symbols, seed=seed + 1)
    q -= q.mean()
    rc = 6.0
    sigma, kc = rc / 3, 9.0 / rc
    ew = (rc, kc, kcells(s.cell, kc), sigma, 0.00552635)
    return s, q, ew


def test_comb_under_ewald(eng, oracle):
    s, q, ew = _ewald_case(51)
    sp, pd = _model(s, True)
    cuts = np.full(len(s), ew[0] + 0.5)
    print("case: COMB + Ewald n=%d" % len(s), flush=True)
    load(eng, sp, s.positions, q, s.cell, pd, cuts, ew)
    compare(eng, oracle_ref(oracle, eng, sp, s.positions, q, s.cell, pd, cuts, ew), len(s),
            "ewald")


def test_comb_relaxation_matches_oracle(eng, oracle):
    """The damped charge dynamics over Ewald, EEM self terms and the COMB
    charge terms: 40 steps (no convergence test reached) from the same
    start, charges and rates against the oracle's."""
    s, q, ew = _ewald_case(61)
    names = sorted(set(s.symbols))
    sp, pd = _model(s, False)
    eem = {"Si": (4.77, 16.0), "O": (8.74, 20.0)}
    rows = []
    for k, sym in enumerate(names):
        chi0, J = eem[sym]
        rows.append([2, 1, k, 0, chi0, 1.0, 0.0, 0.0, -1, 0, 0, 0])
        rows.append([2, 1, k, 0, 0.5 * J, 2.0, 0.0, 0.0, -1, 0, 0, 0])
    pd = np.vstack([pd, np.array(rows, dtype=np.float64)])
    n = len(s)
    cuts = np.full(n, ew[0] + 0.5)
    print("case: COMB relaxation n=%d" % n, flush=True)
    load(eng, sp, s.positions, q, s.cell, pd, cuts, ew)
    steps, dt, inertia, friction, tol = 40, 0.02, 1.0, 4.0, 1e-14
    qg, rg, ng, _ = eng.relax_charges(q, np.zeros(n), steps, dt, inertia, friction, tol)
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cuts)
    ewa = np.array([1, ew[0], ew[1], *ew[2], ew[3], ew[4]])
    qo, ro, no, _ = oracle.relax_dynamic(sp, s.positions, q, s.cell, nl, pd, ewa, steps, dt,
                                         inertia, friction, tol)
    print("  steps gpu %d oracle %d  max|dq| %.2e" % (ng, no, np.abs(qg - qo).max()), flush=True)
    assert ng == no
    assert np.abs(qg - qo).max() <= 1e-9
    assert np.abs(rg - ro).max() <= 1e-9
