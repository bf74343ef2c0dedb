"""GPU parity of COMB's bond-bending term (bond_bend): energy, forces and
virial alone and with COMB's charge terms, on cells small enough that one
atom appears at several images in a list, and through the MD loop.

Contract (FP64 on the device): |dE|/N <= 1e-9 eV; forces and virial
max|d| / max|ref| <= 1e-9; MD positions within 1e-9 A of velocity Verlet
on the oracle after 10 steps.
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import (assert_same_pairs, comb_bend_terms, comb_charge_model, comb_charges,
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


def load(eng, sp, pos, q, cell, pd, cuts, masses=None, momenta=None):
    eng.create_atoms(sp, pos, charges=q, masses=masses, momenta=momenta)
    eng.create_cell(cell, (True, True, True))
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)))
    eng.ewald_off()
    eng.generate_neighbor_lists(cuts)


def compare(eng, ref, n, what):
    e_ref, f_ref, w_ref, x_ref = ref
    e = eng.energy()
    f, w = eng.forces()
    x = eng.electronegativities()
    print("  %s: |dE|/N %.2e  rel|dF| %.2e  rel|dW| %.2e" % (
        what, abs(e - e_ref) / n, rel_maxnorm(f, f_ref), rel_maxnorm(w, w_ref)), flush=True)
    assert abs(e - e_ref) / n <= 1e-9
    assert rel_maxnorm(f, f_ref) <= 1e-9
    assert rel_maxnorm(w, w_ref) <= 1e-9
    assert np.abs(x - x_ref).max() <= 1e-9 * max(1.0, np.abs(x_ref).max())


@pytest.mark.parametrize("name", ["sio2_48_comb_bend", "sio2_48_comb_bend_charge"])
def test_bend_golden(eng, name):
    g = load_golden(name)
    print("case: " + name, flush=True)
    load(eng, g["species"], g["positions"], g["charges"], g["cell"], g["potdesc"], g["cutoffs"])
    compare(eng, (float(g["energy"]), g["forces"], g["stress"], g["enegs"]),
            len(g["positions"]), "golden")


def _case(n_si, seed, charged):
    s = st.amorphous_sio2(n_si, 2.2, seed=seed)
    q = comb_charges(s.symbols, seed=seed + 1)
    names = sorted(set(s.symbols))
    pots = comb_bend_terms() + (comb_charge_model() if charged else [])
    pd, _, _ = build_descriptors(pots, {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    return s, q, sp, pd, np.full(len(s), 3.55)


def oracle_ref(oracle, eng, sp, pos, q, cell, pd, cuts):
    nl = oracle.neighbor_list(pos, cell, [1, 1, 1], cuts)
    assert_same_pairs(eng, nl)
    args = (sp, pos, q, cell, nl, pd)
    f, w = oracle.evaluate("forces", *args)
    return (oracle.evaluate("energy", *args), f, w, oracle.evaluate("enegs", *args))


# 512 Si: 1536 atoms; 4 Si: 12 atoms in a 5.7 A box, so a neighbor appears
# at several images within the 3.55 A list radius
@pytest.mark.parametrize("n_si,charged", [(512, False), (512, True), (4, False), (5, True)])
def test_bend_matches_oracle(eng, oracle, n_si, charged):
    s, q, sp, pd, cuts = _case(n_si, 91 + n_si, charged)
    print("case: amorphous SiO2 n=%d charged=%s" % (len(s), charged), flush=True)
    load(eng, sp, s.positions, q, s.cell, pd, cuts)
    compare(eng, oracle_ref(oracle, eng, sp, s.positions, q, s.cell, pd, cuts), len(s),
            "n=%d" % len(s))


def test_bend_md_matches_oracle(eng, oracle):
    """Bond bending and COMB's charge terms through the MD loop (10 steps of
    0.5 fs, lists rebuilt every step) against velocity Verlet on the
    oracle."""
    s, q, sp, pd, cuts = _case(16, 97, True)
    st.maxwell_boltzmann(s, 300.0, seed=3)
    n = len(s)
    print("case: bend MD n=%d" % n, flush=True)
    load(eng, sp, s.positions, q, s.cell, pd, cuts, masses=s.masses, momenta=s.momenta)
    dt, steps = 0.5 * st.ASE_FS, 10
    eng.md_run(steps, dt, rebuild=1)
    pos_g, mom_g = eng.state()
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cuts)
    f, _ = oracle.evaluate("forces", sp, r, q, s.cell, nl, pd)
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cuts)
        f, _ = oracle.evaluate("forces", sp, r, q, s.cell, nl, pd)
        p = p + 0.5 * dt * f
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  |dr| %.2e A  rel|dp| %.2e" % (dr, dp), flush=True)
    assert dr < 1e-9
    assert dp < 1e-9
