"""GPU parity of the Ewald summation, the EEM electronegativities and the
damped charge relaxation (all FP64 on the device).

Contract (checked per case): energy |dE|/N <= 1e-9 eV; forces and
electronegativities max|d| / max|ref| <= 1e-9; stress max|d| / N <= 1e-9 eV;
relaxation: the same step count as the oracle and charges within 1e-9 e.
"""
import math

import numpy as np
import pytest

from conftest import EWALD_CASES, load_golden
from helpers import assert_same_pairs, ewald_descriptors, rel_maxnorm
from pysic_amd import structures as st

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def kcells(cell, kc):
    rec = 2 * math.pi * np.linalg.inv(cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
            for a in range(3)]


def load_ewald(eng, species, pos, q, cell, pd, ew, scaler=None):
    n = len(pos)
    eng.create_atoms(species, pos, charges=q)
    eng.create_cell(cell, (True, True, True))
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)))
    _, rc, kc, k1, k2, k3, sigma, eps = ew
    eng.set_ewald(rc, kc, [int(k1), int(k2), int(k3)], sigma, eps, scaler=scaler)
    eng.generate_neighbor_lists(np.full(n, rc + 0.5))


def check(eng, oracle, species, pos, q, cell, pd, ew, ref=None, scaler=None):
    n = len(pos)
    print("case: n=%d" % n, flush=True)
    load_ewald(eng, species, pos, q, cell, pd, ew, scaler)
    nl = oracle.neighbor_list(pos, cell, [1, 1, 1], np.full(n, ew[1] + 0.5))
    assert_same_pairs(eng, nl)
    if ref is None:
        args = (species, pos, q, cell, nl, pd, None, ew, scaler)
        e_ref = oracle.evaluate("energy", *args)
        f_ref, s_ref = oracle.evaluate("forces", *args)
        x_ref = oracle.evaluate("enegs", *args)
    else:
        e_ref, f_ref, s_ref, x_ref = ref
    e = eng.energy()
    f, s = eng.forces()
    x = eng.electronegativities()
    print("  |dE|/N %.2e  rel|dF| %.2e  |dW|/N %.2e  rel|dchi| %.2e" % (
        abs(e - e_ref) / n, rel_maxnorm(f, f_ref), np.abs(s - s_ref).max() / n,
        rel_maxnorm(x, x_ref)), flush=True)
    assert abs(e - e_ref) / n <= 1e-9
    assert rel_maxnorm(f, f_ref) <= 1e-9
    assert np.abs(s - s_ref).max() / n <= 1e-9
    assert rel_maxnorm(x, x_ref) <= 1e-9


@pytest.mark.parametrize("name", EWALD_CASES)
def test_golden_cases(eng, oracle, name):
    g = load_golden(name)
    sp, pd, ew = ewald_descriptors(g)
    ref = (float(g["energy"]), g["forces"], g["stress"], g["enegs"])
    check(eng, oracle, sp, g["positions"], g["charges"], g["cell"], pd, ew, ref, g["scaler"])


def test_rocksalt_8x8x8_rc10(eng, oracle):
    # the Ewald workload's parameters (real cutoff 10 A, 1e-6 accuracy:
    # sigma = 2.0443, k_c = 2.5713) on an 8^3 replica of rocksalt
    s = st.rocksalt((8, 8, 8), sigma=0.08, seed=5)
    rc, sigma, kc = 10.0, 2.04430, 2.5713
    K = kcells(s.cell, kc)
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
    sp = (s.charges < 0).astype(np.int32)
    check(eng, oracle, sp, s.positions, s.charges, s.cell, np.zeros((0, 10)), ew)


def test_dense_bin_neighbor_scan(eng, oracle):
    """The neighbor scan taken for large dense cells (Ewald and LJ lists):
    the per-atom FP32-prefiltered scan with its wave-uniform stencil union
    (method 5; picked by default from 2^17 atoms).  Rocksalt 8^3 at
    rc + skin = 10.5 A has 4 bins per axis and 64 ions per bin, which selects
    the dense instance.  The sorted (i, j, offset) set must equal the
    oracle's bit for bit, and E, F, W and chi must agree to 1e-9."""
    s = st.rocksalt((8, 8, 8), sigma=0.08, seed=15)
    rc, sigma, kc = 10.0, 2.04430, 2.5713
    K = kcells(s.cell, kc)
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
    sp = (s.charges < 0).astype(np.int32)
    L = np.linalg.norm(s.cell, axis=1)
    bins = np.floor(L / (rc + 0.5))
    assert bins.min() >= 3 and len(s) / bins.prod() >= 6  # the dense instance's condition
    eng.set_neighbor_method(5)
    try:
        check(eng, oracle, sp, s.positions, s.charges, s.cell, np.zeros((0, 10)), ew)
        # a triclinic replica: bins and wrapped images along sheared axes
        shear = np.array([[1.0, 0.0, 0.0], [0.15, 1.0, 0.0], [-0.1, 0.1, 1.0]])
        cell, pos = s.cell @ shear, s.positions @ shear
        K = kcells(cell, kc)
        ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
        check(eng, oracle, sp, pos, s.charges, cell, np.zeros((0, 10)), ew)
    finally:
        eng.set_neighbor_method(0)


def test_structure_factors_follow_the_state(eng, oracle):
    """S(k) computed for an energy call is reused by a forces call on the
    same configuration; every change of positions, charges or cell between
    calls must reach the result (each step against the oracle)."""
    s = st.rocksalt((4, 4, 4), sigma=0.08, seed=7)
    rc, sigma, kc = 10.0, 2.04430, 2.5713
    K = kcells(s.cell, kc)
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
    sp = (s.charges < 0).astype(np.int32)
    pd = np.zeros((0, 10))
    pos, q, cell = s.positions.copy(), s.charges.copy(), s.cell.copy()
    n = len(pos)
    load_ewald(eng, sp, pos, q, cell, pd, ew)

    def agree(what):
        nl = oracle.neighbor_list(pos, cell, [1, 1, 1], np.full(n, rc + 0.5))
        e_ref = oracle.evaluate("energy", sp, pos, q, cell, nl, pd, None, ew)
        f_ref, s_ref = oracle.evaluate("forces", sp, pos, q, cell, nl, pd, None, ew)
        e = eng.energy()
        f, w = eng.forces()
        print("  %s: |dE|/N %.2e rel|dF| %.2e" % (what, abs(e - e_ref) / n, rel_maxnorm(f, f_ref)),
              flush=True)
        assert abs(e - e_ref) / n <= 1e-9, what
        assert rel_maxnorm(f, f_ref) <= 1e-9, what
        assert np.abs(w - s_ref).max() / n <= 1e-9, what

    agree("initial")
    pos[3] += [0.01, -0.02, 0.005]
    eng.update_coordinates(pos)
    agree("one atom moved")
    q *= 0.9
    eng.update_charges(q)
    agree("charges scaled")
    strain = np.diag([1.01, 0.995, 1.0])
    cell, pos = cell @ strain, pos @ strain
    eng.create_cell(cell, (True, True, True))
    eng.update_coordinates(pos)
    eng.generate_neighbor_lists(np.full(n, rc + 0.5))
    agree("cell strained")


def test_set_ewald_rejects_oversized_kcells(eng):
    from pysic_amd.engine import PysicEngineError
    s = st.rocksalt((1, 1, 1))
    eng.create_atoms((s.charges < 0).astype(np.int32), s.positions, charges=s.charges)
    eng.create_cell(s.cell, (True, True, True))
    with pytest.raises(PysicEngineError, match="reciprocal cells"):
        eng.set_ewald(5.0, 2.0, [4, 1024, 4], 1.5, 0.00552635)


def test_madelung(eng, oracle):
    s = st.rocksalt((4, 4, 4))
    rc, sigma = 10.0, 2.04430
    kc = 2.5713
    K = kcells(s.cell, kc)
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
    load_ewald(eng, (s.charges < 0).astype(np.int32), s.positions, s.charges, s.cell,
               np.zeros((0, 10)), ew)
    e = eng.energy() / len(s)
    ke = 1.0 / (4 * math.pi * 0.00552635)
    assert abs(e - (-1.747564594633 * ke / 5.64)) < 2e-5


# EEM hardness (eV/e^2) with bare-Coulomb Ewald: large enough that the charge
# Hessian restricted to sum(q) = 0 is positive definite (its smallest
# eigenvalue is ~4.6 eV/e^2 on random 1.5 A-minimum SiO2 samples), so the
# damped dynamics converges; chi0 in eV.
EEM = {"Si": (4.77, 16.0), "O": (8.74, 20.0)}


def eem_rows(names):
    rows = []
    for k, sym in enumerate(names):
        chi0, J = EEM[sym]
        rows.append([2, 1, k, 0, chi0, 1.0, 0.0, 0.0, -1, 0])
        rows.append([2, 1, k, 0, 0.5 * J, 2.0, 0.0, 0.0, -1, 0])
    return np.array(rows, dtype=np.float64)


def test_relaxation_matches_oracle(eng, oracle):
    g = load_golden("sio2_48_eem_rc5")
    sp, _, ew = ewald_descriptors(g)
    names = sorted(set(str(x) for x in g["symbols"]))
    pd = eem_rows(names)
    pos, cell = g["positions"], g["cell"]
    n = len(pos)
    q0 = np.zeros(n)
    params = dict(n_steps=400, dt=0.1, inertia=1.0, friction=4.0, tolerance=1e-7)
    load_ewald(eng, sp, pos, q0, cell, pd, ew)
    qg, rg, steps_g, err_g = eng.relax_charges(q0, np.zeros(n), params["n_steps"], params["dt"],
                                               params["inertia"], params["friction"],
                                               params["tolerance"])
    nl = oracle.neighbor_list(pos, cell, [1, 1, 1], np.full(n, ew[1] + 0.5))
    qo, ro, steps_o, err_o = oracle.relax_dynamic(sp, pos, q0, cell, nl, pd, ew,
                                                  params["n_steps"], params["dt"],
                                                  params["inertia"], params["friction"],
                                                  params["tolerance"])
    print("steps gpu %d oracle %d; err %.3e %.3e" % (steps_g, steps_o, err_g, err_o))
    assert steps_g == steps_o < params["n_steps"]
    assert err_g <= params["tolerance"]
    assert np.abs(qg - qo).max() < 1e-9
    assert np.abs(rg - ro).max() < 1e-9
    assert abs(qg.sum()) < 1e-10


def test_relaxation_reaches_eem_solution(eng, oracle):
    # chi is affine in q: chi(q) = chi(0) - H q.  Build H column by column
    # from the oracle, solve H q - nu = chi(0) with sum q = 0 directly, and
    # check the device relaxation converges onto that solution.
    s = st.amorphous_sio2(128, 2.2, seed=3)
    names = ["O", "Si"]
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    n = len(s)
    rc = 8.0
    sigma, kc = rc / 3, 9.0 / rc
    K = kcells(s.cell, kc)
    ew = np.array([1, rc, kc, K[0], K[1], K[2], sigma, 0.00552635])
    pd = eem_rows(names)
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(n, rc + 0.5))
    x0 = oracle.evaluate("enegs", sp, s.positions, np.zeros(n), s.cell, nl, pd, None, ew)
    H = np.zeros((n, n))
    for j in range(n):
        q = np.zeros(n)
        q[j] = 1.0
        H[:, j] = x0 - oracle.evaluate("enegs", sp, s.positions, q, s.cell, nl, pd, None, ew)
    A = np.zeros((n + 1, n + 1))
    A[:n, :n] = H
    A[:n, n] = -1.0
    A[n, :n] = 1.0
    b = np.concatenate([x0, [0.0]])
    q_star = np.linalg.solve(A, b)[:n]
    load_ewald(eng, sp, s.positions, np.zeros(n), s.cell, pd, ew)
    q, r, steps, err = eng.relax_charges(np.zeros(n), np.zeros(n), 2000, 0.1, 1.0, 4.0, 1e-9)
    print("steps %d err %.2e  max|q - q*| %.2e" % (steps, err, np.abs(q - q_star).max()))
    assert steps < 2000
    assert np.abs(q - q_star).max() < 1e-8
