"""GPU parity of the pair-potential kernels (pair.hip): every pair type, a
product, a table and the one-body terms, alone and next to Tersoff, against
the golden vectors of the reference core and the oracle; molecular dynamics
with Lennard-Jones replayed from graphs against eager steps.

Contract (all arithmetic is FP64; the engine sums each pair from both ends,
the reference once, so results agree to rounding):
  energy    |E - E_ref| <= 1e-10 max(1, |E_ref|)
  forces    max|F - F_ref| <= 1e-10 max(1, max|F_ref|)
  stress    max|W - W_ref| <= 1e-10 max(1, max|W_ref|)
With Tersoff in the model its FP32 three-body terms set the tolerance (the
Tersoff contract: |dE|/N <= 1e-7 eV, relative force error <= 1e-5).
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import (MIXTURE_TABLE_ID, assert_same_pairs, mixture_model, mixture_table,
                     rel_maxnorm)
from pysic_amd import structures as st
from pysic_amd.interactions import (Potential, ProductPotential, build_descriptors, lj_argon,
                                    tersoff_si)

pytestmark = pytest.mark.gpu

PAIR_CASES = ["pairs_arkr_mixture_120", "lj_argon_fcc_108", "lj_argon_fcc_108_smooth"]
TOL = 1e-10


@pytest.fixture(scope="module")
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def setup(eng, pos, cell, pbc, sp, pd, cutoffs, tables=None, bd=None, masses=None,
          momenta=None):
    eng.create_atoms(sp, pos, masses=masses, momenta=momenta)
    eng.create_cell(cell, pbc)
    for tid, rows in (tables or {}).items():
        eng.set_potential_table(tid, rows)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)) if bd is None else bd)
    eng.ewald_off()
    eng.generate_neighbor_lists(cutoffs)


def close_enough(e, f, s, e_ref, f_ref, s_ref, tol=TOL):
    de = abs(e - e_ref) / max(1.0, abs(e_ref))
    df = np.abs(f - f_ref).max() / max(1.0, np.abs(f_ref).max())
    ds = np.abs(s - s_ref).max() / max(1.0, np.abs(s_ref).max())
    print("  dE %.2e  dF %.2e  dW %.2e" % (de, df, ds), flush=True)
    assert de <= tol and df <= tol and ds <= tol


@pytest.mark.parametrize("name", PAIR_CASES)
def test_golden_pairs(eng, name):
    g = load_golden(name)
    print("case: %s" % name, flush=True)
    tables = {int(g["table_id"]): g["table"]} if int(g["table_id"]) >= 0 else None
    setup(eng, g["positions"], g["cell"], g["pbc"], g["species"], g["potdesc"], g["cutoffs"],
          tables)
    e = eng.energy()
    f, s = eng.forces()
    close_enough(e, f, s, float(g["energy"]), g["forces"], g["stress"])


def _vs_oracle(eng, oracle, s, model, cutoff, tables=None, tol=TOL):
    names = sorted(set(s.symbols))
    pd, bd, _ = build_descriptors(model, {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    cut = np.full(len(s), cutoff)
    for tid, rows in (tables or {}).items():
        oracle.set_table(tid, rows)
    setup(eng, s.positions, s.cell, s.pbc, sp, pd, cut, tables, bd)
    nl = oracle.neighbor_list(s.positions, s.cell, s.pbc, cut)
    assert_same_pairs(eng, nl)
    args = (sp, s.positions, None, s.cell, nl, pd, bd)
    e_ref = oracle.evaluate("energy", *args)
    f_ref, s_ref = oracle.evaluate("forces", *args)
    e = eng.energy()
    f, w = eng.forces()
    close_enough(e, f, w, e_ref, f_ref, s_ref, tol)
    return e, f, w


def test_mixture_2000_atoms(eng, oracle):
    print("case: Ar/Kr mixture, 2000 atoms, every pair type", flush=True)
    s = st.random_mixture(2000, 46.0, 2.9, ["Kr", "Ar", "Ar"], seed=11)
    _vs_oracle(eng, oracle, s, mixture_model(), 8.5, {MIXTURE_TABLE_ID: mixture_table()})


def test_general_kernel_one_lane_per_atom(eng, oracle):
    """Above 2^18 atoms the general kind runs one thread per atom (below,
    8 lanes per atom, which every smaller case here exercises): Buckingham
    times a power term on 275,684 fcc atoms."""
    s = st.fcc((41, 41, 41), a=5.6, sigma=0.1, seed=3)
    assert len(s) > 1 << 18
    model = [ProductPotential([
        Potential("Buckingham", symbols=[["Ar", "Ar"]], parameters=[500.0, 5e3, 0.35],
                  cutoff=5.0, cutoff_margin=0.5),
        Potential("power", symbols=[["Ar", "Ar"]], parameters=[1.0, 4.0, 0.5])])]
    print("case: general kernel, n=%d" % len(s), flush=True)
    _vs_oracle(eng, oracle, s, model, 5.0)


def test_lj_argon_4000_atoms(eng, oracle):
    print("case: LJ argon fcc 10x10x10, 4000 atoms", flush=True)
    s = st.fcc((10, 10, 10), a=5.26, sigma=0.1, seed=3)
    _vs_oracle(eng, oracle, s, lj_argon(), 2.5 * 3.405 + 0.5)


def test_lj_next_to_tersoff(eng, oracle):
    """Tersoff plus a weak Si-Si LJ tail: the pair kernel adds its forces to
    the Tersoff gather's."""
    print("case: Tersoff + LJ, a-Si 216 atoms", flush=True)
    s = st.amorphous_si(216, 2.3, seed=7)
    model = tersoff_si() + [Potential("LJ", symbols=[["Si", "Si"]], parameters=[0.02, 2.2],
                                      cutoff=5.0, cutoff_margin=0.5)]
    pd, bd, _ = build_descriptors(model, {"Si": 0})
    sp = np.zeros(len(s), np.int32)
    cut = np.full(len(s), 5.5)
    setup(eng, s.positions, s.cell, s.pbc, sp, pd, cut, None, bd)
    nl = oracle.neighbor_list(s.positions, s.cell, s.pbc, cut)
    assert_same_pairs(eng, nl)
    args = (sp, s.positions, None, s.cell, nl, pd, bd)
    e_ref = oracle.evaluate("energy", *args)
    f_ref, w_ref = oracle.evaluate("forces", *args)
    e = eng.energy()
    f, w = eng.forces()
    n = len(s)
    print("  |dE|/N %.2e rel|dF| %.2e |dW|/N %.2e" % (abs(e - e_ref) / n,
                                                     rel_maxnorm(f, f_ref, 1e-2),
                                                     np.abs(w - w_ref).max() / n), flush=True)
    assert abs(e - e_ref) / n <= 1e-7
    assert rel_maxnorm(f, f_ref, 1e-2) <= 1e-5
    assert np.abs(w - w_ref).max() / n <= 1e-6


def _lj_md_engine(s, margin=0.0):
    from pysic_amd.engine import Engine
    pd, _, _ = build_descriptors(lj_argon(margin=margin), {"Ar": 0})
    e = Engine(0)
    setup(e, s.positions, s.cell, s.pbc, np.zeros(len(s), np.int32), pd,
          np.full(len(s), 2.5 * 3.405 + 0.5), masses=s.masses, momenta=s.momenta)
    return e


def test_lj_md_graph_matches_eager():
    s = st.fcc((8, 8, 8), a=5.26, sigma=0.05, seed=4)
    st.maxwell_boltzmann(s, 80.0, seed=5)
    dt = 2.0 * st.ASE_FS
    print("case: LJ argon 2048 atoms, 3 x (10 graph steps vs 10 eager steps)", flush=True)
    a, b = _lj_md_engine(s), _lj_md_engine(s)
    try:
        for rnd in range(3):
            a.md_run(10, dt, rebuild=1)
            b.timing(True)
            b.md_run(10, dt, rebuild=1)
            b.timing(False)
            pa, ma = a.state()
            pb, mb = b.state()
            assert np.array_equal(pa, pb) and np.array_equal(ma, mb), rnd
            assert a.energy() == b.energy()
    finally:
        a.close()
        b.close()


def test_lj_md_matches_oracle_velocity_verlet(oracle):
    """ASE VelocityVerlet semantics with lists rebuilt every 10 steps (the
    reference rebuilds when ASE's skin test says so): 30 steps of 2 fs."""
    s = st.fcc((5, 5, 5), a=5.26, sigma=0.08, seed=8)
    st.maxwell_boltzmann(s, 120.0, seed=9)
    n = len(s)
    print("case: LJ argon %d atoms, 30 VV steps vs the oracle" % n, flush=True)
    rc = 2.5 * 3.405
    cut = np.full(n, rc + 0.5)
    pd, _, _ = build_descriptors(lj_argon(cutoff=rc), {"Ar": 0})
    e = _lj_md_engine(s)
    try:
        dt = 2.0 * st.ASE_FS
        e_gpu = e.md_run(30, dt, rebuild=10, energies=True)
        pos_g, mom_g = e.state()
    finally:
        e.close()
    sp = np.zeros(n, np.int32)
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd)
    e_or = []
    for step in range(30):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        if (step + 1) % 10 == 0:
            nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd)
        p = p + 0.5 * dt * f
        e_or.append(oracle.evaluate("energy", sp, r, None, s.cell, nl, pd))
    print("  max|dr| %.2e  max|dp| %.2e  max|dE| %.2e" % (
        np.abs(pos_g - r).max(), np.abs(mom_g - p).max(),
        np.abs(np.array(e_gpu) - np.array(e_or)).max()), flush=True)
    assert np.abs(pos_g - r).max() < 1e-10
    assert np.abs(mom_g - p).max() < 1e-9 * np.abs(p).max()
    assert np.abs(np.array(e_gpu) - np.array(e_or)).max() <= 1e-10 * abs(e_or[-1])


def test_lj_md_conserves_energy():
    """Velocity Verlet, 2 fs, smoothed LJ cutoff, lists rebuilt every 10
    steps (0.5 A skin): total energy drifts by less than 1e-5 eV/atom over
    400 steps."""
    s = st.fcc((6, 6, 6), a=5.26, sigma=0.05, seed=6)
    st.maxwell_boltzmann(s, 60.0, seed=7)
    dt = 2.0 * st.ASE_FS
    print("case: LJ argon 864 atoms, 400 steps", flush=True)
    e = _lj_md_engine(s, margin=1.0)
    try:
        m = s.masses[:, None]

        def total():
            p = e.state()[1]
            return e.energy() + 0.5 * float(np.sum(p * p / m))

        e0 = total()
        for _ in range(40):
            e.md_run(10, dt, rebuild=10)
        e1 = total()
        print("  drift %.3e eV/atom" % ((e1 - e0) / len(s)), flush=True)
        assert abs(e1 - e0) / len(s) < 1e-5
    finally:
        e.close()


def test_calculator_reads_table_file(oracle, tmp_path, monkeypatch):
    """Pysic reads table_<id>.txt from the working directory, as the
    reference's create_potential does."""
    from pysic_amd import Pysic
    evaluate, neighbor_list, set_table = oracle.evaluate, oracle.neighbor_list, oracle.set_table
    monkeypatch.chdir(tmp_path)
    np.savetxt("table_%04d.txt" % MIXTURE_TABLE_ID, mixture_table())
    s = st.random_mixture(200, 21.2, 2.9, ["Kr", "Ar", "Ar"], seed=2)
    calc = Pysic(atoms=s, potentials=mixture_model())
    try:
        e = calc.get_potential_energy()
        f = calc.get_forces()
    finally:
        calc.engine.close()
    names = sorted(set(s.symbols))
    pd, _, _ = build_descriptors(mixture_model(), {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    set_table(MIXTURE_TABLE_ID, np.loadtxt("table_%04d.txt" % MIXTURE_TABLE_ID))
    nl = neighbor_list(s.positions, s.cell, s.pbc, calc.get_individual_cutoffs())
    e_ref = evaluate("energy", sp, s.positions, None, s.cell, nl, pd)
    f_ref, _ = evaluate("forces", sp, s.positions, None, s.cell, nl, pd)
    assert abs(e - e_ref) <= TOL * max(1.0, abs(e_ref))
    assert np.abs(f - f_ref).max() <= TOL * max(1.0, np.abs(f_ref).max())


def _sheared(s, shear):
    """The structure under a shear of the cell (fractional coordinates kept)."""
    frac = np.linalg.solve(s.cell.T, s.positions.T).T
    cell = s.cell + np.array(shear) * s.cell[0, 0]
    t = s.copy()
    t.cell = cell
    t.set_positions(frac @ cell)
    return t


def test_triclinic_cell(eng, oracle):
    """Pairs and Tersoff together in a triclinic cell: every image shift
    mixes all three cell vectors."""
    print("case: Ar/Kr mixture and Si Tersoff, triclinic cells", flush=True)
    s = _sheared(st.random_mixture(400, 26.0, 2.9, ["Kr", "Ar", "Ar"], seed=21),
                 [[0, 0, 0], [0.31, 0, 0], [0.17, -0.23, 0]])
    _vs_oracle(eng, oracle, s, mixture_model(), 8.5, {MIXTURE_TABLE_ID: mixture_table()})
    si = _sheared(st.diamond_si((4, 4, 4), sigma=0.05, seed=6),
                  [[0, 0, 0], [0.2, 0, 0], [-0.1, 0.15, 0]])
    model = tersoff_si() + [Potential("Buckingham", symbols=[["Si", "Si"]],
                                      parameters=[300.0, 10.0, 0.4], cutoff=4.5)]
    e, f, w = _vs_oracle(eng, oracle, si, model, 5.0, tol=1e-5)
    assert np.isfinite(e)


def test_open_boundaries(eng, oracle):
    """An LJ cluster with no periodic axis and a slab periodic in x only."""
    print("case: LJ argon cluster and slab, open boundaries", flush=True)
    c = st.fcc((4, 4, 4), a=5.26, sigma=0.05, seed=12)
    c.pbc[:] = False
    c.cell = np.diag([40.0, 40.0, 40.0])
    _vs_oracle(eng, oracle, c, lj_argon(margin=1.0), 9.0)
    slab = st.fcc((4, 4, 4), a=5.26, sigma=0.05, seed=13)
    slab.pbc[:] = [True, False, False]
    slab.cell = np.diag([21.04, 40.0, 40.0])
    _vs_oracle(eng, oracle, slab, lj_argon(), 9.0)
