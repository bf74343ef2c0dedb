"""Cells thinner than the neighbor-list radius: a periodic axis whose width is
below cutoff + skin.  The reference hands such cells to ASE's NeighborList
(calculator.py:798-850); the engine scans several bins on each side.

Checks (each case names itself before it runs):
  * neighbor lists bit-exact as sorted (i, j, offset) sets against the
    oracle's enumeration of periodic images (oracle.image_neighbor_list,
    pinned in tests/test_oracle.py against replicated cells), including an
    atom's own images;
  * Tersoff and LJ energy, forces and virial against the oracle fed those
    lists, to the Tersoff contract (|dE|/N <= 1e-7 eV, relative force error
    <= 1e-5, |dW|/N <= 1e-6) and the FP64 pair contract (1e-10);
  * the physics: a thin cell and its 4x4x4 replica (an ordinary cell) have
    the same energy per atom and the same forces;
  * 20 MD steps of the thin 2-atom cell follow the replica's trajectory
    (the FP32 pair arithmetic of the MD force pass sums in another order:
    the bounds are the MD contract's, |dr| <= 1e-5 of the travel and |dp|
    <= 1e-5 of max |p|).
The lists of ASE itself (radii summed, skin added) are not reproduced: ASE
is not installed here, so that list content is parity unpinned; the
energies and forces depend only on the pairs within each cutoff.
"""
import numpy as np
import pytest

from helpers import engine_pair_set, rel_maxnorm, tersoff_descriptors
from pysic_amd import structures as st
from pysic_amd.interactions import build_descriptors, lj_argon

pytestmark = pytest.mark.gpu

A0 = 5.431


@pytest.fixture(scope="module")
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def load(eng, pos, cell, pbc, pd, bd, cutoffs, masses=None, momenta=None):
    n = len(pos)
    eng.create_atoms(np.zeros(n, np.int32), pos, masses=masses, momenta=momenta)
    eng.create_cell(cell, pbc)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)) if bd is None else bd)
    eng.ewald_off()
    eng.generate_neighbor_lists(cutoffs)


def primitive_si():
    """2-atom FCC primitive Si cell (perpendicular widths 3.135 A), one atom
    displaced and one outside the box."""
    cell = 0.5 * A0 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0], [0.25 * A0 + 0.05, 0.25 * A0 - 0.03, 0.25 * A0 + 0.02]])
    pos += 0.1234
    pos[1] += 2.0 * cell[1]
    return pos, cell


def cubic_si(sigma=0.08, seed=5):
    s = st.diamond_si((1, 1, 1), a=A0, sigma=sigma, seed=seed)
    return s.positions.copy(), s.cell.copy()


def replica(pos, cell, k=4):
    reps = np.array([[x, y, z] for x in range(k) for y in range(k) for z in range(k)])
    big = (pos[None, :, :] + (reps @ cell)[:, None, :]).reshape(-1, 3)
    return big, k * cell


CASES = {
    "primitive Si 2 atoms, r 3.5": lambda: (*primitive_si(), 3.5),
    "cubic Si 8 atoms, r 8.0": lambda: (*cubic_si(), 8.0),
    "simple cubic 1 atom a 2.5, r 6.0": lambda: (np.array([[0.3, 0.7, 1.1]]), np.eye(3) * 2.5,
                                                 6.0),
    "Si slab 2x2x1 cells, r 7.0": lambda: (st.diamond_si((2, 2, 1), a=A0, sigma=0.05,
                                                         seed=3).positions,
                                           np.diag([2 * A0, 2 * A0, A0]), 7.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_thin_cell_neighbor_lists(eng, oracle, name):
    pos, cell, rc = CASES[name]()
    n = len(pos)
    print("case: %s" % name, flush=True)
    pd, bd = tersoff_descriptors()
    load(eng, pos, cell, [1, 1, 1], pd, bd, np.full(n, rc))
    got = engine_pair_set(eng)
    want = oracle.image_neighbor_list(pos, cell, [1, 1, 1], np.full(n, rc)).pair_set()
    print("  pairs %d (oracle %d); self images %d" % (len(got), len(want),
                                                     int((want[:, 0] == want[:, 1]).sum())),
          flush=True)
    assert got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.parametrize("which", ["primitive", "cubic"])
def test_thin_cell_tersoff(eng, oracle, which):
    pos, cell = primitive_si() if which == "primitive" else cubic_si()
    n = len(pos)
    print("case: Tersoff on the %s cell, %d atoms, list radius 3.5 A" % (which, n), flush=True)
    pd, bd = tersoff_descriptors()
    cut = np.full(n, 3.5)
    load(eng, pos, cell, [1, 1, 1], pd, bd, cut)
    e = eng.energy()
    f, w = eng.forces()
    nl = oracle.image_neighbor_list(pos, cell, [1, 1, 1], cut)
    sp = np.zeros(n, np.int32)
    e_ref = oracle.evaluate("energy", sp, pos, None, cell, nl, pd, bd)
    f_ref, w_ref = oracle.evaluate("forces", sp, pos, None, cell, nl, pd, bd)
    print("  |dE|/N %.2e rel|dF| %.2e |dW|/N %.2e" % (
        abs(e - e_ref) / n, rel_maxnorm(f, f_ref, 1e-2), np.abs(w - w_ref).max() / n),
        flush=True)
    assert abs(e - e_ref) / n <= 1e-7
    assert rel_maxnorm(f, f_ref, 1e-2) <= 1e-5
    assert np.abs(w - w_ref).max() / n <= 1e-6
    # the 4x4x4 replica is an ordinary cell: same energy per atom, same forces
    big, bcell = replica(pos, cell)
    load(eng, big, bcell, [1, 1, 1], pd, bd, np.full(len(big), 3.5))
    eb = eng.energy()
    fb, wb = eng.forces()
    print("  replica: |E/N - E_big/N| %.2e  max|F - F_big| %.2e" % (
        abs(e / n - eb / len(big)), np.abs(fb[:n] - f).max()), flush=True)
    assert abs(e / n - eb / len(big)) <= 1e-7
    assert np.abs(fb[:n] - f).max() <= 1e-5 * max(1e-2, np.abs(f).max())
    assert np.abs(wb / len(big) - w / n).max() <= 1e-6


def test_thin_cell_lj(eng, oracle):
    pos, cell = cubic_si(0.05, 9)
    n = len(pos)
    print("case: LJ (Ar parameters on a Si lattice), 8 atoms, cutoff 8.5 A", flush=True)
    pd, _, _ = build_descriptors(lj_argon(symbol="Si"), {"Si": 0})
    cut = np.full(n, 9.0)
    load(eng, pos, cell, [1, 1, 1], pd, None, cut)
    e = eng.energy()
    f, w = eng.forces()
    nl = oracle.image_neighbor_list(pos, cell, [1, 1, 1], cut)
    sp = np.zeros(n, np.int32)
    e_ref = oracle.evaluate("energy", sp, pos, None, cell, nl, pd)
    f_ref, w_ref = oracle.evaluate("forces", sp, pos, None, cell, nl, pd)
    de = abs(e - e_ref) / max(1.0, abs(e_ref))
    df = np.abs(f - f_ref).max() / max(1.0, np.abs(f_ref).max())
    dw = np.abs(w - w_ref).max() / max(1.0, np.abs(w_ref).max())
    print("  dE %.2e dF %.2e dW %.2e" % (de, df, dw), flush=True)
    assert de <= 1e-10 and df <= 1e-10 and dw <= 1e-10


def test_thin_cell_md_follows_replica(eng):
    pos, cell = primitive_si()
    n = len(pos)
    print("case: 20 VV steps, primitive Si cell vs its 4x4x4 replica", flush=True)
    pd, bd = tersoff_descriptors()
    rng = np.random.default_rng(4)
    m = np.full(n, 28.0855)
    p = rng.normal(0, 0.3, (n, 3))
    p -= p.mean(axis=0)
    load(eng, pos, cell, [1, 1, 1], pd, bd, np.full(n, 3.5), masses=m, momenta=p)
    dt = st.ASE_FS
    eng.md_run(20, dt, rebuild=1)
    r_thin, p_thin = eng.state()
    big, bcell = replica(pos, cell)
    nb = len(big)
    load(eng, big, bcell, [1, 1, 1], pd, bd, np.full(nb, 3.5), masses=np.full(nb, 28.0855),
         momenta=np.tile(p, (nb // n, 1)))
    eng.md_run(20, dt, rebuild=1)
    r_big, p_big = eng.state()
    dr = np.abs(r_big[:n] - r_thin).max()
    dp = np.abs(p_big[:n] - p_thin).max()
    travel = np.abs(r_thin - pos).max()
    print("  max |dr| %.2e  max |dp| %.2e (travel %.3f A)" % (dr, dp, travel), flush=True)
    assert travel > 1e-3
    assert dr <= 1e-5 * travel and dp <= 1e-5 * np.abs(p_thin).max()
