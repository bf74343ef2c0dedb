"""GPU parity of the Tersoff path: HIP engine vs the oracle and the golden
vectors of the reference core.

Contract (tolerances written here, checked per case):
  neighbor lists   bit-exact as sorted (i, j, offset) sets
  energy           |E - E_ref| / N <= 1e-7 eV/atom
  forces           max|F - F_ref| / max(max|F_ref|, 0.01 eV/A) <= 1e-5
                   (the floor only matters for the perfect lattice, where
                   F_ref is zero by symmetry)
  stress (virial)  max|W - W_ref| / N <= 1e-6 eV/atom (W is extensive; its
                   three-body part carries FP32 rounding that is systematic
                   on a perfect lattice, where W is a sum of +-50 eV terms)
"""
import numpy as np
import pytest

from conftest import TERSOFF_CASES, load_golden
from helpers import assert_same_pairs, rel_maxnorm, tersoff_descriptors
from pysic_amd import structures as st

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pysic_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def load_system(eng, pos, cell, pbc, pd, bd, cutoffs, species=None, masses=None, momenta=None):
    n = len(pos)
    eng.create_atoms(np.zeros(n, np.int32) if species is None else species, pos, masses=masses,
                     momenta=momenta)
    eng.create_cell(cell, pbc)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(bd)
    eng.ewald_off()
    eng.generate_neighbor_lists(cutoffs)


def pair_set(counts, idx, off):
    i = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    rows = np.column_stack([i, idx.astype(np.int64), off.astype(np.int64)])
    return rows[np.lexsort(rows.T[::-1])]


def check_tersoff(eng, oracle, pos, cell, pbc, pd, bd, cutoffs, ref=None, species=None):
    n = len(pos)
    print("case: n=%d" % n, flush=True)
    load_system(eng, pos, cell, pbc, pd, bd, cutoffs, species=species)
    nl = oracle.neighbor_list(pos, cell, pbc, cutoffs)
    counts, idx, off = eng.neighbor_lists()
    assert np.array_equal(pair_set(counts, idx, off), nl.pair_set())
    sp = np.zeros(n, np.int32) if species is None else species
    if ref is None:
        e_ref = oracle.evaluate("energy", sp, pos, None, cell, nl, pd, bd)
        f_ref, s_ref = oracle.evaluate("forces", sp, pos, None, cell, nl, pd, bd)
    else:
        e_ref, f_ref, s_ref = ref
    e = eng.energy()
    f, s = eng.forces()
    err_e = abs(e - e_ref) / n
    err_f = rel_maxnorm(f, f_ref, 1e-2)
    err_s = float(np.abs(s - s_ref).max()) / n
    print("  |dE|/N %.3e  rel|dF| %.3e  |dW|/N %.3e" % (err_e, err_f, err_s), flush=True)
    assert err_e <= 1e-7
    assert err_f <= 1e-5
    assert err_s <= 1e-6
    return e, f, s


@pytest.mark.parametrize("name", TERSOFF_CASES + ["si_diamond_8x8x8_s005_seed0"])
def test_golden_cases(eng, oracle, name):
    g = load_golden(name)
    pd, bd = tersoff_descriptors(g["tersoff"])
    ref = (float(g["energy"]), g["forces"], g["stress"])
    check_tersoff(eng, oracle, g["positions"], g["cell"], g["pbc"], pd, bd, g["cutoffs"], ref)


def test_mixed_species_golden(eng, oracle):
    """Si/C Tersoff (8 triplet factors, Si-Si, Si-C and C-C pair terms)
    against the reference core, in every site-kernel shape."""
    g = load_golden("sic_mixed_192")
    names = sorted(set(str(x) for x in g["symbols"]))
    sp = np.array([names.index(str(x)) for x in g["symbols"]], np.int32)
    ref = (float(g["energy"]), g["forces"], g["stress"])
    for v in range(len(eng.variants())):
        eng.set_variant(v)
        print("  variant %s" % eng.variants()[v], flush=True)
        check_tersoff(eng, oracle, g["positions"], g["cell"], g["pbc"], g["potdesc"],
                      g["bonddesc"], g["cutoffs"], ref, species=sp)
    eng.set_variant(0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_config1_seeds(eng, oracle, seed):
    # Tersoff Si 8x8x8 (4096 atoms), Gaussian displacements 0.05 A
    s = st.diamond_si((8, 8, 8), sigma=0.05, seed=seed)
    pd, bd = tersoff_descriptors()
    check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))


def test_perfect_lattice_closed_form(eng, oracle):
    s = st.diamond_si((8, 8, 8))
    pd, bd = tersoff_descriptors()
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
    e = eng.energy()
    f, _ = eng.forces()
    assert abs(e / len(s) - (-4.629593127667)) <= 1e-7
    assert np.abs(f).max() < 1e-4  # zero by symmetry; FP32 three-body terms


def test_all_kernel_variants_agree(eng, oracle):
    s = st.amorphous_si(512, 2.3, seed=5)
    pd, bd = tersoff_descriptors()
    ref = None
    for v in range(len(eng.variants())):
        eng.set_variant(v)
        out = check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd,
                            np.full(len(s), 3.5))
        if ref is None:
            ref = out
    eng.set_variant(0)


def test_bitwise_deterministic(eng):
    s = st.diamond_si((6, 6, 6), sigma=0.1, seed=9)
    pd, bd = tersoff_descriptors()
    outs = []
    for _ in range(2):
        load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
        e = eng.energy()
        f, w = eng.forces()
        outs.append((e, f, w))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])


def test_neighbor_capacity_growth(eng, oracle):
    # a list capacity far below the real count is grown and the lists rebuilt
    s = st.amorphous_si(300, 2.3, seed=2)
    pd, bd = tersoff_descriptors()
    eng.set_neighbor_capacity(2)
    check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))


def test_short_list_capacity_switch(eng, oracle):
    """A compressed random Si cell (4.5 g/cm^3; up to 15 atoms within the
    3.0 A cutoff) with the 8-entry kernel shape selected: the evaluation
    finds the overflow, switches to a shape that holds it and runs again,
    so energy, forces and virial match the oracle.  A cell with 34 atoms
    within the cutoff is more than any shape holds (32): the call fails
    with that message, and the engine then evaluates a sound cell."""
    from pysic_amd.engine import PysicEngineError
    pd, bd = tersoff_descriptors()
    s = st.amorphous_si(256, 4.5, seed=5, dmin=1.8)
    near = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(len(s), 3.0)).counts
    print("case: compressed a-Si, max %d atoms within 3.0 A" % near.max(), flush=True)
    assert 8 < near.max() <= 16
    eng.set_variant(eng.variants().index("g4_s8"))
    check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
    d = st.amorphous_si(128, 12.0, seed=5, dmin=1.0)
    near = oracle.neighbor_list(d.positions, d.cell, [1, 1, 1], np.full(len(d), 3.0)).counts
    print("case: over-dense Si, max %d atoms within 3.0 A" % near.max(), flush=True)
    assert near.max() > 32
    eng.set_variant(0)
    load_system(eng, d.positions, d.cell, [1, 1, 1], pd, bd, np.full(len(d), 3.5))
    with pytest.raises(PysicEngineError, match="more short-range neighbors"):
        eng.forces()
    eng.set_variant(0)
    g = st.diamond_si((3, 3, 3), sigma=0.05, seed=6)
    check_tersoff(eng, oracle, g.positions, g.cell, [1, 1, 1], pd, bd, np.full(len(g), 3.5))


def test_md_matches_oracle_velocity_verlet(eng, oracle):
    # ASE VelocityVerlet semantics, list rebuilt every step; 20 steps of 1 fs
    s = st.diamond_si((4, 4, 4), sigma=0.1, seed=3)
    st.maxwell_boltzmann(s, 600.0, seed=4)
    pd, bd = tersoff_descriptors()
    n = len(s)
    cut = np.full(n, 3.5)
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, cut, masses=s.masses,
                momenta=s.momenta)
    dt = st.ASE_FS
    steps = 20
    e_gpu = eng.md_run(steps, dt, rebuild=1, energies=True)
    pos_g, mom_g = eng.state()
    # oracle VV
    sp = np.zeros(n, np.int32)
    m = s.masses[:, None]
    r = s.positions.copy()
    p = s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
    e_or = []
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
        p = p + 0.5 * dt * f
        e_or.append(oracle.evaluate("energy", sp, r, None, s.cell, nl, pd, bd))
    # the MD force pass carries the force contract (FP32 pair and three-body
    # arithmetic, 1e-5 relative): held as 1e-5 of the distance travelled and
    # of the momenta (measured: about 1e-6 of each)
    travel = np.abs(r - s.positions).max()
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  travel %.3f A  |dr| %.2e A  rel|dp| %.2e" % (travel, dr, dp), flush=True)
    assert dr < 1e-5 * travel
    assert dp < 1e-5
    assert np.abs(np.array(e_gpu) - np.array(e_or)).max() / n <= 1e-7


def test_md_triclinic_matches_oracle(eng, oracle):
    """The MD loop in a sheared cell (b and c tilted by 0.2 and (-0.1,
    0.15) of the cell length), 20 steps of 1 fs at 600 K with the lists
    rebuilt every step: 44 atoms cross the sheared periodic faces, and the
    drift's binning wraps them along all three cell vectors.  Same bounds as
    the cubic case."""
    s = st.diamond_si((4, 4, 4), sigma=0.1, seed=23)
    frac = np.linalg.solve(s.cell.T, s.positions.T).T
    L = s.cell[0, 0]
    s.cell = s.cell + np.array([[0, 0, 0], [0.2, 0, 0], [-0.1, 0.15, 0]]) * L
    s.set_positions(frac @ s.cell)
    st.maxwell_boltzmann(s, 600.0, seed=24)
    pd, bd = tersoff_descriptors()
    n = len(s)
    cut = np.full(n, 3.5)
    print("case: triclinic Si, %d atoms, 20 VV steps at 600 K" % n, flush=True)
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, cut, masses=s.masses,
                momenta=s.momenta)
    dt, steps = st.ASE_FS, 20
    eng.md_run(steps, dt, rebuild=1)
    pos_g, mom_g = eng.state()
    sp = np.zeros(n, np.int32)
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
        p = p + 0.5 * dt * f
    inv = np.linalg.inv(s.cell)
    crossed = np.any(np.floor(r @ inv) != np.floor(s.positions @ inv), axis=1).sum()
    travel = np.abs(r - s.positions).max()
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  %d atoms crossed a face  travel %.3f A  |dr| %.2e A  rel|dp| %.2e" % (
        crossed, travel, dr, dp), flush=True)
    assert crossed > 0
    assert dr < 1e-5 * travel
    assert dp < 1e-5


def test_md_open_boundaries_matches_oracle(eng, oracle):
    """A 512-atom Si cluster with no periodic axis in a box that fits it
    tightly: 47 atoms start outside the box and about as many stay outside
    while the cluster vibrates (20 steps of 1 fs at 600 K, lists rebuilt
    every step).  The lists equal the all-pairs search, and the trajectory
    follows velocity Verlet on the oracle's forces within the cubic case's
    bounds."""
    s = st.diamond_si((4, 4, 4), sigma=0.1, seed=31)
    st.maxwell_boltzmann(s, 600.0, seed=32)
    pd, bd = tersoff_descriptors()
    n = len(s)
    cut = np.full(n, 3.5)
    print("case: open-boundary Si cluster, %d atoms, 20 VV steps" % n, flush=True)
    load_system(eng, s.positions, s.cell, [0, 0, 0], pd, bd, cut, masses=s.masses,
                momenta=s.momenta)
    counts, idx, off = eng.neighbor_lists()
    d = np.linalg.norm(s.positions[:, None, :] - s.positions[None, :, :], axis=2)
    assert np.array_equal(counts, (d < 3.5).sum(axis=1) - 1)
    assert not off.any()
    dt, steps = st.ASE_FS, 20
    eng.md_run(steps, dt, rebuild=1)
    pos_g, mom_g = eng.state()
    sp = np.zeros(n, np.int32)
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [0, 0, 0], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [0, 0, 0], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
        p = p + 0.5 * dt * f
    L = s.cell[0, 0]
    outside = ((r < 0) | (r >= L)).any(axis=1).sum()
    travel = np.abs(r - s.positions).max()
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  %d atoms outside the box  travel %.3f A  |dr| %.2e A  rel|dp| %.2e" % (
        outside, travel, dr, dp), flush=True)
    assert outside > 0
    assert dr < 1e-5 * travel
    assert dp < 1e-5


@pytest.mark.parametrize("cells,sigma", [((8, 8, 8), 0.05), ((6, 6, 6), 0.1)])
def test_md_force_pass_matches_oracle(eng, oracle, cells, sigma):
    """The force pass of a device MD step (the forces-only site kernel, FP32
    pair and three-body arithmetic) against the oracle's forces at the same
    positions, to the force contract (1e-5 relative max-norm).  One step
    from forces of the API pass F0: p1 = p0 + dt/2 (F0 + F1), so
    F1 = 2 (p1 - p0) / dt - F0 at the step's end positions."""
    s = st.diamond_si(cells, sigma=sigma, seed=17)
    st.maxwell_boltzmann(s, 600.0, seed=18)
    pd, bd = tersoff_descriptors()
    n = len(s)
    cut = np.full(n, 3.5)
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, cut, masses=s.masses,
                momenta=s.momenta)
    f0, _ = eng.forces()
    _, p0 = eng.state()
    dt = st.ASE_FS
    eng.md_run(1, dt, rebuild=1)
    r1, p1 = eng.state()
    f1 = 2.0 * (p1 - p0) / dt - f0
    nl = oracle.neighbor_list(r1, s.cell, [1, 1, 1], cut)
    f_ref, _ = oracle.evaluate("forces", np.zeros(n, np.int32), r1, None, s.cell, nl, pd, bd)
    err = rel_maxnorm(f1, f_ref, 1e-2)
    print("case: MD force pass n=%d  rel|dF| %.2e" % (n, err), flush=True)
    assert err <= 1e-5


def test_mixed_species_md_matches_oracle(eng, oracle):
    """Si/C Tersoff MD through the device loop (graph-replayed steps, force
    gather fused into the next step's drift) against velocity Verlet on the
    oracle: 10 steps of 1 fs, lists rebuilt every step.  Forces in this
    random cell reach ~30 eV/A, so the FP32 three-body rounding (the 1e-5
    force contract) is held as 1e-5 of the distance travelled and of the
    momenta."""
    from helpers import sic_model
    from pysic_amd.interactions import build_descriptors
    s = st.random_mixture(150, (150 / 0.05) ** (1.0 / 3.0), 1.7, ["Si", "Si", "C"], seed=31)
    st.maxwell_boltzmann(s, 300.0, seed=5)
    spmap = {"C": 0, "Si": 1}
    pd, bd, _ = build_descriptors(sic_model(), spmap)
    sp = np.array([spmap[x] for x in s.symbols], np.int32)
    n = len(s)
    cut = np.full(n, 3.5)
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, cut, species=sp, masses=s.masses,
                momenta=s.momenta)
    dt = st.ASE_FS
    steps = 10
    eng.md_run(steps, dt, rebuild=1)
    pos_g, mom_g = eng.state()
    m = s.masses[:, None]
    r = s.positions.copy()
    p = s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
        p = p + 0.5 * dt * f
    travel = np.abs(r - s.positions).max()
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  travel %.3f A  |dr| %.2e A  rel|dp| %.2e" % (travel, dr, dp), flush=True)
    assert dr < 1e-5 * travel
    assert dp < 1e-5


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5])
def test_neighbor_methods_agree(eng, oracle, method):
    # global per-atom scan, LDS-staged bricks, and the bricks' global fallback
    s = st.amorphous_si(700, 2.3, seed=13)
    pd, bd = tersoff_descriptors()
    eng.set_neighbor_method(method)
    try:
        check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
        t = st.diamond_si((5, 4, 3), sigma=0.2, seed=21)  # uneven bins, wrapped halo
        check_tersoff(eng, oracle, t.positions, t.cell, [1, 1, 1], pd, bd, np.full(len(t), 3.5))
    finally:
        eng.set_neighbor_method(0)


def test_full_size_md_workload(oracle):
    """The 1,048,576-atom Tersoff MD workload (64 x 64 x 32 diamond cells,
    sigma 0.1 A, Maxwell-Boltzmann 600 K), set up as the benchmark sets it
    up: the default neighbor scan for a cell of this size, the site-kernel
    shape the autotuner picks, lists rebuilt every step.  Checked against
    the oracle on the host: the sorted (i, j, offset) set, the FP64 energy,
    the forces of the API call, and two device MD steps (the graph-free
    first step and the fused gather + drift + binning of the next), whose
    momenta carry the forces of the MD path's own site kernel."""
    import bench
    from pysic_amd import autotune
    from pysic_amd.engine import Engine
    s = bench.build_workload((64, 64, 32))
    n = len(s)
    assert n == 1 << 20
    pd, bd = tersoff_descriptors()
    cut = np.full(n, 3.5)
    eng = Engine(0)
    try:
        eng.create_atoms(np.zeros(n, np.int32), s.positions, masses=s.masses, momenta=s.momenta)
        eng.create_cell(s.cell, s.pbc)
        eng.set_potentials(pd)
        eng.set_bond_order_factors(bd)
        eng.ewald_off()
        eng.set_neighbor_capacity(16)
        eng.generate_neighbor_lists(cut)
        variant = autotune.best_variant(eng, "tersoff_si")
        eng.set_variant(variant)
        print("case: full-size workload n=%d variant %s" % (n, eng.variants()[variant]),
              flush=True)
        sp = np.zeros(n, np.int32)
        nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
        assert_same_pairs(eng, nl)
        e_ref = oracle.evaluate("energy", sp, s.positions, None, s.cell, nl, pd, bd)
        f0, _ = oracle.evaluate("forces", sp, s.positions, None, s.cell, nl, pd, bd)
        e = eng.energy()
        f, _ = eng.forces()
        err_e, err_f = abs(e - e_ref) / n, rel_maxnorm(f, f0, 1e-2)
        print("  |dE|/N %.2e  rel|dF| %.2e" % (err_e, err_f), flush=True)
        assert err_e <= 1e-7
        assert err_f <= 1e-5
        # two velocity-Verlet steps on the device against the oracle's
        dt = st.ASE_FS
        eng.md_run(2, dt, rebuild=1)
        pos_g, mom_g = eng.state()
        m = s.masses[:, None]
        r, p, f_or = s.positions.copy(), s.momenta.copy(), f0
        for step in range(2):
            p = p + 0.5 * dt * f_or
            r = r + dt * p / m
            nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
            f_or, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
            p = p + 0.5 * dt * f_or
            print("  oracle step %d done" % (step + 1), flush=True)
        # each step's momentum change is dt/2 (F_old + F_new): a force error
        # of 1e-5 max|F| per evaluation bounds the difference by 2e-5 dt max|F|
        dp = np.abs(mom_g - p).max() / (dt * np.abs(f0).max())
        dr = np.abs(pos_g - r).max()
        print("  MD: |dr| %.2e A  |dp| / (dt max|F|) %.2e" % (dr, dp), flush=True)
        assert dp <= 2e-5
        assert dr <= 2e-5 * dt * dt * np.abs(f0).max() / s.masses.min() + 1e-12
    finally:
        eng.close()


@pytest.mark.parametrize("cells", [(1, 1, 1), (2, 1, 1)])
def test_small_cells_column_gather(eng, oracle, cells):
    """Cells narrower than twice the list radius (5.431 A across, 3.5 A
    radius): the engine cannot rule out an atom reaching two images of
    another, so its force gather matches bonds by (index, image) through the
    column-major short lists instead of the index rows.  API forces and a
    device MD run against the oracle."""
    s = st.diamond_si(cells, sigma=0.1, seed=23)
    st.maxwell_boltzmann(s, 600.0, seed=24)
    pd, bd = tersoff_descriptors()
    n = len(s)
    cut = np.full(n, 3.5)
    check_tersoff(eng, oracle, s.positions, s.cell, [1, 1, 1], pd, bd, cut)
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, cut, masses=s.masses,
                momenta=s.momenta)
    dt, steps = st.ASE_FS, 10
    eng.md_run(steps, dt, rebuild=1)
    pos_g, mom_g = eng.state()
    sp = np.zeros(n, np.int32)
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
    for _ in range(steps):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], cut)
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd, bd)
        p = p + 0.5 * dt * f
    travel = np.abs(r - s.positions).max()
    dr, dp = np.abs(pos_g - r).max(), np.abs(mom_g - p).max() / np.abs(p).max()
    print("  small cell MD: travel %.3f A |dr| %.2e A rel|dp| %.2e" % (travel, dr, dp), flush=True)
    assert dr < 1e-5 * travel
    assert dp < 1e-5


def test_coincident_atoms_report_non_finite_forces(eng):
    # two atoms at one position make the bond direction 0/0: the gather flags
    # the non-finite force and the call fails instead of returning NaN
    from pysic_amd.engine import PysicEngineError
    s = st.diamond_si((2, 2, 2), sigma=0.0, seed=0)
    pos = s.positions.copy()
    pos[5] = pos[3]
    pd, bd = tersoff_descriptors()
    print("case: coincident atoms 3 and 5 of a 2x2x2 Si cell", flush=True)
    load_system(eng, pos, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
    with pytest.raises(PysicEngineError, match="non-finite force"):
        eng.forces()
    # the engine recovers: a sound structure evaluates as before
    load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(len(s), 3.5))
    f, _ = eng.forces()
    assert np.isfinite(f).all()


def test_full_size_md_conserves_energy():
    """The 1,048,576-atom workload run NVE for 400 velocity-Verlet steps at
    1 fs and again at 0.5 fs from the same state, lists rebuilt every step.
    The potential energy moves by ~0.16 eV/atom as the displaced lattice
    relaxes, yet the total energy (device FP64 potential energy + kinetic
    energy of the returned momenta) stays within 1e-3 eV/atom, and its error
    shrinks by more than 3x when the step halves: the integrator's O(dt^2)
    error (4x), not a mismatch between forces and energies, which would give
    a drift in proportion to the simulated time (2x)."""
    import bench
    from pysic_amd.engine import Engine
    s = bench.build_workload((64, 64, 32))
    n = len(s)
    pd, bd = tersoff_descriptors()
    m = s.masses[:, None]
    eng = Engine(0)
    try:
        drift = {}
        for dt_fs in (1.0, 0.5):
            eng.create_atoms(np.zeros(n, np.int32), s.positions, masses=s.masses,
                             momenta=s.momenta)
            eng.create_cell(s.cell, s.pbc)
            eng.set_potentials(pd)
            eng.set_bond_order_factors(bd)
            eng.ewald_off()
            eng.set_neighbor_capacity(16)
            eng.generate_neighbor_lists(np.full(n, 3.5))
            eng.set_variant(0)
            e0 = eng.energy() + 0.5 * (s.momenta ** 2 / m).sum()
            print("case: full-size NVE, 400 steps of %.1f fs" % dt_fs, flush=True)
            ep = eng.md_run(400, dt_fs * st.ASE_FS, rebuild=1, energies=True)
            _, mom = eng.state()
            drift[dt_fs] = (ep[-1] + 0.5 * (mom ** 2 / m).sum() - e0) / n
            print("  drift %.3e eV/atom, potential energy range %.3f eV/atom" % (
                drift[dt_fs], (ep.max() - ep.min()) / n), flush=True)
        assert abs(drift[1.0]) < 1e-3
        assert abs(drift[1.0]) > 3.0 * abs(drift[0.5])
    finally:
        eng.close()


def test_full_size_md_is_bitwise_reproducible():
    """Two 50-step runs of the 1,048,576-atom workload from the same state,
    in two handles, end on the same bits: no force, energy or sort path
    depends on the order atomics land in (in-bin order is sorted by the
    original index; every sum is a fixed tree)."""
    import bench
    from pysic_amd.engine import Engine
    s = bench.build_workload((64, 64, 32))
    n = len(s)
    pd, bd = tersoff_descriptors()
    out = []
    for run in range(2):
        eng = Engine(0)
        try:
            eng.create_atoms(np.zeros(n, np.int32), s.positions, masses=s.masses,
                             momenta=s.momenta)
            eng.create_cell(s.cell, s.pbc)
            eng.set_potentials(pd)
            eng.set_bond_order_factors(bd)
            eng.ewald_off()
            eng.set_neighbor_capacity(16)
            eng.generate_neighbor_lists(np.full(n, 3.5))
            eng.set_variant(0)
            print("case: full-size MD, run %d, 50 steps" % run, flush=True)
            ep = eng.md_run(50, st.ASE_FS, rebuild=1, energies=True)
            pos, mom = eng.state()
            out.append((ep, pos, mom))
        finally:
            eng.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_md_switches_shape_at_its_start(eng):
    """md_run's first force evaluation is the API's: with the 8-entry shape
    selected on a compressed cell (up to 15 atoms within 3.0 A) it switches
    to a shape that holds them, and the run ends where the same run with
    the 32-entry shape selected from the start ends.  (An overflow later in
    a run, after the shape is fixed, is reported by md_run.)"""
    s = st.amorphous_si(256, 4.5, seed=5, dmin=1.8)
    pd, bd = tersoff_descriptors()
    n = len(s)
    print("case: MD on a compressed a-Si cell, shapes g4_s8 and g16_s32", flush=True)
    out = []
    for v in ("g4_s8", "g16_s32"):
        load_system(eng, s.positions, s.cell, [1, 1, 1], pd, bd, np.full(n, 3.5),
                    masses=s.masses)
        eng.set_variant(eng.variants().index(v))
        eng.md_run(4, 0.5 * st.ASE_FS, rebuild=1)
        out.append(eng.state())
    eng.set_variant(0)
    dr = np.abs(out[0][0] - out[1][0]).max()
    dp = np.abs(out[0][1] - out[1][1]).max() / np.abs(out[1][1]).max()
    print("  |dr| %.2e A  rel|dp| %.2e" % (dr, dp), flush=True)
    assert np.isfinite(out[0][0]).all()
    assert dr < 1e-9 and dp < 1e-6
