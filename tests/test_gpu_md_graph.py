"""The MD step loop replayed from HIP graphs must give the same trajectory,
bit for bit, as the same steps launched one by one (the evaluation is
deterministic), also when graphs are reused across calls with eager work in
between."""
import numpy as np
import pytest

from pysic_amd import structures as st
from pysic_amd.interactions import build_descriptors, tersoff_si

pytestmark = pytest.mark.gpu


def make_engine(s):
    from pysic_amd.engine import Engine
    pd, bd, _ = build_descriptors(tersoff_si(), {"Si": 0})
    e = Engine(0)
    e.create_atoms(np.zeros(len(s), np.int32), s.positions, masses=s.masses, momenta=s.momenta)
    e.create_cell(s.cell, s.pbc)
    e.set_potentials(pd)
    e.set_bond_order_factors(bd)
    e.ewald_off()
    e.generate_neighbor_lists(np.full(len(s), 3.5))
    return e


def test_graph_md_matches_eager_md():
    s = st.diamond_si((6, 6, 6), sigma=0.1, seed=4)
    st.maxwell_boltzmann(s, 600.0, seed=5)
    dt = st.ASE_FS
    print("case: 1728 atoms, 3 x (10 graph steps vs 10 eager steps)", flush=True)
    a, b = make_engine(s), make_engine(s)
    try:
        for rnd in range(3):
            a.md_run(10, dt, rebuild=1)            # graph replays
            b.timing(True)
            b.md_run(10, dt, rebuild=1)            # launched one by one
            b.timing(False)
            pa, ma = a.state()
            pb, mb = b.state()
            assert np.array_equal(pa, pb) and np.array_equal(ma, mb), rnd
            # eager work between the calls
            assert a.energy() == b.energy()
            fa, wa = a.forces()
            fb, wb = b.forces()
            assert np.array_equal(fa, fb) and np.array_equal(wa, wb)
    finally:
        a.close()
        b.close()
