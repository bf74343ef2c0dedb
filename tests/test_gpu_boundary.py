"""GPU checks of the boundary queries pysic makes besides energy, forces and
electronegativities: the bond-order factor queries of a Coordinator
(PyInterface.f90:579-634) and the stand-alone Ewald energy (1050), against
the reference core's answers (tests/golden/, written by
tests/golden/make_golden.py)."""
import numpy as np
import pytest

from conftest import load_golden
from helpers import tersoff_descriptors

pytestmark = pytest.mark.gpu


def test_tersoff_bond_order_queries_match_the_reference():
    """For a 'tersoff' coordinator the reference's per-atom factors and
    their gradients are zero (it sums level-1 factors only); the engine's
    answers are the same, for the group and an atom of the golden case."""
    from pysic_amd.engine import Engine
    g = load_golden("tersoff_bond_order_queries")
    pos, cell = g["positions"], g["cell"]
    n = len(pos)
    pd, bd = tersoff_descriptors()
    eng = Engine(0)
    try:
        print("case: bond-order queries n=%d" % n, flush=True)
        eng.create_atoms(np.zeros(n, np.int32), pos)
        eng.create_cell(cell, (True, True, True))
        eng.set_potentials(pd)
        eng.set_bond_order_factors(bd)
        eng.ewald_off()
        eng.generate_neighbor_lists(g["cutoffs"])
        grp, atom = int(g["group"]), int(g["atom"])
        assert np.array_equal(eng.bond_order_factors(grp), g["factors"])
        assert np.array_equal(eng.bond_order_gradients(grp, atom), g["gradients"])
        assert np.array_equal(eng.bond_order_gradients(grp, atom, of_factor=True),
                              g["gradients_of_factor"])
        from pysic_amd.engine import PysicEngineError
        with pytest.raises(PysicEngineError):
            eng.bond_order_gradients(grp, n)
    finally:
        eng.close()


def test_standalone_ewald_energy_matches_the_reference():
    """get_ewald_energy on the sheared rocksalt cell for two parameter sets,
    neither the one the model's Ewald uses (whose charge scalers are not 1):
    real space, self, background, reciprocal and the dipole term, with unit
    scalers, to 1e-10 relative.  The model's own Ewald energy is unchanged
    by the call (its cached structure factors are left alone)."""
    from pysic_amd.engine import Engine
    g = load_golden("ewald_energy_reference")
    pos, cell, q = g["positions"], g["cell"], g["charges"]
    n = len(pos)
    sp = (q < 0).astype(np.int32)
    eng = Engine(0)
    try:
        print("case: get_ewald_energy n=%d" % n, flush=True)
        eng.create_atoms(sp, pos, charges=q)
        eng.create_cell(cell, (True, True, True))
        eng.set_potentials(np.zeros((0, 12)))
        eng.set_bond_order_factors(np.zeros((0, 14)))
        rc0 = float(g["params"][0][0])
        eng.set_ewald(rc0, 1.8, [4, 4, 4], rc0 / 3.0, 0.00552635, scaler=g["scaler"])
        eng.generate_neighbor_lists(np.full(n, rc0 + 0.5))
        e_model = eng.energy()
        for p, e_ref in zip(g["params"], g["energies"]):
            rc, kc, k1, k2, k3, sigma, eps = p
            e = eng.ewald_energy(rc, kc, [int(k1), int(k2), int(k3)], sigma, eps)
            print("  sigma %.4f  E %.12f  (reference %.12f)" % (sigma, e, e_ref), flush=True)
            assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
        assert eng.energy() == e_model
    finally:
        eng.close()
