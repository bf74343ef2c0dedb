"""The pysic API beyond the calculations, on the device: CoreMirror and the
calculator's re-synchronisation (core.py:488-762, calculator.py:315-372,
998-1046), force_core_initialization (calculator.py:303-312), the
Coordinator's bond-order queries and the numerical bond-order gradient
(bondorder.py:488-532, calculator.py:1721-1778), and the 'potentiostat' and
'optimize' relaxations and relaxation pipes (relaxation.py:116-180,
394-506).

Tolerances: results of one engine computed twice agree exactly; the
potentiostat loop agrees with the same loop driven by the oracle's
electronegativities to 1e-9 e (FP64 throughout); SLSQP and the damped
dynamics reach the same minimum to 1e-6 e.
"""
import warnings

import numpy as np
import pytest

from conftest import load_golden
from pysic_amd import structures as st

pytestmark = pytest.mark.gpu


def _eem_system(n_si=32, seed=41):
    from pysic_amd import CoulombSummation, Pysic
    from pysic_amd.learned import SIO2_EWALD, eem_potentials
    s = st.amorphous_sio2(n_si, 2.2, seed=seed)
    p = dict(SIO2_EWALD, real_cutoff=6.0, k_cutoff=1.8, sigma=2.0)
    calc = Pysic(s, eem_potentials(), coulomb=CoulombSummation(parameters=p))
    return s, calc, p


def test_core_mirror_follows_in_place_changes():
    from pysic_amd import Pysic
    from pysic_amd.interactions import tersoff_si
    s = st.diamond_si((3, 3, 3), sigma=0.08, seed=2)
    pots = tersoff_si()
    calc = Pysic(s, pots)
    e0 = calc.get_potential_energy()
    print("case: CoreMirror on a 216-atom Tersoff cell, E0 = %.6f" % e0, flush=True)
    assert Pysic.core is calc._mirror
    assert Pysic.core.atoms_ready(calc.get_atoms()) and Pysic.core.potentials_ready(pots)
    # a parameter changed in place: the mirror sees it and the energy follows
    pots[0].set_parameter_value("epsilon", pots[0].get_parameter_value("epsilon") * 1.01)
    assert calc.calculation_required(None, "energy")
    e1 = calc.get_potential_energy()
    fresh = Pysic(s, pots)
    assert e1 == fresh.get_potential_energy() and e1 != e0
    # positions changed in place on the calculator's own structure
    a = calc.get_atoms()
    p = a.get_positions()
    p[5] += 0.01
    a.set_positions(p)
    e2 = calc.get_potential_energy()
    fresh.set_atoms(a)
    assert e2 == fresh.get_potential_energy() and e2 != e1
    assert Pysic.core is fresh._mirror   # the last calculator to synchronise


def test_force_core_initialization():
    from pysic_amd import Pysic
    from pysic_amd.interactions import tersoff_si
    s = st.diamond_si((2, 2, 2), sigma=0.05, seed=3)
    calc = Pysic(s, tersoff_si())
    e0 = calc.get_potential_energy()
    assert not calc.core_initialization_is_forced()
    calc.force_core_initialization(True)
    assert calc.core_initialization_is_forced() and "full_initialization=True" in repr(calc)
    a = calc.get_atoms()
    p = a.get_positions()
    p[0] += 0.02
    a.set_positions(p)
    e1 = calc.get_potential_energy()
    calc.force_core_initialization = False     # plain assignment, as the reference does
    assert not calc.core_initialization_is_forced()
    ref = Pysic(a, tersoff_si()).get_potential_energy()
    print("case: forced initialisation, E %.6f -> %.6f" % (e0, e1), flush=True)
    assert e1 == ref


def test_coordinator_queries_match_the_reference_core():
    # the reference's per-atom factors and gradients of the (level-2)
    # 'tersoff' factor are zero: its per-atom pass skips level-2 factors
    # (Core.f90:2093, 2703); the fixture holds the reference core's arrays
    from pysic_amd import Pysic
    from pysic_amd.interactions import tersoff_si
    g = load_golden("tersoff_bond_order_queries")
    s = st.Structure(["Si"] * len(g["positions"]), g["positions"], g["cell"])
    pots = tersoff_si()
    calc = Pysic(s, pots)
    calc.get_potential_energy()
    coord = [p.get_coordinator() for p in pots if p.get_coordinator() is not None][0]
    assert coord.get_group_index() == int(g["group"])
    coord.calculate_bond_order_factors()
    atom = int(g["atom"])
    print("case: Coordinator queries, 64 atoms, atom %d" % atom, flush=True)
    assert np.array_equal(coord.get_bond_order_factors(), g["factors"])
    assert np.array_equal(coord.get_bond_order_gradients(atom), g["gradients"])
    assert np.array_equal(coord.get_bond_order_gradients_of_factor(atom),
                          g["gradients_of_factor"])
    num = calc.get_numerical_bond_order_gradient(coord, atom, 5)
    assert num == [0.0, 0.0, 0.0]
    assert np.array_equal(calc.get_atoms().get_positions(), g["positions"])


def test_potentiostat_matches_the_oracle_loop(oracle):
    from pysic_amd import ChargeRelaxation
    from pysic_amd.interactions import build_descriptors
    from pysic_amd.learned import SPECIES, eem_potentials, kcells
    s, calc, p = _eem_system()
    n = len(s)
    V, dt, steps = -6.6, 0.02, 6
    relax = ChargeRelaxation("potentiostat", parameters=dict(n_steps=steps, timestep=dt,
                                                             inertia=1.0, potential=V))
    calc.set_charge_relaxation(relax)
    q_dev = relax.charge_relaxation()
    # the same loop on the oracle's electronegativities
    sp = np.array([SPECIES.index(x) for x in s.symbols], np.int32)
    pd, _, _ = build_descriptors(eem_potentials(), {x: k for k, x in enumerate(SPECIES)})
    K = kcells(s.cell, p["k_cutoff"])
    ew = np.array([1, p["real_cutoff"], p["k_cutoff"], K[0], K[1], K[2], p["sigma"],
                   p["epsilon"]])
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(n, p["real_cutoff"] + 0.5))

    def chi(q):
        return oracle.evaluate("enegs", sp, s.positions, q, s.cell, nl, pd, None, ew)
    q = np.zeros(n)
    r = np.zeros(n)
    f = chi(q) - V
    for _ in range(steps):
        q = q + r * dt + f * dt * dt
        r = r + 0.5 * f * dt
        f = chi(q) - V
        r = r + 0.5 * f * dt
    print("case: potentiostat %d steps, 96 atoms: max |dq| %.2e, total charge %.4f" % (
        steps, np.abs(q_dev - q).max(), q.sum()), flush=True)
    assert np.abs(q_dev - q).max() <= 1e-9
    assert np.array_equal(calc.get_atoms().get_initial_charges(), q_dev)


def test_optimize_and_pipes_reach_the_dynamic_minimum():
    from pysic_amd import ChargeRelaxation
    from pysic_amd.learned import SIO2_RELAX
    s, calc, _ = _eem_system(seed=43)
    dyn = ChargeRelaxation("dynamic", parameters=dict(n_steps=5000, tolerance=1e-10,
                                                      **SIO2_RELAX))
    calc.set_charge_relaxation(dyn)
    q_star = dyn.charge_relaxation().copy()
    # SLSQP at fixed total charge (zero): the same minimum
    calc.get_atoms().set_initial_charges(np.zeros(len(s)))
    opt = ChargeRelaxation("optimize", parameters=dict(n_steps=500, tolerance=1e-14))
    calc.set_charge_relaxation(opt)
    q_opt = opt.charge_relaxation()
    print("case: optimize vs dynamic, max |dq| %.2e" % np.abs(q_opt - q_star).max(), flush=True)
    assert np.abs(q_opt - q_star).max() <= 1e-6
    assert abs(q_opt.sum()) <= 1e-9
    # a coarse relaxation piped into a fine one ends where the fine one does
    calc.get_atoms().set_initial_charges(np.zeros(len(s)))
    coarse = ChargeRelaxation("dynamic", parameters=dict(n_steps=5000, tolerance=1e-3,
                                                         **SIO2_RELAX))
    fine = ChargeRelaxation("dynamic", parameters=dict(n_steps=5000, tolerance=1e-10,
                                                       **SIO2_RELAX))
    calc.set_charge_relaxation(coarse)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        coarse.set_relaxation_pipe(fine)
    assert fine.get_calculator() is calc
    q_pipe = coarse.charge_relaxation()
    print("  pipe: coarse %d steps, fine %d steps, max |dq| %.2e" % (
        coarse.last_steps, fine.last_steps, np.abs(q_pipe - q_star).max()), flush=True)
    assert fine.last_steps < dyn.last_steps    # the pipe started the fine one closer
    assert fine.last_error <= 1e-10
    assert np.abs(q_pipe - q_star).max() <= 1e-8
