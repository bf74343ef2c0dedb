"""The pysic calculator API over the GPU engine, used the way pysic scripts use
it (doc/run examples.rst), checked against the oracle."""
import math

import numpy as np
import pytest

from helpers import rel_maxnorm, tersoff_descriptors
from pysic_amd import (ChargeRelaxation, CoulombSummation, Potential, structures as st,
                       tersoff_si)

pytestmark = pytest.mark.gpu


def oracle_tersoff(oracle, s):
    pd, bd = tersoff_descriptors()
    n = len(s)
    nl = oracle.neighbor_list(s.get_positions(), s.get_cell(), [1, 1, 1], np.full(n, 3.5))
    sp = np.zeros(n, np.int32)
    e = oracle.evaluate("energy", sp, s.get_positions(), None, s.get_cell(), nl, pd, bd)
    f, w = oracle.evaluate("forces", sp, s.get_positions(), None, s.get_cell(), nl, pd, bd)
    return e, f, w


def test_tersoff_through_the_calculator(oracle):
    from pysic_amd import Pysic
    s = st.diamond_si((4, 4, 4), sigma=0.05, seed=3)
    st.maxwell_boltzmann(s, 300.0, seed=1)
    calc = Pysic(s, tersoff_si())
    e = calc.get_potential_energy()
    f = calc.get_forces()
    stress = calc.get_stress()
    e_ref, f_ref, w_ref = oracle_tersoff(oracle, s)
    n = len(s)
    assert abs(e - e_ref) / n <= 1e-7
    assert rel_maxnorm(f, f_ref) <= 1e-5
    # stress = -(sum m v v + sum r f) / V (calculator.py:517-534)
    p, m = s.get_momenta(), s.get_masses()
    v = p / m[:, None]
    kin = np.array([p[:, 0] @ v[:, 0], p[:, 1] @ v[:, 1], p[:, 2] @ v[:, 2], p[:, 1] @ v[:, 2],
                    p[:, 0] @ v[:, 2], p[:, 0] @ v[:, 1]])
    want = -(kin + w_ref) / s.get_volume()
    assert np.abs(stress - want).max() * s.get_volume() / n <= 1e-6
    # cached: nothing to compute for the same structure
    assert not calc.calculation_required(s, ["energy", "forces", "stress"])


def test_small_moves_keep_the_list_and_stay_exact(oracle):
    from pysic_amd import Pysic
    s = st.amorphous_si(300, 2.3, seed=4)
    calc = Pysic(s, tersoff_si())
    calc.get_forces()
    builds = calc.get_neighbor_lists().nupdates
    rng = np.random.default_rng(0)
    for _ in range(3):
        t = calc.get_atoms().copy()
        t.set_positions(t.get_positions() + rng.normal(0, 0.01, (len(t), 3)))
        f = calc.get_forces(t)
        e = calc.get_potential_energy(t)
        e_ref, f_ref, _ = oracle_tersoff(oracle, t)
        assert rel_maxnorm(f, f_ref) <= 1e-5
        assert abs(e - e_ref) / len(t) <= 1e-7
    assert calc.get_neighbor_lists().nupdates == builds  # moves < skin: no rebuild
    t = calc.get_atoms().copy()
    p = t.get_positions()
    p[0] += 0.6
    t.set_positions(p)
    calc.get_forces(t)
    assert calc.get_neighbor_lists().nupdates == builds + 1


def test_neighbor_list_accessors(oracle):
    from pysic_amd import Pysic
    s = st.diamond_si((3, 3, 3), sigma=0.05, seed=2)
    calc = Pysic(s, tersoff_si())
    calc.get_forces()
    nlist = calc.get_neighbor_lists()
    nb, off = nlist.get_neighbors(5)
    assert len(nb) == 4
    d = nlist.get_neighbor_distances(5, s, sort=True)
    assert np.all(np.diff(d) >= 0) and d.max() < 3.5


def test_numerical_gradient_helper_matches_forces():
    from pysic_amd import Pysic
    s = st.amorphous_si(64, 2.3, seed=8)
    calc = Pysic(s, tersoff_si())
    f = calc.get_forces()
    g = calc.get_numerical_energy_gradient(7)
    assert np.abs(g - f[7]).max() < 1e-4 * max(1.0, np.abs(f[7]).max())


def test_charge_relaxation_through_the_calculator(oracle):
    from pysic_amd import Pysic
    s = st.amorphous_sio2(32, 2.2, seed=6)
    pots = []
    for sym, (chi0, J) in {"Si": (4.77, 16.0), "O": (8.74, 20.0)}.items():
        pots.append(Potential("charge_self", symbols=[[sym]], parameters=[chi0, 1]))
        pots.append(Potential("charge_self", symbols=[[sym]], parameters=[0.5 * J, 2]))
    rc = 6.0
    coul = CoulombSummation(parameters={"real_cutoff": rc, "k_cutoff": 1.5, "sigma": 2.0,
                                        "epsilon": 0.00552635})
    relax = ChargeRelaxation("dynamic", parameters={"n_steps": 1000, "timestep": 0.1,
                                                    "inertia": 1.0, "friction": 4.0,
                                                    "tolerance": 1e-8})
    calc = Pysic(s, pots, coulomb=coul)
    calc.set_charge_relaxation(relax)
    calc.get_potential_energy()
    q = calc.get_atoms().get_initial_charges()
    assert relax.last_steps < 1000 and relax.last_error <= 1e-8
    assert abs(q.sum()) < 1e-9
    # converged: electronegativities are equalised
    dchi = calc.get_electronegativity_differences()
    assert np.abs(dchi).max() < 1e-6
    # and agree with the oracle's electronegativities at those charges
    names = ["O", "Si"]
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    rows = []
    for k, sym in enumerate(names):
        chi0, J = {"Si": (4.77, 16.0), "O": (8.74, 20.0)}[sym]
        rows += [[2, 1, k, 0, chi0, 1.0, 0, 0, -1, 0], [2, 1, k, 0, 0.5 * J, 2.0, 0, 0, -1, 0]]
    rec = 2 * math.pi * np.linalg.inv(s.cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    K = [int(1.5 * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
         for a in range(3)]
    ew = np.array([1, rc, 1.5, K[0], K[1], K[2], 2.0, 0.00552635])
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(len(s), rc + 0.5))
    x_ref = oracle.evaluate("enegs", sp, s.positions, q, s.cell, nl, np.array(rows, float), None,
                            ew)
    assert rel_maxnorm(calc.get_electronegativities(), x_ref) < 1e-9


def test_device_md_through_the_calculator(oracle):
    from pysic_amd import Pysic
    s = st.diamond_si((3, 3, 3), sigma=0.1, seed=7)
    st.maxwell_boltzmann(s, 600.0, seed=8)
    calc = Pysic(s, tersoff_si())
    e0 = calc.get_potential_energy()
    p0 = s.get_momenta()
    k0 = (p0 ** 2 / (2 * s.get_masses()[:, None])).sum()
    calc.run_md(50, 0.5 * st.ASE_FS)
    t = calc.get_atoms()
    e1 = calc.get_potential_energy()
    k1 = (t.get_momenta() ** 2 / (2 * t.get_masses()[:, None])).sum()
    # NVE at 0.5 fs conserves the total energy well
    assert abs((e1 + k1) - (e0 + k0)) / len(s) < 2e-4


def test_calculator_housekeeping_and_core_updates(oracle):
    """The reference's calculator entry points beyond get_* (calculator.py:
    260-305, 703-735, 1048-1637): equality, removal, cutoff bookkeeping,
    warnings, and the update_core_* calls that push one part of the state."""
    import warnings as W
    from pysic_amd import Pysic
    from pysic_amd.interactions import build_descriptors, lj_argon
    s = st.fcc((4, 4, 4), a=5.26, sigma=0.05, seed=2)
    lj = lj_argon()
    extra = Potential("LJ", symbols=[["Ar", "Ar"]], parameters=[0.001, 3.0], cutoff=6.0)
    calc = Pysic(s, lj + [extra])
    try:
        assert calc == calc and not (calc != calc)
        assert repr(calc).startswith("Pysic(atoms=")
        assert calc.core_initialization_is_forced() is False
        assert calc.get_neighbor_list() is None
        e_two = calc.get_potential_energy()
        calc.remove_potential(extra)
        e_one = calc.get_potential_energy()
        pd, _, _ = build_descriptors(lj, {"Ar": 0})
        cut = np.full(len(s), 2.5 * 3.405)
        nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
        e_ref = oracle.evaluate("energy", np.zeros(len(s), np.int32), s.positions, None, s.cell,
                                nl, pd)
        assert abs(e_one - e_ref) <= 1e-10 * abs(e_ref) and e_two != e_one
        # cutoff bookkeeping (calculator.py:1681-1718)
        calc.set_cutoffs([5.0] * len(s))
        assert not calc.neighbor_lists_expanded([5.0] * len(s))
        assert calc.neighbor_lists_expanded([5.0] * (len(s) - 1) + [5.1])
        # push new coordinates and charges explicitly
        t = calc.get_atoms()
        p = t.get_positions()
        p[0, 0] += 0.02
        t.set_positions(p)
        calc.update_core_coordinates()
        nl = oracle.neighbor_list(p, s.cell, [1, 1, 1], cut)
        f_ref, _ = oracle.evaluate("forces", np.zeros(len(s), np.int32), p, None, s.cell, nl, pd)
        calc.forces = None
        f = calc.get_forces()
        assert np.abs(f - f_ref).max() <= 1e-10 * max(1.0, np.abs(f_ref).max())
        q = np.zeros(len(s))
        q[0] = 0.5
        t.set_initial_charges(q)
        with W.catch_warnings(record=True) as rec:
            W.simplefilter("always")
            calc.update_core_charges()
        assert any("non-zero total charge" in str(r.message) for r in rec)
        assert np.array_equal(calc.get_charges(), q)
        # a pair potential without a cutoff warns (calculator.py:1201-1207)
        with W.catch_warnings(record=True) as rec:
            W.simplefilter("always")
            calc.set_potentials(lj + [Potential("LJ", symbols=[["Ar", "Ar"]],
                                                parameters=[0.0, 1.0])])
            calc.initialize_fortran_core()
        assert any("zero cutoff" in str(r.message) for r in rec)
        calc.update_core_potentials()
        calc.update_core_potential_lists()
        calc.update_core_neighbor_lists()
        calc.update_core_supercell()
        calc.update_core_coulomb()
        assert calc.get_potential_energy() == pytest.approx(
            oracle.evaluate("energy", np.zeros(len(s), np.int32), p, None, s.cell, nl, pd),
            rel=1e-10)
        other = Pysic(s, lj)
        try:
            calc.add_calculator(other)
            assert calc.get_calculators() == [other]
            calc.remove_calculator(other)
            assert calc.get_calculators() == []
        finally:
            other.engine.close()
    finally:
        calc.engine.close()


def test_comb_terms_through_the_calculator(oracle):
    """COMB's bond bending and charge-dependent pair terms set up as a pysic
    script does (Potential / ProductPotential objects on a Pysic
    calculator, charges on the atoms), against the oracle."""
    from helpers import comb_bend_terms, comb_charge_model, comb_charges
    from pysic_amd import Pysic
    from pysic_amd.interactions import build_descriptors
    s = st.amorphous_sio2(64, 2.2, seed=111)
    q = comb_charges(s.get_chemical_symbols(), seed=112)
    s.set_initial_charges(q)
    pots = comb_bend_terms() + comb_charge_model()
    calc = Pysic(s, pots)
    e = calc.get_potential_energy()
    f = calc.get_forces()
    x = calc.get_electronegativities()
    names = sorted(set(s.get_chemical_symbols()))
    pd, _, _ = build_descriptors(pots, {y: k for k, y in enumerate(names)})
    sp = np.array([names.index(y) for y in s.get_chemical_symbols()], np.int32)
    n = len(s)
    nl = oracle.neighbor_list(s.get_positions(), s.get_cell(), [1, 1, 1], np.full(n, 3.55))
    args = (sp, s.get_positions(), q, s.get_cell(), nl, pd)
    e_ref = oracle.evaluate("energy", *args)
    f_ref, _ = oracle.evaluate("forces", *args)
    x_ref = oracle.evaluate("enegs", *args)
    print("  |dE|/N %.2e rel|dF| %.2e rel|dchi| %.2e" % (
        abs(e - e_ref) / n, rel_maxnorm(f, f_ref), rel_maxnorm(x, x_ref)), flush=True)
    assert abs(e - e_ref) / n <= 1e-9
    assert rel_maxnorm(f, f_ref) <= 1e-9
    assert rel_maxnorm(x, x_ref) <= 1e-9


def test_empty_single_atom_and_dimer_closed_forms(oracle):
    """Edge cases, with answers fixed by the mathematics: no atoms (zero
    energy and stress, empty forces, as the reference's loops over zero atoms
    give); one atom (no neighbors: zero energy, force and stress); a Si dimer,
    where zeta = 0 so b = 1 and E = A exp(-l1 r) + B exp(-l2 r) times the
    smoothing f_c(r) of tersoff_si (3.0 A cutoff, 0.3 A margin), F = -dE/dr
    along the bond; and a dimer beyond the cutoff (all zero)."""
    from pysic_amd import Pysic
    print("case: empty, single atom and dimer cells", flush=True)
    L = 20.0
    cell = np.eye(3) * L
    calc = Pysic(st.Structure([], np.zeros((0, 3)), cell), tersoff_si())
    assert calc.get_potential_energy() == 0.0
    assert calc.get_forces().shape == (0, 3)
    assert np.array_equal(calc.get_stress(), np.zeros(6))
    assert calc.get_electronegativities().shape == (0,)

    one = st.Structure(["Si"], np.array([[3.1, 4.7, 9.2]]), cell)
    calc = Pysic(one, tersoff_si())
    assert calc.get_potential_energy() == 0.0
    assert np.array_equal(calc.get_forces(), np.zeros((1, 3)))
    assert np.array_equal(calc.get_stress(), np.zeros(6))

    pots = tersoff_si()
    (A, l1), (B, l2) = [p.get_parameter_values() for p in pots]
    hard, soft = 3.0, 2.7
    for r in (2.35, 2.85, 3.4):
        d = np.array([0.6, -0.48, 0.64])
        d = d / np.linalg.norm(d) * r
        pos = np.array([[7.0, 8.0, 9.0], [7.0, 8.0, 9.0] + d])
        dimer = st.Structure(["Si", "Si"], pos, cell)
        calc = Pysic(dimer, tersoff_si())
        e, f = calc.get_potential_energy(), calc.get_forces()
        if r < soft:
            fc, dfc = 1.0, 0.0
        elif r < hard:
            x = math.pi * (r - soft) / (hard - soft)
            fc, dfc = 0.5 * (1 + math.cos(x)), -0.5 * math.pi / (hard - soft) * math.sin(x)
        else:
            fc, dfc = 0.0, 0.0
        v = A * math.exp(-l1 * r) + B * math.exp(-l2 * r)
        dv = -l1 * A * math.exp(-l1 * r) - l2 * B * math.exp(-l2 * r)
        e_ref = fc * v
        g = dfc * v + fc * dv            # dE/dr
        f_ref = np.array([g * d / r, -g * d / r])
        e_o, f_o, _ = oracle_tersoff(oracle, dimer)
        print("  dimer r=%.2f: E %.12f closed form %.12f oracle %.12f; |dF| %.2e" % (
            r, e, e_ref, e_o, np.abs(f - f_ref).max()), flush=True)
        assert abs(e - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
        assert abs(e_o - e_ref) <= 1e-12 * max(1.0, abs(e_ref))
        assert np.abs(f - f_ref).max() <= 1e-5 * max(np.abs(f_ref).max(), 0.01)
        assert np.allclose(f.sum(axis=0), 0.0, atol=1e-12)
