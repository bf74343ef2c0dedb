"""The drop-in for pysic's f2py module driven the way pysic drives it.

Each case replays, through pysic_amd.pysic_fortran.pysic_interface, the call
sequence pysic's calculator makes when it initialises its core and
evaluates (the documented order: the start_* calls of core.py at import;
create_atoms, distribute_mpi, create_cell; the potentials -- each product's
multipliers first, then its master once per distinct target permutation,
then clear_potential_multipliers -- and the bond-order factors; the storage
and the neighbor lists; create_potential_list and
create_bond_order_factor_list; set_ewald_parameters; FastNeighborList's
per-atom list readback; then calculate_energy / _forces /
_electronegativities), with f2py's argument shapes, and checks the answers
against the reference core's golden vectors.

Tolerances: Tersoff per its contract (|dE|/N <= 1e-7 eV, forces 1e-5
relative, stress 1e-6 eV per atom); FP64 paths (Ewald, EEM, COMB) 1e-9."""
from itertools import permutations

import numpy as np
import pytest

from conftest import load_golden
from helpers import comb_bend_terms, comb_charge_model, rel_maxnorm

pytestmark = pytest.mark.gpu

T = dict(A=1830.8, l1=2.4799, B=471.18, l2=1.7322)


@pytest.fixture()
def iface():
    import pysic_amd.pysic_fortran as pf
    I = pf.pysic_interface
    # pysic/core.py at import
    I.start_potentials()
    I.start_bond_order_factors()
    I.start_mpi()
    I.start_rng(5301)
    yield pf, I
    I.release()
    I.finish_mpi()


def _labels(pf, symbols):
    """(2, k) element codes, as calculator.py builds them with str2ints."""
    return np.array([pf._codes(s, 2) for s in symbols]).T


def _system(pf, I, symbols, pos, cell, pbc=(True, True, True), charges=None):
    n = len(pos)
    masses = np.full(n, 28.0855)
    q = np.zeros(n) if charges is None else np.asarray(charges, dtype=np.float64)
    I.create_atoms(masses, q, np.asarray(pos).T, np.zeros((3, n)), np.zeros(n, np.int32),
                   _labels(pf, symbols))
    I.distribute_mpi(n)
    I.create_cell(np.asarray(cell).T, np.linalg.inv(cell).T, np.array(pbc))


def _add_potentials(pf, I, pots):
    """Potential / ProductPotential objects into the core as pysic adds
    them (group -1: these models carry no coordinator)."""
    I.allocate_potentials(len(pots))
    for pot in pots:
        factors = pot.get_potentials()
        nt = pot.get_number_of_targets()
        no = np.full(nt, -9)
        first = pot.get_symbols()[0]
        for m in reversed(factors[1:]):
            ok = I.add_potential(m.get_potential_type(), np.array(m.get_parameter_values()),
                                 pot.get_cutoff(), pot.get_soft_cutoff(), _labels(pf, first),
                                 no, no, _labels(pf, first), no, no, -1, True)
            assert ok
        master = factors[0]
        for targets in pot.get_symbols():
            for perm in sorted(set(permutations(targets))):
                ok = I.add_potential(master.get_potential_type(),
                                     np.array(master.get_parameter_values()), pot.get_cutoff(),
                                     pot.get_soft_cutoff(), _labels(pf, perm), no, no,
                                     _labels(pf, targets), no, no, -1, False)
                assert ok
        I.clear_potential_multipliers()


def _readback(I, n):
    """FastNeighborList.build's per-atom readback (calculator.py:98-104) as
    sorted (i, j, offset) rows."""
    rows = []
    for i in range(n):
        k = I.get_number_of_neighbors_of_atom(i)
        nb, off = I.get_neighbor_list_of_atom(i, k)
        assert off.shape == (3, k)
        for j, o in zip(nb, off.T):
            rows.append((i, int(j), int(o[0]), int(o[1]), int(o[2])))
    rows = np.array(rows, dtype=np.int64).reshape(-1, 5)
    return rows[np.lexsort(rows.T[::-1])]


def test_tersoff_through_the_interface(iface):
    pf, I = iface
    g = load_golden("si_diamond_4x4x4_s005")
    pos, cell = g["positions"], g["cell"]
    n = len(pos)
    print("case: Tersoff through pysic_interface n=%d" % n, flush=True)
    _system(pf, I, ["Si"] * n, pos, cell)
    si2, no2 = _labels(pf, ["Si", "Si"]), np.full(2, -9)
    I.allocate_potentials(2)
    assert I.add_potential("exponential", np.array([T["A"], T["l1"]]), 3.0, 2.7, si2, no2, no2,
                           si2, no2, no2, -1, False)
    I.clear_potential_multipliers()
    assert I.add_potential("exponential", np.array([-T["B"], T["l2"]]), 3.0, 2.7, si2, no2, no2,
                           si2, no2, no2, 0, False)
    I.clear_potential_multipliers()
    I.allocate_bond_order_factors(1)
    si3 = _labels(pf, ["Si"] * 3)
    assert I.add_bond_order_factor("tersoff", np.array(g["tersoff"]), np.array([2, 5]), 3.0, 2.7,
                                   si3, si3, 0)
    assert I.get_number_of_atoms() == n
    I.allocate_bond_order_storage(n, 1, 1)
    I.generate_neighbor_lists(g["cutoffs"])
    I.create_potential_list()
    I.create_bond_order_factor_list()
    # the lists as pysic reads them back, against the reference core's
    ref = np.column_stack([np.repeat(np.arange(n), g["nl_counts"]), g["nl_idx"], g["nl_off"]])
    ref = ref[np.lexsort(ref.T[::-1])]
    assert np.array_equal(_readback(I, n), ref)
    e = I.calculate_energy()
    f, s = I.calculate_forces(n)
    assert f.shape == (3, n)
    err_e, err_f = abs(e - float(g["energy"])) / n, rel_maxnorm(f.T, g["forces"], 1e-2)
    err_s = np.abs(s - g["stress"]).max() / n
    print("  |dE|/N %.2e  rel|dF| %.2e  |dW|/N %.2e" % (err_e, err_f, err_s), flush=True)
    assert err_e <= 1e-7 and err_f <= 1e-5 and err_s <= 1e-6
    # a Coordinator's queries (bondorder.py:488-532): per-atom factors of a
    # per-bond factor are zero in the reference
    assert np.array_equal(I.calculate_bond_order_factors(n, 0), np.zeros(n))
    gr = I.calculate_bond_order_gradients(n, 0, 5)
    assert gr.shape == (3, n) and not gr.any()
    # a moved atom through update_atom_coordinates and a rebuilt list
    p2 = pos.copy()
    p2[7] += np.array([0.05, -0.03, 0.02])
    I.update_atom_coordinates(p2.T, np.zeros((3, n)))
    I.generate_neighbor_lists(g["cutoffs"])
    e2 = I.calculate_energy()
    assert e2 != e


@pytest.mark.parametrize("name", ["nacl_2x2x2_sheared_rc5", "sio2_48_eem_rc5"])
def test_ewald_and_eem_through_the_interface(iface, name):
    pf, I = iface
    g = load_golden(name)
    pos, cell, q = g["positions"], g["cell"], g["charges"]
    symbols = [str(x) for x in g["symbols"]]
    n = len(pos)
    print("case: %s through pysic_interface n=%d" % (name, n), flush=True)
    _system(pf, I, symbols, pos, cell, charges=q)
    rows = list(zip(g["charge_self_symbols"], g["charge_self"]))
    I.allocate_potentials(2 * len(rows))
    no1 = np.full(1, -9)
    for sym, (chi0, J) in rows:
        el = _labels(pf, [str(sym)])
        for params in ([chi0, 1.0], [0.5 * J, 2.0]):
            assert I.add_potential("charge_self", np.array(params), 0.0, 0.0, el, no1, no1, el,
                                   no1, no1, -1, False)
            I.clear_potential_multipliers()
    I.allocate_bond_order_factors(0)
    I.allocate_bond_order_storage(n, 0, 0)
    rc, kc, sigma, eps = g["ewald"]
    I.generate_neighbor_lists(np.full(n, rc + 0.5))
    I.create_potential_list()
    I.create_bond_order_factor_list()
    I.set_ewald_parameters(rc, kc, np.array(g["kcells"]), sigma, eps, g["scaler"])
    e = I.calculate_energy()
    f, s = I.calculate_forces(n)
    x = I.calculate_electronegativities(n)
    errs = (abs(e - float(g["energy"])) / n, rel_maxnorm(f.T, g["forces"]),
            np.abs(s - g["stress"]).max() / n, rel_maxnorm(x, g["enegs"]))
    print("  |dE|/N %.2e  rel|dF| %.2e  |dW|/N %.2e  rel|dchi| %.2e" % errs, flush=True)
    assert max(errs) <= 1e-9


def test_comb_products_through_the_interface(iface):
    """COMB's charge terms (products with multipliers) and bond bending
    through the interface, against the reference core's golden case."""
    pf, I = iface
    g = load_golden("sio2_48_comb_bend_charge")
    pos, cell, q = g["positions"], g["cell"], g["charges"]
    symbols = [str(x) for x in g["symbols"]]
    n = len(pos)
    print("case: COMB products through pysic_interface n=%d" % n, flush=True)
    _system(pf, I, symbols, pos, cell, charges=q)
    _add_potentials(pf, I, comb_bend_terms() + comb_charge_model())
    I.allocate_bond_order_factors(0)
    I.allocate_bond_order_storage(n, 0, 0)
    I.generate_neighbor_lists(g["cutoffs"])
    I.create_potential_list()
    I.create_bond_order_factor_list()
    e = I.calculate_energy()
    f, s = I.calculate_forces(n)
    x = I.calculate_electronegativities(n)
    errs = (abs(e - float(g["energy"])) / n, rel_maxnorm(f.T, g["forces"]),
            rel_maxnorm(s, g["stress"]), np.abs(x - g["enegs"]).max() / max(1.0, np.abs(
                g["enegs"]).max()))
    print("  |dE|/N %.2e  rel|dF| %.2e  rel|dW| %.2e  |dchi| %.2e" % errs, flush=True)
    assert max(errs) <= 1e-9


def test_lists_built_elsewhere(iface):
    """pysic hands over lists built by ASE for a cell too thin for its own
    (calculator.py:1582-1590): the engine builds its own at the model's
    cutoffs plus the skin, and the energy is the same."""
    pf, I = iface
    g = load_golden("si_diamond_4x4x4_s005")
    pos, cell = g["positions"], g["cell"]
    n = len(pos)
    _system(pf, I, ["Si"] * n, pos, cell)
    si2, no2 = _labels(pf, ["Si", "Si"]), np.full(2, -9)
    I.allocate_potentials(2)
    I.add_potential("exponential", np.array([T["A"], T["l1"]]), 3.0, 2.7, si2, no2, no2, si2, no2,
                    no2, -1, False)
    I.add_potential("exponential", np.array([-T["B"], T["l2"]]), 3.0, 2.7, si2, no2, no2, si2,
                    no2, no2, 0, False)
    I.allocate_bond_order_factors(1)
    si3 = _labels(pf, ["Si"] * 3)
    I.add_bond_order_factor("tersoff", np.array(g["tersoff"]), np.array([2, 5]), 3.0, 2.7, si3,
                            si3, 0)
    I.allocate_bond_order_storage(n, 1, 1)
    for i in range(n):   # whatever lists arrive, the engine rebuilds its own
        I.create_neighbor_list(i + 1, np.zeros(0, np.int32), np.zeros((3, 0), np.int32))
    I.create_potential_list()
    I.create_bond_order_factor_list()
    e = I.calculate_energy()
    assert abs(e - float(g["energy"])) / n <= 1e-7
