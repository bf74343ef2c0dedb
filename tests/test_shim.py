"""The f2py-compatible drop-in (pysic_amd.pysic_fortran.pysic_interface),
host side: the registry answers pysic/core.py reads, as integer codes,
against the reference core's own registry (tests/golden/
registry_reference.npz), and the calls that need no GPU."""
import numpy as np

from conftest import load_golden


def _pf():
    import pysic_amd.pysic_fortran as pf
    return pf


def test_registry_through_the_interface_matches_the_reference():
    pf = _pf()
    I = pf.pysic_interface
    g = load_golden("registry_reference")
    ref = {str(n): [str(x) for x in p[:k]] for n, k, p in
           zip(g["pot_names"], g["pot_nparams"], g["pot_param_names"])}
    n = I.number_of_potentials()
    codes = I.list_valid_potentials(n)
    assert codes.shape == (11, n) and codes.flags.f_contiguous
    names = [pf._text(c) for c in codes.T]   # core.py:56-57 decodes them so
    assert sorted(names) == sorted(set(ref) - {"dihedral"})
    for name in names:
        assert I.is_potential(name)
        k = I.number_of_parameters_of_potential(name)
        pc = I.names_of_parameters_of_potential(name)
        assert pc.shape == (10, 12)
        got = [pf._text(c) for c in pc.T][:k]   # core.py:279-288
        assert got == (["e1", "e2", "r1", "r2"] if name == "step" else ref[name])
        dc = I.descriptions_of_parameters_of_potential(name)
        assert dc.shape == (100, 12) and all(pf._text(c) for c in dc.T[:k])
        assert len(I.description_of_potential(name)) == 500
    assert not I.is_potential("dihedral")
    bn = I.number_of_bond_order_factors()
    assert [pf._text(c) for c in I.list_valid_bond_order_factors(bn).T] == ["tersoff"]
    assert I.number_of_targets_of_bond_order_factor("tersoff") == 3
    assert I.level_of_bond_order_factor("tersoff") == 2
    k = list(g["bond_names"]).index("tersoff")
    for cls in (1, 2):
        m = I.number_of_parameters_of_bond_order_factor("tersoff", cls)
        assert m == int(g["bond_nparams"][k][cls - 1])
        got = [pf._text(c) for c in I.names_of_parameters_of_bond_order_factor("tersoff", cls).T]
        assert got[:m] == [str(x) for x in g["bond_param_names"][k][cls - 1][:m]]


def test_codes_round_trip_and_host_only_calls():
    pf = _pf()
    I = pf.pysic_interface
    for text in ("Si", "O", "charge_exp", "R_0", "Rmax1"):
        assert pf._text(pf._codes(text, 11)) == text
    # the two-integer element codes pysic passes (utility/f2py.py str2ints)
    assert list(pf._codes("Si", 2)) == [119, 9]
    I.start_potentials()
    I.start_bond_order_factors()
    I.start_mpi()
    I.start_rng(5301)
    assert I.get_cpu_id() == 0 and I.get_number_of_cpus() == 1
    assert I.get_mpi_list_of_atoms(5).all()
    I.sync_mpi()
    I.finish_mpi()


def test_add_potential_rejects_what_the_engine_cannot_take():
    """Unknown types and tag/index targets return the reference's failure
    flag (pysic then raises InvalidPotentialError, calculator.py:1281-1285)."""
    pf = _pf()
    I = pf.pysic_interface
    si = np.array([pf._codes("Si", 2), pf._codes("Si", 2)]).T
    no = np.array([-9, -9])
    I.allocate_potentials(1)
    assert not I.add_potential("dihedral", np.array([1.0, 2.0]), 3.0, 2.7, si, no, no, si, no,
                               no, -1, False)
    assert not I.add_potential("LJ", np.array([1.0, 2.0]), 3.0, 2.7, si, np.array([1, -9]), no,
                               si, no, no, -1, False)
    assert not I.add_potential("LJ", np.array([1.0]), 3.0, 2.7, si, no, no, si, no, no, -1,
                               False)
    assert I.add_potential("LJ", np.array([1.0, 2.0]), 3.0, 2.7, si, no, no, si, no, no, -1,
                           False)
    assert not I.add_bond_order_factor("c_scale", np.ones(5), np.array([5, 0]), 3.0, 2.7,
                                       np.array([pf._codes("Si", 2)] * 3).T,
                                       np.array([pf._codes("Si", 2)] * 3).T, 0)
    I.release()
