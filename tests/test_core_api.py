"""The Python API around the engine that needs no GPU: the module queries of
pysic_amd.core (pysic/core.py:44-485) against the reference's registry
fixture, CoreMirror (core.py:488-762), BondOrderParameters target handling
(bondorder.py:165-206, 306-344) and ChargeRelaxation pipes and observers
(relaxation.py:116-180, 300-322, 533-537)."""
import warnings

import numpy as np
import pytest

from conftest import load_golden
from pysic_amd import core
from pysic_amd import structures as st
from pysic_amd.coulomb import ChargeRelaxation, CoulombSummation, InvalidRelaxationError
from pysic_amd.interactions import BondOrderParameters, InvalidParametersError, Potential


def test_module_queries_follow_the_reference_registry():
    g = load_golden("registry_reference")
    ref = {str(n): (int(t), [str(x) for x in p[:k]]) for n, t, k, p in
           zip(g["pot_names"], g["pot_targets"], g["pot_nparams"], g["pot_param_names"])}
    pots = core.list_valid_potentials()
    assert core.list_potentials() == pots
    assert sorted(pots) == sorted(set(ref) - {"dihedral"})
    for name in pots:
        assert core.is_potential(name) and core.is_valid_potential(name)
        assert core.number_of_targets(name) == ref[name][0]
        assert core.number_of_parameters(name) == len(ref[name][1])
        assert core.number_of_parameters(name, as_list=True) == [len(ref[name][1])]
        names = core.names_of_parameters(name)
        if name != "step":
            assert names == ref[name][1]
        for k, pname in enumerate(names):
            assert core.index_of_parameter(name, pname) == k
        notes = core.descriptions_of_parameters(name)
        assert len(notes) == len(names) and all(notes)
    assert core.index_of_parameter("LJ", "sigma") == 1
    assert core.index_of_parameter("LJ", "nothing") is None
    # bond-order factors: per target class [single, pair]
    assert core.list_valid_bond_order_factors() == ["tersoff"]
    assert core.list_bond_order_factors() == ["tersoff"]
    assert core.is_bond_order_factor("tersoff") and not core.is_bond_order_factor("LJ")
    assert core.number_of_targets("tersoff") == 3
    assert core.level_of_factor("tersoff") == 2 and core.level_of_factor("LJ") == 0
    assert core.number_of_parameters("tersoff") == [2, 5]
    assert core.names_of_parameters("tersoff") == [["beta", "eta"], ["a", "c", "d", "h", "mu"]]
    assert core.index_of_parameter("tersoff", "mu") == [1, 4]
    assert core.index_of_parameter("tersoff", "eta") == [0, 1]
    assert [len(x) for x in core.descriptions_of_parameters("tersoff")] == [2, 5]
    # modes come from the relaxation and summation classes
    assert core.is_charge_relaxation("dynamic") and not core.is_charge_relaxation("ewald")
    assert core.is_coulomb_summation("ewald") and core.is_valid_coulomb_summation("ewald")
    assert core.number_of_parameters("dynamic") == 5
    assert core.names_of_parameters("ewald") == ["real_cutoff", "k_cutoff", "sigma", "epsilon"]
    assert len(core.descriptions_of_parameters("potentiostat")) == 4
    assert core.index_of_parameter("dynamic", "friction") == 3
    # unknown names
    assert core.number_of_targets("nothing") == 0 and core.number_of_parameters("nothing") == 0
    assert core.names_of_parameters("nothing") == [] and not core.is_potential("nothing")


def test_description_of_potential(capsys):
    msg = core.description_of_potential("LJ", parameter_values=[0.01, 3.4], cutoff=8.0,
                                        elements=["Ar", "Ar"])
    out = capsys.readouterr().out
    assert msg in out
    assert "potential 'LJ'" in msg and "2-body interaction" in msg
    assert "epsilon" in msg and "= 3.4" in msg and "cutoff = 8.0" in msg and " Ar " in msg


def test_core_mirror_readiness():
    s = st.amorphous_sio2(4, 2.2, seed=1)
    m = core.CoreMirror()
    assert not m.atoms_ready(s) and not m.charges_ready(s) and not m.cell_ready(s)
    assert m.potentials_ready(None) and not m.potentials_ready([])
    m.set_atoms(s)
    assert m.atoms_ready(s) and m.charges_ready(s) and m.cell_ready(s)
    # deep copies: changing the original is seen as a difference
    t = s.copy()
    p = t.get_positions()
    p[0, 0] += 1e-9
    t.set_positions(p)
    assert not m.atoms_ready(t)
    m.set_atomic_positions(t)
    assert m.atoms_ready(t)
    t.set_initial_charges(np.full(len(t), 0.1))
    assert not m.charges_ready(t)
    m.set_charges(t.get_initial_charges())
    assert m.charges_ready(t)
    t.set_cell(t.get_cell() * 1.01)
    assert not m.cell_ready(t)
    m.set_cell(t)
    assert m.cell_ready(t)
    pots = [Potential("LJ", symbols=[["Si", "O"]], parameters=[0.1, 2.0], cutoff=5.0)]
    m.set_potentials(pots)
    assert m.potentials_ready(pots)
    pots[0].set_parameter_value("epsilon", 0.2)
    assert not m.potentials_ready(pots)
    cs = CoulombSummation(parameters=[8.0, 2.0, 2.5, 0.00552635])
    assert not m.coulomb_summation_ready(cs)
    m.set_coulomb(cs)
    assert m.coulomb_summation_ready(cs)
    assert repr(m) == "CoreMirror()"


def test_bond_order_parameter_targets():
    b = BondOrderParameters("tersoff", cutoff=3.0, cutoff_margin=0.3,
                            symbols=[["Si", "Si", "O"]])
    assert b.accepts_target_list([["Si", "O", "O"], ["O", "O", "O"]])
    assert not b.accepts_target_list([["Si", "O"]]) and not b.accepts_target_list(["Si"])
    assert not b.accepts_target_list(5)
    b.add_symbols([["O", "O", "Si"], ["C", "Si", "Si"]])
    b.add_symbols(["Si", "Si", "Si"])
    assert b.get_symbols() == [["Si", "Si", "O"], ["O", "O", "Si"], ["C", "Si", "Si"],
                               ["Si", "Si", "Si"]]
    assert b.get_different_symbols() == ["Si", "O", "C"]
    with pytest.raises(InvalidParametersError):
        b.add_symbols(["Si", "Si"])
    b.set_soft_cutoff(2.5)
    assert b.get_cutoff_margin() == pytest.approx(0.5) and b.get_soft_cutoff() == 2.5
    b.set_cutoff(3.2)   # the margin is kept: the soft cutoff moves with the hard one
    assert b.get_soft_cutoff() == pytest.approx(2.7)
    with pytest.raises(InvalidParametersError):
        b.set_cutoff_margin(4.0)
    assert BondOrderParameters("tersoff").get_different_symbols() == []


def test_relaxation_pipes_and_observers():
    a, b, c = ChargeRelaxation(), ChargeRelaxation(), ChargeRelaxation()
    assert a.get_relaxation_pipe() == [] and a.get_relaxation_pipe(False) is None
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        a.set_relaxation_pipe(b)
        b.set_relaxation_pipe(c)
    assert a.get_relaxation_pipe() == [b, c] and a.get_relaxation_pipe(False) is b
    # loops are refused: c -> a would close a -> b -> c -> a
    with pytest.raises(InvalidRelaxationError):
        c.set_relaxation_pipe(a)
    with pytest.raises(InvalidRelaxationError):
        a.set_relaxation_pipe(a)
    a.clear_relaxation_pipe()
    assert a.get_relaxation_pipe() == [] and b.get_relaxation_pipe() == [c]
    seen = []
    a.add_observer(lambda: seen.append(1), 2)
    a.call_observers(4)
    a.call_observers(3)
    assert seen == [1]
    a.clear_observers()
    a.call_observers(4)
    assert seen == [1] and a.observers == []
    assert ChargeRelaxation.relaxation_parameter_descriptions["dynamic"][3].startswith("friction")
    assert len(CoulombSummation.summation_parameter_descriptions["ewald"]) == 4
    assert repr(a).startswith("ChargeRelaxation('dynamic'")
