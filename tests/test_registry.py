"""The C descriptor registry (include/pysic_registry.h) against the
reference's own registry (tests/golden/registry_reference.npz, written by the
reference core through PyInterface.f90:706-1011), and the Python
interaction objects against the C registry.  Host code only: no GPU."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden

# the engine takes every potential type of the reference but 'dihedral', and
# the 'tersoff' bond-order factor; 'step' is the one documented name
# difference (the reference's registry leaves its second name blank)
NOT_BUILT = {"dihedral"}
STEP_NAMES = ["e1", "e2", "r1", "r2"]


@pytest.fixture(scope="module")
def L():
    from pysic_amd.engine import load
    lib = load()
    return lib


def _s(b):
    return None if b is None else b.decode()


def test_potential_registry_matches_the_reference(L):
    g = load_golden("registry_reference")
    ref = {str(n): (int(t), [str(x) for x in p[:k]]) for n, t, k, p in
           zip(g["pot_names"], g["pot_targets"], g["pot_nparams"], g["pot_param_names"])}
    ours = [_s(L.pysic_potential_name(k)) for k in range(L.pysic_number_of_potentials())]
    assert L.pysic_potential_name(len(ours)) is None
    assert sorted(ours) == sorted(set(ref) - NOT_BUILT)
    for name in ours:
        nb = name.encode()
        nt, names = ref[name]
        assert L.pysic_is_potential(nb) == 1
        assert L.pysic_number_of_targets_of_potential(nb) == nt
        assert L.pysic_number_of_parameters_of_potential(nb) == len(names)
        got = [_s(L.pysic_parameter_name_of_potential(nb, k)) for k in range(len(names))]
        assert got == (STEP_NAMES if name == "step" else names), name
        assert L.pysic_parameter_name_of_potential(nb, len(names)) is None
        assert all(L.pysic_parameter_description_of_potential(nb, k) for k in range(len(names)))
        assert L.pysic_description_of_potential(nb)
    # the reference's 'step' entry as it is (its deviation, checked to stay known)
    assert ref["step"][1] == ["e2", "", "r1", "r2"]
    for bad in (b"dihedral", b"lj", b"", b"exponential "):
        assert L.pysic_is_potential(bad) == 0
        assert L.pysic_number_of_targets_of_potential(bad) == -1


def test_bond_order_registry_matches_the_reference(L):
    g = load_golden("registry_reference")
    names = [str(x) for x in g["bond_names"]]
    k = names.index("tersoff")
    assert L.pysic_number_of_bond_order_factors() == 1
    assert _s(L.pysic_bond_order_factor_name(0)) == "tersoff"
    assert L.pysic_bond_order_factor_name(1) is None
    t = b"tersoff"
    assert L.pysic_is_bond_order_factor(t) == 1
    assert L.pysic_number_of_targets_of_bond_order_factor(t) == int(g["bond_targets"][k])
    assert L.pysic_level_of_bond_order_factor(t) == int(g["bond_levels"][k])
    for cls in (1, 2):
        n = int(g["bond_nparams"][k][cls - 1])
        assert L.pysic_number_of_parameters_of_bond_order_factor(t, cls) == n
        got = [_s(L.pysic_parameter_name_of_bond_order_factor(t, cls, j)) for j in range(n)]
        assert got == [str(x) for x in g["bond_param_names"][k][cls - 1][:n]]
        assert all(L.pysic_parameter_description_of_bond_order_factor(t, cls, j)
                   for j in range(n))
    assert L.pysic_number_of_parameters_of_bond_order_factor(t, 3) == 0
    assert L.pysic_description_of_bond_order_factor(t)
    for other in names:
        if other != "tersoff":
            assert L.pysic_is_bond_order_factor(other.encode()) == 0


def test_python_objects_follow_the_registry(L):
    """The Python Potential / BondOrderParameters tables name the same types,
    targets, parameters and descriptor codes as the C registry."""
    from pysic_amd.interactions import BOND_ORDER_TYPES, POTENTIAL_CODES, POTENTIAL_TYPES
    names = [_s(L.pysic_potential_name(k)) for k in range(L.pysic_number_of_potentials())]
    assert sorted(names) == sorted(POTENTIAL_TYPES)
    for name, (nt, params) in POTENTIAL_TYPES.items():
        nb = name.encode()
        assert L.pysic_number_of_targets_of_potential(nb) == nt
        assert [_s(L.pysic_parameter_name_of_potential(nb, k)) for k in range(len(params))] == \
            params
        assert L.pysic_potential_type_code(nb) == POTENTIAL_CODES[name]
    nt, level, groups = BOND_ORDER_TYPES["tersoff"]
    assert (nt, level) == (3, 2)
    for cls, params in enumerate(groups, start=1):
        assert [_s(L.pysic_parameter_name_of_bond_order_factor(b"tersoff", cls, k))
                for k in range(len(params))] == params
