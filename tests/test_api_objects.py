"""CPU tests of the pysic API objects and their translation to descriptors."""
import numpy as np
import pytest

from pysic_amd import (BondOrderParameters, CoulombSummation, Coordinator, Potential,
                       estimate_ewald_parameters, structures as st, tersoff_si)
from pysic_amd.interactions import (InvalidParametersError, InvalidPotentialError,
                                    build_descriptors)


def test_unknown_potential_raises():
    with pytest.raises(InvalidPotentialError):
        Potential("no_such_potential")


def test_parameter_names_and_values():
    p = Potential("exponential", symbols=[["Si", "O"]], parameters=[2.0, 1.5], cutoff=4.0,
                  cutoff_margin=0.5)
    assert p.get_parameter_names() == ["epsilon", "zeta"]
    assert p.get_soft_cutoff() == 3.5
    p.set_parameter_value("zeta", 2.5)
    assert p.get_parameter_value("zeta") == 2.5
    with pytest.raises(InvalidParametersError):
        p.set_parameter_values([1.0])


def test_permutations_expand_like_update_core_potentials():
    # (Si,O) is added for both target orders, (Si,Si) once (calculator.py:1166-1192)
    pots = [Potential("exponential", symbols=[["Si", "O"]], parameters=[1, 1], cutoff=3),
            Potential("exponential", symbols=[["Si", "Si"]], parameters=[2, 2], cutoff=3)]
    pd, bd, ng = build_descriptors(pots, {"O": 0, "Si": 1})
    pairs = sorted((int(r[2]), int(r[3])) for r in pd)
    assert pairs == [(0, 1), (1, 0), (1, 1)]
    assert ng == 0 and len(bd) == 0


def test_coordinator_groups_and_unpermuted_bonds():
    pots = tersoff_si()
    pd, bd, ng = build_descriptors(pots, {"Si": 0})
    assert ng == 1
    assert [int(r[8]) for r in pd] == [-1, 0]   # repulsive plain, attractive coordinated
    assert bd.shape == (1, 14)
    assert list(bd[0, 3:10]) == [1.1e-6, 0.78734, 0.0, 100390.0, 16.217, -0.59825, 3.0]
    assert bd[0, 10] == 3.0 and abs(bd[0, 11] - 2.7) < 1e-15
    # two potentials sharing equal bond parameters share the group index
    b = BondOrderParameters("tersoff", cutoff=3, cutoff_margin=0.3,
                            parameters=[[1e-6, 0.8], [0, 1e5, 16, -0.6, 3]],
                            symbols=[["Si", "O", "O"]])
    p1 = Potential("exponential", symbols=[["Si", "O"]], parameters=[1, 1], cutoff=3,
                   coordinator=Coordinator([b]))
    p2 = Potential("exponential", symbols=[["Si", "O"]], parameters=[2, 1], cutoff=3,
                   coordinator=Coordinator([b]))
    pd2, bd2, ng2 = build_descriptors([p1, p2], {"O": 0, "Si": 1})
    assert ng2 == 1 and set(int(r[8]) for r in pd2) == {0}
    assert [int(x) for x in bd2[0, :3]] == [1, 0, 0]


def test_bond_parameter_shape_checked():
    with pytest.raises(InvalidParametersError):
        BondOrderParameters("tersoff", parameters=[[1.0], [1, 2, 3, 4, 5]])


def test_estimate_ewald_parameters():
    rc, kc, sigma, eps = estimate_ewald_parameters(9.0, "normal")
    assert (rc, sigma, eps) == (9.0, 3.0, 0.00552635) and kc == 1.0
    rc, kc, sigma, eps = estimate_ewald_parameters(8.0, "high")
    assert sigma == 2.0 and kc == 2.0


def test_coulomb_summation_parameters():
    c = CoulombSummation(parameters={"real_cutoff": 10.0, "k_cutoff": 2.5, "sigma": 2.0,
                                     "epsilon": 0.00552635})
    assert c.get_realspace_cutoff() == 10.0
    assert c.get_parameters() == [10.0, 2.5, 2.0, 0.00552635]


def test_structures():
    s = st.diamond_si((2, 3, 4))
    assert len(s) == 8 * 24
    assert np.allclose(s.get_volume(), 5.431 ** 3 * 24)
    r = st.rocksalt((2, 2, 2))
    assert len(r) == 64 and r.get_initial_charges().sum() == 0
    a = st.amorphous_si(100, 2.3, seed=1)
    rho = 100 * st.AMU_SI / st.AVOGADRO / (a.get_volume() * 1e-24)
    assert abs(rho - 2.3) < 1e-9
    st.maxwell_boltzmann(a, 600.0, seed=2)
    assert np.abs(a.get_momenta().sum(0)).max() < 1e-9
