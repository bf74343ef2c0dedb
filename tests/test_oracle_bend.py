"""CPU tests that pin the oracle's bond-bending term (oracle/pysic_oracle.c:
orc_bend, orc_triplet and the triplet enumeration in orc_evaluate).

Pins, each chosen so that a plausible slip fails one of them:
  * golden vectors of the reference Fortran core on amorphous SiO2 with
    random charges: COMB's Si-O-Si and O-Si-O bending alone
    (tests/golden/sio2_48_comb_bend.npz) and with COMB's charge terms
    (sio2_48_comb_bend_charge.npz): energy, forces, stress, electronegativities;
  * the closed form of one angle, E = eps (cos^n theta - cos^n theta_0)^m,
    with the vertex atom first, second or last in the atom order;
  * nothing acts when the vertex is not the potential's second target;
  * each angle counted once: on a Si/O rocksalt lattice with 1.6 A bonds
    every O vertex has 15 angles, 12 of 90 and 3 of 180 degrees;
  * F = -dE/dr by central differences (smooth cutoffs included) and the
    stress as the strain derivative of the energy.
"""
import math

import numpy as np
import pytest

from conftest import load_golden
from helpers import comb_bend_terms, comb_charges, ref_entries, rel_maxnorm
from pysic_amd import structures as st
from pysic_amd.interactions import Potential, build_descriptors

NAMES = ["sio2_48_comb_bend", "sio2_48_comb_bend_charge"]


def _golden(O, name):
    g = load_golden(name)
    nl = O.neighbor_list(g["positions"], g["cell"], g["pbc"], g["cutoffs"])
    return g, nl, (g["species"], g["positions"], g["charges"], g["cell"], nl, g["potdesc"])


@pytest.mark.parametrize("name", NAMES)
def test_bend_matches_reference_core(oracle, name):
    g, _, args = _golden(oracle, name)
    e = oracle.evaluate("energy", *args)
    f, s = oracle.evaluate("forces", *args)
    x = oracle.evaluate("enegs", *args)
    assert abs(e - float(g["energy"])) <= 1e-12 * abs(float(g["energy"]))
    assert rel_maxnorm(f, g["forces"]) <= 1e-12
    assert rel_maxnorm(s, g["stress"]) <= 1e-12
    assert np.abs(x - g["enegs"]).max() <= 1e-12 * max(1.0, np.abs(g["enegs"]).max())


def _triplet(O, pots, order, theta, r=(1.6, 1.7)):
    """Si-O-Si with the O vertex: atom positions in the given order of
    ('a', 'v', 'b'), legs r[0] and r[1], angle theta, in a large open box."""
    v = np.array([10.0, 10.0, 10.0])
    where = {"v": v, "a": v + [r[0], 0.0, 0.0],
             "b": v + r[1] * np.array([math.cos(theta), math.sin(theta), 0.0])}
    sym = {"v": "O", "a": "Si", "b": "Si"}
    names = ["O", "Si"]
    pos = np.array([where[k] for k in order])
    sp = np.array([names.index(sym[k]) for k in order], np.int32)
    pd, _, _ = build_descriptors(pots, {s: k for k, s in enumerate(names)})
    cell = np.eye(3) * 40.0
    nl = O.neighbor_list(pos, cell, [0, 0, 0], np.full(3, 3.55))
    return O.evaluate("energy", sp, pos, np.zeros(3), cell, nl, pd)


@pytest.mark.parametrize("order", ["avb", "vab", "abv", "bva"])
def test_single_angle_closed_form(oracle, order):
    eps, th0, theta = 2.6, 2.508561734, math.radians(117.0)
    pot = Potential("bond_bend", symbols=[["Si", "O", "Si"]], cutoff=3.05, cutoff_margin=0.5,
                    parameters=[eps, th0, 1, 2])
    e = _triplet(oracle, [pot], order, theta)
    want = eps * (math.cos(theta) - math.cos(th0)) ** 2
    assert abs(e - want) <= 1e-13 * want
    # n = 3, m = 1: eps (cos^3 theta - cos^3 theta_0)
    pot3 = Potential("bond_bend", symbols=[["Si", "O", "Si"]], cutoff=3.05, cutoff_margin=0.5,
                     parameters=[eps, th0, 3, 1])
    e3 = _triplet(oracle, [pot3], order, theta)
    want3 = eps * (math.cos(theta) ** 3 - math.cos(th0) ** 3)
    assert abs(e3 - want3) <= 1e-13 * abs(want3)


def test_vertex_must_be_second_target(oracle):
    # an O-Si-O potential on a Si-O-Si angle: the vertex O is not its
    # second target; nor is it for Si-Si-O (Si would have to be the vertex;
    # at 150 degrees the Si-Si distance, 3.19 A, is past the cutoff, so no
    # angle has a Si vertex)
    for syms in (["O", "Si", "O"], ["Si", "Si", "O"]):
        pot = Potential("bond_bend", symbols=[syms], cutoff=3.05, cutoff_margin=0.5,
                        parameters=[1.0, 2.0, 1, 2])
        assert _triplet(oracle, [pot], "avb", math.radians(150.0)) == 0.0


def test_each_angle_counted_once(oracle):
    """Rocksalt-like simple cubic lattice of alternating Si/O at spacing
    1.6 A (cutoff 2.0 A: only the 6 nearest neighbors).  Every O vertex
    has C(6,2) = 15 angles: 12 of 90 degrees and 3 of 180."""
    s = st.rocksalt((2, 2, 2), a=3.2)
    names = ["O", "Si"]
    symbols = ["Si" if q > 0 else "O" for q in s.charges]
    sp = np.array([names.index(x) for x in symbols], np.int32)
    eps, th0 = 1.0, 2.0
    pot = Potential("bond_bend", symbols=[["Si", "O", "Si"]], cutoff=2.0, parameters=[eps, th0, 1, 2])
    pd, _, _ = build_descriptors([pot], {x: k for k, x in enumerate(names)})
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(len(s), 2.0))
    e = oracle.evaluate("energy", sp, s.positions, np.zeros(len(s)), s.cell, nl, pd)
    n_o = symbols.count("O")
    want = n_o * eps * (12 * (0.0 - math.cos(th0)) ** 2 + 3 * (-1.0 - math.cos(th0)) ** 2)
    assert abs(e - want) <= 1e-12 * want


def _comb_case(O, seed, n_si=16):
    s = st.amorphous_sio2(n_si, 2.2, seed=seed)
    q = comb_charges(s.symbols, seed=seed + 1)
    names = sorted(set(s.symbols))
    pd, _, _ = build_descriptors(comb_bend_terms(), {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    cut = np.full(len(s), 3.55)
    nl = O.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
    return s, sp, q, pd, nl


def test_bend_forces_are_energy_gradient(oracle):
    s, sp, q, pd, nl = _comb_case(oracle, 81)
    f, _ = oracle.evaluate("forces", sp, s.positions, q, s.cell, nl, pd)
    h = 1e-5
    for i in range(0, len(s), 5):
        for a in range(3):
            pp, pm = s.positions.copy(), s.positions.copy()
            pp[i, a] += h
            pm[i, a] -= h
            d = (oracle.evaluate("energy", sp, pp, q, s.cell, nl, pd)
                 - oracle.evaluate("energy", sp, pm, q, s.cell, nl, pd)) / (2 * h)
            assert abs(f[i, a] + d) <= 1e-6 * max(1.0, abs(d)), (i, a)


def test_bend_stress_is_strain_derivative(oracle):
    """W_ab = -dE/de_ab at zero strain (positions and cell strained
    together), in the component order xx yy zz yz xz xy."""
    s, sp, q, pd, nl = _comb_case(oracle, 83)
    _, w = oracle.evaluate("forces", sp, s.positions, q, s.cell, nl, pd)
    h = 1e-6
    comps = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for c, (a, b) in enumerate(comps):
        def energy(eps):
            m = np.eye(3)
            m[a, b] += eps
            return oracle.evaluate("energy", sp, s.positions @ m, q, s.cell @ m, nl, pd)
        d = (energy(h) - energy(-h)) / (2 * h)
        assert abs(w[c] + d) <= 1e-5 * max(1.0, abs(d)), c


@pytest.mark.parametrize("n_si", [16, 4])
def test_reference_core_agrees_on_fresh_bend(oracle, n_si):
    """Against the reference core (where built) on fresh structures; with
    4 Si the 12 atoms sit in a 5.7 A box, so a neighbor appears at several
    images within the 3.55 A list radius."""
    if not oracle.ref_available():
        pytest.skip("reference checker not built (oracle/build_ref.sh)")
    s, sp, q, pd, nl = _comb_case(oracle, 85, n_si)
    names = sorted(set(s.symbols))
    R = oracle.RefCore()
    R.setup(s.symbols, s.positions, s.cell, s.pbc, charges=q, potentials=ref_entries(pd, names),
            cutoffs=np.full(len(s), 3.55))
    args = (sp, s.positions, q, s.cell, nl, pd)
    e = oracle.evaluate("energy", *args)
    f, w = oracle.evaluate("forces", *args)
    fr, wr = R.forces()
    assert abs(e - R.energy()) <= 1e-12 * abs(e)
    assert rel_maxnorm(f, fr) <= 1e-12 and rel_maxnorm(w, wr) <= 1e-12
