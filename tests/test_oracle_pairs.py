"""CPU tests that pin the oracle's pair potentials, products and one-body
terms (oracle/pysic_oracle.c: orc_pair_energy, orc_pair_force,
orc_product_*, the singlet loop).

Pins, each chosen so that a plausible slip fails one of them:
  * golden vectors of the reference Fortran core (tests/golden/pairs_*,
    lj_argon_*), every type, a 3-factor product and a table;
  * the known answer of the reference's documentation: two He atoms 3 A
    apart under LJ(epsilon 0.1, sigma 2.5) (doc/run examples.rst:106-145);
  * the LJ energy of a perfect fcc crystal as a direct lattice sum;
  * F = -dE/dr by central differences for every type whose reference force
    is the gradient of its energy (all but 'step');
  * the stress as the strain derivative of the energy;
  * Newton's third law: the pair forces sum to zero.
"""
import os

import numpy as np
import pytest

from conftest import load_golden
from helpers import MIXTURE_TABLE_ID, mixture_model, mixture_table, rel_maxnorm, ref_entries
from pysic_amd import structures as st
from pysic_amd.interactions import ARGON_LJ, Potential, build_descriptors, lj_argon

PAIR_CASES = ["pairs_arkr_mixture_120", "lj_argon_fcc_108", "lj_argon_fcc_108_smooth"]


def _case(O, name):
    g = load_golden(name)
    if int(g["table_id"]) >= 0:
        O.set_table(int(g["table_id"]), g["table"])
    nl = O.neighbor_list(g["positions"], g["cell"], g["pbc"], g["cutoffs"])
    return g, nl


@pytest.mark.parametrize("name", PAIR_CASES)
def test_pairs_match_reference_core(oracle, name):
    g, nl = _case(oracle, name)
    args = (g["species"], g["positions"], None, g["cell"], nl, g["potdesc"])
    e = oracle.evaluate("energy", *args)
    f, s = oracle.evaluate("forces", *args)
    assert abs(e - float(g["energy"])) <= 1e-12 * abs(float(g["energy"]))
    assert np.abs(f - g["forces"]).max() <= 1e-12 * max(1.0, np.abs(g["forces"]).max())
    assert np.abs(s - g["stress"]).max() <= 1e-12 * max(1.0, np.abs(g["stress"]).max())


def _two_atoms(O, pots, r=3.0, box=40.0):
    names = sorted(set(x for p in pots for t in p.get_symbols() for x in t))
    pd, _, _ = build_descriptors(pots, {s: k for k, s in enumerate(names)})
    pos = np.array([[5.0, 5.0, 5.0], [5.0, 5.0, 5.0 + r]])
    cell = np.eye(3) * box
    nl = O.neighbor_list(pos, cell, [0, 0, 0], np.full(2, max(p.get_cutoff() for p in pots)))
    sp = np.zeros(2, np.int32)
    return (sp, pos, None, cell, nl, pd)


def test_he2_known_answer(oracle):
    """doc/run examples.rst:136-141: E = -0.022274132189576905 eV,
    F = (0, 0, +-0.02211693) eV/A."""
    pot = Potential("LJ", symbols=[["He", "He"]], parameters=[0.1, 2.5], cutoff=10.0)
    args = _two_atoms(oracle, [pot])
    e = oracle.evaluate("energy", *args)
    f, _ = oracle.evaluate("forces", *args)
    assert abs(e - (-0.022274132189576905)) < 1e-15
    assert np.allclose(f, [[0, 0, 0.02211693], [0, 0, -0.02211693]], atol=5e-9)


def test_lj_fcc_lattice_sum(oracle):
    """Perfect fcc argon: E/N = 1/2 sum over lattice vectors |R| < rc of v(|R|)."""
    a, rc = 5.26, 2.5 * ARGON_LJ["sigma"]
    s = st.fcc((4, 4, 4), a=a)
    pd, _, _ = build_descriptors(lj_argon(cutoff=rc), {"Ar": 0})
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], np.full(len(s), rc + 0.5))
    e = oracle.evaluate("energy", np.zeros(len(s), np.int32), s.positions, None, s.cell, nl, pd)
    # lattice vectors: fcc basis + integer translations of the cube
    m = 4
    g = np.array([[i, j, k] for i in range(-m, m + 1) for j in range(-m, m + 1)
                  for k in range(-m, m + 1)], float)
    R = ((g[:, None, :] + st.FCC_BASIS[None, :, :]).reshape(-1, 3)) * a
    d = np.sqrt((R * R).sum(1))
    d = d[(d > 0) & (d < rc)]
    eps, sig = ARGON_LJ["epsilon"], ARGON_LJ["sigma"]
    e_closed = 0.5 * np.sum(eps * ((sig / d) ** 12 - (sig / d) ** 6))
    assert abs(e / len(s) - e_closed) < 1e-12 * abs(e_closed)


def _mixture(O, without_step=True):
    s = st.random_mixture(48, 12.5, 2.9, ["Kr", "Ar", "Ar"], seed=5)
    O.set_table(MIXTURE_TABLE_ID, mixture_table())
    model = [p for p in mixture_model()
             if not (without_step and p.get_potentials()[0].get_potential_type() == "step")]
    names = sorted(set(s.symbols))
    pd, _, _ = build_descriptors(model, {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    cut = np.full(len(s), 8.5)
    return s, sp, pd, cut


def test_mixture_forces_are_energy_gradient(oracle):
    s, sp, pd, cut = _mixture(oracle)
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, s.positions, None, s.cell, nl, pd)
    h = 1e-5
    rng = np.random.default_rng(3)
    for i in rng.choice(len(s), 6, replace=False):
        for c in range(3):
            ep = []
            for sgn in (1, -1):
                p = s.positions.copy()
                p[i, c] += sgn * h
                ep.append(oracle.evaluate("energy", sp, p, None, s.cell, nl, pd))
            fd = -(ep[0] - ep[1]) / (2 * h)
            assert abs(fd - f[i, c]) < 2e-6 * max(1.0, abs(f[i, c])), (i, c, fd, f[i, c])


def test_mixture_stress_is_strain_derivative(oracle):
    s, sp, pd, cut = _mixture(oracle)
    # one-body 'force' terms make E depend on absolute positions; drop them
    rows = pd[pd[:, 0] != 11]
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
    _, st6 = oracle.evaluate("forces", sp, s.positions, None, s.cell, nl, rows)
    h = 1e-6
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for k, (a, b) in enumerate(pairs):
        es = []
        for sgn in (1, -1):
            eps = np.zeros((3, 3))
            eps[a, b] += sgn * h
            if a != b:
                eps[b, a] += sgn * h
            def_ = np.eye(3) + (eps if a == b else 0.5 * eps)
            es.append(oracle.evaluate("energy", sp, s.positions @ def_.T, None,
                                      s.cell @ def_.T, nl, rows))
        dE = (es[0] - es[1]) / (2 * h)
        # W = sum r (x) f = -dE/d(strain)
        assert abs(st6[k] + dE) < 1e-6 * max(1.0, abs(dE)), (k, st6[k], dE)


def test_pair_forces_sum_to_one_body_forces(oracle):
    s, sp, pd, cut = _mixture(oracle, without_step=False)
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
    f, _ = oracle.evaluate("forces", sp, s.positions, None, s.cell, nl, pd)
    nkr = sum(1 for x in s.symbols if x == "Kr")
    assert np.allclose(f.sum(0), nkr * np.array([0.01, -0.02, 0.03]), atol=1e-10)


def test_reference_core_agrees_on_a_fresh_mixture(oracle):
    """The C restatement against the reference Fortran core on a structure
    none of the fixtures holds (only where the reference build exists)."""
    if not oracle.ref_available():
        pytest.skip("reference checker not built (oracle/build_ref.sh)")
    s, sp, pd, cut = _mixture(oracle, without_step=False)
    names = sorted(set(s.symbols))
    R = oracle.RefCore()
    R.setup(s.symbols, s.positions, s.cell, s.pbc, potentials=ref_entries(pd, names), cutoffs=cut,
            tables={MIXTURE_TABLE_ID: mixture_table()})
    nl = oracle.neighbor_list(s.positions, s.cell, [1, 1, 1], cut)
    e = oracle.evaluate("energy", sp, s.positions, None, s.cell, nl, pd)
    f, st6 = oracle.evaluate("forces", sp, s.positions, None, s.cell, nl, pd)
    fr, sr = R.forces()
    assert abs(e - R.energy()) <= 1e-12 * abs(e)
    assert rel_maxnorm(f, fr) <= 1e-12 and rel_maxnorm(st6, sr) <= 1e-12


def test_lj_rebuild_interval_keeps_every_pair(oracle):
    """The LJ workload's list policy (0.5 A skin, rebuilt every 10 steps of
    2 fs, fcc argon at 80 K): velocity Verlet on the oracle, 864 atoms, 30
    steps.  No atom moves more than skin/2 between rebuilds, and no pair
    inside the cutoff is ever missing from the current list."""
    s = st.fcc((6, 6, 6), a=5.26, sigma=0.05, seed=0)
    st.maxwell_boltzmann(s, 80.0, seed=1)
    n, rc, skin, every = len(s), 2.5 * ARGON_LJ["sigma"], 0.5, 10
    pd, _, _ = build_descriptors(lj_argon(cutoff=rc), {"Ar": 0})
    sp, m, box = np.zeros(n, np.int32), s.masses[:, None], s.cell[0, 0]
    dt = 2.0 * st.ASE_FS

    def within(r, d):
        dx = r[:, None, :] - r[None, :, :]
        dx -= box * np.round(dx / box)
        d2 = (dx ** 2).sum(-1)
        np.fill_diagonal(d2, np.inf)
        return d2 < d * d

    r, p = s.positions.copy(), s.momenta.copy()
    nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], np.full(n, rc + skin))
    listed, r_ref = within(r, rc + skin), r.copy()
    f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd)
    worst = 0.0
    for step in range(1, 31):
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        if step % every == 0:
            nl = oracle.neighbor_list(r, s.cell, [1, 1, 1], np.full(n, rc + skin))
            listed, r_ref = within(r, rc + skin), r.copy()
        f, _ = oracle.evaluate("forces", sp, r, None, s.cell, nl, pd)
        p = p + 0.5 * dt * f
        worst = max(worst, float(np.sqrt(((r - r_ref) ** 2).sum(1)).max()))
        assert not (within(r, rc) & ~listed).any(), step
    assert worst < 0.5 * skin
