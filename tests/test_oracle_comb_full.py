"""CPU tests that pin the oracle's bond-order-scaled products: COMB as the
reference's builder makes it (comb.py:630-990), each attractive
charge_exp x charge_abs x exponential product scaled by its coordinator's
'tersoff' bond-order factors, three coordinator groups (orc_pair_bond,
orc_postproc and the doublet loop of oracle/pysic_oracle.c).

Pins:
  * the reference Fortran core's energy, forces, stress and
    electronegativities on amorphous SiO2 with random charges
    (tests/golden/sio2_48_comb_full.npz, sio2_96_comb_full.npz);
  * F = -dE/dr and chi = -dE/dq by central differences of the oracle's
    energy (a dropped gradient term of the bond factor, or an unscaled
    electronegativity, fails these);
  * the coordinators matter: without them the energy differs, and setting
    every bond-order beta to zero (b = 1 for eta = 1... rather, zeta -> 0
    gives b = 1) reduces the scaled model to the unscaled one.
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import rel_maxnorm

NAMES = ["sio2_48_comb_full", "sio2_96_comb_full"]


def _golden(O, name):
    g = load_golden(name)
    nl = O.neighbor_list(g["positions"], g["cell"], g["pbc"], g["cutoffs"])
    return g, nl, (g["species"], g["positions"], g["charges"], g["cell"], nl, g["potdesc"],
                   g["bonddesc"])


@pytest.mark.parametrize("name", NAMES)
def test_comb_full_matches_reference_core(oracle, name):
    g, _, args = _golden(oracle, name)
    e = oracle.evaluate("energy", *args)
    f, s = oracle.evaluate("forces", *args)
    x = oracle.evaluate("enegs", *args)
    print("%s: |dE| %.2e rel|dF| %.2e rel|dW| %.2e rel|dchi| %.2e" % (
        name, abs(e - float(g["energy"])), rel_maxnorm(f, g["forces"]),
        rel_maxnorm(s, g["stress"]), rel_maxnorm(x, g["enegs"])))
    assert abs(e - float(g["energy"])) <= 1e-12 * abs(float(g["energy"]))
    assert rel_maxnorm(f, g["forces"]) <= 1e-12
    assert rel_maxnorm(s, g["stress"]) <= 1e-12
    assert rel_maxnorm(x, g["enegs"]) <= 1e-12


def test_comb_full_gradients(oracle):
    g, nl, args = _golden(oracle, "sio2_48_comb_full")
    sp, pos, q, cell, _, pd, bd = args
    f, _ = oracle.evaluate("forces", *args)
    x = oracle.evaluate("enegs", *args)
    h = 1e-5
    rng = np.random.default_rng(3)
    for a in rng.choice(len(pos), 6, replace=False):
        for c in range(3):
            e = []
            for sg in (1, -1):
                p = pos.copy()
                p[a, c] += sg * h
                e.append(oracle.evaluate("energy", sp, p, q, cell, nl, pd, bd))
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fd - f[a, c]) <= 1e-6 * max(1.0, abs(f[a, c])), (a, c, fd, f[a, c])
        e = []
        for sg in (1, -1):
            qq = q.copy()
            qq[a] += sg * h
            e.append(oracle.evaluate("energy", sp, pos, qq, cell, nl, pd, bd))
        fd = -(e[0] - e[1]) / (2 * h)
        assert abs(fd - x[a]) <= 1e-6 * max(1.0, abs(x[a])), (a, fd, x[a])


def test_coordinators_change_the_energy(oracle):
    from pysic_amd.interactions import build_descriptors, comb_sio2
    g, _, args = _golden(oracle, "sio2_48_comb_full")
    sp, pos, q, cell, nl, pd, bd = args
    names = sorted(set(str(x) for x in g["symbols"]))
    pd0, bd0, _ = build_descriptors(comb_sio2(coordinators=False),
                                    {x: k for k, x in enumerate(names)})
    e_scaled = oracle.evaluate("energy", sp, pos, q, cell, nl, pd, bd)
    e_plain = oracle.evaluate("energy", sp, pos, q, cell, nl, pd0, bd0)
    assert len(bd0) == 0 and abs(e_scaled - e_plain) > 1.0
    # beta = 0: zeta^eta -> 0, every bond factor is 1 and the scaled model
    # is the unscaled one
    bz = np.array(bd, dtype=np.float64).copy()
    bz[:, 3] = 0.0
    e_b0 = oracle.evaluate("energy", sp, pos, q, cell, nl, pd, bz)
    assert abs(e_b0 - e_plain) <= 1e-10 * abs(e_plain)
