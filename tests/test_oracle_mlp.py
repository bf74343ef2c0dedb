"""Pins of the learned-part oracle (oracle/mlp_oracle.py) on the CPU.

Each check is independent of the oracle's own formulas: torch's BF16
conversion and float64 Linear/tanh layers, central finite differences of the
loss, and the closed form of an isolated pair."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(M.NP)
    P[M.OFF_W1:M.OFF_B1] = rng.normal(0, scale / math.sqrt(32), 32 * 64)
    P[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, 4 * 64)
    P[M.OFF_W2:M.OFF_B2] = rng.normal(0, scale / math.sqrt(64), 64 * 64)
    P[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, 64)
    P[M.OFF_W3:M.OFF_B3] = rng.normal(0, 0.3, 64)
    P[M.OFF_B3:] = rng.normal(0, 0.5, 4)
    return P


def test_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000) , rng.normal(0, 1e4, 1000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, -1.00390625, 65504.0]])
    x = x.astype(np.float32)
    # exact ties: low 16 bits 0x8000 with even and odd kept parts
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x), want)


def test_forward_without_rounding_matches_torch_layers():
    import torch
    P = random_params(1)
    W1, b1, W2, b2, w3, b3 = M.unpack(P)
    rng = np.random.default_rng(2)
    feat = rng.normal(0, 2, (50, 32))
    sp = rng.integers(0, 2, 50)
    mean = rng.normal(0, 1, 32)
    isc = rng.uniform(0.5, 2, 32)
    y, _ = M.forward(P, feat, sp, mean, isc, rounding=False)
    l1 = torch.nn.Linear(32, 64, bias=False).double()
    l2 = torch.nn.Linear(64, 64).double()
    l3 = torch.nn.Linear(64, 1, bias=False).double()
    with torch.no_grad():
        l1.weight.copy_(torch.from_numpy(W1.T))
        l2.weight.copy_(torch.from_numpy(W2.T))
        l2.bias.copy_(torch.from_numpy(b2))
        l3.weight.copy_(torch.from_numpy(w3[None, :]))
        x = torch.from_numpy((feat - mean) * isc)
        h1 = torch.tanh(l1(x) + torch.from_numpy(b1[sp]))
        h2 = torch.tanh(l2(h1))
        yt = (l3(h2)[:, 0] + torch.from_numpy(b3[sp])).numpy()
    assert np.abs(y - yt).max() < 1e-12


def test_gradient_without_rounding_matches_finite_differences():
    P = random_params(3)
    rng = np.random.default_rng(4)
    n = 40
    feat = rng.normal(0, 1.5, (n, 32))
    sp = rng.integers(0, 2, n)
    targ = rng.normal(0, 0.5, n)
    mean = np.zeros(32)
    isc = np.ones(32)
    g, loss = M.gradients(P, feat, sp, targ, mean, isc, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 20),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_B1 + 128), 10),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 20),
                          np.arange(M.OFF_B2, M.OFF_B2 + 5), np.arange(M.OFF_W3, M.OFF_W3 + 5),
                          [M.OFF_B3, M.OFF_B3 + 1]])
    h = 1e-6
    for p in idx:
        a = P.copy()
        b = P.copy()
        a[p] += h
        b[p] -= h
        la = M.gradients(a, feat, sp, targ, mean, isc, rounding=False)[1]
        lb = M.gradients(b, feat, sp, targ, mean, isc, rounding=False)[1]
        fd = (la - lb) / (2 * h)
        assert abs(fd - g[p]) <= 1e-7 * max(1.0, abs(g[p])) + 1e-9, (p, fd, g[p])
    # unused species slots get no gradient
    assert g[M.OFF_B3 + 2] == 0.0 and g[M.OFF_B3 + 3] == 0.0


def test_descriptor_dimer_closed_form():
    rc = 6.0
    mu = np.linspace(1.0, 5.5, 16)
    eta = np.full(16, 1.0 / (2 * 0.3 ** 2))
    r = 2.345
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
    cell = np.eye(3) * 40.0
    G = M.descriptors(pos, cell, [True] * 3, [0, 1], 2, rc, mu, eta)
    fc = 0.5 * (math.cos(math.pi * r / rc) + 1)
    expect = np.array([math.exp(-e * (r - m) ** 2) * fc for m, e in zip(mu, eta)])
    # atom 0 sees one species-1 neighbor, atom 1 one species-0 neighbor
    assert np.allclose(G[0, 16:], expect, rtol=0, atol=1e-15) and not G[0, :16].any()
    assert np.allclose(G[1, :16], expect, rtol=0, atol=1e-15) and not G[1, 16:].any()
    # beyond the cutoff: nothing
    pos2 = np.array([[10.0, 10.0, 10.0], [16.5, 10.0, 10.0]])
    assert not M.descriptors(pos2, cell, [True] * 3, [0, 1], 2, rc, mu, eta).any()


def test_descriptor_invariances_and_images():
    from pysic_amd import structures as st
    s = st.amorphous_sio2(16, 2.2, seed=1)
    names = ["O", "Si"]
    sp = np.array([names.index(x) for x in s.symbols])
    mu = np.linspace(1.0, 5.5, 16)
    eta = np.full(16, 5.0)
    G = M.descriptors(s.positions, s.cell, [True] * 3, sp, 2, 6.0, mu, eta)
    # translation (with wrapping) and permutation leave each atom's row unchanged
    shifted = (s.positions + np.array([3.1, -7.2, 0.4])) % s.cell[0, 0]
    assert np.allclose(M.descriptors(shifted, s.cell, [True] * 3, sp, 2, 6.0, mu, eta), G,
                       atol=1e-12)
    perm = np.random.default_rng(0).permutation(len(s))
    Gp = M.descriptors(s.positions[perm], s.cell, [True] * 3, sp[perm], 2, 6.0, mu, eta)
    assert np.allclose(Gp, G[perm], atol=1e-12)
    # the box (L ~ 8.5 A) is smaller than 2 rc: periodic images must count; a
    # 2x2x2 replica has the same per-atom descriptors
    big = np.concatenate([s.positions + s.cell[0, 0] * np.array([a, b, c])
                          for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    Gb = M.descriptors(big, 2 * s.cell, [True] * 3, np.tile(sp, 8), 2, 6.0, mu, eta)
    assert np.allclose(Gb[:len(s)], G, atol=1e-10)


def test_corrector_dimer_closed_form():
    rc = 6.0
    mu = np.linspace(1.0, 5.5, 16)
    eta = np.full(16, 1.0 / (2 * 0.3 ** 2))
    r = 3.21
    pos = np.array([[10.0, 10.0, 10.0], [10.0, 10.0 + r, 10.0]])
    cell = np.eye(3) * 40.0
    chi = np.array([-1.25, 2.5])
    prev = np.array([0.125, -0.5])
    f = M.corrector_features(pos, cell, [True] * 3, [0, 1], rc, mu, eta, chi, prev)
    fc = 0.5 * (math.cos(math.pi * r / rc) + 1)
    phi = np.array([math.exp(-e * (r - m) ** 2) * fc for m, e in zip(mu, eta)])
    # the mixed pair has class 1 for both atoms
    expect = np.zeros((2, M.NCOEF))
    expect[0, 16:32] = phi * 3.75
    expect[0, 64:80] = phi * -0.625
    expect[0, 96] = 0.125
    expect[1, 16:32] = phi * -3.75
    expect[1, 64:80] = phi * 0.625
    expect[1, 96] = -0.5
    assert np.allclose(f, expect, rtol=0, atol=1e-14)
    # same species: class 0 (both species 0) or 2 (both species 1)
    f0 = M.corrector_features(pos, cell, [True] * 3, [0, 0], rc, mu, eta, chi, prev)
    f2 = M.corrector_features(pos, cell, [True] * 3, [1, 1], rc, mu, eta, chi, prev)
    assert np.allclose(f0[0, 0:16], phi * 3.75, atol=1e-14) and not f0[0, 16:48].any()
    assert np.allclose(f2[0, 32:48], phi * 3.75, atol=1e-14) and not f2[0, 0:32].any()


def test_corrector_shift_sum_rule_and_permutation():
    from pysic_amd import structures as st
    s = st.amorphous_sio2(16, 2.2, seed=3)
    sp = np.array([["O", "Si"].index(x) for x in s.symbols])
    rng = np.random.default_rng(5)
    n = len(s)
    chi = rng.normal(size=n)
    prev = rng.normal(size=n) * 0.01
    coef = rng.normal(size=M.NCOEF)
    mu = np.linspace(1.0, 5.5, 16)
    eta = np.full(16, 1.0 / (2 * 0.3 ** 2))
    args = (s.positions, s.cell, [True] * 3, sp, 6.0, mu, eta)
    f = M.corrector_features(*args, chi, prev)
    # chi enters through differences only
    assert np.allclose(M.corrector_features(*args, chi + 3.7, prev), f, rtol=0, atol=1e-12)
    # symmetric pair weights: the pair terms sum to zero over the atoms
    dq = f @ coef
    assert abs(dq.sum() - coef[-1] * prev.sum()) <= 1e-12 * np.abs(dq).sum()
    # relabelling the atoms relabels the result
    perm = rng.permutation(n)
    fp = M.corrector_features(s.positions[perm], s.cell, [True] * 3, sp[perm], 6.0, mu, eta,
                              chi[perm], prev[perm])
    assert np.allclose(fp, f[perm], rtol=0, atol=1e-12)
    # the cell (L ~ 8.5 A) is narrower than 2 rc: a 2x2x2 replica of the same
    # charges gives the same rows
    big = np.concatenate([s.positions + s.cell[0, 0] * np.array([a, b, c])
                          for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    fb = M.corrector_features(big, 2 * s.cell, [True] * 3, np.tile(sp, 8), 6.0, mu, eta,
                              np.tile(chi, 8), np.tile(prev, 8))
    assert np.allclose(fb[:n], f, rtol=0, atol=1e-10)
