"""CPU restatement of the learned charge warm start: descriptors and the
BF16 MLP (forward and backward) in numpy FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product
path).  It is the checker for pysic_amd/csrc/mlp.hip.

The reference has no counterpart for this part (pysic starts its charge
relaxation from the charges on the atoms, relaxation.py:417-450); what is
restated here is the definition in include/pysic_learned.h:

  G[i][16 s + m] = sum_{j != i, r_ij < rc, species_j = s}
                   exp(-eta_m (r_ij - mu_m)^2) (cos(pi r_ij / rc) + 1) / 2
  x  = bf16((G - mean) * inv_scale)            (FP32 arithmetic, then RNE)
  h1 = bf16(tanh(W1^T x + b1[s]))
  h2 = bf16(tanh(W2^T h1 + b2))
  y  = w3 . h2 + b3[s]
with W1, W2, w3 rounded to BF16 and every sum taken exactly (FP64) here.
The gradient is that of mean (y - t)^2 with the roundings taken as the
identity (straight-through), the derivative of tanh written as 1 - h^2 with
the rounded h, as the device computes it.

How it is pinned (tests/test_oracle_mlp.py):
  * bf16(): against torch's float32 -> bfloat16 conversion (round to nearest
    even), including ties and subnormal-free edge values;
  * forward with rounding off: against torch.nn.Linear + tanh in float64;
  * backward with rounding off: against central finite differences of the
    FP64 loss;
  * descriptors: the isolated-dimer closed form, translation and
    permutation invariance, and a periodic image count check.

The residual corrector (include/pysic_learned.h), per stage:
  dq_i = sum_{j != i, r_ij < rc} sum_m phi_m(r_ij) [ a[p][m] (chi_j - chi_i)
                                                  + b[p][m] (d_j - d_i) ] + c d_i
with phi_m the descriptor Gaussians and p = species_i + species_j.  Pinned
by: the isolated-dimer closed form of every feature, invariance of the
features under a uniform shift of chi, sum_i dq_i = c sum_i d_i (symmetric
weights), permutation equivariance, and the charge error of a trained
corrector on held-out cells (relaxed by a dense solve of the linear EEM +
Ewald equations) shrinking stage by stage.
"""
import math

import numpy as np

F, H, S, G = 32, 64, 4, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + S * H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NP = OFF_B3 + S


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even); float32 out."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    assert P.size == NP
    W1 = P[OFF_W1:OFF_B1].reshape(F, H)
    b1 = P[OFF_B1:OFF_W2].reshape(S, H)
    W2 = P[OFF_W2:OFF_B2].reshape(H, H)
    b2 = P[OFF_B2:OFF_W3]
    w3 = P[OFF_W3:OFF_B3]
    b3 = P[OFF_B3:NP]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), np.ravel(b1), np.ravel(W2), np.ravel(b2), np.ravel(w3),
                           np.ravel(b3)]).astype(np.float64)


def _inputs(feat, mean, inv_scale, rounding):
    f32 = np.asarray(feat, dtype=np.float32)
    if rounding:
        x32 = (f32 - np.asarray(mean, np.float32)) * np.asarray(inv_scale, np.float32)
        return bf16(x32).astype(np.float64)
    return (np.asarray(feat, np.float64) - np.asarray(mean, np.float64)) * \
        np.asarray(inv_scale, np.float64)


def forward(P, feat, species, mean, inv_scale, rounding=True):
    """y[n] and the intermediates (x, h1, h2) of the network."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    r = (lambda a: bf16(a).astype(np.float64)) if rounding else (lambda a: a)
    W1r, W2r, w3r = r(W1), r(W2), r(w3)
    sp = np.asarray(species, dtype=np.int64)
    x = _inputs(feat, mean, inv_scale, rounding)
    h1 = r(np.tanh(x @ W1r + b1[sp]))
    h2 = r(np.tanh(h1 @ W2r + b2))
    y = h2 @ w3r + b3[sp]
    return y, dict(x=x, h1=h1, h2=h2, W2=W2r, w3=w3r, sp=sp)


def gradients(P, feat, species, targets, mean, inv_scale, rounding=True):
    """(grad[NP], loss) of mean (y - t)^2 over the given samples."""
    y, c = forward(P, feat, species, mean, inv_scale, rounding)
    n = len(y)
    res = y - np.asarray(targets, np.float64)
    loss = float(np.mean(res * res))
    gy = 2.0 * res / n
    x, h1, h2, W2r, w3r, sp = c["x"], c["h1"], c["h2"], c["W2"], c["w3"], c["sp"]
    g2 = gy[:, None] * w3r[None, :] * (1.0 - h2 * h2)
    g1 = (g2 @ W2r.T) * (1.0 - h1 * h1)
    dW1 = x.T @ g1
    db1 = np.zeros((S, H))
    np.add.at(db1, sp, g1)
    dW2 = h1.T @ g2
    db2 = g2.sum(axis=0)
    dw3 = h2.T @ gy
    db3 = np.zeros(S)
    np.add.at(db3, sp, gy)
    return pack(dW1, db1, dW2, db2, dw3, db3), loss


def descriptors(positions, cell, pbc, species, nspecies, rc, mu, eta):
    """G[n][32] by brute force over periodic images (FP64)."""
    pos = np.asarray(positions, np.float64)
    cell = np.asarray(cell, np.float64)
    n = len(pos)
    sp = np.asarray(species, np.int64)
    mu = np.asarray(mu, np.float64)
    eta = np.asarray(eta, np.float64)
    vol = abs(np.linalg.det(cell))
    reach = []
    for a in range(3):
        if pbc[a]:
            height = vol / np.linalg.norm(np.cross(cell[(a + 1) % 3], cell[(a + 2) % 3]))
            reach.append(int(math.ceil(rc / height)))
        else:
            reach.append(0)
    shifts = np.array([[i, j, k] for i in range(-reach[0], reach[0] + 1)
                       for j in range(-reach[1], reach[1] + 1)
                       for k in range(-reach[2], reach[2] + 1)], dtype=np.float64) @ cell
    out = np.zeros((n, F))
    for i in range(n):
        d = pos[None, :, :] + shifts[:, None, :] - pos[i][None, None, :]
        r = np.sqrt((d * d).sum(axis=2))            # [shift, j]
        ok = (r < rc) & (r > 0.0)
        rr = r[ok]
        jj = np.broadcast_to(np.arange(n)[None, :], r.shape)[ok]
        fc = 0.5 * (np.cos(np.pi * rr / rc) + 1.0)
        g = np.exp(-eta[None, :] * (rr[:, None] - mu[None, :]) ** 2) * fc[:, None]
        for s in range(nspecies):
            out[i, s * G:(s + 1) * G] = g[sp[jj] == s].sum(axis=0)
    return out


NCOEF = 2 * 3 * G + 1


def _pairs(positions, cell, pbc, rc):
    """(i, j, r) of every ordered pair j != i (all periodic images) with r < rc."""
    pos = np.asarray(positions, np.float64)
    cell = np.asarray(cell, np.float64)
    n = len(pos)
    vol = abs(np.linalg.det(cell))
    reach = []
    for a in range(3):
        if pbc[a]:
            height = vol / np.linalg.norm(np.cross(cell[(a + 1) % 3], cell[(a + 2) % 3]))
            reach.append(int(math.ceil(rc / height)))
        else:
            reach.append(0)
    shifts = np.array([[i, j, k] for i in range(-reach[0], reach[0] + 1)
                       for j in range(-reach[1], reach[1] + 1)
                       for k in range(-reach[2], reach[2] + 1)], dtype=np.float64) @ cell
    I, J, R = [], [], []
    for i in range(n):
        d = pos[None, :, :] + shifts[:, None, :] - pos[i][None, None, :]
        r = np.sqrt((d * d).sum(axis=2))
        ok = (r < rc) & (r > 0.0)
        jj = np.broadcast_to(np.arange(n)[None, :], r.shape)[ok]
        I.append(np.full(len(jj), i))
        J.append(jj)
        R.append(r[ok])
    return np.concatenate(I), np.concatenate(J), np.concatenate(R)


def corrector_features(positions, cell, pbc, species, rc, mu, eta, chi, prev):
    """features[n][97] of one corrector stage, so that dq = features @ coef."""
    sp = np.asarray(species, np.int64)
    chi = np.asarray(chi, np.float64)
    prev = np.asarray(prev, np.float64)
    mu = np.asarray(mu, np.float64)
    eta = np.asarray(eta, np.float64)
    n = len(sp)
    I, J, R = _pairs(positions, cell, pbc, rc)
    fc = 0.5 * (np.cos(np.pi * R / rc) + 1.0)
    phi = np.exp(-eta[None, :] * (R[:, None] - mu[None, :]) ** 2) * fc[:, None]
    p = sp[I] + sp[J]
    out = np.zeros((n, NCOEF))
    for m in range(G):
        np.add.at(out, (I, p * G + m), phi[:, m] * (chi[J] - chi[I]))
        np.add.at(out, (I, 3 * G + p * G + m), phi[:, m] * (prev[J] - prev[I]))
    out[:, NCOEF - 1] = prev
    return out


def corrector_stage(positions, cell, pbc, species, rc, mu, eta, chi, prev, coef):
    """dq[n] of one stage with coefficients coef[97]."""
    return corrector_features(positions, cell, pbc, species, rc, mu, eta, chi, prev) @ \
        np.asarray(coef, np.float64)
