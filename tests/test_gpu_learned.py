"""GPU checks of the learned warm start: descriptors, BF16 forward and
backward kernels against the CPU oracle (oracle/mlp_oracle.py), training on
the device, and what the trained network gains over the cold start.

Tolerances (each written where it is checked):
  * descriptors: FP32 Gaussians and sums vs FP64, max |dG| <= 2e-5 * max|G|;
  * forward: the device accumulates in FP32 where the oracle sums exactly,
    so a hidden activation can round to the neighbouring BF16 value in rare
    cases; median |dy| <= 1e-5 * rms(y) and max |dy| <= 1e-2 * rms(y);
  * gradient: max |dg| <= 1e-3 * max |g|, loss within 1e-4 relative.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as MO  # noqa: E402

from pysic_amd import structures as st  # noqa: E402

pytestmark = pytest.mark.gpu

MODEL = os.path.join(ROOT, "pysic_amd", "data", "sio2_eem_mlp.npz")


@pytest.fixture(scope="module")
def pred():
    from pysic_amd.learned import ChargePredictor
    p = ChargePredictor(0)
    yield p
    p.close()


@pytest.fixture(scope="module")
def eem():
    from pysic_amd.learned import EEMSystem
    s = EEMSystem(0)
    yield s
    s.close()


def random_params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(MO.NP, dtype=np.float32)
    P[MO.OFF_W1:MO.OFF_B1] = rng.normal(0, 1 / np.sqrt(32), 2048)
    P[MO.OFF_B1:MO.OFF_W2] = rng.normal(0, 0.2, 256)
    P[MO.OFF_W2:MO.OFF_B2] = rng.normal(0, 1 / np.sqrt(64), 4096)
    P[MO.OFF_B2:MO.OFF_W3] = rng.normal(0, 0.2, 64)
    P[MO.OFF_W3:MO.OFF_B3] = rng.normal(0, 0.3, 64)
    P[MO.OFF_B3:] = rng.normal(0, 0.5, 4)
    return P


@pytest.mark.parametrize("n_si,seed", [(32, 1), (256, 2)])
def test_descriptors_match_oracle(pred, eem, n_si, seed):
    from pysic_amd.learned import DESC_MU, DESC_ETA, DESC_RC
    s = st.amorphous_sio2(n_si, 2.2, seed)
    print("case: descriptors n=%d seed=%d" % (len(s), seed), flush=True)
    eem.load(s)
    G, sp = pred.descriptors(eem.eng)
    spo = np.array([["O", "Si"].index(x) for x in s.symbols])
    assert np.array_equal(sp, spo)
    Gref = MO.descriptors(s.positions, s.cell, [True] * 3, spo, 2, DESC_RC, DESC_MU, DESC_ETA)
    err = np.abs(G - Gref).max() / np.abs(Gref).max()
    print("  max |dG| / max|G| = %.2e" % err, flush=True)
    assert err <= 2e-5


def test_forward_matches_oracle(pred):
    rng = np.random.default_rng(5)
    n = 4000
    feat = rng.gamma(2.0, 1.0, (n, 32)).astype(np.float32)
    sp = rng.integers(0, 2, n).astype(np.int32)
    mean = feat.mean(axis=0)
    isc = 1 / feat.std(axis=0)
    from pysic_amd.learned import DESC_MU, DESC_ETA
    pred.mlp.configure(2, 6.0, DESC_MU, DESC_ETA, mean, isc)
    P = random_params(6)
    pred.mlp.set_params(P)
    pred.mlp.set_training_data(feat, sp, np.zeros(n, np.float32))
    idx = rng.permutation(n).astype(np.int32)
    y = pred.mlp.evaluate(idx).astype(np.float64)
    yref, _ = MO.forward(P, feat[idx], sp[idx], mean, isc, rounding=True)
    d = np.abs(y - yref)
    rms = np.sqrt(np.mean(yref ** 2))
    print("median |dy| %.2e  max |dy| %.2e  rms y %.3f" % (np.median(d), d.max(), rms), flush=True)
    assert np.median(d) <= 1e-5 * rms
    assert d.max() <= 1e-2 * rms


def test_gradients_match_oracle(pred):
    rng = np.random.default_rng(7)
    n = 3000
    feat = rng.gamma(2.0, 1.0, (n, 32)).astype(np.float32)
    sp = rng.integers(0, 2, n).astype(np.int32)
    targ = rng.normal(0, 0.5, n).astype(np.float32)
    mean = feat.mean(axis=0)
    isc = 1 / feat.std(axis=0)
    from pysic_amd.learned import DESC_MU, DESC_ETA
    pred.mlp.configure(2, 6.0, DESC_MU, DESC_ETA, mean, isc)
    P = random_params(8)
    pred.mlp.set_params(P)
    pred.mlp.set_training_data(feat, sp, targ)
    idx = rng.choice(n, 1000, replace=False).astype(np.int32)
    g, loss = pred.mlp.gradients(idx)
    gref, lref = MO.gradients(P, feat[idx], sp[idx], targ[idx], mean, isc, rounding=True)
    err = np.abs(g - gref).max() / np.abs(gref).max()
    print("gradient rel max err %.2e  loss %.6f vs %.6f" % (err, loss, lref), flush=True)
    assert err <= 1e-3
    assert abs(loss - lref) <= 1e-4 * lref
    # blocks: every parameter group carries gradient
    for a, b in [(MO.OFF_W1, MO.OFF_B1), (MO.OFF_B1, MO.OFF_B1 + 128), (MO.OFF_W2, MO.OFF_B2),
                 (MO.OFF_B2, MO.OFF_W3), (MO.OFF_W3, MO.OFF_B3), (MO.OFF_B3, MO.OFF_B3 + 2)]:
        assert np.abs(g[a:b]).max() > 0


def test_training_fits_a_teacher_and_is_deterministic(pred):
    # targets from a fixed teacher network (oracle forward): the device
    # training must drive the loss down by a large factor, and two runs from
    # the same start must agree bit for bit
    rng = np.random.default_rng(9)
    n = 8192
    feat = rng.gamma(2.0, 1.0, (n, 32)).astype(np.float32)
    sp = rng.integers(0, 2, n).astype(np.int32)
    mean = feat.mean(axis=0)
    isc = 1 / feat.std(axis=0)
    from pysic_amd.learned import DESC_MU, DESC_ETA
    pred.mlp.configure(2, 6.0, DESC_MU, DESC_ETA, mean, isc)
    teacher = random_params(10)
    targ = MO.forward(teacher, feat, sp, mean, isc, rounding=False)[0].astype(np.float32)
    start = random_params(11)
    runs = []
    for _ in range(2):
        pred.mlp.set_params(start)
        pred.mlp.set_training_data(feat, sp, targ)
        order = np.concatenate([np.random.default_rng(12 + e).permutation(n) for e in range(30)])
        losses = pred.mlp.train(order.astype(np.int32), 256, lr=3e-3)
        runs.append((losses, pred.mlp.get_params()))
    l0 = runs[0][0][:10].mean()
    l1 = runs[0][0][-30:].mean()
    print("teacher fit: loss %.4e -> %.4e" % (l0, l1), flush=True)
    assert l1 < 0.05 * l0
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][0], runs[1][0])


@pytest.mark.skipif(not os.path.exists(MODEL), reason="trained model not present")
def test_warm_start_gain_on_held_out_config(eem):
    # a configuration outside the training seeds: the warm start reaches the
    # same converged charges in fewer steps of the unchanged relaxation
    from pysic_amd.learned import ChargePredictor
    p = ChargePredictor.load(MODEL, 0)
    s = st.amorphous_sio2(256, 2.2, seed=7)
    eem.load(s)
    tol = 1e-6
    q_star, _, _, _ = eem.relax(np.zeros(len(s)), 1e-11)
    q_cold, n_cold, e_cold, _ = eem.relax(np.zeros(len(s)), tol)
    q0 = p.predict_engine(eem.eng)
    q_warm, n_warm, e_warm, _ = eem.relax(q0, tol)
    rmse0 = np.sqrt(np.mean((q0 - q_star) ** 2))
    print("steps cold %d warm %d; predicted rmse %.3e e; |q-q*| cold %.2e warm %.2e" % (
        n_cold, n_warm, rmse0, np.abs(q_cold - q_star).max(), np.abs(q_warm - q_star).max()),
        flush=True)
    assert abs(q0.sum()) < 1e-9
    assert e_cold <= tol and e_warm <= tol
    assert np.abs(q_warm - q_star).max() <= 1e-6 and np.abs(q_cold - q_star).max() <= 1e-6
    assert rmse0 < 0.25 * np.sqrt(np.mean(q_star ** 2))
    assert n_warm < n_cold
    p.close()


@pytest.mark.skipif(not os.path.exists(MODEL), reason="trained model not present")
def test_warm_start_through_charge_relaxation():
    from pysic_amd import ChargeRelaxation, CoulombSummation, Pysic
    from pysic_amd.learned import ChargePredictor, SIO2_EWALD, SIO2_RELAX, eem_potentials
    p = ChargePredictor.load(MODEL, 0)
    s = st.amorphous_sio2(64, 2.2, seed=11)
    out = []
    for guess in (None, p):
        t = s.copy()
        relax = ChargeRelaxation("dynamic", parameters=dict(n_steps=5000, tolerance=1e-8,
                                                            **SIO2_RELAX))
        if guess is not None:
            relax.set_initial_guess(guess)
        calc = Pysic(t, eem_potentials(), coulomb=CoulombSummation(parameters=SIO2_EWALD))
        calc.set_charge_relaxation(relax)
        calc.get_potential_energy()
        out.append((calc.get_atoms().get_initial_charges().copy(), relax.last_steps))
    (qc, nc), (qw, nw) = out
    print("calculator path: steps cold %d warm %d" % (nc, nw), flush=True)
    assert np.abs(qc - qw).max() < 1e-7
    assert nw < nc
    p.close()
