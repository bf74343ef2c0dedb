"""Learned warm start for pysic's damped charge relaxation.

pysic relaxes charges by damped dynamics from whatever charges the atoms
carry (relaxation.py:417-450).  A ChargePredictor gives those starting
charges from the local environment: radial Gaussian descriptors (16 per
neighbor species, 6 A cutoff) into a 32 -> 64 -> 64 -> 1 network with BF16
weights and FP32 accumulation.  The network is trained on the GPU by the
library's own forward/backward kernels, on charges that the unchanged
relaxation produces; after the warm start that same relaxation runs to the
same tolerance, so the result does not depend on the network's accuracy,
only the number of steps does.

    python -m pysic_amd.learned train --configs 2000 --out model.npz

builds the training set (seeded random-addition SiO2 cells relaxed on the
GPU), trains, and saves the model (FP32 master parameters, their BF16
images as uint16, the descriptor set-up and the normalisation) as .npz.

The damped dynamics shrinks the charge error by about the same factor per
step in every mode (all of them are underdamped at the workload's friction),
so a start k times closer saves only ln k / ln(e0 / tol) of the steps.  The
warm start therefore continues the network's prediction with a residual
corrector: a few learned stages, each fed by the electronegativities of the
current charges (one evaluation of the model), that move the charges along
a local, symmetric, zero-sum kernel of the electronegativity differences.

    python -m pysic_amd.learned corrector --model model.npz --out model.npz

fits the stages on further relaxed cells by linear least squares over
features the device computes.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import structures as st
from .interactions import Potential, build_descriptors

# EEM parameters of the SiO2 workload: chi0 (eV) and hardness J (eV/e^2).
# With bare-Coulomb Ewald and 1.5 A minimum separations the charge Hessian
# on sum(q) = 0 is positive definite for these hardness values (smallest
# eigenvalue ~4.6 eV/e^2 on random samples); smaller J lets the damped
# dynamics run away.
SIO2_EEM = {"Si": (4.77, 16.0), "O": (8.74, 20.0)}
# Ewald at 1e-6 relative accuracy with a 10 A real-space cutoff:
# erfc(rc / (sqrt(2) sigma)) ~ 1e-6 and exp(-sigma^2 kc^2 / 2) ~ 1e-6
SIO2_EWALD = {"real_cutoff": 10.0, "k_cutoff": 2.5713, "sigma": 2.0443, "epsilon": 0.00552635}
# Damped-dynamics parameters (ChargeRelaxation 'dynamic').  The stiffest
# charge mode of a periodic Coulomb system is the longest charge-density
# wave, with a Hessian eigenvalue ~ rho / (eps0 k_min^2), k_min = 2 pi / L:
# it grows with the box edge squared, and the explicit integrator needs
# dt < 2 / sqrt(lambda_max / inertia).  For the 71.9 A, 24,576-atom cell
# dt = 0.04 runs away after ~400 steps and dt = 0.035 converges, so the
# workload uses 0.03; the 22.6 A training cells are stable at 0.1.  The
# friction is near critical damping of the softest mode (lambda_min ~ 4.6).
SIO2_RELAX = {"timestep": 0.03, "inertia": 1.0, "friction": 4.0}
TRAIN_RELAX = {"timestep": 0.1, "inertia": 1.0, "friction": 4.0}
SPECIES = ("O", "Si")      # engine species codes: element names in sorted order
DENSITY = 2.2              # g/cm^3

DESC_RC = 6.0
DESC_MU = np.linspace(1.0, 5.5, 16)
DESC_ETA = np.full(16, 1.0 / (2.0 * 0.3 ** 2))

NPARAM = 6532
# Residual corrector (include/pysic_learned.h): coefficients per stage and
# the number of stages trained.  Each stage cuts the charge error of the
# held-out cells by ~2.7x for one evaluation of chi, where one step of the
# damped dynamics cuts it by ~1.06x.
CORR_NCOEF = 97
CORR_STAGES = 16
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 2048, 2304, 6400, 6464, 6528


def eem_potentials(table=SIO2_EEM):
    pots = []
    for sym, (chi0, hard) in sorted(table.items()):
        pots.append(Potential("charge_self", symbols=[[sym]], parameters=[chi0, 1]))
        pots.append(Potential("charge_self", symbols=[[sym]], parameters=[0.5 * hard, 2]))
    return pots


def kcells(cell, kc):
    """Reciprocal cells per axis needed for |k| < kc (calculator.py:1465-1473)."""
    rec = 2.0 * math.pi * np.linalg.inv(cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
            for a in range(3)]


class EEMSystem:
    """One GPU system holding an SiO2 structure with the EEM + Ewald model."""

    def __init__(self, device=0, ewald=SIO2_EWALD, table=SIO2_EEM):
        from .engine import Engine
        self.eng = Engine(device)
        self.ewald = dict(ewald)
        self.pd, _, _ = build_descriptors(eem_potentials(table), {s: k for k, s in enumerate(SPECIES)})

    def load(self, structure):
        sp = np.array([SPECIES.index(x) for x in structure.symbols], dtype=np.int32)
        e, n = self.eng, len(structure)
        e.create_atoms(sp, structure.positions, charges=np.zeros(n))
        e.create_cell(structure.cell, structure.pbc)
        e.set_potentials(self.pd)
        e.set_bond_order_factors(np.zeros((0, 14)))
        p = self.ewald
        e.set_ewald(p["real_cutoff"], p["k_cutoff"], kcells(structure.cell, p["k_cutoff"]),
                    p["sigma"], p["epsilon"])
        e.generate_neighbor_lists(np.full(n, max(p["real_cutoff"], DESC_RC) + 0.5))
        self.n = n
        return self

    def relax(self, q0, tolerance, n_steps=20000, relax=SIO2_RELAX):
        """Damped relaxation from q0 (zero rates) on the device; returns
        (charges, steps, final max |chi - mean chi|, wall seconds)."""
        t0 = time.perf_counter()
        q, _, steps, err = self.eng.relax_charges(q0, np.zeros(self.n), n_steps, relax["timestep"],
                                                  relax["inertia"], relax["friction"], tolerance)
        return q, steps, err, time.perf_counter() - t0

    def close(self):
        self.eng.close()


def generate_structures(n_configs, n_si, seed0):
    """Seeded random-addition SiO2 cells (seeds seed0 .. seed0 + n_configs - 1)."""
    return [st.amorphous_sio2(n_si, DENSITY, seed0 + k) for k in range(n_configs)]


class ChargePredictor:
    """The warm-start network on one GPU.  Callable on a Pysic calculator's
    engine (predict_engine) or used as ChargeRelaxation.set_initial_guess."""

    def __init__(self, device=0, species=SPECIES, rc=DESC_RC, mu=DESC_MU, eta=DESC_ETA):
        from .engine import Mlp
        self.mlp = Mlp(device)
        assert self.mlp.nparams == NPARAM
        self.species = tuple(species)
        self.rc = float(rc)
        self.mu = np.asarray(mu, dtype=np.float64)
        self.eta = np.asarray(eta, dtype=np.float64)
        self.mean = np.zeros(32, dtype=np.float32)
        self.inv_scale = np.ones(32, dtype=np.float32)
        self.meta = {}
        self.corrector = np.zeros((0, CORR_NCOEF))
        self._configure()

    def _configure(self):
        self.mlp.configure(len(self.species), self.rc, self.mu, self.eta, self.mean,
                           self.inv_scale)

    def set_normalisation(self, feat):
        f = np.asarray(feat, dtype=np.float64)
        self.mean = f.mean(axis=0).astype(np.float32)
        self.inv_scale = (1.0 / np.maximum(f.std(axis=0), 1e-3)).astype(np.float32)
        self._configure()

    def init_params(self, seed, species_mean=None):
        rng = np.random.default_rng(seed)
        P = np.zeros(NPARAM, dtype=np.float32)
        P[OFF_W1:OFF_B1] = rng.normal(0, 1.0 / math.sqrt(32), 2048)
        P[OFF_W2:OFF_B2] = rng.normal(0, 1.0 / math.sqrt(64), 4096)
        P[OFF_W3:OFF_B3] = rng.normal(0, 0.1 / math.sqrt(64), 64)
        if species_mean is not None:
            P[OFF_B3:OFF_B3 + len(species_mean)] = species_mean
        self.mlp.set_params(P)

    def params(self):
        return self.mlp.get_params()

    def descriptors(self, eng):
        return self.mlp.descriptors(eng)

    def predict_engine(self, eng, neutralize=True):
        """Network charges for the system an Engine holds (its lists must
        reach rc)."""
        return self.mlp.predict_charges(eng, neutralize)

    def set_corrector(self, coef):
        self.corrector = np.asarray(coef, dtype=np.float64).reshape(-1, CORR_NCOEF)
        self.mlp.set_corrector(self.corrector)

    def warm_start(self, eng, stages=None):
        """Starting charges for the relaxation of the system an Engine holds:
        the network's charges made neutral, then the corrector's stages (all
        of them by default), each fed by the electronegativities of the
        engine's own model at the current charges."""
        return self.mlp.warm_start(eng, stages)

    def predict_for_calculator(self, calc):
        calc.set_core()
        names = tuple(sorted(calc._species_map()))
        if names != self.species[:len(names)]:
            raise ValueError("the network was trained for species %s, the system has %s"
                             % (self.species, names))
        return self.warm_start(calc.engine)

    def fit(self, feat, species, targets, epochs=60, batch=1024, lr=2e-3, lr_final=5e-5, seed=0,
            log=None):
        """Adam on the device; returns the mean batch loss per epoch."""
        self.mlp.set_training_data(feat, species, targets)
        rng = np.random.default_rng(seed)
        n = len(targets)
        hist = []
        for ep in range(epochs):
            lr_ep = lr * (lr_final / lr) ** (ep / max(1, epochs - 1))
            order = rng.permutation(n).astype(np.int32)
            losses = self.mlp.train(order, batch, lr=lr_ep)
            hist.append(float(losses.mean()))
            if log:
                log("epoch %3d  lr %.2e  loss %.4e" % (ep, lr_ep, hist[-1]))
        return hist

    def save(self, path):
        P = self.params()
        bits = P.view(np.uint32)
        rounded = ((bits.astype(np.uint64) + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)
        np.savez(path, params=P, params_bf16=rounded, species=np.array(self.species),
                 rc=self.rc, mu=self.mu, eta=self.eta, mean=self.mean, inv_scale=self.inv_scale,
                 corrector=self.corrector, meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path, device=0):
        z = np.load(path, allow_pickle=False)
        p = cls(device, species=tuple(str(x) for x in z["species"]), rc=float(z["rc"]),
                mu=z["mu"], eta=z["eta"])
        p.mean = z["mean"].astype(np.float32)
        p.inv_scale = z["inv_scale"].astype(np.float32)
        p._configure()
        p.mlp.set_params(z["params"])
        if "corrector" in z.files:
            p.set_corrector(z["corrector"])
        p.meta = json.loads(str(z["meta"]))
        return p

    def close(self):
        self.mlp.close()


def build_training_set(structures, device=0, tolerance=1e-9, log=None, system=None,
                       predictor=None, relax=TRAIN_RELAX):
    """Relax every structure from zero charges with the plain method and
    collect (descriptors, species, relaxed charge) per atom."""
    sysm = system or EEMSystem(device)
    pred = predictor or ChargePredictor(device)
    feats, sps, qs, steps = [], [], [], []
    t0 = time.perf_counter()
    for k, s in enumerate(structures):
        sysm.load(s)
        q, nst, err, _ = sysm.relax(np.zeros(len(s)), tolerance, relax=relax)
        if err > tolerance:
            raise RuntimeError("training relaxation %d did not converge (%.3e)" % (k, err))
        f, sp = pred.descriptors(sysm.eng)
        feats.append(f)
        sps.append(sp)
        qs.append(q)
        steps.append(nst)
        if log and (k + 1) % max(1, len(structures) // 20) == 0:
            log("relaxed %d/%d configs (%.1f s, mean %d steps)"
                % (k + 1, len(structures), time.perf_counter() - t0, np.mean(steps)))
    if system is None:
        sysm.close()
    return (np.concatenate(feats), np.concatenate(sps), np.concatenate(qs).astype(np.float32),
            np.array(steps))


def train_corrector(pred, structures, n_stages=CORR_STAGES, device=0, tolerance=1e-10,
                    relax=TRAIN_RELAX, log=None, system=None):
    """Fit the corrector's stages one after another by linear least squares.

    Every structure is relaxed from zero charges by the plain method to
    tolerance (q*); its start is the network's neutral prediction.  Stage k's
    features are computed on the device at each cell's current charges and
    previous change (pysic_mlp_corrector_features); its coefficients minimise
    sum over all atoms of (q* - q - features . coef)^2; the stage is then
    applied to every cell before stage k + 1 is fitted.  Returns the charge
    RMSE after each stage."""
    sysm = system or EEMSystem(device)
    cells = []
    t0 = time.perf_counter()
    for k, s in enumerate(structures):
        sysm.load(s)
        q_star, _, err, _ = sysm.relax(np.zeros(len(s)), tolerance, relax=relax)
        if err > tolerance:
            raise RuntimeError("corrector label relaxation %d did not converge (%.3e)" % (k, err))
        q0 = pred.mlp.warm_start(sysm.eng, 0)
        cells.append([s, q_star, q0, np.zeros(len(s))])
    rmse = [float(np.sqrt(np.mean(np.concatenate([c[1] - c[2] for c in cells]) ** 2)))]
    if log:
        log("corrector: %d cells relaxed (%.1f s), start rmse %.3e e"
            % (len(cells), time.perf_counter() - t0, rmse[0]))
    coefs = []
    for k in range(n_stages):
        feats = []
        for c in cells:
            sysm.load(c[0])
            f, _ = pred.mlp.corrector_features(sysm.eng, c[2], c[3])
            feats.append(f)
        X = np.concatenate(feats)
        y = np.concatenate([c[1] - c[2] for c in cells])
        # columns scaled to unit norm: chi differences (eV) and charge
        # changes (e) differ by orders of magnitude
        scale = np.sqrt((X * X).sum(axis=0))
        scale[scale == 0.0] = 1.0
        sol, *_ = np.linalg.lstsq(X / scale, y, rcond=None)
        coef = sol / scale
        coefs.append(coef)
        for c, f in zip(cells, feats):
            dq = f @ coef
            c[2] = c[2] + dq
            c[3] = dq
        rmse.append(float(np.sqrt(np.mean(np.concatenate([c[1] - c[2] for c in cells]) ** 2))))
        if log:
            log("corrector stage %d: rmse %.3e e" % (k, rmse[-1]))
    if system is None:
        sysm.close()
    pred.set_corrector(np.array(coefs))
    return rmse


def train_model(n_configs=2000, n_si=256, seed0=100000, epochs=60, batch=1024, device=0,
                tolerance=1e-9, log=print):
    t0 = time.perf_counter()
    structs = generate_structures(n_configs, n_si, seed0)
    log("generated %d structures of %d atoms (%.1f s)" % (n_configs, 3 * n_si,
                                                          time.perf_counter() - t0))
    pred = ChargePredictor(device)
    feat, sp, q, steps = build_training_set(structs, device, tolerance, log, predictor=pred)
    nval = max(1, n_configs // 20) * 3 * n_si
    tr, va = slice(0, len(q) - nval), slice(len(q) - nval, len(q))
    pred.set_normalisation(feat[tr])
    pred.init_params(seed0, species_mean=[float(q[tr][sp[tr] == k].mean()) for k in range(2)])
    hist = pred.fit(feat[tr], sp[tr], q[tr], epochs=epochs, batch=batch, seed=seed0, log=log)
    # validation: held-out configurations
    pred.mlp.set_training_data(feat[va], sp[va], q[va])
    yv = pred.mlp.evaluate(np.arange(nval, dtype=np.int32)).astype(np.float64)
    err = yv - q[va]
    pred.meta = {"configs": n_configs, "atoms_per_config": 3 * n_si, "seed0": seed0,
                 "density": DENSITY, "label_tolerance": tolerance, "epochs": epochs,
                 "batch": batch, "train_atoms": int(len(q) - nval), "val_atoms": int(nval),
                 "final_train_loss": hist[-1], "val_rmse": float(np.sqrt(np.mean(err ** 2))),
                 "val_maxerr": float(np.abs(err).max()), "label_steps_mean": float(steps.mean()),
                 "charge_std": float(q.std()), "eem": SIO2_EEM, "ewald": SIO2_EWALD,
                 "label_relax": TRAIN_RELAX, "seconds": time.perf_counter() - t0}
    log("validation: rmse %.4e e, max %.4e e (charge std %.3f)" % (
        pred.meta["val_rmse"], pred.meta["val_maxerr"], pred.meta["charge_std"]))
    return pred


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pysic_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--configs", type=int, default=2000)
    t.add_argument("--n-si", type=int, default=256)
    t.add_argument("--seed0", type=int, default=100000)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--tolerance", type=float, default=1e-9)
    t.add_argument("--out", required=True)
    c = sub.add_parser("corrector", help="fit the residual corrector of a trained model")
    c.add_argument("--model", required=True)
    c.add_argument("--configs", type=int, default=200)
    c.add_argument("--n-si", type=int, default=256)
    c.add_argument("--seed0", type=int, default=300000)
    c.add_argument("--stages", type=int, default=CORR_STAGES)
    c.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "train":
        pred = train_model(a.configs, a.n_si, a.seed0, a.epochs, a.batch, 0, a.tolerance, log)
    else:
        t0 = time.perf_counter()
        pred = ChargePredictor.load(a.model, 0)
        structs = generate_structures(a.configs, a.n_si, a.seed0)
        rmse = train_corrector(pred, structs, a.stages, log=log)
        pred.meta["corrector"] = {"configs": a.configs, "atoms_per_config": 3 * a.n_si,
                                  "seed0": a.seed0, "stages": a.stages,
                                  "label_tolerance": 1e-10, "label_relax": TRAIN_RELAX,
                                  "rmse_per_stage": rmse,
                                  "seconds": time.perf_counter() - t0}
    pred.save(a.out)
    log("saved %s: %s" % (a.out, json.dumps(pred.meta)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
