"""Benchmark: Tersoff Si molecular dynamics on one MI355X.

Workload (the headline configuration): diamond Si, 1,048,576 atoms
(64 x 64 x 32 conventional 8-atom cells; the stated "64x64x64 cells" would
hold 2,097,152 atoms, the stated atom count is kept), a = 5.431 A, Gaussian
displacements sigma = 0.1 A
(seeded), Maxwell-Boltzmann momenta at 600 K (seeded), velocity Verlet with
dt = 1 fs, neighbor list (cutoff 3.0 + skin 0.5 A) rebuilt every step.
A "step" is one MD step: drift, cell-list rebuild, Tersoff energy + forces
(+ virial), second half-kick.  Metric: atom-steps/s = N * steps / wall time
of the timed steps, with all inputs resident in HBM.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]

Prints one JSON line.  The cpu_baseline leg times the reference Fortran core
(oracle/_ref, built from the reference sources) on a bounded sample on the
host, single core.  The "qeq" object holds the second half of the metric:
QEq time-to-tolerance on 24,576-atom amorphous SiO2, learned warm start
against cold start (see qeq_benchmark).  Secondary objects: "batched"
(many a-Si cells), "ewald" (rocksalt Coulomb), "lj" (argon Lennard-Jones
molecular dynamics through the pair kernel).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def build_workload(cells=(64, 64, 32), seed=0):
    from pysic_amd import structures as st
    s = st.diamond_si(tuple(cells), a=5.431, sigma=0.1, seed=seed)
    st.maxwell_boltzmann(s, 600.0, seed=seed + 1)
    return s


def tersoff_bytes_per_atom(n_nb):
    # algorithmic bytes of one Tersoff E+F evaluation per atom: position in
    # (3 x f64), neighbor list in (4 B per entry), force out (3 x f64),
    # site energy out (f64)
    return 24.0 + 4.0 * n_nb + 24.0 + 8.0


# one wave64 VALU instruction per SIMD every 4 cycles: 256 CUs x 4 SIMDs x
# 2.4 GHz / 4
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def measured_counters(cells, variant_name, field):
    """A per-launch counter of the tersoff_site kernel recorded in profiles/
    from rocprofv3 --pmc: "traffic_bytes_per_launch" (HBM bytes: FETCH_SIZE
    and WRITE_SIZE, scaled by what the same counters read for a plain device
    copy of known size) or "valu_wave_instructions_per_launch"
    (SQ_INSTS_VALU).  None unless the recorded workload and kernel variant
    are the ones run here."""
    path = os.path.join(ROOT, "profiles", "tersoff_site_si1m_pmc.txt")
    if not os.path.exists(path):
        return None
    rec = {}
    with open(path) as fh:
        for line in fh:
            if ":" in line and line.split(":", 1)[0].strip() in ("cells", "kernel_variant", field):
                k, v = line.split(":", 1)
                rec[k.strip()] = v.strip()
    try:
        if [int(x) for x in rec["cells"].split()] != list(cells):
            return None
        if rec["kernel_variant"] != variant_name:
            return None
        return int(rec[field])
    except (KeyError, ValueError):
        return None


def copy_calibration(nbytes=256 << 20):
    """One device-to-device copy of nbytes (HIP's blit copy kernel): the
    reference reading for FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def qeq_benchmark(model_path, reps=2):
    """QEq time-to-tolerance on amorphous SiO2, 24,576 atoms (8192 Si +
    16384 O, 2.2 g/cm^3, random addition with 1.5 A minimum separation),
    EEM + Ewald (rc 10 A, 1e-6 accuracy), damped charge dynamics stopped at
    max(chi - mean chi) <= 1e-6 eV/e: cold start (q = 0) against the learned
    warm start (descriptors + BF16 network on the device, then the residual
    corrector's stages, each one evaluation of chi, then the same
    relaxation).  The network alone (no corrector stages) is timed beside it.
    The charge residual |q - q*|_inf of each is measured against a 1e-11
    solve."""
    from pysic_amd import structures as st
    from pysic_amd.learned import ChargePredictor, EEMSystem, SIO2_RELAX
    s = st.amorphous_sio2(8192, 2.2, seed=777)
    n = len(s)
    sysm = EEMSystem(0).load(s)
    pred = ChargePredictor.load(model_path, 0)
    stages = pred.mlp.corrector_stages()
    pred.warm_start(sysm.eng)                          # warm-up of every kernel
    sysm.relax(np.zeros(n), 1e-6, n_steps=3)
    q_star, n_star, _, _ = sysm.relax(np.zeros(n), 1e-11)
    tol = 1e-6
    runs = {"cold": [], "warm": [], "network_only": []}
    for _ in range(reps):
        q, steps, err, sec = sysm.relax(np.zeros(n), tol)
        runs["cold"].append((sec, steps, float(np.abs(q - q_star).max()), err))
        for key, k in (("warm", stages), ("network_only", 0)):
            t0 = time.perf_counter()
            q0 = pred.warm_start(sysm.eng, k)
            tp = time.perf_counter() - t0
            q, steps, err, sec = sysm.relax(q0, tol)
            runs[key].append((tp + sec, steps, float(np.abs(q - q_star).max()), err, tp,
                              float(np.sqrt(np.mean((q0 - q_star) ** 2)))))
    sysm.eng.timing(True)
    sysm.relax(np.zeros(n), tol)
    tim = sysm.eng.timing_read()
    sysm.eng.timing(False)
    c = min(runs["cold"])

    def started(w, k):
        return {"seconds": round(w[0], 4), "steps": w[1], "corrector_stages": k,
                "q_residual_inf_e": w[2], "start_seconds": round(w[4], 5),
                "start_rmse_e": w[5]}

    w = min(runs["warm"])
    w0 = min(runs["network_only"])
    out = {"workload": "amorphous SiO2, 24,576 atoms, L = 71.89 A, EEM (chi0 Si/O 4.77/8.74 eV, "
                       "J 16/20 eV/e^2) + Ewald rc 10 A sigma 2.0443 A kc 2.5713 1/A, damped "
                       "dynamics dt %.3g inertia %g friction %g" % (
                           SIO2_RELAX["timestep"], SIO2_RELAX["inertia"], SIO2_RELAX["friction"]),
           "atoms": n, "stop": "max(chi - mean chi) <= 1e-6 eV/e", "reference_steps_1e-11": n_star,
           "cold": {"seconds": round(c[0], 4), "steps": c[1], "q_residual_inf_e": c[2]},
           "warm": started(w, stages),
           "warm_over_cold": round(w[0] / c[0], 4), "target_warm_over_cold": 0.5,
           "network_only": started(w0, 0),
           "network_only_over_cold": round(w0[0] / c[0], 4),
           "ms_per_relaxation_step": round(c[0] / c[1] * 1e3, 3),
           "kernels_us_per_step": {k: round(v[0] / max(v[1], 1) * 1e3, 1) for k, v in tim.items()},
           "model": {k: pred.meta.get(k) for k in ("configs", "atoms_per_config", "val_rmse",
                                                    "val_maxerr", "epochs")},
           "corrector": pred.meta.get("corrector")}
    pred.close()
    sysm.close()
    out["reference_sample"] = qeq_reference_sample(n, c[1])
    return out


def _n_kvectors(cell, kc):
    K = ewald_kcells(cell, kc)
    rec = 2.0 * np.pi * np.linalg.inv(cell).T
    g = np.array(np.meshgrid(np.arange(-K[0], K[0] + 1), np.arange(-K[1], K[1] + 1),
                             np.arange(0, K[2] + 1), indexing="ij")).reshape(3, -1).T
    k2 = ((g @ rec) ** 2).sum(axis=1)
    return K, int(((k2 < kc ** 2) & (k2 > 0)).sum())


def qeq_reference_sample(n_full, steps_full, n_si=128):
    """The reference core's cost of one relaxation step (one electronegativity
    evaluation, relaxation.py:417-450 calls it once per step) on a 384-atom
    sample of the same SiO2 model, scaled to the full system by N * N_k (the
    reciprocal sum, which dominates) and by the cold start's step count."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if not O.ref_available():
            return None
        from pysic_amd import structures as st
        from pysic_amd.learned import SIO2_EEM, SIO2_EWALD
        r = st.amorphous_sio2(n_si, 2.2, seed=5)
        ew = SIO2_EWALD
        K, nk = _n_kvectors(r.cell, ew["k_cutoff"])
        pots = []
        for sym, (chi0, hard) in sorted(SIO2_EEM.items()):
            pots.append((2, [chi0, 1.0], 0.0, 0.0, (sym,), -1))
            pots.append((2, [0.5 * hard, 2.0], 0.0, 0.0, (sym,), -1))
        q = np.random.default_rng(1).normal(0.0, 0.2, len(r))
        q -= q.mean()
        R = O.RefCore()
        R.setup(r.symbols, r.positions, r.cell, r.pbc, charges=q, potentials=pots,
                cutoffs=np.full(len(r), ew["real_cutoff"] + 0.5),
                ewald=(ew["real_cutoff"], ew["k_cutoff"], K, ew["sigma"], ew["epsilon"]))
        R.electronegativities()
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 3.0 and k < 20:
            R.electronegativities()
            k += 1
        t_step = (time.perf_counter() - t0) / k
        big = st.amorphous_sio2(n_full // 3, 2.2, seed=777).cell
        _, nk_full = _n_kvectors(big, ew["k_cutoff"])
        t_full = t_step * (n_full * nk_full) / (len(r) * nk)
        return {"atoms": len(r), "k_vectors": nk, "seconds_per_step": round(t_step, 4),
                "cores": 1, "kind": "reference",
                "extrapolated_seconds_per_step_by_N_Nk": round(t_full, 2),
                "extrapolated_seconds_to_tolerance": round(t_full * steps_full, 1)}
    except Exception as exc:  # the checker is optional on the GPU host
        return {"error": str(exc)}


def build_batched_cells(ncells=512):
    """The 512 cells (random addition on the host)."""
    from pysic_amd.batch import batched_amorphous_si
    t0 = time.perf_counter()
    cells = batched_amorphous_si(ncells=ncells, seed=0)
    return cells, time.perf_counter() - t0


def batched_benchmark(cells, t_build, reps=20, seq_cells=32):
    """Batched Tersoff on 512 independent a-Si cells (atoms uniform in
    [64, 2048], density uniform in 2.25-2.35 g/cm^3, random addition with a
    2.1 A minimum separation, cubic periodic boxes): one evaluation = list
    rebuild + forces + virials + FP64 energies of every cell, replayed from a
    captured HIP graph.  For scale, seq_cells of the same cells are evaluated
    one calculator at a time (calculate_forces + calculate_energy per cell,
    as a pysic user would loop over structures) and the time is scaled to the
    whole batch by atom count."""
    from pysic_amd.batch import BatchedCells
    from pysic_amd.engine import Engine
    from pysic_amd.interactions import build_descriptors, tersoff_si
    ncells = len(cells)
    b = BatchedCells(cells, tersoff_si())
    tune = b.autotune(reps=3, grids=(0, 1024, 2048))
    b.evaluate(use_graph=True)
    b.results()
    b.eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.evaluate(use_graph=True)
    b.eng.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    E, W, F = b.results()   # also checks that no list overflowed
    t0 = time.perf_counter()
    for _ in range(reps):
        b.evaluate(use_graph=False)
    b.eng.synchronize()
    t_eager = (time.perf_counter() - t0) / reps
    b.eng.timing(True)
    b.evaluate(use_graph=False)
    b.eng.synchronize()
    tim = b.eng.timing_read()
    b.eng.timing(False)
    n = b.n
    b.close()
    # one calculator per cell, evaluated in turn
    pd, bd, _ = build_descriptors(tersoff_si(), {"Si": 0})
    engs = []
    for s in cells[:seq_cells]:
        e = Engine(0)
        e.create_atoms(np.zeros(len(s), np.int32), s.positions, masses=s.masses)
        e.create_cell(s.cell, s.pbc)
        e.set_potentials(pd)
        e.set_bond_order_factors(bd)
        e.ewald_off()
        e.generate_neighbor_lists(np.full(len(s), 3.5))
        e.forces()
        engs.append(e)
    t0 = time.perf_counter()
    for e in engs:
        e.generate_neighbor_lists(np.full(e.n, 3.5))
        e.forces()
        e.energy()
    t_seq = time.perf_counter() - t0
    seq_atoms = sum(e.n for e in engs)
    for e in engs:
        e.close()
    return {"workload": "Tersoff a-Si, %d independent cubic periodic cells, %d atoms "
                        "(64..2048 per cell), density 2.25-2.35 g/cm^3, RSA dmin 2.1 A" % (
                            ncells, n),
            "cells": ncells, "atoms": n, "evaluation": "list rebuild + forces + virials + "
                                                      "FP64 energies, per-cell sums",
            "graph_ms": round(t_graph * 1e3, 4), "eager_ms": round(t_eager * 1e3, 4),
            "atom_evaluations_per_s": round(n / t_graph, 1),
            "one_calculator_per_cell_ms_scaled": round(t_seq * n / seq_atoms * 1e3, 3),
            "batch_speedup_vs_one_per_cell": round(t_seq * n / seq_atoms / t_graph, 2),
            "autotune": tune, "kernels_us": {k: round(v[0] / max(v[1], 1) * 1e3, 1)
                                             for k, v in tim.items()},
            "structure_build_s": round(t_build, 2),
            "energy_per_atom_mean_eV": float(E.sum() / n)}


def ewald_kcells(cell, kc):
    rec = 2.0 * np.pi * np.linalg.inv(cell).T
    vol = rec[0] @ np.cross(rec[1], rec[2])
    return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
            for a in range(3)]


EWALD_RC, EWALD_SIGMA, EWALD_KC, EPS0 = 10.0, 2.0443, 2.5713, 0.00552635


def ewald_benchmark(reps=3, ref_reps=(4, 4, 4)):
    """Ewald Coulomb on rocksalt: charges +-1 e, a = 5.64 A, 32^3 conventional
    cells (262,144 ions), sigma = 0.08 A seeded displacements, real-space
    cutoff 10 A, Gaussian width 2.0443 A and k-cutoff 2.5713 1/A (1e-6
    relative error each), FP64 throughout.  One evaluation = energy +
    forces + stress of a new configuration (the energy and forces calls are
    separate, as calculate_energy and calculate_forces; the engine computes
    S(k) once per configuration and the forces call reuses it).  Each timed
    evaluation first moves every ion by another 1e-4 A (outside the timed
    region), so nothing carries over between evaluations.  The reference
    core is timed on a 4^3 replica (512 ions) and scaled by N * N_k, the
    reciprocal work."""
    from pysic_amd import structures as st
    from pysic_amd.engine import Engine
    s = st.rocksalt((32, 32, 32), a=5.64, sigma=0.08, seed=0)
    n = len(s)
    K = ewald_kcells(s.cell, EWALD_KC)
    eng = Engine(0)
    eng.create_atoms((s.charges < 0).astype(np.int32), s.positions, charges=s.charges)
    eng.create_cell(s.cell, (True, True, True))
    eng.set_potentials(np.zeros((0, 10)))
    eng.set_bond_order_factors(np.zeros((0, 14)))
    eng.set_ewald(EWALD_RC, EWALD_KC, K, EWALD_SIGMA, EPS0)
    eng.set_neighbor_capacity(256)
    eng.generate_neighbor_lists(np.full(n, EWALD_RC + 0.5))
    e = eng.energy()
    f, w = eng.forces()
    eng.synchronize()
    eng.timing(True)
    t_eval = 0.0
    for r in range(reps):
        eng.update_coordinates(s.positions + 1e-4 * (r + 1))
        eng.synchronize()
        t0 = time.perf_counter()
        e = eng.energy()
        f, w = eng.forces()
        eng.synchronize()
        t_eval += time.perf_counter() - t0
    t_eval /= reps
    tim = eng.timing_read()
    eng.timing(False)
    eng.close()
    # the number of k-vectors in the half space |k| < kc
    rec = 2.0 * np.pi * np.linalg.inv(s.cell).T
    g = np.array(np.meshgrid(np.arange(-K[0], K[0] + 1), np.arange(-K[1], K[1] + 1),
                             np.arange(0, K[2] + 1), indexing="ij")).reshape(3, -1).T
    kv = g @ rec
    k2 = (kv ** 2).sum(axis=1)
    nk = int(((k2 < EWALD_KC ** 2) & (k2 > 0)).sum())
    sk_us = tim.get("ewald_sk", (0.0, 1))
    sk_avg = sk_us[0] / max(sk_us[1], 1) / 1e3
    # algorithmic work of one structure-factor pass: per (k, ion) one complex
    # multiply-accumulate, 4 FMAs = 8 FP64 flops (the phase generation, one
    # sincos per 16 k, is not counted)
    flops = 8.0 * nk * n
    out = {"workload": "rocksalt 32^3 cells, 262,144 ions, sigma 0.08 A, Ewald rc 10 A "
                       "sigma 2.0443 A kc 2.5713 1/A (1e-6), FP64",
           "ions": n, "k_vectors_half_space": nk, "evaluation": "energy + forces + stress",
           "ms_per_evaluation": round(t_eval * 1e3, 2),
           "energy_eV_per_ion": e / n,
           "kernels_us": {k: round(v[0] / max(v[1], 1) * 1e3, 1) for k, v in tim.items()},
           "structure_factor_roofline": {"bound": "fp64", "achieved": round(flops / sk_avg / 1e12, 2),
                                         "peak": 78.6, "unit": "TFLOP/s",
                                         "frac": round(flops / sk_avg / 1e12 / 78.6, 4)}}
    # the reference core on a small replica
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if O.ref_available():
            r = st.rocksalt(ref_reps, a=5.64, sigma=0.08, seed=0)
            Kr = ewald_kcells(r.cell, EWALD_KC)
            R = O.RefCore()
            R.setup(r.symbols, r.positions, r.cell, [True] * 3, charges=r.charges,
                    cutoffs=np.full(len(r), EWALD_RC + 0.5),
                    ewald=(EWALD_RC, EWALD_KC, Kr, EWALD_SIGMA, EPS0))
            R.energy()
            t0 = time.perf_counter()
            R.energy()
            R.forces()
            t_ref = time.perf_counter() - t0
            recr = 2.0 * np.pi * np.linalg.inv(r.cell).T
            gr = np.array(np.meshgrid(np.arange(-Kr[0], Kr[0] + 1), np.arange(-Kr[1], Kr[1] + 1),
                                      np.arange(0, Kr[2] + 1), indexing="ij")).reshape(3, -1).T
            k2r = ((gr @ recr) ** 2).sum(axis=1)
            nkr = int(((k2r < EWALD_KC ** 2) & (k2r > 0)).sum())
            out["reference_sample"] = {
                "ions": len(r), "k_vectors": nkr, "seconds": round(t_ref, 4), "cores": 1,
                "kind": "reference",
                "extrapolated_seconds_by_N_Nk": round(t_ref * (n * nk) / (len(r) * nkr), 1)}
    except Exception as exc:  # the checker is optional on the GPU host
        out["reference_sample"] = {"error": str(exc)}
    return out


LJ_CELLS, LJ_A, LJ_T, LJ_SKIN, LJ_REBUILD = 64, 5.26, 80.0, 0.5, 10


def lj_benchmark(steps=100, warmup=10, ref_cells=6, reference=True):
    """Lennard-Jones argon molecular dynamics (the pair kernel, pair.hip):
    fcc argon, 64^3 conventional cells = 1,048,576 atoms, a = 5.26 A,
    sigma = 0.05 A seeded displacements, Maxwell-Boltzmann momenta at 80 K,
    LJ epsilon/k_B = 119.8 K, sigma = 3.405 A cut at 2.5 sigma (8.51 A),
    lists with a 0.5 A skin rebuilt every 10 steps, velocity Verlet with
    dt = 2 fs, FP64.  A step: drift, [list rebuild], forces with the second
    half-kick.  The reference core is timed on a 6^3-cell sample (864
    atoms): list build + forces per step, its own loop order."""
    from pysic_amd import structures as st
    from pysic_amd.engine import Engine
    from pysic_amd.interactions import build_descriptors, lj_argon
    s = st.fcc((LJ_CELLS,) * 3, a=LJ_A, sigma=0.05, seed=0)
    st.maxwell_boltzmann(s, LJ_T, seed=1)
    n = len(s)
    rc = 2.5 * 3.405
    pd, _, _ = build_descriptors(lj_argon(cutoff=rc), {"Ar": 0})
    eng = Engine(0)
    eng.create_atoms(np.zeros(n, np.int32), s.positions, masses=s.masses, momenta=s.momenta)
    eng.create_cell(s.cell, s.pbc)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(np.zeros((0, 14)))
    eng.ewald_off()
    eng.set_neighbor_capacity(112)
    eng.generate_neighbor_lists(np.full(n, rc + LJ_SKIN))
    counts, _, _ = eng.neighbor_lists()
    n_list = float(counts.mean())
    dt = 2.0 * st.ASE_FS
    eng.md_run(warmup, dt, rebuild=LJ_REBUILD)
    m = s.masses[:, None]
    e0 = eng.energy() + float((eng.state()[1] ** 2 / (2 * m)).sum())
    eng.synchronize()
    t0 = time.perf_counter()
    eng.md_run(steps, dt, rebuild=LJ_REBUILD)
    eng.synchronize()
    wall = time.perf_counter() - t0
    e1 = eng.energy() + float((eng.state()[1] ** 2 / (2 * m)).sum())
    eng.timing(True)
    eng.md_run(2 * LJ_REBUILD, dt, rebuild=LJ_REBUILD)
    eng.synchronize()
    tim = eng.timing_read()
    eng.timing(False)
    eng.close()
    pf_ms, pf_calls = tim.get("pair_forces", (0.0, 1))
    pf_s = pf_ms / max(pf_calls, 1) / 1e3
    # algorithmic bytes of one forces launch per atom: position 32, list
    # entries 8 each (j + image), momentum read + write 64, force write 32
    bytes_launch = (128.0 + 8.0 * n_list) * n
    out = {"workload": "LJ argon fcc 64^3 cells, 1,048,576 atoms, rc 2.5 sigma = 8.51 A, "
                       "skin 0.5 A, lists rebuilt every 10 steps, VV dt = 2 fs, 80 K, FP64",
           "atoms": n, "list_entries_per_atom": round(n_list, 2), "steps": steps,
           "value": round(n * steps / wall, 1), "unit": "atom-steps/s",
           "ms_per_step": round(wall / steps * 1e3, 4),
           "total_energy_drift_eV_per_atom": (e1 - e0) / n,
           "kernels_us": {k: round(v[0] / max(v[1], 1) * 1e3, 1) for k, v in tim.items()},
           "pair_forces_roofline": {
               "bound": "hbm", "achieved": round(bytes_launch / pf_s / 1e9, 1) if pf_s else 0.0,
               "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(bytes_launch / pf_s / 1e9 / HBM_PEAK_GBS, 4) if pf_s else 0.0,
               "algorithmic_bytes_per_launch": int(bytes_launch)}}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if reference and O.ref_available():
            r = st.fcc((ref_cells,) * 3, a=LJ_A, sigma=0.05, seed=0)
            cut = np.full(len(r), rc + LJ_SKIN)
            R = O.RefCore()
            R.setup(r.symbols, r.positions, r.cell, r.pbc,
                    potentials=[(3, list(pd[0, [4, 5]]), rc, rc, ("Ar", "Ar"), -1)],
                    cutoffs=cut)
            R.forces()
            k = 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < 5.0 and k < 20:
                R.rebuild(r.positions, cut)
                R.forces()
                k += 1
            t_ref = (time.perf_counter() - t0) / k
            out["reference_sample"] = {
                "atoms": len(r), "seconds_per_step": round(t_ref, 4), "cores": 1,
                "kind": "reference", "atom_steps_per_s": round(len(r) / t_ref, 1),
                "note": "list rebuild + forces every step (the reference has no "
                        "rebuild interval); serial amdflang -O2 build"}
    except Exception as exc:  # the checker is optional on the GPU host
        out["reference_sample"] = {"error": str(exc)}
    return out


def comb_benchmark(n_si=8192, steps=100, warmup=10, ref_si=128):
    """COMB's local SiO2 terms as the reference's builder makes them (bond
    bending, charge-dependent repulsion, and charge-dependent attraction
    scaled by three coordinators' 'tersoff' bond-order factors; bend.hip
    and pair.hip) in molecular dynamics: random amorphous SiO2 at
    2.2 g/cm^3, 3 x 8192 = 24,576 atoms, fixed seeded charges (Si in
    [0.5, 2.2] e, O in [-1.3, -0.2] e), Maxwell-Boltzmann momenta at 300 K,
    velocity Verlet with dt = 0.5 fs, lists (3.05 + 0.5 A) rebuilt every
    step, FP64.  The reference core is timed on a 384-atom sample of the
    same kind: list build + forces per step."""
    from pysic_amd import structures as st
    from pysic_amd.engine import Engine
    from pysic_amd.interactions import build_descriptors, comb_sio2

    def system(k, seed):
        s = st.amorphous_sio2(k, 2.2, seed=seed)
        rng = np.random.default_rng(seed + 1)
        q = np.where(np.array(s.symbols) == "Si", rng.uniform(0.5, 2.2, len(s)),
                     rng.uniform(-1.3, -0.2, len(s)))
        return s, q

    s, q = system(n_si, 5)
    st.maxwell_boltzmann(s, 300.0, seed=6)
    n = len(s)
    names = ["O", "Si"]
    pd, bd, _ = build_descriptors(comb_sio2(), {x: k for k, x in enumerate(names)})
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    cut = np.full(n, 3.55)
    eng = Engine(0)
    eng.create_atoms(sp, s.positions, masses=s.masses, charges=q, momenta=s.momenta)
    eng.create_cell(s.cell, s.pbc)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(bd)
    eng.ewald_off()
    eng.generate_neighbor_lists(cut)
    dt = 0.5 * st.ASE_FS
    eng.md_run(warmup, dt, rebuild=1)
    m = s.masses[:, None]
    e0 = eng.energy() + float((eng.state()[1] ** 2 / (2 * m)).sum())
    eng.synchronize()
    t0 = time.perf_counter()
    eng.md_run(steps, dt, rebuild=1)
    eng.synchronize()
    wall = time.perf_counter() - t0
    e1 = eng.energy() + float((eng.state()[1] ** 2 / (2 * m)).sum())
    eng.timing(True)
    eng.md_run(10, dt, rebuild=1)
    eng.synchronize()
    tim = eng.timing_read()
    eng.timing(False)
    eng.close()
    out = {"workload": "COMB local SiO2 terms as comb.py builds them (bond_bend, charge_exp, "
                       "charge_abs, attraction scaled by three 'tersoff' coordinators), "
                       "amorphous SiO2 24,576 atoms, fixed charges, VV dt = 0.5 fs, 300 K, "
                       "lists rebuilt every step, FP64",
           "atoms": n, "steps": steps, "value": round(n * steps / wall, 1),
           "unit": "atom-steps/s", "ms_per_step": round(wall / steps * 1e3, 4),
           "total_energy_drift_eV_per_atom": (e1 - e0) / n,
           "kernels_us": {k: round(v[0] / max(v[1], 1) * 1e3, 1) for k, v in tim.items()}}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if O.ref_available():
            r, qr = system(ref_si, 7)
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from helpers import ref_bonds, ref_entries
            pr, br, ng = build_descriptors(comb_sio2(), {x: k for k, x in enumerate(names)})
            rcut = np.full(len(r), 3.55)
            R = O.RefCore()
            R.setup(r.symbols, r.positions, r.cell, r.pbc, charges=qr,
                    potentials=ref_entries(pr, names), tersoff=ref_bonds(br, names),
                    ngroups=ng, cutoffs=rcut)
            R.forces()
            k = 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < 5.0 and k < 50:
                R.rebuild(r.positions, rcut)
                R.forces()
                k += 1
            t_ref = (time.perf_counter() - t0) / k
            out["reference_sample"] = {
                "atoms": len(r), "seconds_per_step": round(t_ref, 5), "cores": 1,
                "kind": "reference", "atom_steps_per_s": round(len(r) / t_ref, 1),
                "note": "list rebuild + forces per step; serial amdflang -O2 build"}
    except Exception as exc:  # the checker is optional on the GPU host
        out["reference_sample"] = {"error": str(exc)}
    return out


def cpu_baseline(seconds_budget=20.0, ranks=16, mpi_steps=20):
    """Reference Fortran core on the host, on a 16x16x16-cell sample (32,768
    atoms) of the same workload: VV steps with neighbor rebuild + forces each.
    The MPI build (-DMPI, MPICH) runs on `ranks` processes under mpiexec, as
    pysic runs under mpirun; the serial build is timed beside it in this
    process.  The MPI figure is the baseline when it ran."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not O.ref_available():
        return None
    import ref_md
    R, s, cut = ref_md.setup(16)
    n = len(s)
    steps, el = ref_md.run_vv(R, s, cut, seconds=seconds_budget)
    serial = {"value": n * steps / el, "unit": "atom-steps/s", "cores": 1, "kind": "reference",
              "sample": "reference Fortran core (serial, amdflang -O2), %d atoms (16^3 cells of "
                        "the same Si workload), %d velocity-Verlet steps with neighbor rebuild + "
                        "forces each step, %.1f s; its force loop is O(N^2) in N, so this "
                        "sample overstates its 1M-atom rate" % (n, steps, el)}
    mpi = reference_mpi_sample(ranks, mpi_steps)
    if "error" in mpi:
        serial["mpi"] = mpi
        return serial
    out = {"value": mpi["atoms"] * mpi["steps"] / mpi["seconds"], "unit": "atom-steps/s",
           "cores": mpi["ranks"], "kind": "reference",
           "sample": "reference Fortran core built with -DMPI (amdflang -O2, MPICH 3.3.2), "
                     "mpiexec -n %d, one rank per core, %d atoms (16^3 cells of the same Si "
                     "workload), %d velocity-Verlet steps with neighbor rebuild + forces each "
                     "step, %.2f s; its force loop is O(N^2) in N, so this sample overstates "
                     "its 1M-atom rate" % (mpi["ranks"], mpi["atoms"], mpi["steps"],
                                           mpi["seconds"]),
           "serial": serial}
    return out


def reference_mpi_sample(ranks, steps, timeout=300):
    """oracle/ref_md.py under mpiexec (child processes that never touch the
    GPU); its rank-0 JSON line, or {"error": ...}."""
    import shutil
    import subprocess
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    lib = os.path.join(ROOT, "oracle", "_ref", "libpysic_ref_mpi.so")
    if not (os.path.exists(mpiexec) and os.path.exists(lib)):
        return {"error": "no MPI build of the reference core (oracle/_ref) or no mpiexec"}
    env = dict(os.environ, OMP_NUM_THREADS="1")
    # the reference's mpi_stack keeps a 4 x max-neighbors x N integer array
    # on the stack (MPI.f90:416), 50 MB here: each rank raises its stack
    # limit to the hard limit first
    rank_cmd = "ulimit -s $(ulimit -H -s) && exec %s %s --reps 16 --steps %d" % (
        sys.executable, os.path.join(ROOT, "oracle", "ref_md.py"), steps)
    cmd = [mpiexec, "-n", str(ranks), "-launcher", "fork", "bash", "-c", rank_cmd]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env,
                           cwd="/tmp")
    except subprocess.TimeoutExpired:
        return {"error": "mpiexec timed out after %d s" % timeout}
    lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "mpiexec exit %d: %s" % (r.returncode, (r.stderr or r.stdout)[-300:])}
    return json.loads(lines[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, nargs=3, default=[64, 64, 32])
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-qeq", action="store_true", help="skip the QEq warm/cold measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-cells measurement")
    ap.add_argument("--no-ewald", action="store_true", help="skip the Ewald rocksalt measurement")
    ap.add_argument("--no-lj", action="store_true", help="skip the LJ argon measurement")
    ap.add_argument("--no-comb", action="store_true", help="skip the COMB SiO2 measurement")
    ap.add_argument("--only", choices=["qeq", "batch", "ewald", "lj", "comb"],
                    help="run one secondary measurement alone and print its JSON (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device copy (PMC traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only (--gpus 1); the multi-GPU decomposition of this "
              "path is designed but not built", file=sys.stderr)
        sys.exit(2)

    if args.only:
        if args.only == "batch":
            res = batched_benchmark(*build_batched_cells())
        elif args.only == "qeq":
            res = qeq_benchmark(os.path.join(ROOT, "pysic_amd", "data", "sio2_eem_mlp.npz"))
        elif args.only == "comb":
            res = comb_benchmark()
        elif args.only == "lj":
            res = lj_benchmark()
        else:
            res = ewald_benchmark()
        print(json.dumps({args.only: res}))
        return

    # CPU-side structures first (random addition on the host)
    batch_cells = None if args.no_batch else build_batched_cells()

    from pysic_amd.engine import Engine
    from pysic_amd.interactions import build_descriptors, tersoff_si
    from pysic_amd import autotune, structures as st

    s = build_workload(args.cells)
    n = len(s)
    pd, bd, _ = build_descriptors(tersoff_si(), {"Si": 0})
    eng = Engine(0)
    eng.create_atoms(np.zeros(n, np.int32), s.positions, masses=s.masses, momenta=s.momenta)
    eng.create_cell(s.cell, s.pbc)
    eng.set_potentials(pd)
    eng.set_bond_order_factors(bd)
    eng.ewald_off()
    eng.set_neighbor_capacity(16)
    eng.generate_neighbor_lists(np.full(n, 3.5))
    counts, _, _ = eng.neighbor_lists()
    n_nb = float(counts.mean())
    variant = args.variant if args.variant >= 0 else autotune.best_variant(eng, "tersoff_si")
    eng.set_variant(variant)
    dt = st.ASE_FS
    eng.md_run(args.warmup, dt, rebuild=1)
    # total energy before and after the timed steps (FP64 energy kernel; an
    # observer, outside the timed region: a VV step needs forces only)
    epot0 = eng.energy()
    _, mom0 = eng.state()
    ekin0 = float((mom0 ** 2 / (2 * s.masses[:, None])).sum())
    eng.synchronize()
    # the timed steps: device-resident, the step loop replayed from HIP graphs
    t0 = time.perf_counter()
    eng.md_run(args.steps, dt, rebuild=1)
    eng.synchronize()
    t1 = time.perf_counter()
    epot1 = eng.energy()
    _, mom1 = eng.state()
    ekin1 = float((mom1 ** 2 / (2 * s.masses[:, None])).sum())
    drift = ((epot1 + ekin1) - (epot0 + ekin0)) / n
    wall = t1 - t0
    value = n * args.steps / wall
    # per-kernel times: the same steps launched one by one, each kernel
    # bracketed by HIP events on the engine's stream
    eng.timing(True)
    eng.md_run(min(args.steps, 20), dt, rebuild=1)
    eng.synchronize()
    timers = eng.timing_read()
    eng.timing(False)

    site_ms, site_calls = timers.get("tersoff_site", (0.0, 1))
    site_avg_s = site_ms / max(site_calls, 1) / 1e3
    bytes_site = tersoff_bytes_per_atom(n_nb) * n
    achieved = bytes_site / site_avg_s / 1e9 if site_avg_s > 0 else 0.0
    variant_name = eng.variants()[variant]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_counters(args.cells, variant_name, "traffic_bytes_per_launch"),
                "kernel": "tersoff_site", "avg_launch_us": round(site_avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": int(bytes_site)}
    # the bound that holds this kernel: VALU issue (instruction count per
    # launch from the recorded SQ_INSTS_VALU, over the live launch time)
    valu = measured_counters(args.cells, variant_name, "valu_wave_instructions_per_launch")
    if valu and site_avg_s > 0:
        rate = valu / site_avg_s
        roofline["valu_issue"] = {"achieved": round(rate / 1e9, 1), "peak": VALU_ISSUE_PEAK / 1e9,
                                  "unit": "G wave-instructions/s",
                                  "frac": round(rate / VALU_ISSUE_PEAK, 4),
                                  "wave_instructions_per_launch": valu}
    kernels = {k: {"avg_us": round(v[0] / max(v[1], 1) * 1e3, 2), "calls": v[1]}
               for k, v in timers.items()}
    if args.copy_calibration:
        copy_calibration()
    eng.close()
    model = os.path.join(ROOT, "pysic_amd", "data", "sio2_eem_mlp.npz")
    qeq = None if args.no_qeq else qeq_benchmark(model)
    batched = None if batch_cells is None else batched_benchmark(*batch_cells)
    # the reference core is one per process and its Ewald switch, once set,
    # stays on (Core.f90:87, 8106): time the Tersoff baseline before the
    # Ewald sample
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    lj = None if args.no_lj else lj_benchmark()
    ewald = None if args.no_ewald else ewald_benchmark()
    comb = None if args.no_comb else comb_benchmark()
    out = {"metric": "atom-steps/s Tersoff 1M-atom Si; QEq time-to-1e-6 learned vs cold start",
           "value": round(value, 1),
           "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 per-bond arithmetic, f64 positions/separations/force gather",
           "data": "synthetic (seeded diamond Si + Gaussian displacements, MB 600 K)",
           "config": {"workload": "Tersoff Si %dx%dx%d diamond cells, %d atoms, VV dt=1 fs, "
                                  "rebuild every step" % (tuple(args.cells) + (n,)),
                      "atoms": n, "neighbors_per_atom": round(n_nb, 3),
                      "kernel_variant": variant_name,
                      "total_energy_drift_eV_per_atom": drift},
           "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels, "qeq": qeq,
           "batched": batched, "ewald": ewald, "lj": lj, "comb": comb}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
