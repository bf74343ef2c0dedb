"""The Tersoff Si velocity-Verlet sample of the CPU baseline, run on the
reference Fortran core: serially in the calling process (si_md_serial), or
as one rank of an mpiexec launch of this file, which loads the core built
with -DMPI.

CPU-BASELINE AND TEST INFRASTRUCTURE ONLY: bench.py's CPU baseline runs it; nothing
in the product path imports it.

Under MPI every rank runs this same script, as pysic itself runs under
mpirun (pysic/calculator.py drives the core from Python on every rank): the
core gives each rank its share of the atoms (MPI.f90:144-174
mpi_distribute) and sums the forces over the ranks (Core.f90 calculate_forces,
mpi_allreduce).  Each VV step rebuilds the neighbor lists and evaluates the
forces, as the GPU workload does.

    mpiexec -n 16 python3 oracle/ref_md.py --reps 16 --steps 6

prints one JSON line from rank 0: atoms, ranks, steps, seconds.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402

# Tersoff Si (Tersoff 1988, the parameters of the GPU workload)
TERSOFF_PAIR = dict(A=1830.8, l1=2.4799, B=471.18, l2=1.7322)
TERSOFF_BOND = [1.1e-6, 0.78734, 0.0, 100390.0, 16.217, -0.59825, 3.0]


def setup(reps):
    """A reps^3-cell diamond Si sample (0.1 A random displacements, 600 K
    momenta) on the reference core; returns (core, structure, cutoffs)."""
    from pysic_amd import structures as st
    s = st.diamond_si((reps, reps, reps), a=5.431, sigma=0.1, seed=0)
    st.maxwell_boltzmann(s, 600.0, seed=1)
    n = len(s)
    T = TERSOFF_PAIR
    pots = [(1, [T["A"], T["l1"]], 3.0, 2.7, ("Si", "Si"), -1),
            (1, [-T["B"], T["l2"]], 3.0, 2.7, ("Si", "Si"), 0)]
    cut = np.full(n, 3.5)
    R = O.RefCore()
    R.setup(s.symbols, s.positions, s.cell, s.pbc, masses=s.masses, potentials=pots,
            tersoff=[(TERSOFF_BOND, 3.0, 2.7, ("Si", "Si", "Si"), 0)], ngroups=1, cutoffs=cut)
    return R, s, cut


def run_vv(R, s, cut, steps=None, seconds=None):
    """VV steps (neighbor rebuild + forces each); a fixed count, or as many
    as fit in `seconds` (serial use only: under MPI every rank must take the
    same number of steps).  Returns (steps, seconds)."""
    from pysic_amd import structures as st
    dt = st.ASE_FS
    m = s.masses[:, None]
    r, p = s.positions.copy(), s.momenta.copy()
    f, _ = R.forces()
    k = 0
    t0 = time.perf_counter()
    while True:
        p = p + 0.5 * dt * f
        r = r + dt * p / m
        R.rebuild(r, cut)
        f, _ = R.forces()
        p = p + 0.5 * dt * f
        k += 1
        el = time.perf_counter() - t0
        if steps is not None and k >= steps:
            break
        if seconds is not None and (el > seconds or k >= 50):
            break
    return k, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=16)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--forces-out", help="rank 0 saves the first forces and energy (.npz)")
    a = ap.parse_args()
    O.use_mpi_reference()
    R, s, cut = setup(a.reps)
    if a.forces_out:
        f, stress = R.forces()
        e = R.energy()
        if os.environ.get("PMI_RANK", "0") == "0":
            np.savez(a.forces_out, forces=f, stress=stress, energy=e)
    k, el = run_vv(R, s, cut, steps=a.steps)
    if os.environ.get("PMI_RANK", "0") == "0":
        print(json.dumps({"atoms": len(s), "ranks": int(os.environ.get("PMI_SIZE", "1")),
                          "steps": k, "seconds": el}), flush=True)


if __name__ == "__main__":
    main()
