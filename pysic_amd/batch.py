"""Many independent periodic cells evaluated together (include/pysic_batch.h).

pysic evaluates one structure per calculator; screening M structures means M
calculators in turn.  BatchedCells puts the atoms of all M cells into one
device system: one launch per kernel family covers every cell, per-cell
energies and virials are summed on the device, and the whole evaluation
(list rebuild + Tersoff forces + energies) can be captured once into a HIP
graph and replayed.

    cells = BatchedCells(structures, potentials)      # pysic Potential objects
    cells.evaluate(use_graph=True)
    energies, virials, forces = cells.results()
"""
import time

import numpy as np

from .interactions import build_descriptors
from . import structures as st


def batched_amorphous_si(ncells=512, nmin=64, nmax=2048, dmin=2.1, dens=(2.25, 2.35), seed=0):
    """The batched workload's cells (sizes, densities and seeds drawn from one
    seeded generator), built in this process."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=ncells)
    densities = rng.uniform(dens[0], dens[1], size=ncells)
    seeds = rng.integers(0, 2 ** 31 - 1, size=ncells)
    return [st.amorphous_si(int(a), float(b), int(c), dmin)
            for a, b, c in zip(sizes, densities, seeds)]


class BatchedCells:
    """A batch of structures with one interaction model on one GPU."""

    def __init__(self, structures, potentials, device=0, skin=0.5, capacity=None):
        from .engine import Engine
        self.structures = list(structures)
        names = sorted(set(x for s in self.structures for x in s.symbols))
        self.spmap = {x: k for k, x in enumerate(names)}
        pd, bd, _ = build_descriptors(potentials, self.spmap)
        self.natoms = np.array([len(s) for s in self.structures], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms)])
        pos = np.concatenate([s.positions for s in self.structures])
        species = np.array([self.spmap[x] for s in self.structures for x in s.symbols],
                           dtype=np.int32)
        masses = np.concatenate([s.masses for s in self.structures])
        self.n = len(pos)
        e = self.eng = Engine(device)
        e.create_atoms(species, pos, masses=masses)
        e.set_cells(self.natoms, np.array([s.cell for s in self.structures]),
                    np.array([[1 if p else 0 for p in s.pbc] for s in self.structures]))
        e.set_potentials(pd)
        e.set_bond_order_factors(bd)
        e.ewald_off()
        if capacity:
            e.set_neighbor_capacity(capacity)
        # list radius: the largest cutoff of any term plus the skin
        rc = max([row[6] for row in pd] + [row[10] for row in bd])
        self.list_radius = rc + skin
        e.generate_neighbor_lists(np.full(self.n, self.list_radius))
        self.fit_variant()

    def fit_variant(self):
        """Pick the smallest site-kernel shape whose short list holds every
        atom's short neighbors (amorphous cells reach ~10 at 3 A)."""
        names = self.eng.variants()
        caps = {v: int(name.split("_s")[1]) for v, name in enumerate(names)}
        lanes = {v: int(name.split("_")[0][1:]) for v, name in enumerate(names)}
        self.set_variant(0)
        self.eng.evaluate_cells(False, True)
        need = int(self.eng.check_cells()[1])
        if need > 0:
            ok = [v for v in caps if caps[v] >= need]
            if not ok:
                raise ValueError("%d short-range neighbors exceed every kernel shape" % need)
            self.set_variant(min(ok, key=lambda v: (caps[v], lanes[v])))
        return names[self.variant]

    def update_positions(self, positions):
        self.eng.update_coordinates(np.asarray(positions).reshape(-1, 3))

    def set_variant(self, v):
        self.eng.set_variant(v)
        self.variant = v

    def set_site_grid(self, blocks):
        self.eng.set_site_grid(blocks)

    def evaluate(self, use_graph=False, rebuild=True):
        self.eng.evaluate_cells(use_graph, rebuild)

    def results(self):
        """(energies[M], virials[M, 6], forces[N, 3]) of the last evaluation."""
        return self.eng.cell_results()

    def cell_slice(self, c):
        return slice(int(self.offsets[c]), int(self.offsets[c + 1]))

    def autotune(self, reps=3, grids=(0, 1024, 2048)):
        """Time every site-kernel shape (lanes per atom, short-list capacity)
        and persistent grid size on this batch; keep the fastest."""
        from .engine import PysicEngineError
        best = None
        table = {}
        names = self.eng.variants()
        for v, name in enumerate(names):
            for grid in grids:
                try:
                    self.set_variant(v)
                    self.eng.set_site_grid(grid)
                    self.evaluate()
                    self.results()
                except PysicEngineError:
                    continue
                self.eng.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    self.evaluate()
                self.eng.synchronize()
                dt = (time.perf_counter() - t0) / reps
                table[(name, grid)] = dt
                if best is None or dt < best[0]:
                    best = (dt, v, grid)
        if best is None:
            raise PysicEngineError("no site-kernel shape holds this batch's neighbor lists")
        self.set_variant(best[1])
        self.eng.set_site_grid(best[2])
        return {"variant": names[best[1]], "site_grid": best[2],
                "ms": {"%s/grid%d" % k: round(t * 1e3, 3) for k, t in table.items()}}

    def close(self):
        self.eng.close()
