"""Seeded synthetic structures for the Tersoff, pair-potential, Ewald and EEM
workloads.

The reference obtains geometries from ASE (absent here), so the engine owns
its own generators.  All randomness comes from numpy's PCG64 generator with
an explicit seed, so every workload is reproducible bit for bit.
"""
import math

import numpy as np

AMU_SI = 28.0855
AMU_O = 15.999
AMU_NA = 22.98976928
AMU_CL = 35.453
AMU_AR = 39.948
AMU_KR = 83.798
AVOGADRO = 6.02214076e23

# standard atomic weights (amu) of the elements a structure may name without
# explicit masses
ATOMIC_MASS = {
    "H": 1.008, "He": 4.002602, "Li": 6.94, "Be": 9.0121831, "B": 10.81, "C": 12.011,
    "N": 14.007, "O": AMU_O, "F": 18.998403163, "Ne": 20.1797, "Na": AMU_NA, "Mg": 24.305,
    "Al": 26.9815385, "Si": AMU_SI, "P": 30.973761998, "S": 32.06, "Cl": AMU_CL,
    "Ar": AMU_AR, "K": 39.0983, "Ca": 40.078, "Ti": 47.867, "Fe": 55.845, "Ni": 58.6934,
    "Cu": 63.546, "Zn": 65.38, "Ga": 69.723, "Ge": 72.630, "As": 74.921595, "Se": 78.971,
    "Br": 79.904, "Kr": AMU_KR, "Ag": 107.8682, "Sn": 118.710, "Au": 196.966569,
}

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
DIAMOND_BASIS = np.array([
    [0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],
    [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25],
])
ROCKSALT_BASIS = np.array([
    [0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],   # cations
    [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5], [0.5, 0.5, 0.5],   # anions
])


class Structure:
    """Minimal stand-in for ASE Atoms: what the pysic calculator reads.

    positions (N,3) in Angstrom, cell rows = cell vectors, pbc (3,), symbols,
    masses (amu), charges (e), momenta (amu*Angstrom/ASE-time).
    """

    def __init__(self, symbols, positions, cell, pbc=(True, True, True), masses=None,
                 charges=None, momenta=None):
        self.symbols = list(symbols)
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        self.pbc = np.array([bool(p) for p in pbc])
        n = len(self.positions)
        if masses is None:
            missing = sorted(set(s for s in self.symbols if s not in ATOMIC_MASS))
            if missing:
                raise ValueError("no standard mass for %s: pass masses" % ", ".join(missing))
            masses = [ATOMIC_MASS[s] for s in self.symbols]
        self.masses = np.array(masses, dtype=np.float64)
        self.charges = np.zeros(n) if charges is None else np.array(charges, dtype=np.float64)
        self.momenta = np.zeros((n, 3)) if momenta is None else np.array(momenta, dtype=np.float64)

    # ASE-compatible accessors used by the calculator
    def __len__(self):
        return len(self.positions)

    def get_number_of_atoms(self):
        return len(self.positions)

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, p):
        self.positions = np.array(p, dtype=np.float64).reshape(-1, 3)

    def get_cell(self):
        return self.cell.copy()

    def set_cell(self, cell):
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)

    def get_pbc(self):
        return self.pbc.copy()

    def set_pbc(self, pbc):
        self.pbc = np.array([bool(p) for p in np.broadcast_to(pbc, (3,))])

    def get_chemical_symbols(self):
        return list(self.symbols)

    def get_masses(self):
        return self.masses.copy()

    def get_initial_charges(self):
        return self.charges.copy()

    def set_initial_charges(self, q):
        self.charges = np.array(q, dtype=np.float64)

    def get_charges(self):
        return self.charges.copy()

    def set_charges(self, q):
        self.set_initial_charges(q)

    def get_momenta(self):
        return self.momenta.copy()

    def set_momenta(self, p):
        self.momenta = np.array(p, dtype=np.float64).reshape(-1, 3)

    def get_tags(self):
        return np.zeros(len(self.positions), dtype=np.int64)

    def get_volume(self):
        return abs(float(np.linalg.det(self.cell)))

    def copy(self):
        return Structure(self.symbols, self.positions, self.cell, self.pbc, self.masses,
                         self.charges, self.momenta)


def _lattice(basis, a, reps):
    nx, ny, nz = reps
    cells = np.array([[i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)],
                     dtype=np.float64)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    return frac * a


def diamond_si(reps=(8, 8, 8), a=5.431, sigma=0.0, seed=0):
    """Diamond Si, a x reps conventional cells, Gaussian displacements."""
    pos = _lattice(DIAMOND_BASIS, a, reps)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    cell = np.diag([a * reps[0], a * reps[1], a * reps[2]])
    return Structure(["Si"] * len(pos), pos, cell)


def fcc(reps=(4, 4, 4), a=5.26, sigma=0.0, seed=0, symbol="Ar"):
    """Face-centred cubic crystal (argon: a = 5.26 A near its triple point),
    a x reps conventional cells, Gaussian displacements."""
    pos = _lattice(FCC_BASIS, a, reps)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    cell = np.diag([a * reps[0], a * reps[1], a * reps[2]])
    return Structure([symbol] * len(pos), pos, cell)


def random_mixture(n, box, dmin, symbols, seed):
    """Random-addition cubic cell of n atoms; atom k takes
    symbols[k % len(symbols)]."""
    rng = np.random.default_rng(seed)
    pos = random_sequential_addition(n, box, dmin, rng)
    sym = [symbols[k % len(symbols)] for k in range(n)]
    return Structure(sym, pos, np.diag([box] * 3))


def rocksalt(reps=(4, 4, 4), a=5.64, sigma=0.0, seed=0, cation="Na", anion="Cl"):
    """Rocksalt with charges +1 / -1, Gaussian displacements."""
    nx, ny, nz = reps
    cells = np.array([[i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)],
                     dtype=np.float64)
    frac = (cells[:, None, :] + ROCKSALT_BASIS[None, :, :]).reshape(-1, 3)
    pos = frac * a
    q = np.tile(np.array([1.0] * 4 + [-1.0] * 4), len(cells))
    sym = [cation if x > 0 else anion for x in q]
    if sigma > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, sigma, size=pos.shape)
    cell = np.diag([a * nx, a * ny, a * nz])
    return Structure(sym, pos, cell, charges=q)


def random_sequential_addition(n, box, dmin, rng, max_tries=2000000, batch=256):
    """Uniform random positions in a cubic periodic box with a hard minimum
    separation: candidates p = rng.random(3) * box are tried one after another
    and kept when no kept position lies closer than dmin (minimum image).

    Candidates are drawn and tested in batches against a cell grid of the
    kept positions, then resolved in order among themselves, so the result
    and the generator's final state are those of trying them one at a time."""
    ncell = max(1, int(box // dmin))
    w = box / ncell
    d2 = dmin * dmin
    cap = 4
    grid = np.zeros((ncell, ncell, ncell, cap, 3))
    count = np.zeros((ncell, ncell, ncell), dtype=np.int64)
    offs = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)])
    pos = np.zeros((n, 3))
    placed = 0
    tries = 0
    while placed < n:
        if tries >= max_tries:
            raise RuntimeError("random sequential addition jammed at %d of %d" % (placed, n))
        state = rng.bit_generator.state
        nb = min(batch, max_tries - tries)
        cand = rng.random((nb, 3)) * box
        cells = (cand // w).astype(int) % ncell
        around = (cells[:, None, :] + offs[None, :, :]) % ncell            # [nb, 27, 3]
        gx, gy, gz = around[..., 0], around[..., 1], around[..., 2]
        kept = grid[gx, gy, gz]                                            # [nb, 27, cap, 3]
        valid = np.arange(cap)[None, None, :] < count[gx, gy, gz][..., None]
        d = kept - cand[:, None, None, :]
        d -= box * np.round(d / box)
        clash = ((d * d).sum(axis=-1) < d2) & valid
        free = np.nonzero(~clash.any(axis=(1, 2)))[0]
        # candidates that pass the kept positions, resolved in order
        fc = cand[free]
        dd = fc[None, :, :] - fc[:, None, :]
        dd -= box * np.round(dd / box)
        near = (dd * dd).sum(axis=-1) < d2                                 # [m, m]
        blocked = np.zeros(len(free), dtype=bool)
        used = nb
        for a in range(len(free)):
            if blocked[a]:
                continue
            blocked |= near[a]
            k = free[a]
            cx, cy, cz = cells[k]
            slot = count[cx, cy, cz]
            if slot == cap:
                grid = np.concatenate([grid, np.zeros_like(grid)], axis=3)
                cap *= 2
            grid[cx, cy, cz, slot] = cand[k]
            count[cx, cy, cz] = slot + 1
            pos[placed] = cand[k]
            placed += 1
            if placed == n:
                used = k + 1
                break
        tries += used
        if used < nb:
            # leave the generator where one-at-a-time trials would stop
            rng.bit_generator.state = state
            rng.random((used, 3))
    return pos


def amorphous_si(n, density_gcc, seed, dmin=2.1):
    """Random-addition a-Si cell: n atoms at the given density (g/cm^3)."""
    rng = np.random.default_rng(seed)
    vol = n * AMU_SI / AVOGADRO / density_gcc * 1e24  # Angstrom^3
    box = vol ** (1.0 / 3.0)
    pos = random_sequential_addition(n, box, dmin, rng)
    return Structure(["Si"] * n, pos, np.diag([box] * 3))


def amorphous_sio2(n_si, density_gcc, seed, dmin=1.5):
    """Random-addition SiO2 (1:2) cell at the given density."""
    rng = np.random.default_rng(seed)
    n = 3 * n_si
    mass = n_si * (AMU_SI + 2 * AMU_O)
    vol = mass / AVOGADRO / density_gcc * 1e24
    box = vol ** (1.0 / 3.0)
    pos = random_sequential_addition(n, box, dmin, rng)
    species = np.array(["Si"] * n_si + ["O"] * (2 * n_si))
    perm = rng.permutation(n)
    return Structure(list(species[perm]), pos, np.diag([box] * 3))


# ASE unit conventions (eV, Angstrom, amu): 1 fs in ASE time units and k_B
ASE_FS = 1e-15 * math.sqrt(1.602176634e-19 / 1.66053906660e-27) * 1e10
KB_EV = 8.617333262e-5


def maxwell_boltzmann(struct, temperature_k, seed):
    """Momenta drawn from N(0, sqrt(m kT)) with zero net momentum."""
    rng = np.random.default_rng(seed)
    m = struct.masses[:, None]
    p = rng.normal(size=(len(struct), 3)) * np.sqrt(m * KB_EV * temperature_k)
    p -= p.sum(axis=0) / len(struct) * (m / m.mean())
    struct.set_momenta(p)
    return struct
