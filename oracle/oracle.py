"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (the FP64 C restatement in pysic_oracle.c) and,
when it has been built, ``oracle/_ref/libpysic_ref.so`` (the reference
Fortran core compiled from its own sources by build_ref.sh).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use this module;
the engine in pysic_amd never imports it.

Descriptor formats shared with the engine's C-ABI:
  potential rows (12 doubles): type, n_targets, e1, e2, p1, p2, cutoff,
      soft_cutoff, group, p3, p4, link; type 1 exponential, 2 charge_self,
      3 LJ, 4 spring, 5 Buckingham, 6 power, 7 tabulated, 8 shift_power,
      9 step, 10 constant, 11 force; link 1 = multiplier of the preceding
      potential.  Rows of 10 (p3 = p4 = link = 0) are accepted and padded.
  bond rows (14 doubles, 'tersoff'): e1 e2 e3 (center, j, k), beta, eta,
      a, c, d, h, mu, cutoff, soft_cutoff, group, 0
  ewald (8 doubles): on, real_cutoff, k_cutoff, K1, K2, K3, sigma, epsilon
"""
import ctypes
import math
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpysic_ref.so")
# the reference core built with -DMPI (build_ref.sh), for the multi-rank
# CPU baseline (ref_md.py); selected by use_mpi_reference()
REF_MPI_PATH = os.path.join(HERE, "_ref", "libpysic_ref_mpi.so")


def use_mpi_reference():
    """Load the MPI build of the reference core in this process (before the
    first RefCore); every rank of an mpiexec launch calls it."""
    global REF_PATH
    if _ref is not None:
        raise RuntimeError("the reference core is already loaded")
    REF_PATH = REF_MPI_PATH

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)

_lib = None
_ref = None


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_neighbor_list.restype = ctypes.c_long
        L.orc_neighbor_list.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _dp, ctypes.c_long,
                                        _ip, _ip, _ip, _lp]
        L.orc_evaluate.restype = ctypes.c_int
        L.orc_evaluate.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _ip, _ip,
                                   _ip, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                   _dp, _dp]
        L.orc_relax_dynamic.restype = ctypes.c_int
        L.orc_relax_dynamic.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _ip, _ip, _ip,
                                        ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _dp, _dp]
        L.orc_set_table.restype = ctypes.c_int
        L.orc_set_table.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_pick.restype = ctypes.c_int
        L.orc_pick.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        _lib = L
    return _lib


def pot_rows(potdesc):
    """Potential descriptor rows as a flat array of 12-double rows."""
    pd = np.asarray(potdesc, dtype=np.float64)
    if pd.size == 0:
        return np.zeros(0)
    if pd.ndim != 2:
        pd = pd.reshape(-1, 12)
    if pd.shape[1] == 10:
        pd = np.hstack([pd, np.zeros((len(pd), 2))])
    assert pd.shape[1] == 12
    return np.ascontiguousarray(pd.reshape(-1))


def set_table(table_id, values):
    """Registers a 'tabulated' potential table (rows of value, derivative)."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, 2)
    rc = lib().orc_set_table(int(table_id), len(v), _d(v))
    if rc != 0:
        raise RuntimeError("orc_set_table failed: %d" % rc)


def ref_available():
    return os.path.exists(REF_PATH)


class NeighborLists:
    """CSR neighbor lists: start[N+1], idx[nnz] (0-based), off[nnz,3]."""

    def __init__(self, counts, idx, off):
        self.counts = np.ascontiguousarray(counts, dtype=np.int32)
        self.start = np.zeros(len(counts) + 1, dtype=np.int32)
        np.cumsum(self.counts, out=self.start[1:])
        self.idx = np.ascontiguousarray(idx, dtype=np.int32)
        self.off = np.ascontiguousarray(off, dtype=np.int32).reshape(-1, 3)

    def pair_set(self):
        """Sorted (i, j, o1, o2, o3) rows: the list as a set."""
        i = np.repeat(np.arange(len(self.counts), dtype=np.int64), self.counts)
        rows = np.column_stack([i, self.idx.astype(np.int64), self.off.astype(np.int64)])
        order = np.lexsort(rows.T[::-1])
        return rows[order]


def inverse_of(cell):
    """Row-major inverse of the matrix M whose columns are the cell vectors,
    with the bits the reference's caller produces (calculator.py:1557-1561
    passes numpy's inv(cell).T, i.e. M^-1, in Fortran order)."""
    return np.ascontiguousarray(np.linalg.inv(np.asarray(cell, dtype=np.float64)).T)


def image_neighbor_list(pos, cell, pbc, cutoffs):
    """Neighbor lists by enumeration of periodic images, for any cell: every
    (i, j, n) with |r_j - r_i + n . cell| < cutoffs[i] and (j, n) != (i, 0),
    the offsets n relative to the given (unwrapped) positions.  Where a cell
    is thinner than the cutoff the reference hands the search to ASE's
    NeighborList (calculator.py:798-850), whose lists hold the pairs within
    the sum of two radii plus the skin; the energies and forces only need the
    pairs within each potential's cutoff, which this list contains.
    O(N^2 * images): small cells only."""
    pos = np.asarray(pos, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    cut = np.asarray(cutoffs, dtype=np.float64)
    n = len(pos)
    inv = np.linalg.inv(cell)
    frac = pos @ inv
    wrap = np.where(np.asarray(pbc, bool)[None, :], np.floor(frac), 0.0)
    vol = abs(np.linalg.det(cell))
    reach = []
    for a in range(3):
        if pbc[a]:
            width = vol / np.linalg.norm(np.cross(cell[(a + 1) % 3], cell[(a + 2) % 3]))
            reach.append(int(math.ceil(cut.max() / width)) + 1)
        else:
            reach.append(0)
    shifts = np.array([[x, y, z] for x in range(-reach[0], reach[0] + 1)
                       for y in range(-reach[1], reach[1] + 1)
                       for z in range(-reach[2], reach[2] + 1)], dtype=np.int64)
    counts = np.zeros(n, dtype=np.int32)
    idx, off = [], []
    for i in range(n):
        # images of every j around i's wrapped position
        base = (wrap[i][None, :] - wrap).astype(np.int64)            # [j, 3]
        o = base[None, :, :] + shifts[:, None, :]                     # [shift, j, 3]
        # separation_vector's order (Geometry.f90:449-466): (r_j - r_i),
        # then + o_1 v_1, + o_2 v_2, + o_3 v_3
        d = np.broadcast_to(pos[None, :, :] - pos[i][None, None, :], o.shape).copy()
        of = o.astype(np.float64)
        for a in range(3):
            d = d + of[:, :, a:a + 1] * cell[a][None, None, :]
        r2 = d[:, :, 0] * d[:, :, 0] + d[:, :, 1] * d[:, :, 1] + d[:, :, 2] * d[:, :, 2]
        ok = r2 < cut[i] * cut[i]
        ok &= ~((np.arange(n)[None, :] == i) & (o == 0).all(axis=2))
        sidx, jidx = np.nonzero(ok)
        counts[i] = len(jidx)
        idx.append(jidx)
        off.append(o[sidx, jidx])
    return NeighborLists(counts, np.concatenate(idx) if n else np.zeros(0, np.int32),
                         np.concatenate(off) if n else np.zeros((0, 3), np.int32))


def neighbor_list(pos, cell, pbc, cutoffs):
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    inv = inverse_of(cell)
    pbc = np.ascontiguousarray(pbc, dtype=np.int32)
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    n = len(pos)
    counts = np.zeros(max(n, 1), dtype=np.int32)
    cap = max(64, 16 * n)
    while True:
        idx = np.zeros(cap, dtype=np.int32)
        off = np.zeros(3 * cap, dtype=np.int32)
        need = ctypes.c_long(0)
        tot = L.orc_neighbor_list(n, _d(pos), _d(cell), _d(inv), _i(pbc), _d(cut), cap,
                                  _i(counts), _i(idx), _i(off), ctypes.byref(need))
        if tot == -2:
            cap = int(need.value)
            continue
        if tot < 0:
            raise RuntimeError("orc_neighbor_list failed: %d" % tot)
        return NeighborLists(counts[:n], idx[:tot], off[:3 * tot])


def evaluate(calc, species, pos, charges, cell, nl, potdesc, bonddesc=None, ewald=None,
             scaler=None):
    """calc: 'energy' -> float; 'forces' -> (F[N,3], stress[6]); 'enegs' -> chi[N]."""
    L = lib()
    kind = {"energy": 1, "forces": 2, "enegs": 3}[calc]
    sp = np.ascontiguousarray(species, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    q = np.ascontiguousarray(charges if charges is not None else np.zeros(n), dtype=np.float64)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    inv = inverse_of(cell)
    pd = pot_rows(potdesc)
    npot = len(pd) // 12
    bd = None if bonddesc is None else np.ascontiguousarray(bonddesc, dtype=np.float64).reshape(-1)
    nbond = 0 if bd is None else len(bd) // 14
    ew = None if ewald is None else np.ascontiguousarray(ewald, dtype=np.float64)
    sc = None if scaler is None else np.ascontiguousarray(scaler, dtype=np.float64)
    e = np.zeros(1)
    f = np.zeros((n, 3))
    st = np.zeros(6)
    x = np.zeros(n)
    rc = L.orc_evaluate(kind, n, _i(sp), _d(pos), _d(q), _d(cell), _d(inv), _i(nl.start),
                        _i(nl.idx), _i(nl.off), npot, _d(pd), nbond, _d(bd), _d(ew), _d(sc),
                        _d(e), _d(f), _d(st), _d(x))
    if rc != 0:
        raise RuntimeError("orc_evaluate failed: %d" % rc)
    if kind == 1:
        return float(e[0])
    if kind == 2:
        return f, st
    return x


def relax_dynamic(species, pos, charges, cell, nl, potdesc, ewald, n_steps, dt, inertia,
                  friction, tolerance, rates=None, scaler=None):
    """Damped charge dynamics; returns (charges, rates, steps, error)."""
    L = lib()
    sp = np.ascontiguousarray(species, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    q = np.array(charges, dtype=np.float64)
    r = np.zeros(n) if rates is None else np.array(rates, dtype=np.float64)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    inv = inverse_of(cell)
    pd = pot_rows(potdesc)
    ew = None if ewald is None else np.ascontiguousarray(ewald, dtype=np.float64)
    sc = None if scaler is None else np.ascontiguousarray(scaler, dtype=np.float64)
    err = np.zeros(1)
    steps = L.orc_relax_dynamic(n, _i(sp), _d(pos), _d(q), _d(cell), _d(inv), _i(nl.start),
                                _i(nl.idx), _i(nl.off), len(pd) // 12, _d(pd), _d(ew), _d(sc),
                                n_steps, dt, inertia, friction, tolerance, _d(r), _d(err))
    if steps < 0:
        raise RuntimeError("orc_relax_dynamic failed: %d" % steps)
    return q, r, steps, float(err[0])


# --------------------------------------------------------------------------
# the reference Fortran core (checker build)
# --------------------------------------------------------------------------

def _label_code(sym):
    """Two-integer element code of pysic/utility/f2py.py (str2ints(s, 2))."""
    table = {}
    for k, ch in enumerate("abcdefghijklmnopqrstuvwxyz"):
        table[ch] = k + 1
    for k, ch in enumerate("ABCDEFGHIJKLMNOPQRSTUVWXYZ"):
        table[ch] = 101 + k
    out = [table.get(c, 0) for c in sym]
    while len(out) < 2:
        out.append(0)
    return out[:2]


def _labels(syms):
    return np.asfortranarray(np.array([_label_code(s) for s in syms], dtype=np.int32).T)


class RefCore:
    """The reference core, driven in the order of calculator.py:1595-1637.

    The reference core is a process-wide singleton; one RefCore at a time.
    Once Ewald parameters are set they stay in force for the process: the
    reference only ever switches evaluate_ewald on (Core.f90:87, 8106).
    """

    def __init__(self):
        global _ref
        if _ref is None:
            if not ref_available():
                raise RuntimeError("reference checker not built (oracle/build_ref.sh)")
            R = ctypes.CDLL(REF_PATH)
            R.ref_add_potential.restype = ctypes.c_int
            R.ref_add_potential.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                            ctypes.c_double, ctypes.c_double, _ip, ctypes.c_int,
                                            ctypes.c_int]
            R.ref_add_potential_orig.restype = ctypes.c_int
            R.ref_add_potential_orig.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                                 ctypes.c_double, ctypes.c_double, _ip, _ip,
                                                 ctypes.c_int, ctypes.c_int]
            R.ref_add_tersoff.restype = ctypes.c_int
            R.ref_add_tersoff.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _ip,
                                          ctypes.c_int]
            R.ref_create_system.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip, _dp,
                                            _dp, _ip]
            R.ref_update_coordinates.argtypes = [ctypes.c_int, _dp, _dp]
            R.ref_update_charges.argtypes = [ctypes.c_int, _dp]
            for name in ("ref_allocate_potentials", "ref_allocate_bond_order_factors"):
                getattr(R, name).argtypes = [ctypes.c_int]
            R.ref_allocate_bond_order_storage.argtypes = [ctypes.c_int] * 3
            R.ref_generate_neighbor_lists.argtypes = [ctypes.c_int, _dp]
            R.ref_neighbor_count.argtypes = [ctypes.c_int, _ip]
            R.ref_neighbor_list.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
            R.ref_set_ewald.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip,
                                        ctypes.c_double, ctypes.c_double, _dp]
            R.ref_clear_multipliers.argtypes = []
            R.ref_energy.argtypes = [_dp]
            R.ref_forces.argtypes = [ctypes.c_int, _dp, _dp]
            R.ref_electronegativities.argtypes = [ctypes.c_int, _dp]
            R.ref_number_of_types.argtypes = [_ip, _ip]
            R.ref_list_potentials.argtypes = [ctypes.c_int, _ip]
            R.ref_list_bond_order_factors.argtypes = [ctypes.c_int, _ip]
            R.ref_potential_info.argtypes = [ctypes.c_int, ctypes.c_char_p, _ip, _ip, _ip]
            R.ref_bond_order_factor_info.argtypes = [ctypes.c_int, ctypes.c_char_p, _ip, _ip,
                                                     _ip, _ip]
            R.ref_bond_order_factors.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
            R.ref_bond_order_gradients.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
            R.ref_bond_order_gradients_of_factor.argtypes = [ctypes.c_int, ctypes.c_int,
                                                             ctypes.c_int, _dp]
            R.ref_get_ewald_energy.argtypes = [ctypes.c_double, ctypes.c_double, _ip,
                                               ctypes.c_double, ctypes.c_double, _dp]
            R.ref_start()
            _ref = R
        self.R = _ref

    def setup(self, symbols, pos, cell, pbc, charges=None, masses=None, potentials=(),
              tersoff=(), ngroups=0, cutoffs=None, ewald=None, scaler=None, tables=None,
              momenta=None):
        """potentials: list of (kind, params, cutoff, soft, symbols_tuple, group
        [, multipliers[, original_symbols]]) with symbols in target order (all
        permutations already expanded; original_symbols, default the same, are
        the unpermuted targets, calculator.py:1247-1277) and multipliers a list
        of (kind, params[, original_symbols]) in the order the reference stores
        them; tersoff: list of (params7, cutoff, soft,
        (si, sj, sk), group); tables: {id: rows of (value, derivative)}."""
        R = self.R
        n = len(pos)
        self.n = n
        pos_f = np.asfortranarray(np.asarray(pos, dtype=np.float64).T)
        mom = np.zeros((3, n), order="F") if momenta is None else \
            np.asfortranarray(np.asarray(momenta, dtype=np.float64).T)
        q = np.zeros(n) if charges is None else np.asarray(charges, dtype=np.float64).copy()
        m = np.ones(n) if masses is None else np.asarray(masses, dtype=np.float64).copy()
        tags = np.zeros(n, dtype=np.int32)
        el = np.asfortranarray(np.array([_label_code(s) for s in symbols], dtype=np.int32).T)
        cellv = np.ascontiguousarray(cell, dtype=np.float64)  # C rows -> Fortran columns
        inv_f = np.ascontiguousarray(inverse_of(cellv).T)  # Fortran(r,c) = (M^-1)(r,c)
        pbc_i = np.asarray([1 if p else 0 for p in pbc], dtype=np.int32)
        R.ref_create_system(n, _d(m), _d(q), _d(pos_f), _d(mom), _i(tags), _i(el), _d(cellv),
                            _d(inv_f), _i(pbc_i))
        R.ref_allocate_potentials(len(potentials))
        here = os.getcwd()
        if tables:
            # read_table opens table_<id>.txt in the working directory
            tdir = tempfile.mkdtemp(prefix="pysic_tables_")
            for tid, vals in tables.items():
                np.savetxt(os.path.join(tdir, "table_%04d.txt" % tid),
                           np.asarray(vals, dtype=np.float64).reshape(-1, 2), fmt="%.17g")
            os.chdir(tdir)
        try:
            for entry in potentials:
                kind, params, cut, soft, syms, group = entry[:6]
                mults = entry[6] if len(entry) > 6 else ()
                orig = entry[7] if len(entry) > 7 else syms
                e = _labels(syms)
                # multipliers first (is_multiplier), then the potential itself
                for m in mults:
                    mp = np.asarray(m[1], dtype=np.float64)
                    mo = _labels(m[2] if len(m) > 2 else syms)
                    if not R.ref_add_potential_orig(m[0], len(syms), len(mp), _d(mp), cut, soft,
                                                    _i(e), _i(mo), group, 1):
                        raise RuntimeError("reference add_potential (multiplier) failed")
                p = np.asarray(params, dtype=np.float64)
                o = _labels(orig)
                ok = R.ref_add_potential_orig(kind, len(syms), len(p), _d(p), cut, soft, _i(e),
                                              _i(o), group, 0)
                R.ref_clear_multipliers()
                if not ok:
                    raise RuntimeError("reference add_potential failed")
        finally:
            os.chdir(here)
        R.ref_allocate_bond_order_factors(len(tersoff))
        for params, cut, soft, syms, group in tersoff:
            p = np.asarray(params, dtype=np.float64)
            e = np.asfortranarray(np.array([_label_code(s) for s in syms], dtype=np.int32).T)
            if not R.ref_add_tersoff(_d(p), cut, soft, _i(e), group):
                raise RuntimeError("reference add_bond_order_factor failed")
        R.ref_allocate_bond_order_storage(n, ngroups, len(tersoff))
        cut = np.asarray(cutoffs, dtype=np.float64)
        R.ref_generate_neighbor_lists(n, _d(cut))
        R.ref_create_lists()
        if ewald is not None:
            rc, kc, K, sigma, eps = ewald
            kk = np.asarray(K, dtype=np.int32)
            sc = np.ones(n) if scaler is None else np.asarray(scaler, dtype=np.float64)
            R.ref_set_ewald(n, rc, kc, _i(kk), sigma, eps, _d(sc))

    def rebuild(self, pos, cutoffs):
        pos_f = np.asfortranarray(np.asarray(pos, dtype=np.float64).T)
        mom = np.zeros((3, self.n), order="F")
        self.R.ref_update_coordinates(self.n, _d(pos_f), _d(mom))
        cut = np.asarray(cutoffs, dtype=np.float64)
        self.R.ref_generate_neighbor_lists(self.n, _d(cut))

    def set_charges(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.R.ref_update_charges(self.n, _d(q))

    def neighbor_lists(self):
        counts = np.zeros(self.n, dtype=np.int32)
        idxs, offs = [], []
        c = ctypes.c_int(0)
        for i in range(self.n):
            self.R.ref_neighbor_count(i, ctypes.byref(c))
            counts[i] = c.value
            if c.value > 0:
                nb = np.zeros(c.value, dtype=np.int32)
                of = np.zeros((c.value, 3), dtype=np.int32)  # Fortran (3,n) == C (n,3)
                self.R.ref_neighbor_list(i, c.value, _i(nb), _i(of))
                idxs.append(nb)
                offs.append(of)
        idx = np.concatenate(idxs) if idxs else np.zeros(0, np.int32)
        off = np.concatenate(offs) if offs else np.zeros((0, 3), np.int32)
        return NeighborLists(counts, idx, off)

    def energy(self):
        e = np.zeros(1)
        self.R.ref_energy(_d(e))
        return float(e[0])

    def forces(self):
        f = np.zeros((self.n, 3))  # Fortran (3,n) == C (n,3)
        st = np.zeros(6)
        self.R.ref_forces(self.n, _d(f), _d(st))
        return f, st

    def electronegativities(self):
        x = np.zeros(self.n)
        self.R.ref_electronegativities(self.n, _d(x))
        return x

    # -- the reference's descriptor registry (PyInterface.f90:706-1011) --
    def registry(self):
        """{'potentials': [(name, n_targets, [parameter names])],
        'bond_order_factors': [(name, n_targets, level, [[names] per target
        class 1, 2])]}, as the reference's own registry reports them."""
        R = self.R
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        R.ref_number_of_types(ctypes.byref(a), ctypes.byref(b))
        pc = np.zeros((a.value, 11), np.int32)
        R.ref_list_potentials(a.value, _i(pc))
        bc = np.zeros((b.value, 11), np.int32)
        R.ref_list_bond_order_factors(b.value, _i(bc))
        pots, bonds = [], []
        for code in pc:
            name = ints2str(code)
            nm = name.encode()
            nt, npar = ctypes.c_int(0), ctypes.c_int(0)
            names = np.zeros((12, 10), np.int32)  # Fortran (10,12)
            R.ref_potential_info(len(nm), nm, ctypes.byref(nt), ctypes.byref(npar), _i(names))
            pots.append((name, nt.value, [ints2str(c) for c in names[:npar.value]]))
        for code in bc:
            name = ints2str(code)
            nm = name.encode()
            nt, lv = ctypes.c_int(0), ctypes.c_int(0)
            npar = np.zeros(2, np.int32)
            names = np.zeros((2, 12, 10), np.int32)  # Fortran (10,12,2)
            R.ref_bond_order_factor_info(len(nm), nm, ctypes.byref(nt), ctypes.byref(lv),
                                         _i(npar), _i(names))
            bonds.append((name, nt.value, lv.value,
                          [[ints2str(c) for c in names[k][:npar[k]]] for k in range(2)]))
        return {"potentials": pots, "bond_order_factors": bonds}

    # -- bond-order factor queries (PyInterface.f90:579-634) --
    def bond_order_factors(self, group):
        x = np.zeros(self.n)
        self.R.ref_bond_order_factors(self.n, group, _d(x))
        return x

    def bond_order_gradients(self, group, atom):
        """Gradients (N,3) of every factor of the group with respect to
        moving atom (0-based)."""
        g = np.zeros((self.n, 3))  # Fortran (3,n) == C (n,3)
        self.R.ref_bond_order_gradients(self.n, group, atom + 1, _d(g))
        return g

    def bond_order_gradients_of_factor(self, group, atom):
        """Gradients (N,3) of atom's factor with respect to moving every atom."""
        g = np.zeros((self.n, 3))
        self.R.ref_bond_order_gradients_of_factor(self.n, group, atom + 1, _d(g))
        return g

    def ewald_energy(self, rc, kc, kcells, sigma, eps):
        """get_ewald_energy (PyInterface.f90:1050): the whole Ewald sum with
        the given parameters, unit scalers and the dipole term."""
        e = np.zeros(1)
        kk = np.asarray(kcells, dtype=np.int32)
        self.R.ref_get_ewald_energy(rc, kc, _i(kk), sigma, eps, _d(e))
        return float(e[0])


_I2S = {-10: "0", 201: "_"}
for _k in range(1, 10):
    _I2S[-_k] = str(_k)
for _k, _ch in enumerate("abcdefghijklmnopqrstuvwxyz"):
    _I2S[_k + 1] = _ch
for _k, _ch in enumerate("ABCDEFGHIJKLMNOPQRSTUVWXYZ"):
    _I2S[101 + _k] = _ch


def ints2str(codes):
    """The reference's integer codes back to text (Utility.f90:205-250;
    0 and unknown codes are blanks), stripped."""
    return "".join(_I2S.get(int(c), " ") for c in codes).strip()
