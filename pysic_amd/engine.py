"""ctypes binding of libpysic_hip.so and a thin object over one handle.

This is the only place the Python side touches native code.  It plays the
role of the reference's f2py module (`pysic.pysic_fortran.pysic_interface`,
imported at pysic/calculator.py:16): every method maps one-to-one onto an
entry point of include/pysic_hip.h.  There is no fallback: if the library is
missing or cannot find a GPU, constructing an Engine raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpysic_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

_lib = None

# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "pysic_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "pysic_destroy": (None, [_vp]),
    "pysic_last_error": (ctypes.c_char_p, [_vp]),
    "pysic_create_atoms": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]),
    "pysic_create_cell": (ctypes.c_int, [_vp, _dp, _dp, _ip]),
    "pysic_update_atom_coordinates": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "pysic_update_atom_charges": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_get_number_of_atoms": (ctypes.c_int, [_vp, _ip]),
    "pysic_set_potentials": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_set_potential_table": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "pysic_set_bond_order_factors": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_set_ewald": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip,
                                       ctypes.c_double, ctypes.c_double, _dp]),
    "pysic_generate_neighbor_lists": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_get_neighbor_counts": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "pysic_get_neighbor_lists": (ctypes.c_int, [_vp, ctypes.c_long, _ip, _ip]),
    "pysic_calculate_energy": (ctypes.c_int, [_vp, _dp]),
    "pysic_calculate_energy_noalloc": (ctypes.c_int, [_vp, _dp]),
    "pysic_calculate_forces": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "pysic_calculate_electronegativities": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_relax_charges": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _ip, _dp]),
    "pysic_md_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
    "pysic_get_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "pysic_set_kernel_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pysic_kernel_variant_count": (ctypes.c_int, []),
    "pysic_kernel_variant_name": (ctypes.c_char_p, [ctypes.c_int]),
    "pysic_set_neighbor_capacity": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pysic_set_neighbor_method": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pysic_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pysic_timing_read": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _dp,
                                         _lp]),
    "pysic_synchronize": (ctypes.c_int, [_vp]),
    "pysic_calculate_bond_order_factors": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "pysic_calculate_bond_order_gradients": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                                            ctypes.c_int, _dp]),
    "pysic_calculate_bond_order_gradients_of_factor": (ctypes.c_int, [_vp, ctypes.c_int,
                                                                      ctypes.c_int, ctypes.c_int,
                                                                      _dp]),
    "pysic_get_ewald_energy": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _ip,
                                              ctypes.c_double, ctypes.c_double, _dp]),
    # descriptor registry (include/pysic_registry.h), host only
    "pysic_number_of_potentials": (ctypes.c_int, []),
    "pysic_number_of_bond_order_factors": (ctypes.c_int, []),
    "pysic_potential_name": (ctypes.c_char_p, [ctypes.c_int]),
    "pysic_bond_order_factor_name": (ctypes.c_char_p, [ctypes.c_int]),
    "pysic_is_potential": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_is_bond_order_factor": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_potential_type_code": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_number_of_targets_of_potential": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_number_of_targets_of_bond_order_factor": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_level_of_bond_order_factor": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_number_of_parameters_of_potential": (ctypes.c_int, [ctypes.c_char_p]),
    "pysic_number_of_parameters_of_bond_order_factor": (ctypes.c_int, [ctypes.c_char_p,
                                                                       ctypes.c_int]),
    "pysic_parameter_name_of_potential": (ctypes.c_char_p, [ctypes.c_char_p, ctypes.c_int]),
    "pysic_parameter_name_of_bond_order_factor": (ctypes.c_char_p, [ctypes.c_char_p, ctypes.c_int,
                                                                    ctypes.c_int]),
    "pysic_parameter_description_of_potential": (ctypes.c_char_p, [ctypes.c_char_p,
                                                                   ctypes.c_int]),
    "pysic_parameter_description_of_bond_order_factor": (ctypes.c_char_p, [ctypes.c_char_p,
                                                                           ctypes.c_int,
                                                                           ctypes.c_int]),
    "pysic_description_of_potential": (ctypes.c_char_p, [ctypes.c_char_p]),
    "pysic_description_of_bond_order_factor": (ctypes.c_char_p, [ctypes.c_char_p]),
    # batches of cells (include/pysic_batch.h)
    "pysic_set_cells": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _ip]),
    "pysic_get_number_of_cells": (ctypes.c_int, [_vp, _ip]),
    "pysic_set_site_grid": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pysic_evaluate_cells": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "pysic_check_cells": (ctypes.c_int, [_vp, _ip]),
    "pysic_get_cell_results": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]),
    # learned charge warm start (include/pysic_learned.h)
    "pysic_mlp_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "pysic_mlp_destroy": (None, [_vp]),
    "pysic_mlp_last_error": (ctypes.c_char_p, [_vp]),
    "pysic_mlp_param_count": (ctypes.c_int, []),
    "pysic_mlp_configure": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _fp,
                                           _fp]),
    "pysic_mlp_set_params": (ctypes.c_int, [_vp, _fp]),
    "pysic_mlp_get_params": (ctypes.c_int, [_vp, _fp]),
    "pysic_mlp_descriptors": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _fp, _ip]),
    "pysic_mlp_predict_charges": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _dp]),
    "pysic_mlp_set_corrector": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pysic_mlp_corrector_stages": (ctypes.c_int, [_vp]),
    "pysic_mlp_corrector_features": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "pysic_mlp_warm_start": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _dp]),
    "pysic_mlp_set_training_data": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _ip, _fp]),
    "pysic_mlp_evaluate": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _fp]),
    "pysic_mlp_gradients": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _fp, _fp]),
    "pysic_mlp_train": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]),
}


class PysicEngineError(RuntimeError):
    pass


def load():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PysicEngineError(
                "libpysic_hip.so is not built; run __graft_entry__.build() or "
                "python -m pysic_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Engine:
    """One device-resident system (one pysic handle)."""

    def __init__(self, device=0):
        self.L = load()
        h = _vp()
        rc = self.L.pysic_create(ctypes.byref(h), device)
        if rc != 0:
            raise PysicEngineError("pysic_create failed (%d): no usable HIP device %d" % (rc, device))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.pysic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.pysic_last_error(self.h)
            raise PysicEngineError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # --- system ----------------------------------------------------------
    def create_atoms(self, species, positions, masses=None, charges=None, momenta=None):
        pos = _f64(positions, (-1, 3))
        n = len(pos)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        m = None if masses is None else _f64(masses)
        q = None if charges is None else _f64(charges)
        p = None if momenta is None else _f64(momenta, (-1, 3))
        self._check(self.L.pysic_create_atoms(self.h, n, _d(m), _d(q), _d(pos), _d(p), _i(sp)),
                    "create_atoms")
        self.n = n

    def create_cell(self, cell, pbc=(True, True, True)):
        cell = _f64(cell, (3, 3))
        inv = np.ascontiguousarray(np.linalg.inv(cell).T)  # M^-1, M = vectors as columns
        pb = np.array([1 if p else 0 for p in pbc], dtype=np.int32)
        self._check(self.L.pysic_create_cell(self.h, _d(cell), _d(inv), _i(pb)), "create_cell")

    def update_coordinates(self, positions, momenta=None):
        pos = _f64(positions, (-1, 3))
        p = None if momenta is None else _f64(momenta, (-1, 3))
        self._check(self.L.pysic_update_atom_coordinates(self.h, len(pos), _d(pos), _d(p)),
                    "update_atom_coordinates")

    def update_charges(self, charges):
        q = _f64(charges)
        self._check(self.L.pysic_update_atom_charges(self.h, len(q), _d(q)), "update_atom_charges")

    def set_potentials(self, potdesc):
        """Descriptor rows of 12 doubles (include/pysic_hip.h); rows of 10
        (no p3, p4, link) are padded with zeros."""
        pd = np.asarray(potdesc, dtype=np.float64)
        if pd.size == 0:
            pd = np.zeros((0, 12))
        elif pd.ndim != 2:
            pd = pd.reshape(-1, 12)
        if pd.shape[1] == 10:
            pd = np.hstack([pd, np.zeros((len(pd), 2))])
        if pd.shape[1] != 12:
            raise ValueError("potential descriptor rows hold 12 doubles")
        pd = np.ascontiguousarray(pd.reshape(-1))
        self._check(self.L.pysic_set_potentials(self.h, len(pd) // 12, _d(pd)), "set_potentials")

    def set_potential_table(self, table_id, rows):
        """Table of the 'tabulated' potentials with this id: rows of
        (value, derivative)."""
        t = _f64(rows, (-1, 2))
        self._check(self.L.pysic_set_potential_table(self.h, int(table_id), len(t), _d(t)),
                    "set_potential_table")

    def set_bond_order_factors(self, bonddesc):
        bd = _f64(bonddesc).reshape(-1)
        self._check(self.L.pysic_set_bond_order_factors(self.h, len(bd) // 14, _d(bd)),
                    "set_bond_order_factors")

    def set_ewald(self, real_cutoff, k_cutoff, kcells, sigma, epsilon, scaler=None, on=True):
        K = np.ascontiguousarray(kcells, dtype=np.int32)
        sc = None if scaler is None else _f64(scaler)
        self._check(self.L.pysic_set_ewald(self.h, 1 if on else 0, real_cutoff, k_cutoff, _i(K),
                                           sigma, epsilon, _d(sc)), "set_ewald")

    def ewald_off(self):
        K = np.zeros(3, dtype=np.int32)
        self._check(self.L.pysic_set_ewald(self.h, 0, 0.0, 0.0, _i(K), 0.0, 0.0, None), "set_ewald")

    # --- neighbor lists --------------------------------------------------
    def generate_neighbor_lists(self, cutoffs):
        c = _f64(cutoffs)
        self._check(self.L.pysic_generate_neighbor_lists(self.h, len(c), _d(c)),
                    "generate_neighbor_lists")

    def neighbor_lists(self):
        """(counts[N], idx[nnz], offsets[nnz,3]) in atom order."""
        counts = np.zeros(self.n, dtype=np.int32)
        self._check(self.L.pysic_get_neighbor_counts(self.h, self.n, _i(counts)),
                    "get_neighbor_counts")
        nnz = int(counts.sum())
        idx = np.zeros(max(nnz, 1), dtype=np.int32)
        off = np.zeros((max(nnz, 1), 3), dtype=np.int32)
        self._check(self.L.pysic_get_neighbor_lists(self.h, nnz, _i(idx), _i(off)),
                    "get_neighbor_lists")
        return counts, idx[:nnz], off[:nnz]

    def bond_order_factors(self, group):
        """calculate_bond_order_factors (PyInterface.f90:579)."""
        x = np.zeros(self.n)
        self._check(self.L.pysic_calculate_bond_order_factors(self.h, self.n, int(group), _d(x)),
                    "calculate_bond_order_factors")
        return x

    def bond_order_gradients(self, group, atom, of_factor=False):
        """calculate_bond_order_gradients[_of_factor] (PyInterface.f90:605,
        634): (N,3); atom 0-based."""
        g = np.zeros((self.n, 3))
        fn = (self.L.pysic_calculate_bond_order_gradients_of_factor if of_factor else
              self.L.pysic_calculate_bond_order_gradients)
        self._check(fn(self.h, self.n, int(group), int(atom), _d(g)),
                    "calculate_bond_order_gradients")
        return g

    def ewald_energy(self, real_cut, k_cut, kcells, sigma, epsilon):
        """get_ewald_energy (PyInterface.f90:1050)."""
        k = np.ascontiguousarray(kcells, dtype=np.int32)
        e = np.zeros(1)
        self._check(self.L.pysic_get_ewald_energy(self.h, float(real_cut), float(k_cut), _i(k),
                                                  float(sigma), float(epsilon), _d(e)),
                    "get_ewald_energy")
        return float(e[0])

    def set_neighbor_method(self, method):
        self._check(self.L.pysic_set_neighbor_method(self.h, int(method)), "set_neighbor_method")

    def set_neighbor_capacity(self, cap):
        self._check(self.L.pysic_set_neighbor_capacity(self.h, int(cap)), "set_neighbor_capacity")

    # --- evaluation ------------------------------------------------------
    def energy(self):
        e = np.zeros(1)
        self._check(self.L.pysic_calculate_energy(self.h, _d(e)), "calculate_energy")
        return float(e[0])

    def energy_noalloc(self):
        """pysic_calculate_energy_noalloc: the energy with no device
        allocation (fixed library scratch; see include/pysic_hip.h)."""
        e = np.zeros(1)
        self._check(self.L.pysic_calculate_energy_noalloc(self.h, _d(e)),
                    "calculate_energy_noalloc")
        return float(e[0])

    def forces(self):
        f = np.zeros((self.n, 3))
        st = np.zeros(6)
        self._check(self.L.pysic_calculate_forces(self.h, self.n, _d(f), _d(st)), "calculate_forces")
        return f, st

    def electronegativities(self):
        x = np.zeros(self.n)
        self._check(self.L.pysic_calculate_electronegativities(self.h, self.n, _d(x)),
                    "calculate_electronegativities")
        return x

    def relax_charges(self, charges, rates, n_steps, timestep, inertia, friction, tolerance):
        q = np.array(charges, dtype=np.float64)
        r = np.array(rates, dtype=np.float64)
        steps = ctypes.c_int(0)
        err = np.zeros(1)
        self._check(self.L.pysic_relax_charges(self.h, self.n, _d(q), _d(r), int(n_steps),
                                               timestep, inertia, friction, tolerance,
                                               ctypes.byref(steps), _d(err)), "relax_charges")
        return q, r, steps.value, float(err[0])

    def md_run(self, nsteps, dt, rebuild=1, energies=False):
        e = np.zeros(max(nsteps, 1)) if energies else None
        self._check(self.L.pysic_md_run(self.h, int(nsteps), float(dt), int(rebuild), _d(e)),
                    "md_run")
        return None if e is None else e[:nsteps]

    def state(self):
        pos = np.zeros((self.n, 3))
        mom = np.zeros((self.n, 3))
        self._check(self.L.pysic_get_state(self.h, self.n, _d(pos), _d(mom)), "get_state")
        return pos, mom

    # --- batches of cells -------------------------------------------------
    def set_cells(self, natoms, cells, pbc=None):
        """Cell c owns the next natoms[c] atoms (in create_atoms order)."""
        na = np.ascontiguousarray(natoms, dtype=np.int32)
        cv = _f64(cells, (-1, 3, 3))
        inv = np.ascontiguousarray(np.transpose(np.linalg.inv(cv), (0, 2, 1)))
        pb = np.ones((len(na), 3), dtype=np.int32) if pbc is None else \
            np.ascontiguousarray(pbc, dtype=np.int32).reshape(-1, 3)
        self._check(self.L.pysic_set_cells(self.h, len(na), _i(na), _d(cv), _d(inv), _i(pb)),
                    "set_cells")
        self.ncells = len(na)

    def set_site_grid(self, blocks):
        self._check(self.L.pysic_set_site_grid(self.h, int(blocks)), "set_site_grid")

    def evaluate_cells(self, use_graph=False, rebuild=True):
        self._check(self.L.pysic_evaluate_cells(self.h, 1 if use_graph else 0, 1 if rebuild else 0),
                    "evaluate_cells")

    def check_cells(self):
        f = np.zeros(4, dtype=np.int32)
        self._check(self.L.pysic_check_cells(self.h, _i(f)), "check_cells")
        return f

    def cell_results(self, forces=True):
        e = np.zeros(self.ncells)
        w = np.zeros((self.ncells, 6))
        f = np.zeros((self.n, 3)) if forces else None
        self._check(self.L.pysic_get_cell_results(self.h, self.ncells, _d(e), _d(w), self.n, _d(f)),
                    "get_cell_results")
        return e, w, f

    # --- tuning / measurement -------------------------------------------
    def set_variant(self, v):
        self._check(self.L.pysic_set_kernel_variant(self.h, int(v)), "set_kernel_variant")

    def variants(self):
        return [self.L.pysic_kernel_variant_name(v).decode()
                for v in range(self.L.pysic_kernel_variant_count())]

    def timing(self, on):
        self._check(self.L.pysic_timing(self.h, 1 if on else 0), "timing")

    def timing_read(self):
        k = 32
        names = ctypes.create_string_buffer(k * 32)
        ms = np.zeros(k)
        calls = np.zeros(k, dtype=np.int64)
        cnt = self.L.pysic_timing_read(self.h, k, names, 32, _d(ms),
                                       calls.ctypes.data_as(_lp))
        if cnt < 0:
            self._check(cnt, "timing_read")
        out = {}
        raw = names.raw
        for i in range(cnt):
            nm = raw[i * 32:(i + 1) * 32].split(b"\0")[0].decode()
            out[nm] = (float(ms[i]), int(calls[i]))
        return out

    def synchronize(self):
        self._check(self.L.pysic_synchronize(self.h), "synchronize")


class Mlp:
    """The charge warm-start network on one device (include/pysic_learned.h)."""

    def __init__(self, device=0):
        self.L = load()
        m = _vp()
        rc = self.L.pysic_mlp_create(ctypes.byref(m), device)
        if rc != 0:
            raise PysicEngineError("pysic_mlp_create failed (%d): no usable HIP device %d"
                                   % (rc, device))
        self.m = m
        self.nparams = self.L.pysic_mlp_param_count()

    def close(self):
        if getattr(self, "m", None):
            self.L.pysic_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.pysic_mlp_last_error(self.m)
            raise PysicEngineError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def configure(self, nspecies, rc, mu, eta, mean, inv_scale):
        mu = _f64(mu)
        eta = _f64(eta)
        mean = np.ascontiguousarray(mean, dtype=np.float32)
        isc = np.ascontiguousarray(inv_scale, dtype=np.float32)
        assert len(mu) == 16 and len(eta) == 16 and len(mean) == 32 and len(isc) == 32
        self._check(self.L.pysic_mlp_configure(self.m, int(nspecies), float(rc), _d(mu), _d(eta),
                                               _f(mean), _f(isc)), "mlp_configure")

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.nparams
        self._check(self.L.pysic_mlp_set_params(self.m, _f(p)), "mlp_set_params")

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.L.pysic_mlp_get_params(self.m, _f(p)), "mlp_get_params")
        return p

    def descriptors(self, eng):
        feat = np.zeros((eng.n, 32), dtype=np.float32)
        sp = np.zeros(eng.n, dtype=np.int32)
        self._check(self.L.pysic_mlp_descriptors(self.m, eng.h, eng.n, _f(feat), _i(sp)),
                    "mlp_descriptors")
        return feat, sp

    def predict_charges(self, eng, neutralize=True):
        q = np.zeros(eng.n)
        self._check(self.L.pysic_mlp_predict_charges(self.m, eng.h, eng.n, 1 if neutralize else 0,
                                                     _d(q)), "mlp_predict_charges")
        return q

    CORRECTOR_NCOEF = 97

    def set_corrector(self, coef):
        """Corrector stages, coef[n_stages][97] (empty: none)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1, self.CORRECTOR_NCOEF)
        self._check(self.L.pysic_mlp_set_corrector(self.m, len(c), _d(c)), "mlp_set_corrector")

    def corrector_stages(self):
        return int(self.L.pysic_mlp_corrector_stages(self.m))

    def corrector_features(self, eng, charges, previous):
        """(features[n][97], chi[n]) of one corrector stage at the given
        charges and previous change; leaves the engine's charges at `charges`."""
        q = _f64(charges)
        d = _f64(previous)
        assert len(q) == eng.n and len(d) == eng.n
        feat = np.zeros((eng.n, self.CORRECTOR_NCOEF))
        chi = np.zeros(eng.n)
        self._check(self.L.pysic_mlp_corrector_features(self.m, eng.h, eng.n, _d(q), _d(d),
                                                        _d(feat), _d(chi)),
                    "mlp_corrector_features")
        return feat, chi

    def warm_start(self, eng, n_stages=None):
        """Neutral predicted charges followed by n_stages corrector stages
        (all held ones by default), computed on the device."""
        k = self.corrector_stages() if n_stages is None else int(n_stages)
        q = np.zeros(eng.n)
        self._check(self.L.pysic_mlp_warm_start(self.m, eng.h, eng.n, k, _d(q)), "mlp_warm_start")
        return q

    def set_training_data(self, feat, species, targets):
        f = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 32)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        assert len(f) == len(sp) == len(t)
        self._check(self.L.pysic_mlp_set_training_data(self.m, len(f), _f(f), _i(sp), _f(t)),
                    "mlp_set_training_data")

    def evaluate(self, samples):
        s = np.ascontiguousarray(samples, dtype=np.int32)
        y = np.zeros(len(s), dtype=np.float32)
        self._check(self.L.pysic_mlp_evaluate(self.m, len(s), _i(s), _f(y)), "mlp_evaluate")
        return y

    def gradients(self, samples):
        s = np.ascontiguousarray(samples, dtype=np.int32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.L.pysic_mlp_gradients(self.m, len(s), _i(s), _f(g), _f(loss)),
                    "mlp_gradients")
        return g, float(loss[0])

    def train(self, order, batch, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        o = np.ascontiguousarray(order, dtype=np.int32)
        nsteps = len(o) // batch
        assert nsteps >= 1
        losses = np.zeros(nsteps, dtype=np.float32)
        self._check(self.L.pysic_mlp_train(self.m, nsteps, int(batch), _i(o[:nsteps * batch]),
                                           lr, beta1, beta2, eps, _f(losses)), "mlp_train")
        return losses
