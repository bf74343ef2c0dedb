"""Drop-in for pysic's f2py module: ``pysic_fortran.pysic_interface``.

pysic imports its core as ``import pysic.pysic_fortran as pf`` (pysic/
calculator.py:16, core.py:3, interactions/bondorder.py:5, utility/mpi.py:4)
and calls ``pf.pysic_interface.<name>(...)``.  This module provides an object
``pysic_interface`` with every name pysic calls on the Tersoff / Ewald /
charge-relaxation path, with the calling conventions f2py gives the
reference's routines (fortran/PyInterface.f90): array dimensions inferred
from the arrays, ``intent(out)`` arguments returned, 3-vectors as Fortran
(3, N) arrays, element labels as two-integer codes and names as integer
codes (Utility.f90:205-250), atom indices 1-based where the f2py call takes
them so.  Behind it is one engine handle of libpysic_hip.so, created when
the atoms are, and a module-level core that holds what the reference's
Fortran module globals hold.

To use it, pysic's package gets a ``pysic/pysic_fortran.py`` containing

    from pysic_amd.pysic_fortran import pysic_interface

Limits are the engine's (one cell, up to four elements; no tags or indices
as targets: add_potential reports failure for them, and pysic raises its
InvalidPotentialError).
"""
import numpy as np

from . import engine as _engine
from .interactions import POTENTIAL_TYPES, factor_rows

# --- the reference's character codes (Utility.f90:205-250) ------------------
_S2I = {"_": 201}
for _k in range(1, 10):
    _S2I[str(_k)] = -_k
_S2I["0"] = -10
for _k, _c in enumerate("abcdefghijklmnopqrstuvwxyz"):
    _S2I[_c] = _k + 1
for _k, _c in enumerate("ABCDEFGHIJKLMNOPQRSTUVWXYZ"):
    _S2I[_c] = 101 + _k
_I2S = {v: k for k, v in _S2I.items()}


def _codes(text, length):
    """A string as `length` integer codes, blank padded (str2int)."""
    t = (text or "")[:length].ljust(length)
    return np.array([_S2I.get(c, 0) for c in t], dtype=np.int32)


def _text(codes):
    """Integer codes back to a stripped string (int2str)."""
    return "".join(_I2S.get(int(c), " ") for c in np.asarray(codes).ravel()).strip()


def _label(codes):
    return _text(codes)


def _s(b):
    return None if b is None else b.decode()


class _Core:
    """The state the reference keeps in its Fortran module globals
    (Core.f90:13-100): atoms, cell, interactions, Ewald parameters and the
    neighbor lists -- here one engine handle, plus the interactions as they
    were added (re-expressed as descriptor rows for the engine whenever the
    atoms or the interactions change)."""

    def __init__(self):
        self.eng = None
        self.cell = None
        self.pbc = None
        self.reset()

    def reset(self, keep_model=False):
        if self.eng is not None:
            self.eng.close()
        self.eng = None
        self.n = 0
        self.labels = []            # element label per atom
        self.species_of = {}        # label -> engine species id
        if not keep_model:
            self.pots = []          # (name, params, n_targets, labels, orig labels, cutoff,
                                    #  soft, group, multipliers [(name, params, orig)])
            self.mults = []         # pending multipliers
            self.bonds = []         # (params, cutoff, soft, labels, group)
            self.ewald = None       # (rc, kc, K, sigma, eps, scaler)
        self.pushed = False
        self.lists = None           # CSR (counts, start, idx, off) of the current lists
        self.external_lists = False
        self.seed = None

    def require_atoms(self):
        if self.eng is None:
            raise RuntimeError("pysic_interface: create_atoms has not been called")

    def species(self, labels):
        """Engine species of the labels, or None if one is absent from the
        system (such an entry can act on no atom)."""
        out = []
        for lab in labels:
            if lab not in self.species_of:
                return None
            out.append(self.species_of[lab])
        return out

    def push_model(self):
        """Hands the interactions to the engine (the reference's
        create_potential_list / create_bond_order_factor_list)."""
        self.require_atoms()
        rows, brows = [], []
        for name, params, nt, labels, olabels, cut, soft, group, mults in self.pots:
            sp = self.species(labels)
            if sp is None:
                continue
            osp = [self.species_of.get(lab, -1) for lab in olabels]
            rows += factor_rows(name, params, nt, sp, cut, soft, group, 0, osp)
            for mname, mparams, molabels in mults:
                mosp = [self.species_of.get(lab, -1) for lab in molabels]
                rows += factor_rows(mname, mparams, nt, sp, cut, soft, group, 1, mosp)
        for params, cut, soft, labels, group in self.bonds:
            sp = self.species(labels)
            if sp is not None:
                brows.append(sp + list(params) + [cut, soft, group, 0.0])
        self.eng.set_potentials(np.array(rows, dtype=np.float64).reshape(-1, 12))
        self.eng.set_bond_order_factors(np.array(brows, dtype=np.float64).reshape(-1, 14))
        self.pushed = True

    def model_cutoffs(self, skin=0.5):
        """Per-atom neighbor radius: the Coulomb real-space cutoff and the
        largest cutoff of any potential or bond-order factor naming the
        atom's element (get_individual_cutoffs, calculator.py:860-915), plus
        FastNeighborList's skin (calculator.py:48); used when pysic handed
        over neighbor lists built elsewhere."""
        base = self.ewald[0] if self.ewald is not None else 0.0
        cut = {lab: base for lab in self.species_of}
        for name, params, nt, labels, olabels, c, soft, group, mults in self.pots:
            for lab in labels:
                if lab in cut:
                    cut[lab] = max(cut[lab], c)
        for params, c, soft, labels, group in self.bonds:
            for lab in labels:
                if lab in cut:
                    cut[lab] = max(cut[lab], c)
        return np.array([cut[lab] + skin for lab in self.labels])

    def ensure_ready(self):
        self.require_atoms()
        if not self.pushed:
            self.push_model()
        if self.lists is None and self.external_lists:
            self.generate(self.model_cutoffs())
        if self.lists is None:
            raise RuntimeError("pysic_interface: neighbor lists have not been generated")

    def generate(self, cutoffs):
        self.eng.generate_neighbor_lists(np.asarray(cutoffs, dtype=np.float64))
        counts, idx, off = self.eng.neighbor_lists()
        start = np.zeros(len(counts) + 1, dtype=np.int64)
        np.cumsum(counts, out=start[1:])
        self.lists = (counts, start, idx, off)


_core = _Core()


class _PysicInterface:
    """The functions of PyInterface.f90 that pysic calls, each with the
    argument order and return values of its f2py wrapper."""

    # -- lifecycle and MPI (PyInterface.f90:51-195, 1023) -------------------
    def start_rng(self, seed):
        _core.seed = int(seed)   # pysic's random numbers come from numpy here

    def start_mpi(self):
        pass

    def finish_mpi(self):
        pass

    def sync_mpi(self):
        pass

    def distribute_mpi(self, n_atoms):
        pass   # one GPU holds every atom

    def get_cpu_id(self):
        return 0

    def get_number_of_cpus(self):
        return 1

    def get_mpi_list_of_atoms(self, n_atoms):
        return np.ones(int(n_atoms), dtype=bool)

    def start_potentials(self):
        pass   # the registry is static in libpysic_hip.so

    def start_bond_order_factors(self):
        pass

    def release(self):
        _core.reset()

    # -- system (PyInterface.f90:208-311) ------------------------------------
    def create_cell(self, vectors, inverse, periodicity):
        """vectors: (3,3) with the cell vectors as columns (calculator.py
        passes cell.T); inverse: M^-1 of that matrix (inv(cell).T)."""
        v = np.asarray(vectors, dtype=np.float64)
        _core.cell = np.ascontiguousarray(v.T)
        _core.pbc = [bool(p) for p in periodicity]
        if _core.eng is not None:
            _core.eng.create_cell(_core.cell, _core.pbc)
            _core.lists = None

    def get_cell_vectors(self):
        return np.asfortranarray(_core.cell.T)

    def create_atoms(self, masses, charges, positions, momenta, tags, elements, n_atoms=None):
        """positions, momenta: (3, N); elements: (2, N) label codes.  The
        interactions and the cell stay, as in the reference's core."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).T)
        mom = np.ascontiguousarray(np.asarray(momenta, dtype=np.float64).T)
        el = np.asarray(elements).reshape(2, -1)
        n = pos.shape[0]
        labels = [_label(el[:, i]) for i in range(n)]
        names = []
        for lab in labels:
            if lab not in names:
                names.append(lab)
        if len(names) > 4:
            raise ValueError("pysic_interface: the engine takes at most 4 elements")
        _core.reset(keep_model=True)
        _core.n = n
        _core.labels = labels
        _core.species_of = {lab: k for k, lab in enumerate(names)}
        sp = np.array([_core.species_of[lab] for lab in labels], dtype=np.int32)
        _core.eng = _engine.Engine(0)
        _core.eng.create_atoms(sp, pos, masses=np.asarray(masses, dtype=np.float64),
                               charges=np.asarray(charges, dtype=np.float64), momenta=mom)
        if _core.cell is not None:
            _core.eng.create_cell(_core.cell, _core.pbc)
        if _core.ewald is not None and len(_core.ewald[5]) == n:
            self._apply_ewald()

    def _apply_ewald(self):
        rc, kc, K, sigma, eps, scaler = _core.ewald
        _core.eng.set_ewald(rc, kc, K, sigma, eps, scaler=scaler)

    def update_atom_coordinates(self, positions, momenta, n_atoms=None):
        _core.require_atoms()
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).T)
        mom = np.ascontiguousarray(np.asarray(momenta, dtype=np.float64).T)
        _core.eng.update_coordinates(pos, mom)

    def update_atom_charges(self, charges, n_atoms=None):
        _core.require_atoms()
        _core.eng.update_charges(np.asarray(charges, dtype=np.float64))

    def get_number_of_atoms(self):
        return _core.n

    def examine_atoms(self):
        print("pysic_interface: %d atoms, elements %s" % (_core.n, sorted(_core.species_of)))

    def examine_cell(self):
        print("pysic_interface: cell (rows)\n%s\nperiodic %s" % (_core.cell, _core.pbc))

    def examine_potentials(self):
        print("pysic_interface: %d potential rows" % len(_core.pot_rows))

    def examine_bond_order_factors(self):
        print("pysic_interface: %d bond-order factor rows" % len(_core.bond_rows))

    # -- interactions (PyInterface.f90:370-559, 1098) ------------------------
    def allocate_potentials(self, n_pots):
        _core.pots = []
        _core.mults = []
        _core.pushed = False

    def allocate_bond_order_factors(self, n_bonds):
        _core.bonds = []
        _core.pushed = False

    def allocate_bond_order_storage(self, n_atoms, n_groups, n_factors):
        pass   # the engine sizes its own storage

    def add_potential(self, pot_name, parameters, cutoff, smooth_cut, elements, tags, indices,
                      orig_elements, orig_tags, orig_indices, pot_index, is_multiplier,
                      n_targets=None, n_params=None):
        """One target permutation of a potential, or a multiplier that the
        next potentials take until clear_potential_multipliers
        (Core.f90:880-960, 347-355).  Returns the reference's success flag."""
        name = pot_name.decode() if isinstance(pot_name, bytes) else str(pot_name)
        if name not in POTENTIAL_TYPES:
            return False
        nt = POTENTIAL_TYPES[name][0]
        el = np.asarray(elements).reshape(2, -1)
        orig = np.asarray(orig_elements).reshape(2, -1)
        params = [float(x) for x in np.ravel(parameters)]
        if el.shape[1] != nt or len(params) != len(POTENTIAL_TYPES[name][1]):
            return False
        # tags and indices as targets are not supported (-9: unused)
        if any(int(x) != -9 for x in np.ravel(tags)) or \
                any(int(x) != -9 for x in np.ravel(indices)):
            return False
        labels = [_label(el[:, k]) for k in range(nt)]
        olabels = [_label(orig[:, k]) for k in range(nt)]
        if is_multiplier:
            _core.mults.append((name, params, olabels))
            return True
        _core.pots.append((name, params, nt, labels, olabels, float(cutoff), float(smooth_cut),
                           int(pot_index), list(_core.mults)))
        _core.pushed = False
        return True

    def clear_potential_multipliers(self):
        _core.mults = []

    def add_bond_order_factor(self, bond_name, parameters, param_split, cutoff, smooth_cut,
                              elements, orig_elements, group_index, n_targets=None,
                              n_params=None, n_split=None):
        name = bond_name.decode() if isinstance(bond_name, bytes) else str(bond_name)
        split = [int(x) for x in np.ravel(param_split)]
        params = [float(x) for x in np.ravel(parameters)]
        if name != "tersoff" or split != [2, 5] or len(params) != 7:
            return False
        el = np.asarray(elements).reshape(2, -1)
        if el.shape[1] != 3:
            return False
        _core.bonds.append((params, float(cutoff), float(smooth_cut),
                            [_label(el[:, k]) for k in range(3)], int(group_index)))
        _core.pushed = False
        return True

    def create_potential_list(self):
        _core.push_model()

    def create_bond_order_factor_list(self):
        _core.push_model()

    def set_ewald_parameters(self, real_cut, k_radius, reciprocal_cut, sigma, epsilon, scaler,
                             n_atoms=None):
        _core.require_atoms()
        _core.ewald = (float(real_cut), float(k_radius),
                       [int(k) for k in np.ravel(reciprocal_cut)], float(sigma), float(epsilon),
                       np.array(scaler, dtype=np.float64))
        self._apply_ewald()

    def get_ewald_energy(self, real_cut, k_cut, reciprocal_cut, sigma, epsilon):
        _core.ensure_ready()
        return _core.eng.ewald_energy(real_cut, k_cut, np.ravel(reciprocal_cut), sigma, epsilon)

    # -- neighbor lists (PyInterface.f90:528, 1062-1092) ---------------------
    def generate_neighbor_lists(self, cutoffs, n_atoms=None):
        _core.require_atoms()
        _core.external_lists = False
        _core.generate(cutoffs)

    def get_number_of_neighbors_of_atom(self, atom_index):
        """atom_index 0-based (the wrapper adds one, PyInterface.f90:1075-1080)."""
        counts = _core.lists[0]
        return int(counts[int(atom_index)])

    def get_neighbor_list_of_atom(self, atom_index, n_neighbors):
        """(neighbors 0-based, offsets (3, n)) of atom atom_index (0-based)."""
        counts, start, idx, off = _core.lists
        i = int(atom_index)
        a, b = int(start[i]), int(start[i + 1])
        n = int(n_neighbors)
        return idx[a:a + n].copy(), np.asfortranarray(off[a:a + n].T)

    def create_neighbor_list(self, atom_index, neighbors, offsets, n_nbs=None):
        """A list pysic built elsewhere (ASE, calculator.py:1582-1590).  The
        engine builds its own at the model's cutoffs plus the skin instead:
        its lists contain every pair inside every interaction cutoff, so the
        results are the same."""
        _core.external_lists = True
        _core.lists = None

    # -- evaluation (PyInterface.f90:579-691) --------------------------------
    def calculate_energy(self):
        _core.ensure_ready()
        return _core.eng.energy()

    def calculate_forces(self, n_atoms):
        """(forces (3, N), stress (6))."""
        _core.ensure_ready()
        f, s = _core.eng.forces()
        return np.asfortranarray(f.T), s

    def calculate_electronegativities(self, n_atoms):
        _core.ensure_ready()
        return _core.eng.electronegativities()

    def calculate_bond_order_factors(self, n_atoms, group_index):
        _core.ensure_ready()
        return _core.eng.bond_order_factors(group_index)

    def calculate_bond_order_gradients(self, n_atoms, group_index, atom_index):
        """(3, N); atom_index 1-based as bondorder.py:518 passes it."""
        _core.ensure_ready()
        return np.asfortranarray(_core.eng.bond_order_gradients(group_index,
                                                                int(atom_index) - 1).T)

    def calculate_bond_order_gradients_of_factor(self, n_atoms, group_index, atom_index):
        _core.ensure_ready()
        return np.asfortranarray(_core.eng.bond_order_gradients(group_index,
                                                                int(atom_index) - 1,
                                                                of_factor=True).T)

    # -- descriptor registry (PyInterface.f90:706-1011) ----------------------
    def number_of_potentials(self):
        return _engine.load().pysic_number_of_potentials()

    def number_of_bond_order_factors(self):
        return _engine.load().pysic_number_of_bond_order_factors()

    def is_potential(self, string):
        return bool(_engine.load().pysic_is_potential(str(string).strip().encode()))

    def is_bond_order_factor(self, string):
        return bool(_engine.load().pysic_is_bond_order_factor(str(string).strip().encode()))

    def list_valid_potentials(self, n_pots):
        L = _engine.load()
        return np.asfortranarray(np.array([_codes(_s(L.pysic_potential_name(k)), 11)
                                           for k in range(int(n_pots))]).T)

    def list_valid_bond_order_factors(self, n_bonds):
        L = _engine.load()
        return np.asfortranarray(np.array([_codes(_s(L.pysic_bond_order_factor_name(k)), 11)
                                           for k in range(int(n_bonds))]).T)

    def number_of_targets_of_potential(self, pot_name):
        return _engine.load().pysic_number_of_targets_of_potential(pot_name.encode())

    def number_of_targets_of_bond_order_factor(self, bond_name):
        return _engine.load().pysic_number_of_targets_of_bond_order_factor(bond_name.encode())

    def level_of_bond_order_factor(self, bond_name):
        return _engine.load().pysic_level_of_bond_order_factor(bond_name.encode())

    def number_of_parameters_of_potential(self, pot_name):
        return _engine.load().pysic_number_of_parameters_of_potential(pot_name.encode())

    def number_of_parameters_of_bond_order_factor(self, bond_name, n_targets):
        return _engine.load().pysic_number_of_parameters_of_bond_order_factor(
            bond_name.encode(), int(n_targets))

    @staticmethod
    def _table(get, count, width):
        out = np.zeros((12, width), dtype=np.int32)
        for k in range(count):
            out[k] = _codes(_s(get(k)), width)
        return np.asfortranarray(out.T)

    def names_of_parameters_of_potential(self, pot_name):
        L, nb = _engine.load(), pot_name.encode()
        return self._table(lambda k: L.pysic_parameter_name_of_potential(nb, k),
                           max(L.pysic_number_of_parameters_of_potential(nb), 0), 10)

    def names_of_parameters_of_bond_order_factor(self, bond_name, n_targets):
        L, nb, t = _engine.load(), bond_name.encode(), int(n_targets)
        return self._table(lambda k: L.pysic_parameter_name_of_bond_order_factor(nb, t, k),
                           max(L.pysic_number_of_parameters_of_bond_order_factor(nb, t), 0), 10)

    def descriptions_of_parameters_of_potential(self, pot_name):
        L, nb = _engine.load(), pot_name.encode()
        return self._table(lambda k: L.pysic_parameter_description_of_potential(nb, k),
                           max(L.pysic_number_of_parameters_of_potential(nb), 0), 100)

    def descriptions_of_parameters_of_bond_order_factor(self, bond_name, n_targets):
        L, nb, t = _engine.load(), bond_name.encode(), int(n_targets)
        return self._table(
            lambda k: L.pysic_parameter_description_of_bond_order_factor(nb, t, k),
            max(L.pysic_number_of_parameters_of_bond_order_factor(nb, t), 0), 100)

    def description_of_potential(self, pot_name):
        return (_s(_engine.load().pysic_description_of_potential(pot_name.encode())) or
                "").ljust(500)

    def description_of_bond_order_factor(self, bond_name):
        return (_s(_engine.load().pysic_description_of_bond_order_factor(bond_name.encode())) or
                "").ljust(500)


pysic_interface = _PysicInterface()
