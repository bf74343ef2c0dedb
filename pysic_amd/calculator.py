"""The Pysic calculator and FastNeighborList on the MI355X engine.

Mirrors pysic/calculator.py (Pysic :192-1832, FastNeighborList :29-189): the
ASE-calculator methods get_potential_energy / get_forces / get_stress /
get_electronegativities with the same caching and units, over one device
engine per calculator (the reference keeps one process-global Fortran core,
Core.f90:19-97; here every calculator owns its own handle).
"""
import math
import warnings

import numpy as np

from .core import CoreMirror
from .coulomb import CoulombSummation
from .engine import Engine
from .interactions import build_descriptors, read_table, table_ids


class MissingAtomsError(Exception):
    pass


class LockedCoreError(Exception):
    pass


def _symbols(atoms):
    return list(atoms.get_chemical_symbols())


def _warn_charge(q):
    """calculator.py:1545-1547"""
    total = float(np.sum(q))
    if total > 0.001 or total < -0.001:
        warnings.warn("There is non-zero total charge in the system: {tc}".format(tc=total))


def _warn_potentials(potentials):
    """calculator.py:1201-1207"""
    for pot in potentials or []:
        if pot.get_number_of_targets() > 1 and pot.get_cutoff() < 0.01:
            warnings.warn("Potential with zero cutoff present: \n\n" + str(pot))
        if pot.get_symbols() is None:
            warnings.warn("Potential with no targets present: \n\n" + str(pot))


def _charges(atoms):
    try:
        return np.array(atoms.get_initial_charges(), dtype=np.float64)
    except AttributeError:
        return np.array(atoms.get_charges(), dtype=np.float64)


class FastNeighborList:
    """Neighbor lists built by the engine's cell list (calculator.py:29-189).

    cutoffs are per atom; the search radius is cutoff + skin, and the lists
    are rebuilt when an atom has moved more than skin since the last build
    (the ASE NeighborList rule the reference inherits).
    """

    neighbor_marginal = 0.5

    def __init__(self, cutoffs, skin=None):
        self.skin = FastNeighborList.neighbor_marginal if skin is None else float(skin)
        self.cutoffs = np.asarray(cutoffs, dtype=np.float64) + self.skin
        self.positions = None
        self.nupdates = 0
        self._counts = None
        self._start = None
        self._idx = None
        self._off = None
        self._engine = None

    def _attach(self, engine):
        self._engine = engine

    def update(self, atoms):
        pos = atoms.get_positions()
        if (self.positions is None or len(pos) != len(self.positions)
                or ((pos - self.positions) ** 2).sum(1).max() > self.skin ** 2):
            self.build(atoms)
            return True
        return False

    def build(self, atoms):
        if self._engine is None:
            raise MissingAtomsError("neighbor list is not attached to a calculator")
        self.positions = atoms.get_positions().copy()
        self._engine.generate_neighbor_lists(self.cutoffs)
        self._counts = None
        self.nupdates += 1

    def _fetch(self):
        if self._counts is None:
            c, i, o = self._engine.neighbor_lists()
            self._counts = c
            self._start = np.concatenate([[0], np.cumsum(c)])
            self._idx = i
            self._off = o

    def get_neighbors(self, index, atoms=None, sort=False):
        self._fetch()
        a, b = self._start[index], self._start[index + 1]
        nb, off = self._idx[a:b].copy(), self._off[a:b].copy()
        if sort and atoms is not None and len(nb):
            d = self.get_neighbor_distances(index, atoms)
            o = np.argsort(d, kind="stable")
            nb, off = nb[o], off[o]
        return nb, off

    def get_neighbor_separations(self, index, atoms, sort=False):
        nb, off = self.get_neighbors(index)
        pos = atoms.get_positions()
        sep = pos[nb] + off @ atoms.get_cell() - pos[index]
        if sort:
            sep = sep[np.argsort(np.sqrt((sep * sep).sum(1)), kind="stable")]
        return sep

    def get_neighbor_distances(self, index, atoms, sort=False):
        sep = self.get_neighbor_separations(index, atoms)
        d = np.sqrt((sep * sep).sum(1))
        return np.sort(d) if sort else d


class _InitMode:
    """The full-initialisation switch.  The reference keeps it in an
    attribute named like the method that sets it (calculator.py:214, 303-312);
    this object reads as the flag and calls as the setter, so both
    `if calc.force_core_initialization:` and
    `calc.force_core_initialization(True)` work."""

    def __init__(self, value):
        self.value = bool(value)

    def __call__(self, new_mode):
        self.value = bool(new_mode)

    def __bool__(self):
        return self.value

    def __eq__(self, other):
        return bool(other) == self.value

    def __repr__(self):
        return str(self.value)


class Pysic:
    """ASE calculator over the MI355X engine (calculator.py:192-1832).

    Each calculator owns an engine and a CoreMirror of what it holds;
    Pysic.core is the mirror of the calculator that synchronised last."""

    core = CoreMirror()

    def __init__(self, atoms=None, potentials=None, charge_relaxation=None, coulomb=None,
                 full_initialization=False, device=0):
        self.engine = Engine(device)
        self.structure = None
        self.potentials = None
        self.charge_relaxation = None
        self.coulomb = None
        self.neighbor_list = None
        self.force_core_initialization = _InitMode(full_initialization)
        self.extra_calculators = []
        self.forces = self.stress = self.energy = self.electronegativities = None
        self._mirror = CoreMirror()   # what this calculator's engine holds
        self._mirror.engine = self.engine
        self.set_atoms(atoms)
        self.set_potentials(potentials)
        self.set_charge_relaxation(charge_relaxation)
        self.set_coulomb_summation(coulomb)

    def __eq__(self, other):
        """Same structure (and charges), neighbor list, potentials and extra
        calculators (calculator.py:260-287)."""
        try:
            return (isinstance(other, Pysic)
                    and (self.structure is other.structure or
                         (self.structure is not None and other.structure is not None and
                          self._same_structure(other.structure)))
                    and self.potentials == other.potentials
                    and self.extra_calculators == other.extra_calculators)
        except Exception:
            return False

    def __ne__(self, other):
        return not self.__eq__(other)

    def __repr__(self):
        return "Pysic(atoms={atoms},potentials={pots},full_initialization={init})".format(
            atoms=str(self.structure), pots=str(self.potentials),
            init=str(self.force_core_initialization))

    def core_initialization_is_forced(self):
        """True if the engine is fully re-initialised before every
        calculation (calculator.py:297-300)."""
        return bool(self.force_core_initialization)

    # --- configuration (calculator.py:537-820) -------------------------------
    def _invalidate(self):
        self.forces = self.stress = self.energy = self.electronegativities = None

    def set_atoms(self, atoms=None):
        if atoms is None:
            return
        if self.structure is not None and self._same_structure(atoms):
            return
        self.structure = atoms.copy() if hasattr(atoms, "copy") else atoms
        self._invalidate()

    def _same_structure(self, atoms):
        s = self.structure
        try:
            return (len(s) == len(atoms) and _symbols(s) == _symbols(atoms)
                    and np.array_equal(s.get_positions(), atoms.get_positions())
                    and np.array_equal(s.get_cell(), atoms.get_cell())
                    and np.array_equal(s.get_pbc(), atoms.get_pbc())
                    and np.array_equal(_charges(s), _charges(atoms))
                    and np.array_equal(s.get_momenta(), atoms.get_momenta()))
        except Exception:
            return False

    def get_atoms(self):
        return self.structure

    def set_potentials(self, potentials):
        self.potentials = None if potentials is None else list(potentials)
        self._invalidate()

    def add_potential(self, potential):
        if self.potentials is None:
            self.potentials = []
        self.potentials.append(potential)
        self._invalidate()

    def get_potentials(self):
        return self.potentials

    def set_coulomb_summation(self, coulomb):
        self.coulomb = coulomb
        self._invalidate()

    def get_coulomb_summation(self):
        return self.coulomb

    def set_charge_relaxation(self, charge_relaxation):
        self.charge_relaxation = charge_relaxation
        if charge_relaxation is not None:
            charge_relaxation.set_calculator(self)
        self._invalidate()

    def get_charge_relaxation(self):
        return self.charge_relaxation

    def add_calculator(self, calc):
        self.extra_calculators.append(calc)

    def remove_calculator(self, calc):
        """calculator.py:703-711"""
        self.extra_calculators.remove(calc)

    def remove_potential(self, potential):
        """calculator.py:714-735: the potential list is replaced, so cutoffs
        and the engine's model are recomputed."""
        new_pots = list(self.get_potentials() or [])
        new_pots.remove(potential)
        self.set_potentials(new_pots)

    def get_calculators(self):
        return self.extra_calculators

    def get_neighbor_lists(self):
        return self.neighbor_list

    def get_neighbor_list(self):
        """calculator.py:854-857"""
        return self.neighbor_list

    def get_charges(self, system=None):
        """The initial charges of system, or of the calculator's structure
        (calculator.py:1515-1529)."""
        return _charges(system if system is not None else self.structure)

    def set_cutoffs(self, cutoffs):
        """calculator.py:1681-1689"""
        self.saved_cutoffs = None if cutoffs is None else list(cutoffs)

    def neighbor_lists_expanded(self, cutoffs):
        """True if any atom's cutoff grew since set_cutoffs (or nothing was
        saved): the lists must then be rebuilt (calculator.py:1692-1718)."""
        saved = getattr(self, "saved_cutoffs", None)
        if saved is None or cutoffs is None or len(saved) != len(cutoffs):
            return True
        return any(old < new for old, new in zip(saved, cutoffs))

    def get_individual_cutoffs(self, scaler=1.0):
        """Per-atom maximum cutoff over the potentials, bond factors and the
        Coulomb real-space cutoff that target the atom (calculator.py:860-915)."""
        if self.structure is None:
            return None
        base = self.coulomb.get_realspace_cutoff() if self.coulomb is not None else 0.0
        cuts = []
        for sym in _symbols(self.structure):
            c = base
            for pot in self.potentials or []:
                if sym in pot.get_different_symbols():
                    c = max(c, pot.get_cutoff())
                coord = pot.get_coordinator()
                if coord is not None:
                    for bond in coord.get_bond_order_parameters():
                        if any(sym == t[0] for t in (bond.get_symbols() or [])) or \
                                any(sym in t for t in (bond.get_symbols() or [])):
                            c = max(c, bond.get_cutoff())
            cuts.append(c * scaler)
        return cuts

    def create_neighbor_lists(self, cutoffs=None, marginal=None):
        if cutoffs is None:
            cutoffs = self.get_individual_cutoffs(1.0)
        self.neighbor_list = FastNeighborList(cutoffs, skin=marginal)
        self.neighbor_list._attach(self.engine)

    # --- core synchronisation (calculator.py:998-1046, 1450-1637) -----------
    def _species_map(self):
        names = sorted(set(_symbols(self.structure)))
        if len(names) > 4:
            raise LockedCoreError("at most 4 elements per system")
        return {s: k for k, s in enumerate(names)}

    def _kcells(self):
        kc = self.coulomb.parameters["k_cutoff"]
        rec = 2.0 * math.pi * np.linalg.inv(self.structure.get_cell()).T
        vol = np.dot(rec[0], np.cross(rec[1], rec[2]))
        return [int(kc * np.linalg.norm(np.cross(rec[(a + 1) % 3], rec[(a + 2) % 3])) / vol + 0.5)
                for a in range(3)]

    def _full_init(self):
        s = self.structure
        if s is None:
            raise MissingAtomsError("no atoms assigned to the calculator")
        spmap = self._species_map()
        species = np.array([spmap[x] for x in _symbols(s)], dtype=np.int32)
        eng = self.engine
        eng.create_atoms(species, s.get_positions(), masses=s.get_masses(), charges=_charges(s),
                         momenta=s.get_momenta())
        eng.create_cell(s.get_cell(), s.get_pbc())
        _warn_potentials(self.potentials)
        pd, bd, _ = build_descriptors(self.potentials or [], spmap)
        for tid in table_ids(self.potentials):
            eng.set_potential_table(tid, read_table(tid))
        eng.set_potentials(pd)
        eng.set_bond_order_factors(bd)
        if self.coulomb is not None:
            p = self.coulomb.parameters
            sc = self.coulomb.get_scaling_factors()
            if sc is not None and len(sc) != len(s):
                raise ValueError("length of the scaling factor vector does not match the atoms")
            eng.set_ewald(p["real_cutoff"], p["k_cutoff"], self._kcells(), p["sigma"], p["epsilon"],
                          sc)
        else:
            eng.ewald_off()
        self.create_neighbor_lists()
        self.neighbor_list.build(s)
        m = self._mirror
        m.set_atoms(s)
        m.set_potentials(self.potentials)
        m.set_coulomb(self.coulomb)
        m.set_neighbor_lists(self.neighbor_list)
        m.potential_lists_ready = m.bond_order_factor_lists_ready = True
        m.cutoffs = list(self.neighbor_list.cutoffs)
        Pysic.core = m

    def set_core(self):
        """Brings the engine up to date with the calculator
        (calculator.py:998-1046): from scratch when forced, on the first call,
        or when the atom count, elements or cell changed; otherwise it
        pushes only the coordinates, charges, potentials or Coulomb
        summation that differ from the mirror."""
        s = self.structure
        if s is None:
            raise MissingAtomsError("no atoms assigned to the calculator")
        m = self._mirror
        if (self.force_core_initialization or m.structure is None or len(m.structure) != len(s)
                or _symbols(m.structure) != _symbols(s) or not m.cell_ready(s)
                or not m.potentials_ready(self.potentials)
                or (self.coulomb is not None and not m.coulomb_summation_ready(self.coulomb))
                or (self.coulomb is None and m.coulomb is not None)):
            # a new model or cell changes the species tables, the cutoffs
            # and the cell lists: the engine is set up again
            self._full_init()
            return
        if not m.atoms_ready(s):
            self.update_core_coordinates()
        if not m.charges_ready(s):
            self.update_core_charges()
        if not m.neighbor_lists_ready(self.neighbor_list):
            self.update_core_neighbor_lists()
        Pysic.core = m

    # The reference pushes each part of the calculator's state into its core
    # with one call per part (calculator.py:998-1482); set_core() does what
    # is needed before every calculation, and these entry points do one part
    # each on demand.
    def initialize_fortran_core(self):
        """Everything, from scratch (calculator.py:1595-1637)."""
        self._full_init()

    def _core_ready(self):
        if self._mirror.structure is None:
            self._full_init()
            return False
        return True

    def update_core_coordinates(self):
        """Positions and momenta (calculator.py:1485-1513)."""
        if self._core_ready():
            s = self.structure
            self.engine.update_coordinates(s.get_positions(), s.get_momenta())
            self._mirror.set_atomic_positions(s)
            self._mirror.set_atomic_momenta(s)
            self.neighbor_list.update(s)

    def update_core_charges(self):
        """Charges (calculator.py:1532-1553), with its warning for a charged
        system."""
        if self._core_ready():
            q = _charges(self.structure)
            _warn_charge(q)
            self.engine.update_charges(q)
            self._mirror.set_charges(q)
            self._invalidate()

    def update_core_supercell(self):
        """The cell (calculator.py:1555-1565): the cell lists depend on it,
        so the engine is set up again."""
        self._full_init()

    def update_core_potentials(self):
        """Potentials, tables and bond-order factors (calculator.py:1063-1446)."""
        if self._core_ready():
            spmap = self._species_map()
            _warn_potentials(self.potentials)
            pd, bd, _ = build_descriptors(self.potentials or [], spmap)
            for tid in table_ids(self.potentials):
                self.engine.set_potential_table(tid, read_table(tid))
            self.engine.set_potentials(pd)
            self.engine.set_bond_order_factors(bd)
            self._mirror.set_potentials(self.potentials)
            new_cut = self.get_individual_cutoffs(1.0)
            nl = self.neighbor_list
            if nl is None or not np.array_equal(nl.cutoffs, np.asarray(new_cut) + nl.skin):
                # the list radius follows the cutoffs (calculator.py:1431-1446)
                self.create_neighbor_lists(new_cut)
                self.neighbor_list.build(self.structure)
                self._mirror.set_neighbor_lists(self.neighbor_list)
            self._mirror.potential_lists_ready = True
            self._invalidate()

    def update_core_potential_lists(self):
        """The per-atom potential lists (calculator.py:1048-1061): the engine
        compiles its species tables with the potentials, so nothing is left
        to do once update_core_potentials has run."""
        if self._core_ready():
            self._mirror.potential_lists_ready = True
            self._mirror.bond_order_factor_lists_ready = True

    def update_core_coulomb(self):
        """Ewald parameters and scaling factors (calculator.py:1450-1483)."""
        if self._core_ready():
            if self.coulomb is not None:
                p = self.coulomb.parameters
                sc = self.coulomb.get_scaling_factors()
                self.engine.set_ewald(p["real_cutoff"], p["k_cutoff"], self._kcells(),
                                      p["sigma"], p["epsilon"], sc)
            else:
                self.engine.ewald_off()
            self._mirror.set_coulomb(self.coulomb)
            self._invalidate()

    def update_core_neighbor_lists(self):
        """Rebuild the neighbor lists (calculator.py:1567-1593)."""
        if self._core_ready():
            self.neighbor_list.build(self.structure)
            self._mirror.set_neighbor_lists(self.neighbor_list)

    # --- calculations (calculator.py:315-534, 918-990) ---------------------
    def calculation_required(self, atoms=None, quantities=("forces", "energy", "stress",
                                                           "electronegativities")):
        """True if a quantity is not cached, or if the structure, charges,
        cell or potentials differ from what the engine holds
        (calculator.py:315-372): a calculator whose atoms were changed in
        place recomputes."""
        if isinstance(quantities, str):
            quantities = [quantities]
        table = {"energy": self.energy, "forces": self.forces, "stress": self.stress,
                 "electronegativities": self.electronegativities}
        if any(table.get(q, 0) is None for q in quantities):
            return True
        m, s = self._mirror, self.structure
        if s is None:
            return True
        return not (m.atoms_ready(s) and m.charges_ready(s) and m.cell_ready(s)
                    and m.potentials_ready(self.potentials))

    def relax_charges(self, charges, rates, n_steps, timestep, inertia, friction, tolerance):
        self.set_core()
        out = self.engine.relax_charges(charges, rates, n_steps, timestep, inertia, friction,
                                        tolerance)
        self._mirror.set_charges(out[0].copy())   # the engine holds the relaxed charges
        return out

    def _relax_if_needed(self, skip):
        if self.charge_relaxation is not None and not skip:
            self.charge_relaxation.charge_relaxation()
            self._invalidate()

    def _empty(self):
        # a structure without atoms: the reference's loops over zero atoms
        # give zero energy and stress and empty forces and electronegativities;
        # nothing is sent to the engine
        return self.structure is not None and len(self.structure) == 0

    def calculate_forces(self, skip_charge_relaxation=False):
        if self._empty():
            self.forces, self.stress = np.zeros((0, 3)), np.zeros(6)
            return
        self.set_core()
        self._relax_if_needed(skip_charge_relaxation)
        self.set_core()
        f, st = self.engine.forces()
        for calc in self.extra_calculators:
            f = f + calc.get_forces(self.structure)
            st = st + calc.get_stress(self.structure)
        self.forces, self.stress = f, st

    def calculate_energy(self, skip_charge_relaxation=False):
        if self._empty():
            self.energy = 0.0
            return
        self.set_core()
        self._relax_if_needed(skip_charge_relaxation)
        self.set_core()
        e = self.engine.energy()
        for calc in self.extra_calculators:
            e += calc.get_potential_energy(self.structure)
        self.energy = e

    def calculate_stress(self, skip_charge_relaxation=False):
        self.calculate_forces(skip_charge_relaxation)

    def calculate_electronegativities(self):
        if self._empty():
            self.electronegativities = np.zeros(0)
            return
        self.set_core()
        self.electronegativities = self.engine.electronegativities()

    def get_forces(self, atoms=None, skip_charge_relaxation=False):
        self.set_atoms(atoms)
        if self.calculation_required(atoms, "forces"):
            self.calculate_forces(skip_charge_relaxation)
        return np.copy(self.forces)

    def get_potential_energy(self, atoms=None, force_consistent=False,
                             skip_charge_relaxation=False):
        self.set_atoms(atoms)
        if self.calculation_required(atoms, "energy"):
            self.calculate_energy(skip_charge_relaxation)
        return self.energy

    def get_stress(self, atoms=None, skip_charge_relaxation=False):
        """-(sum m v (x) v + sum r (x) f) / V in the order xx yy zz yz xz xy
        (calculator.py:468-534)."""
        self.set_atoms(atoms)
        if self.calculation_required(atoms, "stress"):
            self.calculate_stress(skip_charge_relaxation)
        s = self.structure
        p = s.get_momenta()
        m = s.get_masses()
        v = p / m[:, None]
        kin = np.array([p[:, 0] @ v[:, 0], p[:, 1] @ v[:, 1], p[:, 2] @ v[:, 2],
                        p[:, 1] @ v[:, 2], p[:, 0] @ v[:, 2], p[:, 0] @ v[:, 1]])
        return np.copy(-(kin + self.stress) / s.get_volume())

    def get_electronegativities(self, atoms=None):
        self.set_atoms(atoms)
        if self.calculation_required(atoms, "electronegativities"):
            self.calculate_electronegativities()
        return np.copy(self.electronegativities)

    def get_electronegativity_differences(self, atoms=None):
        x = self.get_electronegativities(atoms)
        return x - x.sum() / len(x)

    # --- finite-difference debug helpers (calculator.py:1639-1832) ----------
    def get_numerical_energy_gradient(self, atom_index, shift=0.0001, atoms=None):
        s = (atoms if atoms is not None else self.structure).copy()
        out = np.zeros(3)
        for c in range(3):
            e = []
            for sg in (1, -1):
                t = s.copy()
                p = t.get_positions()
                p[atom_index, c] += sg * shift
                t.set_positions(p)
                e.append(self.get_potential_energy(t, skip_charge_relaxation=True))
            out[c] = -(e[0] - e[1]) / (2 * shift)
        self.set_atoms(s)
        return out

    def get_numerical_electronegativity(self, atom_index, shift=0.001, atoms=None):
        s = (atoms if atoms is not None else self.structure).copy()
        e = []
        for sg in (1, -1):
            t = s.copy()
            q = _charges(t)
            q[atom_index] += sg * shift
            t.set_initial_charges(q)
            e.append(self.get_potential_energy(t, skip_charge_relaxation=True))
        self.set_atoms(s)
        return -(e[0] - e[1]) / (2 * shift)

    def get_numerical_bond_order_gradient(self, coordinator, atom_index, moved_index,
                                          shift=0.001, atoms=None):
        """Central differences of atom atom_index's bond-order factor in
        coordinator's group with respect to the position of atom
        moved_index (calculator.py:1721-1778); a debugging aid for the
        analytic gradients.  The calculator's structure is restored."""
        system = (atoms if atoms is not None else self.structure).copy()
        orig = system.copy()
        out = []
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                t = system.copy()
                p = t.get_positions()
                p[moved_index, c] += sg * shift
                t.set_positions(p)
                self.set_atoms(t)
                self.set_core()
                coordinator.calculate_bond_order_factors()
                vals.append(coordinator.get_bond_order_factors()[atom_index])
            out.append((vals[0] - vals[1]) / (2.0 * shift))
        self.set_atoms(orig)
        self.set_core()
        return out

    # --- device-resident dynamics -----------------------------------------
    def run_md(self, steps, dt, rebuild=1, energies=False):
        """Velocity Verlet on the device (ASE VelocityVerlet semantics);
        updates the calculator's structure with the final state."""
        self.set_core()
        e = self.engine.md_run(steps, dt, rebuild=rebuild, energies=energies)
        pos, mom = self.engine.state()
        self.structure.set_positions(pos)
        self.structure.set_momenta(mom)
        self._mirror.set_atomic_positions(self.structure)
        self._mirror.set_atomic_momenta(self.structure)
        self.neighbor_list.positions = pos.copy()
        self._invalidate()
        return e
