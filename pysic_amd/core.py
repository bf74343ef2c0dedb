"""Queries about the interaction types the engine knows, and the mirror of
what a calculator's engine holds (pysic/core.py:44-762).

The query functions ask the descriptor registry of libpysic_hip.so
(include/pysic_registry.h), as the reference's ask its Fortran core through
the f2py module, so their answers follow the engine: the potential types it
evaluates and the 'tersoff' bond-order factor.  Charge relaxation modes and
Coulomb summation modes come from the ChargeRelaxation and CoulombSummation
classes, as in the reference.

CoreMirror keeps deep copies of what was last pushed to an engine; the
calculator compares its own state with it before every calculation and
pushes only what changed (calculator.py:998-1046).
"""
import copy

import numpy as np

from .coulomb import ChargeRelaxation, CoulombSummation


def _lib():
    from . import engine
    return engine.load()


def _s(ptr):
    return ptr.decode() if ptr is not None else None


# -- potential and bond-order factor types (core.py:44-126) -------------------
def list_valid_potentials():
    """Names of the potential types the engine evaluates (core.py:49-62)."""
    L = _lib()
    return [_s(L.pysic_potential_name(k)) for k in range(L.pysic_number_of_potentials())]


def list_potentials():
    """Same as list_valid_potentials (core.py:44-47)."""
    return list_valid_potentials()


def list_valid_bond_order_factors():
    """Names of the bond-order factor types (core.py:70-83)."""
    L = _lib()
    return [_s(L.pysic_bond_order_factor_name(k))
            for k in range(L.pysic_number_of_bond_order_factors())]


def list_bond_order_factors():
    return list_valid_bond_order_factors()


def is_valid_potential(potential_name):
    """True if the name is a potential type (core.py:96-104)."""
    return bool(_lib().pysic_is_potential(str(potential_name).encode()))


def is_potential(potential_name):
    return is_valid_potential(potential_name)


def is_valid_bond_order_factor(bond_order_name):
    """True if the name is a bond-order factor type (core.py:117-125)."""
    return bool(_lib().pysic_is_bond_order_factor(str(bond_order_name).encode()))


def is_bond_order_factor(bond_order_name):
    return is_valid_bond_order_factor(bond_order_name)


def is_valid_charge_relaxation(relaxation_name):
    """True if the name is a ChargeRelaxation mode (core.py:142-153)."""
    return relaxation_name in ChargeRelaxation.relaxation_modes


def is_charge_relaxation(relaxation_name):
    return is_valid_charge_relaxation(relaxation_name)


def is_valid_coulomb_summation(summation_name):
    """True if the name is a CoulombSummation mode (core.py:171-184)."""
    return summation_name in CoulombSummation.summation_modes


def is_coulomb_summation(summation_name):
    return is_valid_coulomb_summation(summation_name)


# -- targets and parameters (core.py:187-426) ---------------------------------
def number_of_targets(potential_name):
    """Atoms a potential or bond-order factor acts on; 0 for an unknown name
    (core.py:187-201)."""
    L, nb = _lib(), str(potential_name).encode()
    if is_potential(potential_name):
        return L.pysic_number_of_targets_of_potential(nb)
    if is_bond_order_factor(potential_name):
        return L.pysic_number_of_targets_of_bond_order_factor(nb)
    return 0


def level_of_factor(bond_order_name):
    """1 for a per-atom, 2 for a per-pair bond-order factor; 0 otherwise
    (core.py:204-215)."""
    if is_bond_order_factor(bond_order_name):
        return _lib().pysic_level_of_bond_order_factor(str(bond_order_name).encode())
    return 0


def number_of_parameters(potential_name, as_list=False):
    """Parameter count of a potential (a list of one with as_list), per
    target class [single, pair] of a bond-order factor, or of a charge
    relaxation or Coulomb summation mode (core.py:218-260)."""
    L, nb = _lib(), str(potential_name).encode()
    if is_potential(potential_name):
        n = L.pysic_number_of_parameters_of_potential(nb)
        return [n] if as_list else n
    if is_bond_order_factor(potential_name):
        return [L.pysic_number_of_parameters_of_bond_order_factor(nb, t) for t in (1, 2)]
    if is_charge_relaxation(potential_name):
        n = len(ChargeRelaxation.relaxation_parameters[potential_name])
        return [n] if as_list else n
    if is_coulomb_summation(potential_name):
        n = len(CoulombSummation.summation_parameters[potential_name])
        return [n] if as_list else n
    return 0


def names_of_parameters(potential_name):
    """Parameter names: a list for a potential or mode, a list per target
    class for a bond-order factor (core.py:263-320)."""
    L, nb = _lib(), str(potential_name).encode()
    if is_potential(potential_name):
        return [_s(L.pysic_parameter_name_of_potential(nb, k))
                for k in range(L.pysic_number_of_parameters_of_potential(nb))]
    if is_bond_order_factor(potential_name):
        return [[_s(L.pysic_parameter_name_of_bond_order_factor(nb, t, k))
                 for k in range(L.pysic_number_of_parameters_of_bond_order_factor(nb, t))]
                for t in (1, 2)]
    if is_charge_relaxation(potential_name):
        return list(ChargeRelaxation.relaxation_parameters[potential_name])
    if is_coulomb_summation(potential_name):
        return list(CoulombSummation.summation_parameters[potential_name])
    return []


def index_of_parameter(potential_name, parameter_name):
    """Index of a parameter: an int for a potential or relaxation mode, or
    [target class index, index] for a bond-order factor; None when absent
    (core.py:323-363)."""
    if is_potential(potential_name) or is_charge_relaxation(potential_name):
        names = names_of_parameters(potential_name)
        return names.index(parameter_name) if parameter_name in names else None
    if is_bond_order_factor(potential_name):
        for t, names in enumerate(names_of_parameters(potential_name)):
            if parameter_name in names:
                return [t, names.index(parameter_name)]
    return None


def descriptions_of_parameters(potential_name):
    """Short physical descriptions of the parameters, shaped as
    names_of_parameters (core.py:366-426)."""
    L, nb = _lib(), str(potential_name).encode()
    if is_potential(potential_name):
        return [_s(L.pysic_parameter_description_of_potential(nb, k))
                for k in range(L.pysic_number_of_parameters_of_potential(nb))]
    if is_bond_order_factor(potential_name):
        return [[_s(L.pysic_parameter_description_of_bond_order_factor(nb, t, k))
                 for k in range(L.pysic_number_of_parameters_of_bond_order_factor(nb, t))]
                for t in (1, 2)]
    if is_charge_relaxation(potential_name):
        return list(ChargeRelaxation.relaxation_parameter_descriptions[potential_name])
    if is_coulomb_summation(potential_name):
        return list(CoulombSummation.summation_parameter_descriptions[potential_name])
    return []


def description_of_potential(potential_name, parameter_values=None, cutoff=None,
                             elements=None, tags=None, indices=None):
    """Prints a description of a potential type, or of one potential with
    the given parameters, cutoff and targets (core.py:428-485); returns the
    text as well."""
    names = names_of_parameters(potential_name)
    notes = descriptions_of_parameters(potential_name)
    text = _s(_lib().pysic_description_of_potential(str(potential_name).encode())) or ""
    msg = "\npotential '%s'\n%d-body interaction\n%s\n\nparameters (%d):\n" % (
        potential_name, number_of_targets(potential_name), text,
        number_of_parameters(potential_name))
    if parameter_values is None:
        for name, note in zip(names, notes):
            msg += "%s : %s\n" % (name, note)
    else:
        for name, note, val in zip(names, notes, parameter_values):
            msg += "%s : %s = %s\n" % (name, note, val)
    if cutoff is not None:
        msg += "\ncutoff = %s\n" % cutoff
    for label, seq in (("symbols", elements), ("tags", tags), ("indices", indices)):
        if seq is not None:
            msg += "\ntargeted %s:%s\n" % (label, "".join(" %s " % x for x in seq))
    print(msg)
    return msg


# -- the mirror of an engine (core.py:488-762) --------------------------------
def _numbers(atoms):
    if hasattr(atoms, "get_atomic_numbers"):
        return list(atoms.get_atomic_numbers())
    return list(atoms.get_chemical_symbols())


def _charges(atoms):
    try:
        return np.asarray(atoms.get_initial_charges())
    except AttributeError:
        return np.asarray(atoms.get_charges())


class CoreMirror:
    """Deep copies of the structure, potentials, neighbor lists and Coulomb
    summation last pushed to an engine (core.py:488-762).  Each Pysic
    calculator owns an engine and a CoreMirror of it; Pysic.core is the
    mirror of the calculator that synchronised its engine last, which is
    what the reference's single core would hold."""

    def __init__(self):
        self.structure = None
        self.potentials = None
        self.neighbor_lists = None
        self.coulomb = None
        self.cutoffs = None
        self.potential_lists_ready = False
        self.bond_order_factor_lists_ready = False
        self.mpi_ready = True
        self.engine = None

    def __repr__(self):
        return "CoreMirror()"

    def get_atoms(self):
        return self.structure

    def view_fortran(self):
        """Prints what the engine holds (core.py:536-550 prints the Fortran
        core's arrays): atoms, cell, potentials and Coulomb summation."""
        s = self.structure
        lines = ["engine: %s" % ("none" if self.engine is None else
                                 "device %d" % getattr(self.engine, "device", 0))]
        if s is None:
            lines.append("atoms: none")
        else:
            lines.append("atoms: %d (%s)" % (len(s), " ".join(sorted(set(
                s.get_chemical_symbols())))))
            lines.append("cell:\n%s\npbc: %s" % (np.asarray(s.get_cell()),
                                                 list(np.asarray(s.get_pbc()))))
        lines.append("potentials: %d" % (0 if self.potentials is None else len(self.potentials)))
        for p in self.potentials or []:
            lines.append("  %r" % (p,))
        lines.append("coulomb: %r" % (self.coulomb,))
        print("\n".join(lines))

    def set_atoms(self, atoms):
        self.structure = copy.deepcopy(atoms)
        if hasattr(self.structure, "constraints"):
            del self.structure.constraints
        self.potential_lists_ready = False

    def set_charges(self, charges):
        try:
            self.structure.set_initial_charges(charges)
        except AttributeError:
            self.structure.set_charges(charges)

    def set_atomic_positions(self, atoms):
        self.structure.set_positions(atoms.get_positions())

    def set_atomic_momenta(self, atoms):
        self.structure.set_momenta(atoms.get_momenta())

    def set_cell(self, atoms):
        self.structure.set_cell(atoms.get_cell())
        self.structure.set_pbc(atoms.get_pbc())

    def set_potentials(self, potentials):
        self.potentials = copy.deepcopy(potentials)
        self.potential_lists_ready = False

    def set_neighbor_lists(self, lists):
        # the lists' cutoffs and skin (the engine holds the lists themselves)
        self.neighbor_lists = None if lists is None else (
            list(lists.cutoffs), lists.skin)

    def set_coulomb(self, coulomb):
        self.coulomb = copy.deepcopy(coulomb)

    def atoms_ready(self, atoms):
        """Same atoms, positions and momenta as the engine (core.py:648-668)."""
        s = self.structure
        if s is None or len(s) != len(atoms):
            return False
        return (_numbers(s) == _numbers(atoms)
                and np.array_equal(s.get_positions(), atoms.get_positions())
                and np.array_equal(s.get_momenta(), atoms.get_momenta()))

    def charges_ready(self, atoms):
        s = self.structure
        if s is None or len(s) != len(atoms):
            return False
        return np.array_equal(_charges(s), _charges(atoms))

    def cell_ready(self, atoms):
        s = self.structure
        if s is None:
            return False
        return (np.array_equal(s.get_cell(), atoms.get_cell())
                and np.array_equal(np.asarray(s.get_pbc()), np.asarray(atoms.get_pbc())))

    def potentials_ready(self, pots):
        if self.potentials is None:
            return pots is None
        return self.potentials == pots

    def neighbor_lists_ready(self, lists):
        if self.neighbor_lists is None or lists is None:
            return False
        return self.neighbor_lists == (list(lists.cutoffs), lists.skin)

    def coulomb_summation_ready(self, coulomb):
        if self.coulomb is None:
            return False
        return self.coulomb == coulomb
