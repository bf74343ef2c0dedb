"""Interaction objects of the pysic API: Potential, ProductPotential,
BondOrderParameters, Coordinator, and their translation into engine
descriptors.

Mirrors pysic/interactions/local.py (ProductPotential local.py:7-282,
Potential :284-776) and pysic/interactions/bondorder.py (BondOrderParameters
:7-375, Coordinator :379-532) for the pair and one-body potential types
(exponential, LJ, spring, Buckingham, power, tabulated, shift_power, step,
charge_self, constant, force) and the 'tersoff' bond-order factor.
Parameter names and orders follow the reference's descriptor registry
(Potentials.f90:2909-4990, 5685-5736).
"""
import math
import os
from itertools import permutations

import numpy as np


class InvalidPotentialError(Exception):
    """Raised for unknown potential types or bad targets (utility/error.py)."""


class InvalidParametersError(Exception):
    """Raised for bad parameter lists (utility/error.py)."""


# name -> (n_targets, parameter names)  (Potentials.f90 characterizers:
# LJ 2916-2961, spring 3016-3033, force 3096-3115, exponential 3612-3630,
# constant 3917-3932, Buckingham 3980-3999, power 4060-4109, tabulated
# 4169-4217, charge_self 4312-4330, shift_power 4744-4793, step 4858-4907).
# The reference's 'step' registry writes 'e1' and then 'e2' into the first
# name slot and leaves the second blank; the names below are the ones its
# description gives (V = e1 below r1, e2 above r2).
POTENTIAL_TYPES = {
    "exponential": (2, ["epsilon", "zeta"]),
    "charge_self": (1, ["epsilon", "n"]),
    "LJ": (2, ["epsilon", "sigma"]),
    "spring": (2, ["k", "R_0"]),
    "Buckingham": (2, ["A", "C", "sigma"]),
    "power": (2, ["epsilon", "a", "n"]),
    "tabulated": (2, ["id", "range", "scale"]),
    "shift_power": (2, ["epsilon", "r1", "r2", "n"]),
    "step": (2, ["e1", "e2", "r1", "r2"]),
    "constant": (1, ["V"]),
    "force": (1, ["Fx", "Fy", "Fz"]),
    # charge-dependent pair terms of COMB (Potentials.f90:3691-3903, 4384-4731)
    "charge_exp": (2, ["epsilon", "Rmax1", "Rmin1", "Qmax1", "Qmin1", "Rmax2", "Rmin2",
                       "Qmax2", "Qmin2", "xi1", "xi2"]),
    "charge_abs": (2, ["a1", "b1", "Q1", "n1", "a2", "b2", "Q2", "n2"]),
    "charge_pair": (2, ["epsilon", "n1", "n2"]),
    # the angle term of COMB (Potentials.f90:3150-3328)
    "bond_bend": (3, ["epsilon", "theta_0", "n", "m"]),
}
# descriptor type codes (include/pysic_hip.h)
POTENTIAL_CODES = {"exponential": 1, "charge_self": 2, "LJ": 3, "spring": 4, "Buckingham": 5,
                   "power": 6, "tabulated": 7, "shift_power": 8, "step": 9, "constant": 10,
                   "force": 11, "charge_exp": 12, "charge_abs": 13,
                   "charge_pair": 14, "bond_bend": 15}
# types whose energy depends on the charges and on which atom of the pair
# the parameters belong to; their rows are followed by a continuation row
CHARGE_PAIR_TYPES = ("charge_exp", "charge_abs", "charge_pair")

# name -> (n_targets, level, [[scaling params], [local params]])
BOND_ORDER_TYPES = {
    "tersoff": (3, 2, [["beta", "eta"], ["a", "c", "d", "h", "mu"]]),
}


def is_potential(name):
    return name in POTENTIAL_TYPES


def is_bond_order_factor(name):
    return name in BOND_ORDER_TYPES


def _normalise_targets(symbols, n_targets):
    if symbols is None:
        return None
    if isinstance(symbols, str):
        symbols = [[symbols]]
    out = []
    for t in symbols:
        if isinstance(t, str):
            t = [t]
        t = list(t)
        if len(t) != n_targets:
            raise InvalidPotentialError("target list %s does not have %d symbols" % (t, n_targets))
        out.append(t)
    return out


class BondOrderParameters:
    """Bond-order factor parameters (bondorder.py:7-375).

    parameters: [[beta, eta], [a, c, d, h, mu]] for 'tersoff'.
    """

    def __init__(self, bond_order_type, cutoff=0.0, cutoff_margin=0.0, parameters=None,
                 symbols=None):
        if not is_bond_order_factor(bond_order_type):
            raise InvalidParametersError("unknown bond order factor type: %s" % bond_order_type)
        self.bond_order_type = bond_order_type
        self.n_targets, self.level, self.names = BOND_ORDER_TYPES[bond_order_type]
        self.n_params = [len(x) for x in self.names]
        self.cutoff = float(cutoff)
        self.cutoff_margin = float(cutoff_margin)
        self.parameters = None
        self.set_parameters(parameters if parameters is not None else
                            [[0.0] * k for k in self.n_params])
        self.symbols = _normalise_targets(symbols, self.n_targets)

    def accepts_parameters(self, params):
        if len(params) != len(self.n_params):
            return False
        return all(len(params[i]) == self.n_params[i] for i in range(len(self.n_params)))

    def set_parameters(self, params):
        if not self.accepts_parameters(params):
            raise InvalidParametersError("bond order parameters must have shape %s" % self.n_params)
        self.parameters = [list(map(float, p)) for p in params]

    def get_parameters_as_list(self):
        """Single-element parameters first, then pair parameters
        (bondorder.py:135-145)."""
        return [x for p in self.parameters for x in p]

    def set_parameter_value(self, name, value):
        for i, group in enumerate(self.names):
            if name in group:
                self.parameters[i][group.index(name)] = float(value)
                return
        raise InvalidParametersError("no parameter %s" % name)

    def get_parameter_value(self, name):
        for i, group in enumerate(self.names):
            if name in group:
                return self.parameters[i][group.index(name)]
        raise InvalidParametersError("no parameter %s" % name)

    def get_symbols(self):
        return self.symbols

    def set_symbols(self, symbols):
        self.symbols = _normalise_targets(symbols, self.n_targets)

    def accepts_target_list(self, targets):
        """True if targets is a list of lists of get_number_of_targets()
        symbols each (bondorder.py:321-344); the symbols are not checked."""
        try:
            return all(isinstance(t, list) and len(t) == self.n_targets for t in targets)
        except TypeError:
            return False

    def add_symbols(self, symbols):
        """Adds target lists: a list of them, or a single one
        (bondorder.py:165-184)."""
        if self.accepts_target_list(symbols):
            new = [list(t) for t in symbols]
        elif self.accepts_target_list([symbols]):
            new = [list(symbols)]
        else:
            raise InvalidParametersError("Invalid number of targets.")
        self.symbols = (self.symbols or []) + new

    def get_different_symbols(self):
        """Each symbol the targets name, once, in order of appearance
        (bondorder.py:193-203)."""
        out = []
        for targets in self.symbols or []:
            for sym in targets:
                if sym not in out:
                    out.append(sym)
        return out

    def get_cutoff(self):
        return self.cutoff

    def set_cutoff(self, c):
        self.cutoff = float(c)

    def get_cutoff_margin(self):
        return self.cutoff_margin

    def set_cutoff_margin(self, m):
        """bondorder.py:285-298: the margin may not exceed the cutoff."""
        m = float(m)
        if m > self.cutoff:
            raise InvalidParametersError(
                "A cutoff margin (%s) longer than the hard cutoff (%s) was given."
                % (m, self.cutoff))
        self.cutoff_margin = m

    def set_soft_cutoff(self, cutoff):
        """Sets the soft cutoff through the margin (cutoff - soft), so a later
        change of the hard cutoff moves the soft one too (bondorder.py:
        306-318)."""
        self.set_cutoff_margin(self.cutoff - float(cutoff))

    def get_soft_cutoff(self):
        """cutoff - margin (bondorder.py:241-243)."""
        return self.cutoff - self.cutoff_margin

    def get_level(self):
        return self.level

    def get_number_of_targets(self):
        return self.n_targets

    def get_bond_order_type(self):
        return self.bond_order_type

    def __eq__(self, other):
        return (isinstance(other, BondOrderParameters)
                and self.bond_order_type == other.bond_order_type
                and self.parameters == other.parameters and self.symbols == other.symbols
                and self.cutoff == other.cutoff and self.cutoff_margin == other.cutoff_margin)

    def __repr__(self):
        return "BondOrderParameters('%s',symbols=%s,parameters=%s,cutoff=%s,cutoff_margin=%s)" % (
            self.bond_order_type, self.symbols, self.parameters, self.cutoff, self.cutoff_margin)


class Coordinator:
    """Holds the bond-order parameters that scale a potential
    (bondorder.py:379-532)."""

    def __init__(self, bond_order_parameters=None):
        if bond_order_parameters is None:
            bond_order_parameters = []
        if isinstance(bond_order_parameters, BondOrderParameters):
            bond_order_parameters = [bond_order_parameters]
        self.bond_order_params = list(bond_order_parameters)
        self.bond_order_factors = None
        self.group_index = -1

    def get_bond_order_parameters(self):
        return self.bond_order_params

    def set_bond_order_parameters(self, p):
        self.bond_order_params = [p] if isinstance(p, BondOrderParameters) else list(p)

    def add_bond_order_parameters(self, p):
        self.bond_order_params.append(p)

    def set_group_index(self, g):
        self.group_index = g

    def get_group_index(self):
        return self.group_index

    # The reference's Coordinator asks the one Fortran core
    # (bondorder.py:488-532); here the engine is that of the calculator
    # that synchronised last (Pysic.core), with this coordinator's group.
    @staticmethod
    def _engine():
        from .calculator import Pysic
        eng = Pysic.core.engine
        if eng is None or Pysic.core.structure is None:
            raise InvalidPotentialError("no calculator has set up a system for the coordinator")
        return eng

    def calculate_bond_order_factors(self):
        """Computes and stores the per-atom bond-order factors of this
        coordinator's group (bondorder.py:488-497)."""
        self.bond_order_factors = self._engine().bond_order_factors(self.group_index)

    def get_bond_order_factors(self):
        """The factors stored by calculate_bond_order_factors."""
        return self.bond_order_factors

    def get_bond_order_gradients(self, atom_index):
        """(N, 3): the gradients of every atom's factor with respect to the
        position of atom atom_index (bondorder.py:507-518)."""
        return self._engine().bond_order_gradients(self.group_index, int(atom_index))

    def get_bond_order_gradients_of_factor(self, atom_index):
        """(N, 3): the gradient of atom atom_index's factor with respect to
        the position of every atom (bondorder.py:521-532)."""
        return self._engine().bond_order_gradients(self.group_index, int(atom_index),
                                                   of_factor=True)

    def __eq__(self, other):
        return isinstance(other, Coordinator) and \
            self.bond_order_params == other.bond_order_params

    def __ne__(self, other):
        return not self.__eq__(other)

    def __repr__(self):
        return "Coordinator(bond_order_parameters=%s)" % (self.bond_order_params,)


class Potential:
    """A potential acting on groups of atoms (local.py:284-776)."""

    def __init__(self, potential_type, symbols=None, tags=None, indices=None, parameters=None,
                 cutoff=0.0, cutoff_margin=0.0, coordinator=None):
        if not is_potential(potential_type):
            raise InvalidPotentialError("There is no potential called '%s'." % potential_type)
        if tags is not None or indices is not None:
            raise InvalidPotentialError("only element targets are supported by this engine")
        self.potential_type = potential_type
        self.n_targets, self.names = POTENTIAL_TYPES[potential_type]
        self.symbols = _normalise_targets(symbols, self.n_targets)
        self.parameters = [0.0] * len(self.names)
        if parameters is not None:
            self.set_parameter_values(parameters)
        self.cutoff = float(cutoff)
        self.cutoff_margin = float(cutoff_margin)
        self.coordinator = coordinator

    def set_parameter_values(self, values):
        if len(values) != len(self.names):
            raise InvalidParametersError("'%s' takes %d parameters" % (self.potential_type,
                                                                     len(self.names)))
        self.parameters = [float(v) for v in values]

    def get_parameter_values(self):
        return list(self.parameters)

    def set_parameter_value(self, name, value):
        self.parameters[self.names.index(name)] = float(value)

    def get_parameter_value(self, name):
        return self.parameters[self.names.index(name)]

    def get_parameter_names(self):
        return list(self.names)

    def get_symbols(self):
        return self.symbols

    def set_symbols(self, s):
        self.symbols = _normalise_targets(s, self.n_targets)

    def get_different_symbols(self):
        out = []
        for t in (self.symbols or []):
            for s in t:
                if s not in out:
                    out.append(s)
        return out

    def get_cutoff(self):
        return self.cutoff

    def set_cutoff(self, c):
        self.cutoff = float(c)

    def get_cutoff_margin(self):
        return self.cutoff_margin

    def set_cutoff_margin(self, m):
        self.cutoff_margin = float(m)

    def get_soft_cutoff(self):
        """cutoff - margin (local.py:478-480)."""
        return self.cutoff - self.cutoff_margin

    def get_coordinator(self):
        return self.coordinator

    def set_coordinator(self, c):
        self.coordinator = c

    def get_number_of_targets(self):
        return self.n_targets

    def get_potential_type(self):
        return self.potential_type

    def get_potentials(self):
        """[self], as ProductPotential.get_potentials (local.py:751-757)."""
        return [self]

    def is_multiplier(self):
        """[False], as ProductPotential.is_multiplier (local.py:759-765)."""
        return [False]

    def __eq__(self, other):
        return (isinstance(other, Potential) and self.potential_type == other.potential_type
                and self.symbols == other.symbols and self.parameters == other.parameters
                and self.cutoff == other.cutoff and self.cutoff_margin == other.cutoff_margin
                and self.coordinator == other.coordinator)

    def __repr__(self):
        return "Potential('%s',symbols=%s,parameters=%s,cutoff=%s,cutoff_margin=%s)" % (
            self.potential_type, self.symbols, self.parameters, self.cutoff, self.cutoff_margin)


class ProductPotential:
    """The product of potentials with the same number of targets
    (local.py:7-282).  Targets, cutoffs and coordinator are the first
    potential's; the others only multiply its value."""

    def __init__(self, potentials):
        self.potentials = []
        self.n_targets = None
        for pot in potentials:
            self.add_potential(pot)

    def set_potentials(self, potentials):
        self.potentials = []
        self.n_targets = None
        for pot in potentials:
            self.add_potential(pot)

    def add_potential(self, potential):
        """local.py:53-84: every factor must have the same number of targets."""
        if self.n_targets is None:
            self.n_targets = potential.get_number_of_targets()
        elif self.n_targets != potential.get_number_of_targets():
            raise InvalidPotentialError(
                "You may only multiply potentials with the same number of targets.")
        self.potentials.append(potential)

    def get_potentials(self):
        return self.potentials

    def is_multiplier(self):
        """[False, True, ..., True] (local.py:256-263)."""
        return [False] + [True] * (len(self.potentials) - 1)

    def get_symbols(self):
        return self.potentials[0].get_symbols()

    def set_symbols(self, symbols):
        self.potentials[0].set_symbols(symbols)

    def get_different_symbols(self):
        return self.potentials[0].get_different_symbols()

    def get_cutoff(self):
        return self.potentials[0].get_cutoff()

    def set_cutoff(self, cutoff):
        self.potentials[0].set_cutoff(cutoff)

    def get_cutoff_margin(self):
        return self.potentials[0].get_cutoff_margin()

    def set_cutoff_margin(self, margin):
        self.potentials[0].set_cutoff_margin(margin)

    def get_soft_cutoff(self):
        return self.potentials[0].get_soft_cutoff()

    def get_number_of_targets(self):
        return self.n_targets

    def get_coordinator(self):
        return self.potentials[0].get_coordinator()

    def set_coordinator(self, coordinator):
        self.potentials[0].set_coordinator(coordinator)

    def __eq__(self, other):
        return (isinstance(other, ProductPotential) and self.n_targets == other.n_targets
                and self.potentials == other.potentials)

    def __repr__(self):
        return "ProductPotential(%s)" % self.potentials


def table_ids(potentials):
    """Ids of the tables that the 'tabulated' potentials (and product
    factors) use."""
    ids = set()
    for pots in potentials or []:
        for pot in pots.get_potentials():
            if pot.get_potential_type() == "tabulated":
                ids.add(int(pot.get_parameter_values()[0]))
    return sorted(ids)


def read_table(table_id, directory="."):
    """The file table_<id>.txt that the reference reads when it creates a
    'tabulated' potential (Potentials.f90:2462-2465, read_table
    Utility.f90:384-433): one row per point, value and derivative."""
    name = os.path.join(directory, "table_%04d.txt" % int(table_id))
    if not os.path.exists(name):
        raise InvalidPotentialError("table file %s not found" % name)
    rows = np.loadtxt(name, dtype=np.float64, ndmin=2)
    if rows.shape[1] < 2 or len(rows) < 2:
        raise InvalidPotentialError("%s needs at least 2 rows of value and derivative" % name)
    return np.ascontiguousarray(rows[:, :2])


def factor_rows(name, params, n_targets, species, cutoff, soft_cutoff, group, link, orig):
    """The engine descriptor rows of one potential factor applied to the
    species tuple `species` (include/pysic_hip.h): its 12-double row, and a
    continuation row (link 2) where the row cannot hold the factor -- for a
    charge-dependent factor its type, the species of its original first
    target (its parameters 1 belong to that atom, Potentials.f90:3797-3829)
    and parameters 5-13; for 'bond_bend' its type, the third target and the
    three original targets (the second is the vertex,
    Potentials.f90:3201-3215).  orig: the species of the original targets
    (-1 for an element absent from the system)."""
    v = [float(x) for x in params]
    p = v + [0.0] * (13 - len(v))
    e = list(species) + [0] * (2 - len(species))
    code = POTENTIAL_CODES[name]
    out = [[code, n_targets, e[0], e[1], p[0], p[1], cutoff, soft_cutoff, group, p[2], p[3],
            link]]
    o = list(orig) + [-1] * (3 - len(orig))
    if name in CHARGE_PAIR_TYPES:
        out.append([code, o[0]] + p[4:13] + [2])
    elif name == "bond_bend":
        out.append([code, e[2]] + o[:3] + [0.0] * 6 + [2])
    return out


def build_descriptors(potentials, species_of):
    """Translate potentials into the engine's descriptor rows, the way
    update_core_potentials does for the Fortran core (calculator.py:1063-1446):

    * every potential is added once per distinct permutation of each target
      tuple (calculator.py:1166-1192);
    * a coordinator's group index is the running index of the first
      coordinated potential with equal bond parameters (1227-1233);
    * bond-order factors are added per target tuple without permutation
      (permutate = False, 1372).

    * a ProductPotential adds its first potential per permutation, each
      followed by the other factors as multiplier rows, in the order the
      reference stores them (reversed, calculator.py:1209-1213, 1262-1277).

    species_of maps element symbols to engine species ids; potentials naming
    elements absent from the system are dropped (they can act on no atom).
    Returns (potdesc (k,12), bonddesc (m,14), n_groups).
    """
    pot_rows, bond_rows = [], []
    coord_list = []  # (coordinator, group index)
    pot_index = 0
    for pot in potentials or []:
        coord = pot.get_coordinator()
        if coord is not None:
            if not any(c is coord or c == coord for c, _ in coord_list):
                coord_list.append((coord, pot_index))
            pot_index += 1
    for pot in potentials or []:
        group = -1
        coord = pot.get_coordinator()
        if coord is not None:
            for c, g in coord_list:
                if c.get_bond_order_parameters() == coord.get_bond_order_parameters():
                    group = g
                    break
            coord.set_group_index(group)
        factors = pot.get_potentials()
        master, mults = factors[0], list(reversed(factors[1:]))
        nt = pot.get_number_of_targets()

        def rows(f, e, link, orig):
            return factor_rows(f.get_potential_type(), f.get_parameter_values(), nt, e,
                               pot.get_cutoff(), pot.get_soft_cutoff(), group, link,
                               [species_of.get(x, -1) for x in orig])

        tuples = pot.get_symbols() or []
        for targets in tuples:
            for perm in sorted(set(permutations(targets))):
                if any(s not in species_of for s in perm):
                    continue
                e = [species_of[s] for s in perm] + [0] * (2 - len(perm))
                # the reference keeps each copy's original targets; its
                # multipliers carry the first target tuple (calculator.py:1241-1277)
                pot_rows += rows(master, e, 0, targets)
                for m in mults:
                    pot_rows += rows(m, e, 1, tuples[0])
    for coord, g in coord_list:
        for bond in coord.get_bond_order_parameters():
            if bond.get_bond_order_type() != "tersoff":
                raise InvalidParametersError("only 'tersoff' bond order factors are supported")
            for targets in bond.get_symbols() or []:
                if any(s not in species_of for s in targets):
                    continue
                e = [species_of[s] for s in targets]
                bond_rows.append(e + bond.get_parameters_as_list() +
                                 [bond.get_cutoff(), bond.get_soft_cutoff(), g, 0.0])
    pd = np.array(pot_rows, dtype=np.float64).reshape(-1, 12)
    bd = np.array(bond_rows, dtype=np.float64).reshape(-1, 14)
    return pd, bd, len(coord_list)


# Tersoff (1988) silicon, written with pysic's building blocks: a repulsive
# exponential, an attractive exponential (epsilon = -B) scaled by a
# coordinator holding the 'tersoff' bond-order factor
# (the construction of interactions/comb.py:668-731).
TERSOFF_SI = dict(A=1830.8, lambda1=2.4799, B=471.18, lambda2=1.7322, beta=1.1e-6,
                  eta=0.78734, a=0.0, c=100390.0, d=16.217, h=-0.59825, mu=3.0,
                  cutoff=3.0, margin=0.3)


def tersoff_si(params=None):
    P = dict(TERSOFF_SI)
    if params:
        P.update(params)
    bond = BondOrderParameters("tersoff", cutoff=P["cutoff"], cutoff_margin=P["margin"],
                               parameters=[[P["beta"], P["eta"]],
                                           [P["a"], P["c"], P["d"], P["h"], P["mu"]]],
                               symbols=[["Si", "Si", "Si"]])
    rep = Potential("exponential", symbols=[["Si", "Si"]], parameters=[P["A"], P["lambda1"]],
                    cutoff=P["cutoff"], cutoff_margin=P["margin"])
    att = Potential("exponential", symbols=[["Si", "Si"]], parameters=[-P["B"], P["lambda2"]],
                    cutoff=P["cutoff"], cutoff_margin=P["margin"],
                    coordinator=Coordinator([bond]))
    return [rep, att]


# Argon, Lennard-Jones: epsilon/k_B = 119.8 K, sigma = 3.405 A.  pysic's LJ is
# epsilon ((sigma/r)^12 - (sigma/r)^6) (Potentials.f90:2993-3005), so its
# epsilon is 4 times the well depth.
ARGON_LJ = dict(epsilon=4 * 119.8 * 8.617333262e-5, sigma=3.405)


def lj_argon(cutoff=2.5 * 3.405, margin=0.0, symbol="Ar"):
    """Argon as one pysic 'LJ' potential (cutoff 2.5 sigma by default)."""
    return [Potential("LJ", symbols=[[symbol, symbol]],
                      parameters=[ARGON_LJ["epsilon"], ARGON_LJ["sigma"]],
                      cutoff=cutoff, cutoff_margin=margin)]


# COMB charge-dependent pair terms with the defaults of the reference's COMB
# builder (pysic/interactions/comb.py:126-444, 665-990), without its
# bond-order coordinators: per-element charge-screening radii (rmax, rmin at
# qmax, qmin), attractive alpha and charge_abs a, b, Q, n; attractive B and
# repulsive A, lambda mix as sqrt(B_i B_j), (alpha_i + alpha_j)/2,
# sqrt(A_i A_j), (lambda_i + lambda_j)/2.
COMB = {"Si": dict(rmax=-1.658949, rmin=1.651725, qmax=4.0, qmin=-4.0, alpha=1.7322, a=1.0,
                   b=-9.5367431640625e-07, q=0.0, n=10.0, B=471.18, A=1830.8, lam=2.4799,
                   cut=3.0, margin=0.2),
        "O": dict(rmax=-0.00112, rmin=0.00148, qmax=5.5046, qmin=-1.8349, alpha=2.68, a=1.00098,
                  b=-2.26e-6, q=1.83485, n=10.0, B=260.8931, A=3326.699, lam=5.36, cut=3.0,
                  margin=0.4)}


def comb_charge_model():
    """Repulsive charge_exp x exponential and attractive charge_exp x
    charge_abs x exponential products for Si-Si, O-O and Si-O (the order of
    comb.py's pieces), plus a lone charge_abs O-O term."""
    C = COMB
    P = Potential

    def cexp(i, j, eps, xi, cut, margin):
        return P("charge_exp", symbols=[[i, j]], cutoff=cut, cutoff_margin=margin,
                 parameters=[eps, C[i]["rmax"], C[i]["rmin"], C[i]["qmax"], C[i]["qmin"],
                             C[j]["rmax"], C[j]["rmin"], C[j]["qmax"], C[j]["qmin"],
                             C[i][xi], C[j][xi]])

    def cabs(i, j):
        return P("charge_abs", symbols=[[i, j]],
                 parameters=[C[i]["a"], C[i]["b"], C[i]["q"], C[i]["n"],
                             C[j]["a"], C[j]["b"], C[j]["q"], C[j]["n"]])

    pots = []
    for i, j in (("Si", "Si"), ("O", "O"), ("Si", "O")):
        cut, margin = (C[i]["cut"], C[i]["margin"]) if i == j else (3.05, 0.5)
        B = math.sqrt(C[i]["B"] * C[j]["B"])
        zeta = 0.5 * (C[i]["alpha"] + C[j]["alpha"])
        pots.append(ProductPotential([
            cexp(i, j, B, "alpha", cut, margin), cabs(i, j),
            P("exponential", symbols=[[i, j]], parameters=[-1.0, zeta])]))
    for i, j in (("Si", "Si"), ("O", "O"), ("Si", "O")):
        cut, margin = (C[i]["cut"], C[i]["margin"]) if i == j else (3.05, 0.5)
        A = math.sqrt(C[i]["A"] * C[j]["A"])
        lam = 0.5 * (C[i]["lam"] + C[j]["lam"])
        pots.append(ProductPotential([
            cexp(i, j, A, "lam", cut, margin),
            P("exponential", symbols=[[i, j]], parameters=[1.0, lam])]))
    lone = cabs("O", "O")
    lone.set_cutoff(3.0)
    lone.set_cutoff_margin(0.5)
    pots.append(lone)
    return pots


def comb_bend_terms():
    """COMB's bond bending with the reference builder's defaults
    (comb.py:94-125, 636-662): Si-O-Si epsilon 2.6 eV, theta_0 2.508561734;
    O-Si-O epsilon 0.3122 eV, theta_0 1.910611932; n = 1, m = 2; cutoff
    3.05 A with a 0.5 A margin."""
    P = Potential
    return [P("bond_bend", symbols=[["Si", "O", "Si"]], cutoff=3.05, cutoff_margin=0.5,
              parameters=[2.6, 2.508561734, 1, 2]),
            P("bond_bend", symbols=[["O", "Si", "O"]], cutoff=3.05, cutoff_margin=0.5,
              parameters=[0.3122, 1.910611932, 1, 2])]


# COMB's 'tersoff' bond-order parameters per center element with the
# reference builder's defaults (comb.py:181-221, 279-319, 345-379): beta,
# eta, mu, a, c, d, h; a Si-O bond takes the mean of the two a's.  Cutoffs
# and margins per (center, partner, third) triplet.
COMB_BOND = {"Si": dict(beta=1.0999e-6, eta=0.78734, mu=3.0, a=1.7322, c=100390.0, d=16.218,
                        h=-0.59826),
             "O": dict(beta=2.0, eta=1.0, mu=1.0, a=2.68, c=6.6, d=1.0, h=-0.229)}
COMB_BOND_CUT = {("Si", "Si", "O"): (3.05, 0.5), ("Si", "Si", "Si"): (3.0, 0.2),
                 ("O", "O", "Si"): (3.05, 0.5), ("O", "O", "O"): (3.0, 0.4),
                 ("O", "Si", "Si"): (3.05, 0.5), ("O", "Si", "O"): (3.05, 0.5),
                 ("Si", "O", "Si"): (3.05, 0.5), ("Si", "O", "O"): (3.05, 0.5)}


def comb_coordinators():
    """The three coordinators of COMB's attractive terms, built as
    comb.py:703-729 (Si-Si: Si-Si-O, Si-Si-Si), 773-800 (O-O: O-O-Si,
    O-O-O) and 846-895 (Si-O: O-Si-Si, O-Si-O, Si-O-O, Si-O-Si) do."""
    B = COMB_BOND
    a_sio = 0.5 * (B["Si"]["a"] + B["O"]["a"])

    def bond(triplet, a=None):
        p = B[triplet[0]]
        cut, margin = COMB_BOND_CUT[triplet]
        return BondOrderParameters("tersoff", cutoff=cut, cutoff_margin=margin,
                                   parameters=[[p["beta"], p["eta"]],
                                               [p["a"] if a is None else a, p["c"], p["d"],
                                                p["h"], p["mu"]]],
                                   symbols=[list(triplet)])
    sisi = Coordinator([bond(("Si", "Si", "O")), bond(("Si", "Si", "Si"))])
    oo = Coordinator([bond(("O", "O", "Si")), bond(("O", "O", "O"))])
    sio = Coordinator([bond(("O", "Si", "Si"), a_sio), bond(("O", "Si", "O"), a_sio),
                       bond(("Si", "O", "O"), a_sio), bond(("Si", "O", "Si"), a_sio)])
    return sisi, oo, sio


def comb_sio2(coordinators=True):
    """COMB's local SiO2 terms with the reference builder's defaults
    (comb.py:630-990): the Si-O-Si and O-Si-O bending, the charge-dependent
    attraction -- each attractive product scaled by its coordinator's
    bond-order factor -- and the charge-dependent repulsion.  With
    coordinators=False the attraction is unscaled."""
    pots = comb_bend_terms() + comb_charge_model()
    if coordinators:
        # comb_charge_model's first three products: attraction Si-Si, O-O, Si-O
        for prod, coord in zip(pots[2:5], comb_coordinators()):
            prod.set_coordinator(coord)
    return pots
