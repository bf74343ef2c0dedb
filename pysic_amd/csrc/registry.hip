// registry.hip -- the descriptor registry of libpysic_hip.so (host code).
//
// Answers the queries pysic makes of its core about the interaction types it
// knows (PyInterface.f90:706-1011: number_of_potentials, is_potential,
// list_valid_potentials, number_of_targets_of_potential,
// names_of_parameters_of_potential, ... and their bond-order-factor
// counterparts).  It lists the types this engine evaluates: every potential
// type of the reference except 'dihedral', and the 'tersoff' bond-order
// factor.  Type names, target counts, bond-order levels and parameter names
// and orders are the reference's (its characterizers, Potentials.f90:
// 2909-5075 and 5581-6769), so that pysic's Python objects built from them
// carry over; the descriptions are this engine's own short texts.  One
// deliberate difference: the reference's 'step' registry leaves its second
// parameter name blank (Potentials.f90:4888-4890 writes 'e1' and then 'e2'
// into the first slot); here the four names are e1, e2, r1, r2, as the
// reference's description of 'step' uses them.
#include <cstring>

#include "../../include/pysic_registry.h"

namespace {

constexpr int kMaxParams = 12;

struct PotentialType {
    const char *name;
    int code;        // descriptor type code (include/pysic_hip.h)
    int n_targets;
    int n_params;
    const char *params[kMaxParams];
    const char *notes[kMaxParams];
    const char *description;
};

const PotentialType kPotentials[] = {
    {"LJ", 3, 2, 2, {"epsilon", "sigma"}, {"energy scale", "length scale"},
     "Lennard-Jones pair potential: V(r) = epsilon [(sigma/r)^12 - (sigma/r)^6]"},
    {"spring", 4, 2, 2, {"k", "R_0"}, {"spring constant", "equilibrium distance"},
     "harmonic spring, zero at the cutoff r_c: V(r) = k/2 [(r - R_0)^2 - (r_c - R_0)^2] "
     "where |r - R_0| < |r_c - R_0|, else 0"},
    {"force", 11, 1, 3, {"Fx", "Fy", "Fz"},
     {"x component of the force", "y component of the force", "z component of the force"},
     "constant external force on an atom: V = -F . r"},
    {"bond_bend", 15, 3, 4, {"epsilon", "theta_0", "n", "m"},
     {"energy scale", "equilibrium angle (radians)", "power of the cosine",
      "power of the difference"},
     "angle term at the middle atom of a triplet: V = epsilon (cos^n theta - cos^n theta_0)^m"},
    {"exponential", 1, 2, 2, {"epsilon", "zeta"}, {"energy scale", "inverse decay length"},
     "exponential pair potential: V(r) = epsilon exp(-zeta r)"},
    {"constant", 10, 1, 1, {"V"}, {"the constant energy"},
     "constant energy per atom: V = V"},
    {"Buckingham", 5, 2, 3, {"A", "C", "sigma"},
     {"energy scale of the repulsion", "energy scale of the dispersion", "length scale"},
     "Buckingham pair potential: V(r) = A exp(-r/sigma) - C (sigma/r)^6"},
    {"power", 6, 2, 3, {"epsilon", "a", "n"},
     {"energy scale", "length scale", "exponent"},
     "power-law pair potential: V(r) = epsilon (a/r)^n"},
    {"tabulated", 7, 2, 3, {"id", "range", "scale"},
     {"number of the table file table_<id>.txt", "distance the table spans",
      "energy multiplier"},
     "pair potential interpolated from a table of values and derivatives"},
    {"charge_self", 2, 1, 2, {"epsilon", "n"}, {"energy scale", "power of the charge"},
     "charge self energy of an atom: V(q) = epsilon q^n"},
    {"charge_pair", 14, 2, 3, {"epsilon", "n1", "n2"},
     {"energy scale", "power of the first charge", "power of the second charge"},
     "charge product of two atoms: V = epsilon q1^n1 q2^n2"},
    {"charge_exp", 12, 2, 11,
     {"epsilon", "Rmax1", "Rmin1", "Qmax1", "Qmin1", "Rmax2", "Rmin2", "Qmax2", "Qmin2", "xi1",
      "xi2"},
     {"energy scale", "radius at charge Qmax1", "radius at charge Qmin1", "largest charge 1",
      "smallest charge 1", "radius at charge Qmax2", "radius at charge Qmin2",
      "largest charge 2", "smallest charge 2", "inverse decay length 1",
      "inverse decay length 2"},
     "charge-dependent exponential of COMB: V = epsilon exp((xi1 D1(q1) + xi2 D2(q2))/2)"},
    {"charge_abs", 13, 2, 8, {"a1", "b1", "Q1", "n1", "a2", "b2", "Q2", "n2"},
     {"offset 1", "scale 1", "reference charge 1", "exponent 1", "offset 2", "scale 2",
      "reference charge 2", "exponent 2"},
     "charge-dependent factor of COMB: V = sqrt(f1 f2), f = max(0, a + b |q - Q|^n)"},
    {"shift_power", 8, 2, 4, {"epsilon", "r1", "r2", "n"},
     {"energy scale", "distance where V vanishes", "distance where V = epsilon", "exponent"},
     "shifted power law: V(r) = epsilon ((r1 - r)/(r1 - r2))^n"},
    {"step", 9, 2, 4, {"e1", "e2", "r1", "r2"},
     {"energy below r1", "energy above r2", "inner distance", "outer distance"},
     "smoothed step: V = e1 below r1 and e2 above r2, joined by a cosine between r1 and r2"},
};

struct BondType {
    const char *name;
    int n_targets;
    int level;       // 1: per-atom factor, 2: per-bond factor
    int n_params[2]; // per target class: single atom, pair
    const char *params[2][kMaxParams];
    const char *notes[2][kMaxParams];
    const char *description;
};

const BondType kBonds[] = {
    {"tersoff", 3, 2, {2, 5},
     {{"beta", "eta"}, {"a", "c", "d", "h", "mu"}},
     {{"prefactor of the bond-order sum", "exponent of the bond-order sum"},
      {"decay rate of the distance term", "angle term numerator", "angle term denominator",
       "cosine offset of the angle term", "exponent of the distance term"}},
     "Tersoff bond order: b_ij = (1 + (beta zeta_ij)^eta)^(-1/(2 eta)), zeta_ij = sum_k "
     "f_c(r_ik) exp((a (r_ij - r_ik))^mu) g(theta_ijk), g = 1 + c^2/d^2 - c^2/(d^2 + (h - "
     "cos theta)^2)"},
};

constexpr int kNumPotentials = sizeof(kPotentials) / sizeof(kPotentials[0]);
constexpr int kNumBonds = sizeof(kBonds) / sizeof(kBonds[0]);

const PotentialType *find_potential(const char *name)
{
    if (!name) return nullptr;
    for (const auto &p : kPotentials)
        if (std::strcmp(p.name, name) == 0) return &p;
    return nullptr;
}

const BondType *find_bond(const char *name)
{
    if (!name) return nullptr;
    for (const auto &b : kBonds)
        if (std::strcmp(b.name, name) == 0) return &b;
    return nullptr;
}

}  // namespace

extern "C" {

int pysic_number_of_potentials(void) { return kNumPotentials; }
int pysic_number_of_bond_order_factors(void) { return kNumBonds; }

const char *pysic_potential_name(int k)
{
    return (k >= 0 && k < kNumPotentials) ? kPotentials[k].name : nullptr;
}

const char *pysic_bond_order_factor_name(int k)
{
    return (k >= 0 && k < kNumBonds) ? kBonds[k].name : nullptr;
}

int pysic_is_potential(const char *name) { return find_potential(name) ? 1 : 0; }
int pysic_is_bond_order_factor(const char *name) { return find_bond(name) ? 1 : 0; }

int pysic_potential_type_code(const char *name)
{
    const PotentialType *p = find_potential(name);
    return p ? p->code : PYSIC_REGISTRY_UNKNOWN;
}

int pysic_number_of_targets_of_potential(const char *name)
{
    const PotentialType *p = find_potential(name);
    return p ? p->n_targets : PYSIC_REGISTRY_UNKNOWN;
}

int pysic_number_of_targets_of_bond_order_factor(const char *name)
{
    const BondType *b = find_bond(name);
    return b ? b->n_targets : PYSIC_REGISTRY_UNKNOWN;
}

int pysic_level_of_bond_order_factor(const char *name)
{
    const BondType *b = find_bond(name);
    return b ? b->level : PYSIC_REGISTRY_UNKNOWN;
}

int pysic_number_of_parameters_of_potential(const char *name)
{
    const PotentialType *p = find_potential(name);
    return p ? p->n_params : PYSIC_REGISTRY_UNKNOWN;
}

int pysic_number_of_parameters_of_bond_order_factor(const char *name, int target_class)
{
    const BondType *b = find_bond(name);
    if (!b) return PYSIC_REGISTRY_UNKNOWN;
    if (target_class < 1 || target_class > 2) return 0;
    return b->n_params[target_class - 1];
}

const char *pysic_parameter_name_of_potential(const char *name, int k)
{
    const PotentialType *p = find_potential(name);
    return (p && k >= 0 && k < p->n_params) ? p->params[k] : nullptr;
}

const char *pysic_parameter_name_of_bond_order_factor(const char *name, int target_class, int k)
{
    const BondType *b = find_bond(name);
    if (!b || target_class < 1 || target_class > 2) return nullptr;
    return (k >= 0 && k < b->n_params[target_class - 1]) ? b->params[target_class - 1][k]
                                                          : nullptr;
}

const char *pysic_parameter_description_of_potential(const char *name, int k)
{
    const PotentialType *p = find_potential(name);
    return (p && k >= 0 && k < p->n_params) ? p->notes[k] : nullptr;
}

const char *pysic_parameter_description_of_bond_order_factor(const char *name,
                                                             int target_class, int k)
{
    const BondType *b = find_bond(name);
    if (!b || target_class < 1 || target_class > 2) return nullptr;
    return (k >= 0 && k < b->n_params[target_class - 1]) ? b->notes[target_class - 1][k]
                                                          : nullptr;
}

const char *pysic_description_of_potential(const char *name)
{
    const PotentialType *p = find_potential(name);
    return p ? p->description : nullptr;
}

const char *pysic_description_of_bond_order_factor(const char *name)
{
    const BondType *b = find_bond(name);
    return b ? b->description : nullptr;
}

}  // extern "C"
