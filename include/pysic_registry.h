/*
 * pysic_registry.h -- the descriptor registry of libpysic_hip.so.
 *
 * Replaces the registry queries of the reference's f2py module
 * (fortran/PyInterface.f90:706-1011), which pysic/core.py:44-446 uses to
 * validate and describe Potential and BondOrderParameters objects.  Pure host
 * code: no handle, no GPU.  Strings are static and owned by the library;
 * names are the reference's, without the blank padding and the integer
 * coding of the f2py interface (Utility.f90:205-250).  The registry lists
 * the types this engine evaluates; a name outside it is unknown.
 */
#ifndef PYSIC_REGISTRY_H
#define PYSIC_REGISTRY_H

#ifdef __cplusplus
extern "C" {
#endif

/* returned by the integer queries for a name the registry does not know */
#define PYSIC_REGISTRY_UNKNOWN -1

/* number_of_potentials (PyInterface.f90:706) */
int pysic_number_of_potentials(void);
/* number_of_bond_order_factors (PyInterface.f90:719) */
int pysic_number_of_bond_order_factors(void);
/* list_valid_potentials (PyInterface.f90:763): the k-th name, 0-based;
 * NULL beyond the count */
const char *pysic_potential_name(int k);
/* list_valid_bond_order_factors (PyInterface.f90:783) */
const char *pysic_bond_order_factor_name(int k);
/* is_potential (PyInterface.f90:733): 1 or 0 */
int pysic_is_potential(const char *name);
/* is_bond_order_factor (PyInterface.f90:748) */
int pysic_is_bond_order_factor(const char *name);
/* The descriptor type code of a potential (include/pysic_hip.h: 1
 * 'exponential' ... 15 'bond_bend'); no counterpart in the reference, whose
 * core takes the name itself (add_potential, PyInterface.f90:433). */
int pysic_potential_type_code(const char *name);
/* number_of_targets_of_potential (PyInterface.f90:803) */
int pysic_number_of_targets_of_potential(const char *name);
/* number_of_targets_of_bond_order_factor (PyInterface.f90:818) */
int pysic_number_of_targets_of_bond_order_factor(const char *name);
/* level_of_bond_order_factor (PyInterface.f90:834) */
int pysic_level_of_bond_order_factor(const char *name);
/* number_of_parameters_of_potential (PyInterface.f90:849) */
int pysic_number_of_parameters_of_potential(const char *name);
/* number_of_parameters_of_bond_order_factor (PyInterface.f90:865): the
 * parameters of target class 1 (per atom) or 2 (per pair); 0 for another
 * class */
int pysic_number_of_parameters_of_bond_order_factor(const char *name, int target_class);
/* names_of_parameters_of_potential (PyInterface.f90:886): the k-th, 0-based */
const char *pysic_parameter_name_of_potential(const char *name, int k);
/* names_of_parameters_of_bond_order_factor (PyInterface.f90:914) */
const char *pysic_parameter_name_of_bond_order_factor(const char *name, int target_class,
                                                      int k);
/* descriptions_of_parameters_of_potential (PyInterface.f90:943) */
const char *pysic_parameter_description_of_potential(const char *name, int k);
/* descriptions_of_parameters_of_bond_order_factor (PyInterface.f90:971) */
const char *pysic_parameter_description_of_bond_order_factor(const char *name,
                                                             int target_class, int k);
/* description_of_potential (PyInterface.f90:994) */
const char *pysic_description_of_potential(const char *name);
/* description_of_bond_order_factor (PyInterface.f90:1011) */
const char *pysic_description_of_bond_order_factor(const char *name);

#ifdef __cplusplus
}
#endif

#endif
