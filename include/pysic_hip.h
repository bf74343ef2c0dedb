/*
 * pysic_hip.h -- C-ABI of the MI355X pysic force engine (libpysic_hip.so).
 *
 * Replaces the reference's f2py module `pysic_interface`
 * (fortran/PyInterface.f90) for the Tersoff / Ewald / charge-relaxation
 * path.  Plain pointers and sizes only; all host buffers are copied, none is
 * retained.  Every call returns 0 on success or a negative PYSIC_E* code,
 * with a message available from pysic_last_error().  Arguments are checked
 * before the handle changes: a call refused for its arguments leaves the
 * loaded system as it was.  Indices are 0-based,
 * arrays of 3-vectors are row-major (N,3), cells are given as three vectors
 * in rows, and species are small integers 0..3 chosen by the caller.
 *
 * Potential descriptors (12 doubles per entry, one entry per target
 * permutation as calculator.py:1166-1192 adds them):
 *   type, n_targets, e1, e2, p1, p2, cutoff, soft_cutoff (= cutoff - margin),
 *   group (-1 none), p3, p4, link
 * type: 1 'exponential', 2 'charge_self', 3 'LJ', 4 'spring', 5 'Buckingham',
 *   6 'power', 7 'tabulated', 8 'shift_power', 9 'step', 10 'constant',
 *   11 'force', 12 'charge_exp', 13 'charge_abs', 14 'charge_pair',
 *   15 'bond_bend' (3 targets); p1..p4 are the parameters in the reference's
 *   order (Potentials.f90:2909-4990).  Only a lone 'exponential' may carry a
 *   bond-order group.
 * link: 0 starts a potential; 1 makes the entry a multiplier of the
 *   potential before it (a ProductPotential factor: add_potential with
 *   is_multiplier, calculator.py:1197-1369); a multiplier's targets, cutoffs
 *   and group are ignored (the product uses its first potential's); 2
 *   continues the entry before it, which must be of type 12-15; for 12-14
 *   as
 *     type, s, p5, p6, ..., p13, 2
 *   with p5.. its further parameters and s the species of the potential's
 *   original first target (the element its parameters 1 belong to; -1 for
 *   one absent from the system), calculator.py:1241-1277 and
 *   Potentials.f90:3797-3829.  At most 16 distinct such factors.  A
 *   'bond_bend' row (e1, e2 its first two targets) is followed by
 *     15, e3, o1, o2, o3, 0, 0, 0, 0, 0, 0, 2
 *   with e3 its third target and o1..o3 the potential's original targets
 *   (o2, the vertex, decides where it acts, Potentials.f90:3201-3215).  At
 *   most 12 such rows, no multipliers, no group.
 * Bond-order descriptors ('tersoff', 14 doubles per entry, no permutation as
 * calculator.py:1372):
 *   e1 e2 e3 (center, j, k), beta, eta, a, c, d, h, mu, cutoff, soft_cutoff,
 *   group, 0
 */
#ifndef PYSIC_HIP_H
#define PYSIC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PYSIC_OK 0
#define PYSIC_EINVAL -1
#define PYSIC_EHIP -2
#define PYSIC_ESTATE -3
#define PYSIC_ECAP -4
#define PYSIC_ENOMEM -5
#define PYSIC_ERANGE -6

typedef struct pysic_ctx pysic_ctx;

/* start_potentials / start_bond_order_factors / start_mpi / start_rng
 * (PyInterface.f90:51-195, called at import by pysic/core.py:20-28): one
 * handle owns one device stream and all device buffers of one system. */
int pysic_create(pysic_ctx **out, int device);
/* release (PyInterface.f90:1023) */
void pysic_destroy(pysic_ctx *h);
const char *pysic_last_error(pysic_ctx *h);

/* create_atoms (PyInterface.f90:250); species replaces the int-coded labels */
int pysic_create_atoms(pysic_ctx *h, int n, const double *masses, const double *charges,
                       const double *positions, const double *momenta, const int *species);
/* create_cell (PyInterface.f90:208): vectors in rows; inverse = M^-1
 * row-major where M has the vectors as columns */
int pysic_create_cell(pysic_ctx *h, const double *vectors, const double *inverse,
                      const int *pbc);
/* update_atom_coordinates (PyInterface.f90:277); momenta may be NULL, which
 * keeps the momenta the atoms have */
int pysic_update_atom_coordinates(pysic_ctx *h, int n, const double *positions,
                                  const double *momenta);
/* update_atom_charges (PyInterface.f90:295) */
int pysic_update_atom_charges(pysic_ctx *h, int n, const double *charges);
/* get_number_of_atoms (PyInterface.f90:311) */
int pysic_get_number_of_atoms(pysic_ctx *h, int *n);

/* allocate_potentials + add_potential (PyInterface.f90:370, 433) */
int pysic_set_potentials(pysic_ctx *h, int npot, const double *desc);
/* The table of 'tabulated' potentials with this id: nrows rows of (value,
 * derivative), the file table_<id>.txt that create_potential reads
 * (Potentials.f90:2462-2465, read_table Utility.f90:384-433). */
int pysic_set_potential_table(pysic_ctx *h, int id, int nrows, const double *rows);
/* allocate_bond_order_factors + add_bond_order_factor (PyInterface.f90:384, 482)
 * + allocate_bond_order_storage (402) */
int pysic_set_bond_order_factors(pysic_ctx *h, int nbond, const double *desc);
/* set_ewald_parameters (PyInterface.f90:1039); scaler may be NULL; on = 0
 * switches the Coulomb summation off.  kcells[a] (the reciprocal cells per
 * axis, K_a) must lie in 0..1023, else PYSIC_EINVAL. */
int pysic_set_ewald(pysic_ctx *h, int on, double real_cutoff, double k_cutoff,
                    const int *kcells, double sigma, double epsilon, const double *scaler);

/* generate_neighbor_lists (PyInterface.f90:1062): per-atom search radii
 * (cutoff + skin) */
int pysic_generate_neighbor_lists(pysic_ctx *h, int n, const double *cutoffs);
/* get_number_of_neighbors_of_atom (PyInterface.f90:1075), for all atoms */
int pysic_get_neighbor_counts(pysic_ctx *h, int n, int *counts);
/* get_neighbor_list_of_atom (PyInterface.f90:1085), for all atoms at once:
 * CSR in atom order, idx[nnz], offsets[3*nnz] */
int pysic_get_neighbor_lists(pysic_ctx *h, long nnz, int *idx, int *offsets);

/* calculate_energy (PyInterface.f90:652) */
int pysic_calculate_energy(pysic_ctx *h, double *energy);
/* No counterpart in the reference: pysic_calculate_energy for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Allocates nothing: the energy
 * pass's per-block partial sums and its total live in a fixed device array
 * of the library, 8 doubles per site-kernel tile for PYSIC_NOALLOC_MAX_TILES
 * tiles plus one row (2 MiB per GPU, held until the process ends); a system
 * needing more tiles (more than 2,097,152 atoms at 4 lanes per atom: tiles =
 * ceil(N * G / 256) for kernel variant g<G>_s<S>) gets PYSIC_ERANGE.  Takes
 * one cell of one species with Tersoff and exponential terms only, after its
 * neighbor lists are generated (potentials set before the atoms exist are
 * compiled by the first evaluation: until pysic_calculate_energy has run
 * once, the call refuses with PYSIC_ESTATE); the result equals
 * pysic_calculate_energy's, bit for bit.  Every call on one
 * GPU shares the array (it keeps the last call's contents; nothing in it
 * needs to start at zero): calls that overlap on one GPU race. */
#define PYSIC_NOALLOC_MAX_TILES 32768
int pysic_calculate_energy_noalloc(pysic_ctx *h, double *energy);
/* calculate_forces (PyInterface.f90:668): forces (N,3), stress(6) =
 * sum r (x) f in the order xx yy zz yz xz xy */
int pysic_calculate_forces(pysic_ctx *h, int n, double *forces, double *stress);
/* calculate_electronegativities (PyInterface.f90:684): -dE/dq of the
 * charge_self, Ewald and charge-dependent pair terms */
int pysic_calculate_electronegativities(pysic_ctx *h, int n, double *enegs);

/* calculate_bond_order_factors (PyInterface.f90:579): the per-atom factors
 * of bond-order group `group` (a Coordinator's index).  The reference sums
 * only per-atom (level-1) factors into them; 'tersoff', the factor type
 * this engine takes, is per-bond (level 2), so they are zero, as the
 * reference's are. */
int pysic_calculate_bond_order_factors(pysic_ctx *h, int n, int group, double *factors);
/* calculate_bond_order_gradients (PyInterface.f90:605): gradients (N,3) of
 * every per-atom factor of the group with respect to moving atom `atom`
 * (0-based; the reference's f2py call takes it 1-based); zero here, as
 * above. */
int pysic_calculate_bond_order_gradients(pysic_ctx *h, int n, int group, int atom,
                                         double *gradients);
/* calculate_bond_order_gradients_of_factor (PyInterface.f90:634): gradients
 * (N,3) of atom `atom`'s factor with respect to moving every atom; zero here,
 * as above. */
int pysic_calculate_bond_order_gradients_of_factor(pysic_ctx *h, int n, int group, int atom,
                                                   double *gradients);
/* get_ewald_energy (PyInterface.f90:1050, Core.f90:8077-8089): the Ewald
 * sum of the current charges with these parameters -- real space over the
 * neighbor lists (which must reach real_cut), self, background, reciprocal
 * and the dipole term |sum q r|^2 / (6 V epsilon) -- with unit charge
 * scalers.  The Ewald parameters set by pysic_set_ewald are not changed. */
int pysic_get_ewald_energy(pysic_ctx *h, double real_cut, double k_cut, const int *kcells,
                           double sigma, double epsilon, double *energy);

/* ChargeRelaxation.charge_relaxation 'dynamic' (charges/relaxation.py:417-450)
 * run on the device; charges/rates in and out (original order).  Returns the
 * step count in *steps and the final max(dchi) in *error. */
int pysic_relax_charges(pysic_ctx *h, int n, double *charges, double *rates, int n_steps,
                        double timestep, double inertia, double friction, double tolerance,
                        int *steps, double *error);

/* ASE VelocityVerlet (usage at doc/run examples.rst:177-268) run on the
 * device for nsteps with the neighbor list rebuilt every `rebuild` steps;
 * energies (nsteps, may be NULL) receives the potential energy after each
 * step.  Between rebuilds the lists are not checked: the caller picks an
 * interval over which no atom moves more than half the skin (the list
 * radius minus the largest cutoff).  rebuild = 1 is always exact. */
int pysic_md_run(pysic_ctx *h, int nsteps, double dt, int rebuild, double *energies);
/* positions and momenta back in original order (after pysic_md_run) */
int pysic_get_state(pysic_ctx *h, int n, double *positions, double *momenta);

/* tuning and measurement */
int pysic_set_kernel_variant(pysic_ctx *h, int variant);
int pysic_kernel_variant_count(void);
const char *pysic_kernel_variant_name(int variant);
int pysic_set_neighbor_capacity(pysic_ctx *h, int cap);
/* neighbor scan: 0 auto, 1 per-atom scan from global memory, 2 LDS-staged
 * bricks, 3 bricks with no LDS room (the global-memory fallback of 2),
 * 4 lane-parallel scan over x-contiguous bin columns, 5 per-atom scan with an
 * FP32 prefilter and the exact FP64 test on survivors (0 selects 5) */
int pysic_set_neighbor_method(pysic_ctx *h, int method);
int pysic_timing(pysic_ctx *h, int on);
/* per kernel family: name, accumulated ms, launch count; returns the count */
int pysic_timing_read(pysic_ctx *h, int max, char *names, int name_len, double *ms, long *calls);
int pysic_synchronize(pysic_ctx *h);

#ifdef __cplusplus
}
#endif

#endif
