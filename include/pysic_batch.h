/*
 * pysic_batch.h -- many independent periodic cells in one handle.
 *
 * No counterpart in the reference: a pysic calculator holds one structure
 * (calculator.py:1555-1629 pushes one cell and its atoms into the core), so M
 * structures cost M calculators evaluated one after the other.  Here the
 * atoms of all M cells live in one handle and every kernel family runs once
 * over all of them: the cell-list bins of each cell form their own key range
 * (atoms are sorted within their cell), neighbor and Tersoff kernels look up
 * the geometry of each atom's cell, and per-cell energies and virials are
 * summed in a fixed order.
 *
 * Use: pysic_create_atoms with all atoms, cell by cell (cell c owns the next
 * natoms[c] atoms), then pysic_set_cells, pysic_set_potentials /
 * pysic_set_bond_order_factors, pysic_generate_neighbor_lists as for one
 * cell.  pysic_calculate_energy / _forces then return batch totals; the calls
 * below give per-cell results and a device-only evaluation that can be
 * captured into a HIP graph.
 */
#ifndef PYSIC_BATCH_H
#define PYSIC_BATCH_H

#include "pysic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ncells cells: natoms[c] atoms each (sum = atoms created), vectors and
 * inverses 9 doubles per cell (rows; inverse = M^-1 row-major, as
 * pysic_create_cell), pbc 3 ints per cell.  pysic_create_cell or
 * pysic_create_atoms return the handle to a single cell. */
int pysic_set_cells(pysic_ctx *h, int ncells, const int *natoms, const double *vectors,
                    const double *inverses, const int *pbc);
int pysic_get_number_of_cells(pysic_ctx *h, int *ncells);
/* Persistent launch of the Tersoff site kernels: at most `blocks` workgroups
 * walk all atom tiles (0: one workgroup per tile). */
int pysic_set_site_grid(pysic_ctx *h, int blocks);
/* One device evaluation: [rebuild != 0: cell lists and neighbor lists],
 * Tersoff forces and virials, FP64 energies, per-cell sums.  Nothing is
 * copied to the host and nothing is checked; with use_graph != 0 the launch
 * sequence is captured into a HIP graph once (per position buffer) and
 * replayed after.  Results and overflow checks: pysic_get_cell_results. */
int pysic_evaluate_cells(pysic_ctx *h, int use_graph, int rebuild);
/* The four device flags (max list length, max short-list length, wrap range,
 * unpaired bond) of the last evaluation. */
int pysic_check_cells(pysic_ctx *h, int *flags4);
/* Energies[ncells], virials[ncells][6] (xx yy zz yz xz xy, eV) and forces
 * [n][3] in the caller's atom order (any output may be null).  Fails with
 * PYSIC_ECAP if the last evaluation overflowed a list capacity. */
int pysic_get_cell_results(pysic_ctx *h, int ncells, double *energies, double *virials, int n,
                           double *forces);

#ifdef __cplusplus
}
#endif

#endif
