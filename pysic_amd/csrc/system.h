// system.h -- device-resident system and the host launchers of each kernel
// family.  One System lives inside each pysic handle; every buffer is owned
// by it and sized for the current atom count (grown, never shrunk).
#pragma once

#include "common.h"

#include <vector>

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;  // elements
    void ensure(size_t k)
    {
        if (k <= cap && p) return;
        if (p) HIP_TRY(hipFree(p));
        p = nullptr;
        size_t want = k < 1 ? 1 : k;
        HIP_TRY(hipMalloc((void **)&p, want * sizeof(T)));
        cap = want;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Kernel timing: events recorded around every launch of a kernel family on
// the system's stream, accumulated only while timing is enabled.
struct KernelTimer {
    bool on = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<std::pair<std::string, double>> total;  // name -> ms
    std::vector<std::pair<std::string, long>> calls;
    void begin(const char *name, hipStream_t s, hipEvent_t *a);
    void end(const char *name, hipStream_t s, hipEvent_t a);
    void collect();
    void reset();
};

struct System {
    hipStream_t stream = nullptr;
    int n = 0;
    CellGeom geom{};

    // atom data in the current (cell-sorted) order; *2 are the swap buffers
    DBuf<double4> pos, pos2;    // x, y, z, species (as a double)
    DBuf<double4> mom, mom2;    // px, py, pz, mass
    DBuf<double4> frc, frc2;    // fx, fy, fz, 0
    DBuf<double> q, q2;         // charges
    DBuf<double> cut, cut2;     // neighbor-search radius per atom (cutoff + skin)
    DBuf<int> id, id2;          // original index of the atom in this slot
    DBuf<int> key, key2;        // cell of each atom (current order)
    DBuf<int> wrap, wrap2;      // packed wrap offsets (current order)
    DBuf<int> perm;             // new slot -> old slot
    DBuf<float4> wpos;          // wrapped positions, FP32 (neighbor prefilter)
    DBuf<int> cellcount, cellstart, cellfill, scanblk;
    DBuf<int> prevstart;        // the previous sort's cell starts (its slot ranges)
    DBuf<int> dirty;            // per bin: an atom entered or left it since the last sort
    DBuf<int> scanloc;          // cell starts within blocks of 1024 bins
    bool order_valid = false;   // the slot order is the last sort's (key[] holds its bins)
    bool unique_images = false; // no atom reaches two images of another within the list radius
    bool bins_clean = false;    // cellcount and dirty are all zero between sorts
    bool qc_uniform = false;    // for this MD run: q, q2 all one value, cut, cut2 too
                                // (the sort then leaves them in place)
    DBuf<int> qc_check;
    double max_cut = 0.0;

    // neighbor lists (column-major: entry e of atom i at e*n + i)
    int cap = 0;
    int nb_method = 0;          // 0 auto (= 5), 1 global scan, 2 LDS bricks, 3 bricks w/o LDS room,
                                // 4 lane groups, 5 FP32-prefiltered scan
    DBuf<int> nbr, nimg, ncnt;
    DBuf<int> flags;            // [0] max list length, [1] max short length, [2] wrap range, [3] unpaired bond

    // Tersoff short lists and per-bond gradients
    int scap = 0;
    DBuf<int> sj, simg, scnt;
    DBuf<int> srow;             // short lists as rows: srow[i * scap + t], -1 past the end
    DBuf<float4> dU;            // dU_i / dr_ij per short-list entry
    DBuf<double> partial;       // per-block energy + virial partials
    int partial_blocks = 0;
    int site_grid = 0;          // > 0: persistent site launch with this many blocks
    DBuf<double> totals;        // reduced energy + virial (7 doubles)

    DBuf<double> host_stage;    // scratch for original-order transfers
    DBuf<double4> out4;

    Model model{};
    DBuf<Model> dmodel;
    bool model_dirty = true;

    // pair potentials of pair.hip and their tables
    PairModel pmodel{};
    DBuf<PairModel> dpmodel;
    DBuf<double2> tables;       // 'tabulated' rows (value, derivative), all tables
    DBuf<double> gpartial;      // per-block energy + virial partials of pair.hip
    int gpartial_blocks = 0;
    DBuf<double> gtotals;       // reduced energy + virial (7 doubles)
    // bond_bend: each vertex's neighbors within the largest bend cutoff
    // (column-major, capacity cap, as the full lists) and the forces its
    // angles put on them
    DBuf<int> bsj, bsimg, bscnt;
    DBuf<double4> bG;
    // bond-order-scaled pair potentials: the model, each atom's neighbors
    // within its radius (column-major, capacity cap), per slot the
    // derivative of the atom's site energy with respect to that bond vector
    // (x, y, z) and to the neighbor's charge (w), and per atom the
    // derivative with respect to its own charge
    BopModel bmodel{};
    DBuf<BopModel> dbmodel;
    DBuf<int> osj, osimg, oscnt;
    DBuf<double4> oD;
    DBuf<double> oW;            // per slot: f_c V dB*/dzeta / 2 of the bond
    DBuf<double> oown;

    // Batch of independent periodic cells (ncells > 0).  The atoms of cell c
    // occupy slots [hatom0[c], hatom0[c+1]) in every order: the bins of each
    // cell form their own key range, so the sort moves atoms within cells.
    int ncells = 0;
    std::vector<CellGeom> hcells;
    std::vector<int> hatom0;
    DBuf<CellGeom> cells;
    DBuf<int> cell_of, first_bin, atom0;
    DBuf<double> atom_ev;       // [7][n]: per-atom site energy and virial
    DBuf<double> cell_ev;       // [ncells][8]: energy, virial per cell
    int nbins = 0;              // bins of the sort over all cells

    KernelTimer timer;

    void release_all();
};

// Calls f with the geometry policy of S (one cell or a batch of cells).
template <class F>
void with_geo(System &S, F &&f)
{
    if (S.ncells > 0)
        f(GeoBatch{S.cells.p, S.cell_of.p, S.first_bin.p});
    else
        f(GeoOne{S.geom});
}

// neighbor.hip
void nb_configure_grid(System &S, double max_cut);
void nb_sort_atoms(System &S, bool drift = false, double dt = 0.0,
                   double half_dt = 0.0);  // [drift +] bin + counting sort + permute
// gather of the last step's Tersoff forces from dU + its second half-kick +
// the next first half-kick + drift, then bin + counting sort + permute
// (the forces are not stored)
void nb_gather_drift_sort(System &S, double dt, double half_dt);
void nb_build_lists(System &S);               // cell-list search, reference semantics
void nb_swap_state(System &S);                // the buffers a sort alternates
int nb_read_flags(System &S, int *out4);      // copies the 4 flags to host
void nb_clear_flags(System &S);               // zero the flags (kernel; graph-safe)
// before an MD run: if every charge and every cutoff is one value, copy them
// to the swap buffers and set S.qc_uniform (clear it after the run)
bool nb_begin_uniform_qc(System &S);

// tersoff.hip
// SITE_ENERGY: energy only (FP64); SITE_FORCES: dU/dr_ij + virial (the
// multi-species kernel computes all three in either mode)
// SITE_FORCES_ONLY: dU/dr_ij without the virial (a velocity-Verlet step)
// SITE_ENERGY_SHORT: SITE_ENERGY over the short list a SITE_FORCES pass on
// the same positions and variant left (one species; else as SITE_ENERGY)
enum { SITE_ENERGY = 1, SITE_FORCES = 2, SITE_FORCES_ONLY = 3, SITE_ENERGY_SHORT = 4 };
void ts_site(System &S, int variant, int mode);
void ts_gather(System &S, double half_dt, int kick);  // forces (+ optional momentum kick)
void ts_reduce(System &S);                    // totals[0..6] = E, W(6)
void ts_cell_sums(System &S);                 // batch: cell_ev[c] = E, W(6) of cell c
int ts_energy_tiles(const System &S, int variant);  // partial rows of the energy pass
// energy pass into caller scratch (one cell, one species; allocates nothing):
// partial[ts_energy_tiles * 8], totals[0] = E
void ts_energy_into(System &S, int variant, double *partial, double *totals);
int ts_variant_count();
const char *ts_variant_name(int v);
int ts_variant_capacity(int v);
int ts_variant_for_capacity(int need);

// pair.hip
// PAIR_ENERGY: gtotals[0] = E; PAIR_FORCES: forces into frc, gtotals[1..6] =
// virial; PAIR_FORCES_ONLY: forces into frc, then p += half_dt F if kick
// (the second half-kick of a velocity-Verlet step, with the total force).
// accumulate: add to frc instead of overwriting it.
enum { PAIR_ENERGY = 0, PAIR_FORCES = 1, PAIR_FORCES_ONLY = 2 };
void pr_eval(System &S, int mode, bool accumulate, double half_dt = 0.0, int kick = 0);
// chi += the charge-dependent pair terms' -dE/dq (nothing without them)
void pr_chi(System &S, double *chi);
// bond_bend angles: energy / virial partials into gpartial blocks
// [blk0, blk0 + blocks), forces added into frc (then p += half_dt F if kick)
void bend_eval(System &S, int mode, int blk0, double half_dt, int kick);
int bend_blocks(const System &S);
// bond-order-scaled pair potentials: the same contract as bend_eval, and
// their share of chi
void bop_eval(System &S, int mode, int blk0, double half_dt, int kick);
int bop_blocks(const System &S, int mode);
void bop_chi(System &S, double *chi);

// md.hip
void md_drift(System &S, double dt, double half_dt);  // p += dt/2 F ; r += dt p/m (no rebuild)
void sys_scatter_original(System &S, const double4 *src, double4 *dst);  // dst[id[s]] = src[s]
