// neighbor.hip -- cell-list neighbor build on the MI355X.
//
// Replaces core_create_space_partitioning / core_build_neighbor_lists
// (Core.f90:8130-8280) with a binning pass, a stable counting sort of the
// atoms by cell (atoms are physically permuted so that memory order is
// spatial order), and a per-atom scan of the 27 adjacent cells.
//
// The list produced is the reference's list as a set: every (j, o) with
//   |r_j - r_i + o1 v1 + o2 v2 + o3 v3|^2 < cutoff_i^2,   (j, o) != (i, 0),
// where o = (cell-image shift) - wrap_i + wrap_j exactly as in
// Core.f90:8207-8238, and the squared distance is formed in FP64 in the
// reference's operation order (Geometry.f90:461-464, Quaternions.f90:480)
// with contraction disabled, so the accept/reject decision is bit-identical.
// Bins are at least max_cut wide (so 27 cells are complete) but are not
// capped at 40 per axis as in Geometry.f90:703: the set does not depend on
// the partition, only the cost does.
#include "system.h"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ void frac_of(const CellGeom &g, double x, double y, double z,
                                        double *rel)
{
    // relative_coordinates: matmul(inverse, position) (Geometry.f90:527-534)
    rel[0] = g.inv[0] * x + g.inv[1] * y + g.inv[2] * z;
    rel[1] = g.inv[3] * x + g.inv[4] * y + g.inv[5] * z;
    rel[2] = g.inv[6] * x + g.inv[7] * y + g.inv[8] * z;
}

// Bin one atom: wrap offset -floor(frac) on periodic axes
// (Geometry.f90:597-619) and the bin of the wrapped fractional coordinate.
// Non-finite or far-away coordinates are flagged and binned at 0.
__device__ __forceinline__ void bin_atom(const CellGeom &g, const double4 &p, int *key, int *wrap,
                                         bool *bad_out)
{
    double rel[3];
    frac_of(g, p.x, p.y, p.z, rel);
    int w[3], c[3];
    bool bad = false;
    for (int a = 0; a < 3; ++a) {
        if (g.pbc[a]) {
            double fl = floor(rel[a]);
            bool ok = fabs(fl) <= (double)PYSIC_WRAP_MAX;
            bad |= !ok;
            w[a] = ok ? -(int)fl : 0;
            double fb = ok ? floor((rel[a] + w[a]) * g.nc[a]) : 0.0;
            fb = fmin(fmax(fb, 0.0), (double)(g.nc[a] - 1));
            c[a] = (int)fb;
        } else {
            w[a] = 0;
            double fb = floor(rel[a] * g.nc[a]) + 1.0;
            bool ok = fb == fb;  // not NaN
            bad |= !ok;
            fb = ok ? fmin(fmax(fb, 0.0), (double)(g.nc[a] + 1)) : 0.0;
            c[a] = (int)fb;
        }
    }
    *key = c[0] + g.dim[0] * (c[1] + g.dim[1] * c[2]);
    *wrap = pack_img(w[0], w[1], w[2]);
    *bad_out = bad;
}

// Records the new bin k of the atom in slot i.  With track (the slot order
// is the last sort's, so key[i] still holds the bin the atom was sorted
// into), a change of bin marks both bins dirty: the sort then re-sorts only
// the dirty bins and moves the atoms of the others as blocks.
__device__ __forceinline__ void record_bin(int i, int k, int w, int nbins, int track,
                                           int *__restrict__ key, int *__restrict__ wrap,
                                           int *__restrict__ cellcount, int *__restrict__ dirty)
{
    if (track) {
        const int old = key[i];
        if (old != k) {
            dirty[k] = 1;
            if ((unsigned)old < (unsigned)nbins) dirty[old] = 1;
        }
    }
    key[i] = k;
    wrap[i] = w;
    atomicAdd(cellcount + k, 1);
}

template <class Geo>
__global__ void k_bin(int n, Geo geo, const double4 *__restrict__ pos, int *__restrict__ key,
                      int *__restrict__ wrap, int *__restrict__ cellcount, int *__restrict__ flags,
                      int nbins, int track, int *__restrict__ dirty)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k, w;
    bool bad;
    bin_atom(geo.of(i), pos[i], &k, &w, &bad);
    k += geo.bin0(i);
    if (bad) atomicOr(flags + 2, 1);
    record_bin(i, k, w, nbins, track, key, wrap, cellcount, dirty);
}

// velocity-Verlet first half (p += dt/2 F; r += dt p/m) fused with binning
// One velocity-Verlet step boundary in one pass, for the MD loop of a model
// whose forces are all Tersoff-kernel forces: gather the forces of the step
// that ended from the per-bond gradients (gather_bond_force), its second
// half-kick, the next step's first half-kick, drift, and binning.  The same
// arithmetic in the same order as k_gather (kick as an fma) followed by
// k_drift_bin, so the trajectory is bit-identical; the forces are never
// stored, and momenta and positions are read and written once.
template <class Geo>
__global__ void k_gather_drift_bin(int n, Geo geo, double dt, double half_dt,
                                   const int *__restrict__ sj, const int *__restrict__ simg,
                                   const int *__restrict__ scnt, const int *__restrict__ srow,
                                   int scap, int unique, const float4 *__restrict__ dU,
                                   double4 *__restrict__ mom, double4 *__restrict__ pos,
                                   int *__restrict__ key, int *__restrict__ wrap,
                                   int *__restrict__ cellcount, int *__restrict__ flags,
                                   int nbins, int track, int *__restrict__ dirty)
{
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = mom[i];   // loaded first: in flight during the gather
    double4 r = pos[i];
    bool miss = false;
    const double3 f = unique ? gather_rows_any(i, n, scap, srow, dU, miss)
                             : gather_bond_force(i, n, sj, simg, scnt, dU, miss);
    // flags[3]: bit 0 a bond without its reverse, bit 1 a non-finite force
    const int fbad = (miss ? 1 : 0) |
                     ((isfinite(f.x) && isfinite(f.y) && isfinite(f.z)) ? 0 : 2);
    if (fbad) atomicOr(flags + 3, fbad);
    p.x = fma(half_dt, f.x, p.x);
    p.y = fma(half_dt, f.y, p.y);
    p.z = fma(half_dt, f.z, p.z);
    p.x += half_dt * f.x;
    p.y += half_dt * f.y;
    p.z += half_dt * f.z;
    mom[i] = p;
    double s = dt / p.w;  // p.w carries the mass
    r.x += s * p.x;
    r.y += s * p.y;
    r.z += s * p.z;
    pos[i] = r;
    int k, w;
    bool bad;
    bin_atom(geo.of(i), r, &k, &w, &bad);
    k += geo.bin0(i);
    if (bad) atomicOr(flags + 2, 1);
    record_bin(i, k, w, nbins, track, key, wrap, cellcount, dirty);
}

template <class Geo>
__global__ void k_drift_bin(int n, Geo geo, double dt, double half_dt,
                            const double4 *__restrict__ frc, double4 *__restrict__ mom,
                            double4 *__restrict__ pos, int *__restrict__ key,
                            int *__restrict__ wrap, int *__restrict__ cellcount,
                            int *__restrict__ flags, int nbins, int track,
                            int *__restrict__ dirty)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 f = frc[i];
    double4 p = mom[i];
    p.x += half_dt * f.x;
    p.y += half_dt * f.y;
    p.z += half_dt * f.z;
    mom[i] = p;
    double4 r = pos[i];
    double s = dt / p.w;  // p.w carries the mass
    r.x += s * p.x;
    r.y += s * p.y;
    r.z += s * p.z;
    pos[i] = r;
    int k, w;
    bool bad;
    bin_atom(geo.of(i), r, &k, &w, &bad);
    k += geo.bin0(i);
    if (bad) atomicOr(flags + 2, 1);
    record_bin(i, k, w, nbins, track, key, wrap, cellcount, dirty);
}

// --- exclusive scan of the cell counts (1024 elements per block, then the
// block totals; a bin's start is loc[k] + blk[k / SCAN_BLK]) ---------------
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_BLK = SCAN_T * SCAN_PER;

__device__ int block_exclusive_scan(int v, int *lds, int *total)
{
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < SCAN_T; d <<= 1) {
        int x = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += x;
        __syncthreads();
    }
    int incl = lds[t];
    *total = lds[SCAN_T - 1];
    __syncthreads();
    return incl - v;
}

// Also zeroes the counts it reads and the scatter's fill counters, so the
// next binning and this sort's scatter start from zero without a fill pass.
__global__ void k_scan_local(int m, int *__restrict__ in, int *__restrict__ out,
                             int *__restrict__ blk, int *__restrict__ fill)
{
    __shared__ int lds[SCAN_T];
    int base = blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER], s = 0;
    for (int q = 0; q < SCAN_PER; ++q) {
        v[q] = base + q < m ? in[base + q] : 0;
        if (base + q < m) {
            in[base + q] = 0;
            fill[base + q] = 0;
        }
        s += v[q];
    }
    int tot;
    int ex = block_exclusive_scan(s, lds, &tot);
    for (int q = 0; q < SCAN_PER; ++q) {
        if (base + q < m) out[base + q] = ex;
        ex += v[q];
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// single block: exclusive scan of the block totals, sequential chunks
__global__ void k_scan_blocks(int nb, int *__restrict__ blk, int *__restrict__ out, int m)
{
    __shared__ int lds[SCAN_T];
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_T) {
        int t = base + threadIdx.x;
        int v = t < nb ? blk[t] : 0;
        int tot;
        int ex = block_exclusive_scan(v, lds, &tot);
        if (t < nb) blk[t] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) out[m] = carry;
}

__device__ __forceinline__ int bin_start(int k, const int *__restrict__ loc,
                                         const int *__restrict__ blk)
{
    return loc[k] + blk[k / SCAN_BLK];
}

// zero two bin arrays of length m (a kernel, so that graphs hold no memset)
__global__ void k_fill_int(int m, int *__restrict__ a, int *__restrict__ b)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) a[t] = 0;
    else if (t < 2 * m) b[t - m] = 0;
}

__global__ void k_clear_flags(int *__restrict__ flags)
{
    if (threadIdx.x < 4 && blockIdx.x == 0) flags[threadIdx.x] = 0;
}

// New slot of the atom in old slot i.  A bin no atom entered or left keeps
// its atoms in the same order (sorted by original index), so they move as a
// block: new start + (i - old start).  The atoms of a dirty bin claim slots
// with atomics and k_cellsort orders them.  all_dirty: every bin is dirty
// (the old order is not a previous sort's).
__global__ void k_scatter(int n, const int *__restrict__ key, const int *__restrict__ loc,
                          const int *__restrict__ blk, const int *__restrict__ prevstart,
                          const int *__restrict__ dirty, int all_dirty, int *__restrict__ fill,
                          int *__restrict__ perm, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = key[i];
    int start = bin_start(k, loc, blk);
    int slot = (all_dirty || dirty[k]) ? start + atomicAdd(fill + k, 1)
                                       : start + (i - prevstart[k]);
    if ((unsigned)slot < (unsigned)n) perm[slot] = i;
    else atomicOr(flags + 2, 2);   // an inconsistent sort: reported, never written
}

// Within each dirty cell, order atoms by original index: the atomic slot
// claims above are unordered, this makes the whole permutation
// deterministic (the same as sorting every cell).  Writes the final cell
// starts and clears the dirty flags for the next sort.
__global__ void k_cellsort(int ncell, int n, const int *__restrict__ loc,
                           const int *__restrict__ blk, int *__restrict__ perm,
                           const int *__restrict__ id, int *__restrict__ dirty, int all_dirty,
                           int *__restrict__ start)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int a = bin_start(c, loc, blk), b = c + 1 < ncell ? bin_start(c + 1, loc, blk) : n;
    start[c] = a;
    if (c == ncell - 1) start[ncell] = n;
    if (!(all_dirty || dirty[c])) return;
    dirty[c] = 0;
    for (int x = a + 1; x < b; ++x) {
        int v = perm[x];
        int kv = id[v];
        int y = x - 1;
        while (y >= a && id[perm[y]] > kv) {
            perm[y + 1] = perm[y];
            --y;
        }
        perm[y + 1] = v;
    }
}

// Forces are not carried: after the drift they are dead until recomputed.
// QC: move the charges and cutoffs too.  Without it they stay where they
// are, which is right only while both buffers of each hold one value for
// every atom (checked before a run that uses it).
template <class Geo, bool QC>
__global__ void k_permute(int n, const int *__restrict__ perm, const double4 *__restrict__ pos,
                          const double4 *__restrict__ mom, const double *__restrict__ q,
                          const double *__restrict__ cut, const int *__restrict__ id,
                          const int *__restrict__ key, const int *__restrict__ wrap,
                          double4 *__restrict__ pos2, double4 *__restrict__ mom2,
                          double *__restrict__ q2, double *__restrict__ cut2,
                          int *__restrict__ id2, int *__restrict__ key2, int *__restrict__ wrap2,
                          Geo geo, float4 *__restrict__ wpos, int *__restrict__ flags)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const CellGeom &g = geo.of(s);   // slot s stays in the cell of slot perm[s]
    int o = perm[s];
    if ((unsigned)o >= (unsigned)n) {
        atomicOr(flags + 2, 2);   // an inconsistent sort (reported): keep the slot
        o = s;
    }
    double4 p = pos[o];
    int w = wrap[o];
    pos2[s] = p;
    mom2[s] = mom[o];
    if constexpr (QC) {
        q2[s] = q[o];
        cut2[s] = cut[o];
    }
    id2[s] = id[o];
    key2[s] = key[o];
    wrap2[s] = w;
    // wrapped position in FP32 for the neighbor prefilter
    double wx = img_x(w), wy = img_y(w), wz = img_z(w);
    // w carries the packed wrap offsets (bits), so the prefilter forms the
    // list image without loading wrap[]
    wpos[s] = make_float4((float)(p.x + wx * g.v[0] + wy * g.v[3] + wz * g.v[6]),
                          (float)(p.y + wx * g.v[1] + wy * g.v[4] + wz * g.v[7]),
                          (float)(p.z + wx * g.v[2] + wy * g.v[5] + wz * g.v[8]),
                          __int_as_float(w));
}

// Per-atom scan with an FP32 prefilter.  Every candidate of the 27 bins is
// tested on FP32 wrapped positions (16 B per candidate, the wrap offsets in
// the fourth word); margin2 bounds the FP32 error of the squared distance.
// A candidate with d2 < cutoff^2 - 2 margin2 is a neighbor for sure, one
// with d2 >= cutoff^2 + margin2 is not; the thin shell between takes the
// exact FP64 test in the reference's operation order, so the accepted set is
// the exact one.  Accepted and shell candidates are queued (slot, image,
// shell flag) in this lane's LDS column and written out by drain(): on a
// SIMD, writing (or testing) them inline would enter that branch whenever
// any lane of the wave has one, which is almost every candidate; queued,
// every lane works through its entries together.
constexpr int NB_QUEUE = 12;
#ifndef PYSIC_NB_UNIFORM_SPAN
#define PYSIC_NB_UNIFORM_SPAN 4
#endif

template <class Geo, bool DENSE>
__global__ __launch_bounds__(256) void k_neighbors_pf(int n, Geo geo,
                                                      const double4 *__restrict__ pos,
                                                      const float4 *__restrict__ wpos,
                                                      const int *__restrict__ key,
                                                      const int *__restrict__ wrap,
                                                      const int *__restrict__ start,
                                                      const double *__restrict__ cut, int cap,
                                                      float margin2, int *__restrict__ nbr,
                                                      int *__restrict__ nimg,
                                                      int *__restrict__ ncnt,
                                                      int *__restrict__ flags)
{
    __shared__ int q_slot[NB_QUEUE][256];
    __shared__ int q_shift[NB_QUEUE][256];
    int s = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;   // no barrier below: lanes only touch their own LDS column
    const int tid = threadIdx.x;
    const CellGeom g = geo.of(s);
    const int b0 = geo.bin0(s);
    start += b0;
    const double4 ps = pos[s];
    const float4 qs = wpos[s];
    const int k = key[s] - b0;
    const int c0 = k % g.dim[0];
    const int c1 = (k / g.dim[0]) % g.dim[1];
    const int c2 = k / (g.dim[0] * g.dim[1]);
    const int ws = wrap[s];
    const int wsx = img_x(ws), wsy = img_y(ws), wsz = img_z(ws);
    const double cs = cut[s];
    const double c2lim = cs * cs;
    const float c2f = (float)c2lim + margin2;
    const float v0x = (float)g.v[0], v0y = (float)g.v[1], v0z = (float)g.v[2];
    const float v1x = (float)g.v[3], v1y = (float)g.v[4], v1z = (float)g.v[5];
    const float v2x = (float)g.v[6], v2y = (float)g.v[7], v2z = (float)g.v[8];
    const float c2sure = (float)c2lim - 2.0f * margin2;
    int count = 0, nq = 0;
    constexpr int SHELL = 1 << 30;  // queue flag: exact test needed (images use bits 0-29)
    // write out the queued entries, in queue order.  Sparse bins (a few hits
    // among many candidates): an entry holds the candidate and its bin's
    // periodic shift, and the pair's image, bin shift + wrap(t) - wrap(s), is
    // formed here, once per accepted candidate rather than in the scan loop,
    // where the branch runs whenever any lane of the wave has a hit; the atom
    // itself at image 0 is dropped here.  Dense bins (most candidates hit):
    // the image is formed at the hit, as the drain would otherwise reload
    // each entry's wrap.  Shell entries first pass the exact test.
    auto drain = [&]() {
        for (int u = 0; u < nq; ++u) {
            const int t = q_slot[u][tid];
            const int q = q_shift[u][tid];
            int im = q & ~SHELL;
            if constexpr (!DENSE) {
                const int wt = __float_as_int(wpos[t].w);
                const int ox = img_x(im) - wsx + img_x(wt), oy = img_y(im) - wsy + img_y(wt),
                          oz = img_z(im) - wsz + img_z(wt);
                if (t == s && ox == 0 && oy == 0 && oz == 0) continue;
                im = pack_img(ox, oy, oz);
            }
            if ((q & SHELL) && !(dot64(separation64(ps, pos[t], im, g.v)) < c2lim)) continue;
            if (count < cap) {
                nbr[(size_t)count * n + s] = t;
                nimg[(size_t)count * n + s] = im;
            }
            ++count;
        }
        nq = 0;
    };
    // FP32 test of every candidate of one bin (row + bx) under the bin's
    // periodic shift (sx, sy, sz).  UNIFORM: every lane of the wave scans the
    // same candidates, so a full queue drains every lane's queue at once.
    auto scan_bin = [&](auto uniform_tag, int row, int bx, int sx, int sy, int sz) {
        constexpr bool UNIFORM = decltype(uniform_tag)::value;
        const float sxf = (float)sx, syf = (float)sy, szf = (float)sz;
        const float shx = sxf * v0x + syf * v1x + szf * v2x - qs.x;
        const float shy = sxf * v0y + syf * v1y + szf * v2y - qs.y;
        const float shz = sxf * v0z + syf * v1z + szf * v2z - qs.z;
        const int sh = pack_img(sx, sy, sz);
        const int ox0 = sx - wsx, oy0 = sy - wsy, oz0 = sz - wsz;
        const int t1 = start[row + bx + 1];
        for (int t = start[row + bx]; t < t1; ++t) {
            const float4 qt = wpos[t];
            const float fx = qt.x + shx, fy = qt.y + shy, fz = qt.z + shz;
            const float d2f = fx * fx + fy * fy + fz * fz;
            const bool hit = d2f < c2f;
            if (UNIFORM) {
                if (__ballot(hit && nq == NB_QUEUE)) drain();
            }
            if (hit) {
                if (!UNIFORM && nq == NB_QUEUE) drain();
                const int flag = d2f < c2sure ? 0 : SHELL;
                if constexpr (DENSE) {
                    const int wt = __float_as_int(qt.w);
                    const int ox = ox0 + img_x(wt), oy = oy0 + img_y(wt), oz = oz0 + img_z(wt);
                    if (!(t == s && ox == 0 && oy == 0 && oz == 0)) {
                        q_slot[nq][tid] = t;
                        q_shift[nq][tid] = pack_img(ox, oy, oz) | flag;
                        ++nq;
                    }
                } else {
                    q_slot[nq][tid] = t;
                    q_shift[nq][tid] = sh | flag;
                    ++nq;
                }
            }
        }
    };
    // Wave-uniform scan: when the wave's atoms share one bin row of a
    // periodic cell at least 3 bins wide on every axis, and span at most
    // PYSIC_NB_UNIFORM_SPAN + 1 bins of it (dense bins: the union of their
    // stencils is then at most (SPAN + 3) / 3 times one stencil), every lane
    // scans that union -- the 9 neighboring rows over x from one bin before
    // the wave's first bin to one after its last.  Loop bounds
    // and candidates are then the same for all lanes (scalar loads, no
    // divergence); the extra candidates lie beyond every lane's cutoff or
    // are genuine neighbors, and with 3 or more bins per axis no atom image
    // is reached twice, so the accepted set is the 27-bin stencil's.
    bool uniform = false;
    int ux0 = 0, ux1 = 0, uc1 = 0, uc2 = 0;
    if constexpr (DENSE) {  // one periodic cell, >= 3 bins per axis (checked by the host)
        {
            const unsigned long long act = __ballot(1);
            const int lo = __ffsll((unsigned long long)act) - 1;
            const int hi = 63 - __clzll((long long)act);
            // atoms are sorted by bin: the first active lane holds the
            // wave's lowest bin, the last its highest
            const int kmin = __builtin_amdgcn_readfirstlane(__shfl(k, lo));
            const int kmax = __builtin_amdgcn_readfirstlane(__shfl(k, hi));
            if (kmin / g.dim[0] == kmax / g.dim[0] &&
                kmax - kmin <= PYSIC_NB_UNIFORM_SPAN) {
                uniform = true;
                ux0 = kmin % g.dim[0];
                ux1 = kmax % g.dim[0];
                uc1 = (kmin / g.dim[0]) % g.dim[1];
                uc2 = kmin / (g.dim[0] * g.dim[1]);
            }
        }
    }
    if (uniform) {
        for (int dz = -1; dz <= 1; ++dz) {
            int bz = uc2 + dz, sz = 0;
            if (bz < 0) { bz += g.nc[2]; sz = -1; }
            else if (bz >= g.nc[2]) { bz -= g.nc[2]; sz = 1; }
            for (int dy = -1; dy <= 1; ++dy) {
                int by = uc1 + dy, sy = 0;
                if (by < 0) { by += g.nc[1]; sy = -1; }
                else if (by >= g.nc[1]) { by -= g.nc[1]; sy = 1; }
                const int row = g.dim[0] * (by + g.dim[1] * bz);
                for (int x = ux0 - 1; x <= ux1 + 1; ++x) {
                    int bx = x, sx = 0;
                    if (bx < 0) { bx += g.nc[0]; sx = -1; }
                    else if (bx >= g.nc[0]) { bx -= g.nc[0]; sx = 1; }
                    scan_bin(std::true_type{}, row, bx, sx, sy, sz);
                }
            }
        }
    } else {
        for (int dz = -1; dz <= 1; ++dz) {
            int bz = c2 + dz, sz = 0;
            if (g.pbc[2]) {
                if (bz < 0) { bz += g.nc[2]; sz = -1; }
                else if (bz >= g.nc[2]) { bz -= g.nc[2]; sz = 1; }
            } else if (bz < 0 || bz >= g.dim[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int by = c1 + dy, sy = 0;
                if (g.pbc[1]) {
                    if (by < 0) { by += g.nc[1]; sy = -1; }
                    else if (by >= g.nc[1]) { by -= g.nc[1]; sy = 1; }
                } else if (by < 0 || by >= g.dim[1]) continue;
                const int row = g.dim[0] * (by + g.dim[1] * bz);
                for (int dx = -1; dx <= 1; ++dx) {
                    int bx = c0 + dx, sx = 0;
                    if (g.pbc[0]) {
                        if (bx < 0) { bx += g.nc[0]; sx = -1; }
                        else if (bx >= g.nc[0]) { bx -= g.nc[0]; sx = 1; }
                    } else if (bx < 0 || bx >= g.dim[0]) continue;
                    scan_bin(std::false_type{}, row, bx, sx, sy, sz);
                }
            }
        }
    }
    drain();
    ncnt[s] = count < cap ? count : cap;
    if (count > cap) atomicMax(flags, count);
}

// Per-atom scan of the 27 adjacent bins (Core.f90:8182-8252 as a full list).
template <class Geo>
__global__ __launch_bounds__(256) void k_neighbors(int n, Geo geo, const double4 *__restrict__ pos,
                                                   const int *__restrict__ key,
                                                   const int *__restrict__ wrap,
                                                   const int *__restrict__ start,
                                                   const double *__restrict__ cut, int cap,
                                                   int *__restrict__ nbr, int *__restrict__ nimg,
                                                   int *__restrict__ ncnt, int *__restrict__ flags)
{
    int s = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const CellGeom g = geo.of(s);
    const int b0 = geo.bin0(s);
    start += b0;
    double4 ps = pos[s];
    int k = key[s] - b0;
    int c0 = k % g.dim[0];
    int c1 = (k / g.dim[0]) % g.dim[1];
    int c2 = k / (g.dim[0] * g.dim[1]);
    int ws = wrap[s];
    int wsx = img_x(ws), wsy = img_y(ws), wsz = img_z(ws);
    double cs = cut[s];
    double c2lim = cs * cs;
    int count = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        int bz = c2 + dz, sz = 0;
        if (g.pbc[2]) {
            if (bz < 0) { bz += g.nc[2]; sz = -1; }
            else if (bz >= g.nc[2]) { bz -= g.nc[2]; sz = 1; }
        } else if (bz < 0 || bz >= g.dim[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int by = c1 + dy, sy = 0;
            if (g.pbc[1]) {
                if (by < 0) { by += g.nc[1]; sy = -1; }
                else if (by >= g.nc[1]) { by -= g.nc[1]; sy = 1; }
            } else if (by < 0 || by >= g.dim[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                int bx = c0 + dx, sx = 0;
                if (g.pbc[0]) {
                    if (bx < 0) { bx += g.nc[0]; sx = -1; }
                    else if (bx >= g.nc[0]) { bx -= g.nc[0]; sx = 1; }
                } else if (bx < 0 || bx >= g.dim[0]) continue;
                int cell = bx + g.dim[0] * (by + g.dim[1] * bz);
                int t1 = start[cell + 1];
                for (int t = start[cell]; t < t1; ++t) {
                    int wt = wrap[t];
                    int ox = sx + img_x(wt) - wsx;
                    int oy = sy + img_y(wt) - wsy;
                    int oz = sz + img_z(wt) - wsz;
                    if (t == s && ox == 0 && oy == 0 && oz == 0) continue;
                    int im = pack_img(ox, oy, oz);
                    double d2 = dot64(separation64(ps, pos[t], im, g.v));
                    if (d2 < c2lim) {
                        if (count < cap) {
                            nbr[(size_t)count * n + s] = t;
                            nimg[(size_t)count * n + s] = im;
                        }
                        ++count;
                    }
                }
            }
        }
    }
    ncnt[s] = count < cap ? count : cap;
    if (count > cap) atomicMax(flags, count);
}


// ---------------------------------------------------------------------------
// LDS-staged brick scan.  One workgroup owns a brick of Bx*By*Bz bins; the
// positions of every atom in the brick and its one-bin halo are loaded into
// LDS once (coalesced, each halo atom ~3x instead of once per neighbor
// candidate), then every atom of the brick scans its 27 bins out of LDS.
// Bricks whose halo exceeds the LDS capacity take the same scan from global
// memory, so the result never depends on the capacity.
// ---------------------------------------------------------------------------
struct BrickCfg {
    int B[3];       // bins per brick edge
    int nb[3];      // bricks per axis
    int lcap;       // halo atoms held in LDS
};

__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

__device__ __forceinline__ int upper_bound_idx(const int *arr, int len, int x)
{
    // largest h in [0, len) with arr[h] <= x (arr ascending, arr[0] = 0)
    int lo = 0, hi = len;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (arr[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ int block_scan_inplace(int *arr, int len, int *lds_tmp)
{
    // exclusive scan of arr[0..len) in place; returns the total; blockDim 256
    int carry = 0;
    for (int base = 0; base < len; base += SCAN_T) {
        int t = base + threadIdx.x;
        int v = t < len ? arr[t] : 0;
        int tot;
        int ex = block_exclusive_scan(v, lds_tmp, &tot);
        if (t < len) arr[t] = ex + carry;
        carry += tot;
        __syncthreads();
    }
    return carry;
}

__global__ __launch_bounds__(256) void k_neighbors_brick(int n, CellGeom g, BrickCfg bc,
                                                         const double4 *__restrict__ pos,
                                                         const int *__restrict__ wrap,
                                                         const int *__restrict__ start,
                                                         const double *__restrict__ cut, int cap,
                                                         int *__restrict__ nbr,
                                                         int *__restrict__ nimg,
                                                         int *__restrict__ ncnt,
                                                         int *__restrict__ flags)
{
    extern __shared__ double4 smem4[];
    const int Hx = bc.B[0] + 2, Hy = bc.B[1] + 2, Hz = bc.B[2] + 2;
    const int H = Hx * Hy * Hz;
    const int C = bc.B[0] * bc.B[1] * bc.B[2];
    double4 *s_pos = smem4;                              // [lcap]
    int *s_tag = (int *)(s_pos + bc.lcap);               // [lcap] shift + wrap, packed
    int *s_gid = s_tag + bc.lcap;                        // [lcap]
    int *s_hstart = s_gid + bc.lcap;                     // [H]
    int *s_hcnt = s_hstart + H;                          // [H]
    int *s_hloc = s_hcnt + H;                            // [H + 1]
    int *s_hshift = s_hloc + H + 1;                      // [H]
    int *s_cloc = s_hshift + H;                          // [C + 1]
    __shared__ int s_tmp[SCAN_T];

    int b = blockIdx.x;
    int b0 = b % bc.nb[0], b1 = (b / bc.nb[0]) % bc.nb[1], b2 = b / (bc.nb[0] * bc.nb[1]);
    int o0 = b0 * bc.B[0], o1 = b1 * bc.B[1], o2 = b2 * bc.B[2];

    // ---- halo bins: global bin, image shift, atom range -------------------------
    for (int h = threadIdx.x; h < H; h += blockDim.x) {
        int hx = h % Hx, hy = (h / Hx) % Hy, hz = h / (Hx * Hy);
        int gc[3] = {o0 + hx - 1, o1 + hy - 1, o2 + hz - 1};
        int sh[3] = {0, 0, 0};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            if (g.pbc[a]) {
                sh[a] = floor_div(gc[a], g.nc[a]);
                gc[a] -= sh[a] * g.nc[a];
            } else if (gc[a] < 0 || gc[a] >= g.dim[a]) {
                ok = false;
            }
        }
        int cnt = 0, st0 = 0;
        if (ok) {
            int cell = gc[0] + g.dim[0] * (gc[1] + g.dim[1] * gc[2]);
            st0 = start[cell];
            cnt = start[cell + 1] - st0;
        }
        s_hstart[h] = st0;
        s_hcnt[h] = cnt;
        s_hloc[h] = cnt;
        s_hshift[h] = pack_img(sh[0], sh[1], sh[2]);
    }
    // center bins: real only inside the grid (no wrap)
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        int cx = c % bc.B[0], cy = (c / bc.B[0]) % bc.B[1], cz = c / (bc.B[0] * bc.B[1]);
        bool real = o0 + cx < g.dim[0] && o1 + cy < g.dim[1] && o2 + cz < g.dim[2];
        int h = (cx + 1) + Hx * ((cy + 1) + Hy * (cz + 1));
        s_cloc[c] = real ? 0 : -1;  // marker, filled below
        (void)h;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        int cx = c % bc.B[0], cy = (c / bc.B[0]) % bc.B[1], cz = c / (bc.B[0] * bc.B[1]);
        int h = (cx + 1) + Hx * ((cy + 1) + Hy * (cz + 1));
        s_cloc[c] = s_cloc[c] < 0 ? 0 : s_hcnt[h];
    }
    __syncthreads();
    int T = block_scan_inplace(s_hloc, H, s_tmp);
    int Cn = block_scan_inplace(s_cloc, C, s_tmp);
    if (threadIdx.x == 0) {
        s_hloc[H] = T;
        s_cloc[C] = Cn;
    }
    __syncthreads();
    const bool in_lds = T <= bc.lcap;

    // ---- stage the halo atoms ---------------------------------------------------
    if (in_lds) {
        for (int x = threadIdx.x; x < T; x += blockDim.x) {
            int h = upper_bound_idx(s_hloc, H, x);
            // skip empty bins that share the same offset
            while (h + 1 < H && s_hloc[h + 1] <= x) ++h;
            int t = s_hstart[h] + (x - s_hloc[h]);
            int sh = s_hshift[h], wt = wrap[t];
            s_pos[x] = pos[t];
            s_tag[x] = pack_img(img_x(sh) + img_x(wt), img_y(sh) + img_y(wt), img_z(sh) + img_z(wt));
            s_gid[x] = t;
        }
    }
    __syncthreads();

    // ---- every atom of the brick scans its 27 bins ------------------------------
    for (int x = threadIdx.x; x < Cn; x += blockDim.x) {
        int c = upper_bound_idx(s_cloc, C, x);
        while (c + 1 < C && s_cloc[c + 1] <= x) ++c;
        int cx = c % bc.B[0], cy = (c / bc.B[0]) % bc.B[1], cz = c / (bc.B[0] * bc.B[1]);
        int h = (cx + 1) + Hx * ((cy + 1) + Hy * (cz + 1));
        int q = x - s_cloc[c];
        int s = s_hstart[h] + q;
        double4 ps = in_lds ? s_pos[s_hloc[h] + q] : pos[s];
        int ws = wrap[s];
        int wsx = img_x(ws), wsy = img_y(ws), wsz = img_z(ws);
        double cs = cut[s];
        double c2lim = cs * cs;
        int count = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int hn = h + dx + Hx * (dy + Hy * dz);
                    int m = s_hcnt[hn];
                    if (in_lds) {
                        int l0 = s_hloc[hn];
                        for (int u = 0; u < m; ++u) {
                            int tag = s_tag[l0 + u];
                            int t = s_gid[l0 + u];
                            int ox = img_x(tag) - wsx, oy = img_y(tag) - wsy, oz = img_z(tag) - wsz;
                            if (t == s && ox == 0 && oy == 0 && oz == 0) continue;
                            int im = pack_img(ox, oy, oz);
                            double d2 = dot64(separation64(ps, s_pos[l0 + u], im, g.v));
                            if (d2 < c2lim) {
                                if (count < cap) {
                                    nbr[(size_t)count * n + s] = t;
                                    nimg[(size_t)count * n + s] = im;
                                }
                                ++count;
                            }
                        }
                    } else {
                        int sh = s_hshift[hn];
                        int t0 = s_hstart[hn];
                        for (int u = 0; u < m; ++u) {
                            int t = t0 + u;
                            int wt = wrap[t];
                            int ox = img_x(sh) + img_x(wt) - wsx;
                            int oy = img_y(sh) + img_y(wt) - wsy;
                            int oz = img_z(sh) + img_z(wt) - wsz;
                            if (t == s && ox == 0 && oy == 0 && oz == 0) continue;
                            int im = pack_img(ox, oy, oz);
                            double d2 = dot64(separation64(ps, pos[t], im, g.v));
                            if (d2 < c2lim) {
                                if (count < cap) {
                                    nbr[(size_t)count * n + s] = t;
                                    nimg[(size_t)count * n + s] = im;
                                }
                                ++count;
                            }
                        }
                    }
                }
        ncnt[s] = count < cap ? count : cap;
        if (count > cap) atomicMax(flags, count);
    }
}

// ---------------------------------------------------------------------------
// Lane-parallel scan: G lanes per atom.  The bins of each (dy, dz) column
// are consecutive in the cell-sorted atom array (bin index runs x fastest),
// so a column is a few contiguous segments, one per periodic x shift; the
// lanes of a group test consecutive candidates of a segment in parallel and
// the hits are compacted with a ballot, which keeps the sequential (dz, dy,
// dx, atom) order of the per-atom scan.  The scan covers g.reach bins on
// each side: one, or more when a bin is narrower than the list radius (a
// cell thinner than the cutoff, whose atoms then meet several periodic
// images of one another and of themselves).
// ---------------------------------------------------------------------------

// bin b + shift of a periodic axis with nc bins: b in [0, nc), shift in cells
__device__ __forceinline__ void wrap_bin(int &b, int &shift, int nc)
{
    shift = b >= 0 ? b / nc : -((-b + nc - 1) / nc);
    b -= shift * nc;
}

template <int G>
__global__ __launch_bounds__(256) void k_neighbors_group(int n, CellGeom g,
                                                         const double4 *__restrict__ pos,
                                                         const int *__restrict__ key,
                                                         const int *__restrict__ wrap,
                                                         const int *__restrict__ start,
                                                         const double *__restrict__ cut, int cap,
                                                         int *__restrict__ nbr,
                                                         int *__restrict__ nimg,
                                                         int *__restrict__ ncnt,
                                                         int *__restrict__ flags)
{
    constexpr int APB = 256 / G;
    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int s = blockIdx.x * APB + a;
    const bool valid = s < n;
    const int gbase = ((threadIdx.x & 63) / G) * G;
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    const unsigned long long below = (1ull << l) - 1ull;

    double4 ps = valid ? pos[s] : make_double4(0, 0, 0, 0);
    int k = valid ? key[s] : 0;
    int c0 = k % g.dim[0];
    int c1 = (k / g.dim[0]) % g.dim[1];
    int c2 = k / (g.dim[0] * g.dim[1]);
    int ws = valid ? wrap[s] : 0;
    int wsx = img_x(ws), wsy = img_y(ws), wsz = img_z(ws);
    double csq = valid ? cut[s] : 0.0;
    csq = csq * csq;
    int count = 0;
    // test the candidates [b, e) with x shift sx (all lanes of the group)
    auto scan = [&](int b, int e, int sx, int sy, int sz) {
        for (int t0 = b; t0 < e; t0 += G) {
            int t = t0 + l;
            bool hit = false;
            int im = 0;
            if (t < e) {
                int wt = wrap[t];
                int ox = sx + img_x(wt) - wsx;
                int oy = sy + img_y(wt) - wsy;
                int oz = sz + img_z(wt) - wsz;
                im = pack_img(ox, oy, oz);
                if (!(t == s && ox == 0 && oy == 0 && oz == 0)) {
                    double d2 = dot64(separation64(ps, pos[t], im, g.v));
                    hit = d2 < csq;
                }
            }
            unsigned long long bal = __ballot(hit);
            unsigned long long gb = (bal >> gbase) & gmask;
            int slot = count + __popcll(gb & below);
            if (hit && slot < cap) {
                nbr[(size_t)slot * n + s] = t;
                nimg[(size_t)slot * n + s] = im;
            }
            count += __popcll(gb);
        }
    };
    const int R0 = g.reach[0], R1 = g.reach[1], R2 = g.reach[2];
    for (int dz = valid ? -R2 : R2 + 1; dz <= R2; ++dz) {
        int bz = c2 + dz, sz = 0;
        if (g.pbc[2]) wrap_bin(bz, sz, g.nc[2]);
        else if (bz < 0 || bz >= g.dim[2]) continue;
        for (int dy = -R1; dy <= R1; ++dy) {
            int by = c1 + dy, sy = 0;
            if (g.pbc[1]) wrap_bin(by, sy, g.nc[1]);
            else if (by < 0 || by >= g.dim[1]) continue;
            int row = g.dim[0] * (by + g.dim[1] * bz);
            // consecutive bins sharing one x shift merge into one segment
            int sb = 0, se = 0, ss = 0;
            bool open = false;
            for (int dx = -R0; dx <= R0; ++dx) {
                int bx = c0 + dx, sx = 0;
                if (g.pbc[0]) wrap_bin(bx, sx, g.nc[0]);
                else if (bx < 0 || bx >= g.dim[0]) continue;
                int b0 = start[row + bx], b1 = start[row + bx + 1];
                if (open && se == b0 && ss == sx) {
                    se = b1;
                } else {
                    if (open) scan(sb, se, ss, sy, sz);
                    sb = b0;
                    se = b1;
                    ss = sx;
                    open = true;
                }
            }
            if (open) scan(sb, se, ss, sy, sz);
        }
    }
    if (valid && l == 0) {
        ncnt[s] = count < cap ? count : cap;
        if (count > cap) atomicMax(flags, count);
    }
}

}  // namespace

// Bins per axis: floor(perpendicular width / max_cut) as in
// get_optimal_splitting (Geometry.f90:679-707) without the 40 cap; the total
// is kept below ~2 bins per atom by coarsening (bins only grow, so the
// 27-bin stencil stays complete).
// Every charge equals the first and every cutoff the first: bit 0 of
// *differs is set otherwise.
__global__ void k_qc_differs(int n, const double *__restrict__ q, const double *__restrict__ cut,
                             int *__restrict__ differs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (q[i] != q[0] || cut[i] != cut[0]) atomicOr(differs, 1);
}

bool nb_begin_uniform_qc(System &S)
{
    S.qc_uniform = false;
    if (S.n < 1) return false;
    S.qc_check.ensure(1);
    HIP_TRY(hipMemsetAsync(S.qc_check.p, 0, sizeof(int), S.stream));   // never under capture
    k_qc_differs<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, S.q.p, S.cut.p, S.qc_check.p);
    int differs = 1;
    HIP_TRY(hipMemcpyAsync(&differs, S.qc_check.p, sizeof(int), hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    if (differs) return false;
    // both buffers of each hold the one value, whichever a sort leaves current
    S.q2.ensure(S.n);
    S.cut2.ensure(S.n);
    HIP_TRY(hipMemcpyAsync(S.q2.p, S.q.p, sizeof(double) * S.n, hipMemcpyDeviceToDevice, S.stream));
    HIP_TRY(hipMemcpyAsync(S.cut2.p, S.cut.p, sizeof(double) * S.n, hipMemcpyDeviceToDevice,
                           S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    S.qc_uniform = true;
    return true;
}

void nb_configure_grid(System &S, double max_cut)
{
    PYSIC_REQUIRE(max_cut > 0.0, PYSIC_EINVAL, "neighbor cutoff must be positive");
    auto configure = [&](CellGeom &g, long natoms) {
        double w[3];
        for (int i = 0; i < 3; ++i) {
            const double *a = g.v + 3 * ((i + 1) % 3), *b = g.v + 3 * ((i + 2) % 3),
                         *v = g.v + 3 * i;
            double nx = a[1] * b[2] - a[2] * b[1], ny = a[2] * b[0] - a[0] * b[2],
                   nz = a[0] * b[1] - a[1] * b[0];
            w[i] = std::fabs(v[0] * nx + v[1] * ny + v[2] * nz) /
                   std::sqrt(nx * nx + ny * ny + nz * nz);
            PYSIC_REQUIRE(w[i] > 0.0 && std::isfinite(w[i]), PYSIC_EINVAL,
                          "the cell has no width along one of its axes");
            // a cell thinner than the list radius gets one bin on that axis
            // and a scan over several bins (its periodic images)
            g.nc[i] = std::max(1, (int)std::floor(w[i] / max_cut));
        }
        long limit = std::max<long>(64, 2L * natoms);
        while ((long)g.nc[0] * g.nc[1] * g.nc[2] > limit) {
            int a = 0;
            for (int i = 1; i < 3; ++i)
                if (g.nc[i] > g.nc[a]) a = i;
            g.nc[a] = std::max(1, (int)(g.nc[a] * 0.8));
        }
        for (int i = 0; i < 3; ++i) {
            g.dim[i] = g.pbc[i] ? g.nc[i] : g.nc[i] + 2;
            // every atom within max_cut lies within ceil(max_cut / bin width)
            // bins: the fractional coordinates of two atoms r apart differ by
            // at most r / width along each axis
            const double bw = w[i] / g.nc[i];
            g.reach[i] = bw >= max_cut ? 1 : (int)std::ceil(max_cut / bw);
            PYSIC_REQUIRE(g.reach[i] <= PYSIC_REACH_MAX, PYSIC_ERANGE,
                          "cell more than 64 times thinner than the neighbor cutoff");
        }
        g.ncell = g.dim[0] * g.dim[1] * g.dim[2];
    };
    const int old_nbins = S.nbins;
    const CellGeom old_geom = S.geom;
    if (S.ncells == 0) {
        configure(S.geom, S.n);
        S.nbins = S.geom.ncell;
        if (S.nbins != old_nbins || std::memcmp(S.geom.dim, old_geom.dim, sizeof(int) * 3) != 0)
            S.order_valid = false;
    } else {
        S.order_valid = false;
        std::vector<int> fb(S.ncells);
        long total = 0;
        for (int c = 0; c < S.ncells; ++c) {
            configure(S.hcells[c], S.hatom0[c + 1] - S.hatom0[c]);
            fb[c] = (int)total;
            total += S.hcells[c].ncell;
        }
        PYSIC_REQUIRE(total < (1L << 30), PYSIC_ERANGE, "too many cell-list bins in the batch");
        S.nbins = (int)total;
        for (int c = 0; c < S.ncells; ++c)
            PYSIC_REQUIRE(S.hcells[c].reach[0] == 1 && S.hcells[c].reach[1] == 1 &&
                              S.hcells[c].reach[2] == 1,
                          PYSIC_ERANGE,
                          "cell " + std::to_string(c) +
                              " of the batch is thinner than the neighbor cutoff; a batch "
                              "needs cells at least as wide as the cutoff");
        S.cells.ensure(S.ncells);
        S.first_bin.ensure(S.ncells);
        HIP_TRY(hipMemcpy(S.cells.p, S.hcells.data(), sizeof(CellGeom) * S.ncells,
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(S.first_bin.p, fb.data(), sizeof(int) * S.ncells, hipMemcpyHostToDevice));
    }
    S.max_cut = max_cut;
    if (!S.order_valid) S.bins_clean = false;  // bins changed: counts and flags refilled
    // every periodic perpendicular width above twice the list radius: a
    // lattice vector is at least as long as the smallest such width, so no
    // atom reaches two images of another within the radius and each
    // neighbor appears once per list (the gather then matches rows by index)
    auto unique_in = [&](const CellGeom &g) {
        for (int a = 0; a < 3; ++a) {
            if (!g.pbc[a]) continue;
            const double *u = g.v + 3 * ((a + 1) % 3), *w = g.v + 3 * ((a + 2) % 3),
                         *v = g.v + 3 * a;
            double nx = u[1] * w[2] - u[2] * w[1], ny = u[2] * w[0] - u[0] * w[2],
                   nz = u[0] * w[1] - u[1] * w[0];
            double width = std::fabs(v[0] * nx + v[1] * ny + v[2] * nz) /
                           std::sqrt(nx * nx + ny * ny + nz * nz);
            if (!(width > 2.0 * max_cut)) return false;
        }
        return true;
    };
    bool uq = true;
    if (S.ncells == 0) uq = unique_in(S.geom);
    for (int c = 0; c < S.ncells; ++c) uq = uq && unique_in(S.hcells[c]);
    S.unique_images = uq;
}

// The buffers a sort alternates: the per-atom arrays and the cell starts
// (the last sort's starts are the next sort's old ones).
void nb_swap_state(System &S)
{
    std::swap(S.pos, S.pos2);
    std::swap(S.mom, S.mom2);
    std::swap(S.q, S.q2);
    std::swap(S.cut, S.cut2);
    std::swap(S.id, S.id2);
    std::swap(S.key, S.key2);
    std::swap(S.wrap, S.wrap2);
    std::swap(S.cellstart, S.prevstart);
}

// The counting sort after binning: scan of the bin counts, scatter, per-bin
// sort by original index, permutation of every per-atom array.
static void sort_binned(System &S)
{
    const int n = S.n;
    const int nbins = S.nbins;
    hipStream_t st = S.stream;
    const int nb = div_up(nbins, SCAN_BLK);
    const int all_dirty = S.order_valid ? 0 : 1;
    k_scan_local<<<nb, SCAN_T, 0, st>>>(nbins, S.cellcount.p, S.scanloc.p, S.scanblk.p,
                                        S.cellfill.p);
    k_scan_blocks<<<1, SCAN_T, 0, st>>>(nb, S.scanblk.p, S.scanloc.p, nbins);
    // S.cellstart holds the last sort's starts here (the old slot ranges)
    k_scatter<<<div_up(n, 256), 256, 0, st>>>(n, S.key.p, S.scanloc.p, S.scanblk.p,
                                              S.cellstart.p, S.dirty.p, all_dirty, S.cellfill.p,
                                              S.perm.p, S.flags.p);
    k_cellsort<<<div_up(nbins, 256), 256, 0, st>>>(nbins, n, S.scanloc.p, S.scanblk.p, S.perm.p,
                                                   S.id.p, S.dirty.p, all_dirty, S.prevstart.p);
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        auto kern = S.qc_uniform ? k_permute<Geo, false> : k_permute<Geo, true>;
        kern<<<div_up(n, 256), 256, 0, st>>>(n, S.perm.p, S.pos.p, S.mom.p, S.q.p, S.cut.p, S.id.p,
                                             S.key.p, S.wrap.p, S.pos2.p, S.mom2.p, S.q2.p,
                                             S.cut2.p, S.id2.p, S.key2.p, S.wrap2.p, geo,
                                             S.wpos.p, S.flags.p);
    });
    HIP_TRY(hipGetLastError());
    // the new starts went to prevstart, which nb_swap_state makes cellstart
    S.order_valid = true;
}

static void ensure_sort_buffers(System &S)
{
    const int n = S.n;
    const int nbins = S.nbins;
    S.key.ensure(n); S.key2.ensure(n);
    S.wrap.ensure(n); S.wrap2.ensure(n);
    S.perm.ensure(n);
    const bool fresh = S.cellcount.cap < (size_t)nbins || S.cellfill.cap < (size_t)nbins ||
                       S.dirty.cap < (size_t)nbins || !S.cellcount.p || !S.dirty.p;
    S.cellcount.ensure(nbins);
    S.cellfill.ensure(nbins);
    S.dirty.ensure(nbins);
    S.cellstart.ensure((size_t)nbins + 1);
    S.prevstart.ensure((size_t)nbins + 1);
    S.scanloc.ensure((size_t)nbins + 1);
    S.scanblk.ensure(div_up(nbins, SCAN_BLK) + 1);
    S.pos2.ensure(n); S.mom2.ensure(n); S.q2.ensure(n); S.cut2.ensure(n);
    S.id2.ensure(n);
    S.wpos.ensure(n);
    S.flags.ensure(4);
    if (fresh || !S.bins_clean) {
        // counts and dirty flags start at zero (a kernel: this may be captured)
        k_fill_int<<<div_up(2 * nbins, 256), 256, 0, S.stream>>>(nbins, S.cellcount.p, S.dirty.p);
        S.order_valid = false;
        S.bins_clean = true;
    }
}

void nb_sort_atoms(System &S, bool drift, double dt, double half_dt)
{
    const int n = S.n;
    const int nbins = S.nbins;
    hipStream_t st = S.stream;
    ensure_sort_buffers(S);

    hipEvent_t ev;
    S.timer.begin(drift ? "drift_bin_sort" : "bin_sort", st, &ev);
    const int track = S.order_valid ? 1 : 0;
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        if (drift)
            k_drift_bin<Geo><<<div_up(n, 256), 256, 0, st>>>(n, geo, dt, half_dt, S.frc.p, S.mom.p,
                                                             S.pos.p, S.key.p, S.wrap.p,
                                                             S.cellcount.p, S.flags.p, nbins,
                                                             track, S.dirty.p);
        else
            k_bin<Geo><<<div_up(n, 256), 256, 0, st>>>(n, geo, S.pos.p, S.key.p, S.wrap.p,
                                                       S.cellcount.p, S.flags.p, nbins, track,
                                                       S.dirty.p);
    });
    sort_binned(S);
    S.timer.end(drift ? "drift_bin_sort" : "bin_sort", st, ev);
    nb_swap_state(S);
}

void nb_gather_drift_sort(System &S, double dt, double half_dt)
{
    const int n = S.n;
    const int nbins = S.nbins;
    hipStream_t st = S.stream;
    PYSIC_REQUIRE(S.sj.p && S.simg.p && S.scnt.p && S.srow.p && S.dU.p, PYSIC_ESTATE,
                  "no per-bond gradients to gather");
    ensure_sort_buffers(S);
    hipEvent_t ev;
    S.timer.begin("gather_drift_bin_sort", st, &ev);
    const int track = S.order_valid ? 1 : 0;
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        k_gather_drift_bin<Geo><<<div_up(n, 256), 256, 0, st>>>(
            n, geo, dt, half_dt, S.sj.p, S.simg.p, S.scnt.p, S.srow.p, S.scap,
            S.unique_images ? 1 : 0, S.dU.p, S.mom.p, S.pos.p, S.key.p, S.wrap.p, S.cellcount.p,
            S.flags.p, nbins, track, S.dirty.p);
    });
    sort_binned(S);
    S.timer.end("gather_drift_bin_sort", st, ev);
    nb_swap_state(S);
}

void nb_clear_flags(System &S)
{
    S.flags.ensure(4);
    k_clear_flags<<<1, 64, 0, S.stream>>>(S.flags.p);
    HIP_TRY(hipGetLastError());
}

void nb_build_lists(System &S)
{
    const int n = S.n;
    const CellGeom &g = S.geom;
    hipStream_t st = S.stream;
    S.nbr.ensure((size_t)S.cap * n);
    S.nimg.ensure((size_t)S.cap * n);
    S.ncnt.ensure(n);
    double interior = (double)g.nc[0] * g.nc[1] * g.nc[2];
    double apb = n / interior;  // atoms per bin
    PYSIC_REQUIRE(S.ncells == 0 || S.nb_method == 0 || S.nb_method == 1 || S.nb_method == 5,
                  PYSIC_EINVAL, "a batch of cells takes neighbor method 0, 1 or 5");
    // the default for one cell below 2^17 atoms is the lane-parallel scan:
    // one thread per atom would leave most of the chip idle (measured on
    // 24,576 atoms of SiO2 at 3.55 A: 91 us per thread, 34 us per group)
    // a cell thinner than the cutoff: the lane-parallel scan, the one that
    // searches several bins on each side
    const bool thin = S.ncells == 0 && (g.reach[0] > 1 || g.reach[1] > 1 || g.reach[2] > 1);
    const int method =
        ((S.nb_method == 0 && S.ncells == 0 && n < (1 << 17)) || thin) ? 4 : S.nb_method;
    bool brick = method == 2 || method == 3;
    hipEvent_t ev;
    S.timer.begin("neighbors", st, &ev);
    if (method == 0 || method == 5) {
        // FP32 prefilter margin: bound on the rounding of the FP32 squared
        // distance of wrapped positions (|error| per component <= ~1.2e-6 L,
        // taken 2.5x larger); open axes leave positions unbounded: no
        // prefilter.  A batch takes the largest cell's bound.
        float margin2 = 0.0f;
        auto margin_of = [&](const CellGeom &c) {
            if (!(c.pbc[0] && c.pbc[1] && c.pbc[2])) return 3.0e38f;
            double L = 0.0;
            for (int a = 0; a < 3; ++a)
                L += std::sqrt(c.v[3 * a] * c.v[3 * a] + c.v[3 * a + 1] * c.v[3 * a + 1] +
                               c.v[3 * a + 2] * c.v[3 * a + 2]);
            double delta = 3.0e-6 * L;
            return (float)(4.0 * std::sqrt(3.0) * S.max_cut * delta + 12.0 * delta * delta);
        };
        if (S.ncells == 0) margin2 = margin_of(g);
        for (int c = 0; c < S.ncells; ++c) margin2 = std::max(margin2, margin_of(S.hcells[c]));
        // dense bins of one periodic cell: the kernel with the wave-uniform
        // scan (a separate instance: its second scan path costs registers)
        const bool dense = S.ncells == 0 && g.pbc[0] && g.pbc[1] && g.pbc[2] && g.nc[0] >= 3 &&
                           g.nc[1] >= 3 && g.nc[2] >= 3 && apb >= 6.0;
        with_geo(S, [&](auto geo) {
            using Geo = decltype(geo);
            auto go = [&](auto kern) {
                kern<<<div_up(n, 256), 256, 0, st>>>(n, geo, S.pos.p, S.wpos.p, S.key.p, S.wrap.p,
                                                     S.cellstart.p, S.cut.p, S.cap, margin2,
                                                     S.nbr.p, S.nimg.p, S.ncnt.p, S.flags.p);
            };
            if constexpr (std::is_same<Geo, GeoOne>::value) {
                if (dense) go(k_neighbors_pf<Geo, true>);
                else go(k_neighbors_pf<Geo, false>);
            } else {
                go(k_neighbors_pf<Geo, false>);
            }
        });
    } else if (method == 4) {
        // lane-parallel scan; wider groups when the bins are dense
        if (apb * 3.0 <= 10.0)
            k_neighbors_group<8><<<div_up(n, 256 / 8), 256, 0, st>>>(
                n, g, S.pos.p, S.key.p, S.wrap.p, S.cellstart.p, S.cut.p, S.cap, S.nbr.p, S.nimg.p,
                S.ncnt.p, S.flags.p);
        else
            k_neighbors_group<32><<<div_up(n, 256 / 32), 256, 0, st>>>(
                n, g, S.pos.p, S.key.p, S.wrap.p, S.cellstart.p, S.cut.p, S.cap, S.nbr.p, S.nimg.p,
                S.ncnt.p, S.flags.p);
    } else if (!brick) {
        with_geo(S, [&](auto geo) {
            k_neighbors<decltype(geo)><<<div_up(n, 256), 256, 0, st>>>(
                n, geo, S.pos.p, S.key.p, S.wrap.p, S.cellstart.p, S.cut.p, S.cap, S.nbr.p,
                S.nimg.p, S.ncnt.p, S.flags.p);
        });
    } else {
        // grow the brick one bin-edge at a time (smallest edge first) while
        // its expected atom count stays within one pass of 256 threads
        BrickCfg bc;
        int B[3] = {1, 1, 1};
        for (;;) {
            int a = 0;
            for (int q = 1; q < 3; ++q)
                if (B[q] < B[a]) a = q;
            if (B[a] >= g.dim[a] || B[a] >= 8) break;
            double cexp = (double)B[0] * B[1] * B[2] / B[a] * (B[a] + 1) * apb * 1.15;
            if (cexp > 256.0) break;
            B[a] += 1;
        }
        for (int a = 0; a < 3; ++a) {
            bc.B[a] = B[a];
            bc.nb[a] = div_up(g.dim[a], bc.B[a]);
        }
        int H = (bc.B[0] + 2) * (bc.B[1] + 2) * (bc.B[2] + 2);
        int C = bc.B[0] * bc.B[1] * bc.B[2];
        size_t fixed = (size_t)(4 * H + C + 2) * sizeof(int);
        size_t per_atom = sizeof(double4) + 2 * sizeof(int);
        long room = (long)((60 * 1024 - (long)fixed) / (long)per_atom);  // stay under 64 KB
        int lcap = (int)(H * apb * 1.6) + 64;
        lcap = ((lcap + 63) / 64) * 64;
        lcap = (int)std::max(0L, std::min((long)lcap, room));
        if (method == 3) lcap = 8;  // exercise the global-memory fallback
        bc.lcap = lcap;
        size_t smem = (size_t)lcap * per_atom + fixed;
        int nbricks = bc.nb[0] * bc.nb[1] * bc.nb[2];
        k_neighbors_brick<<<nbricks, 256, smem, st>>>(n, g, bc, S.pos.p, S.wrap.p, S.cellstart.p,
                                                      S.cut.p, S.cap, S.nbr.p, S.nimg.p, S.ncnt.p,
                                                      S.flags.p);
    }
    HIP_TRY(hipGetLastError());
    S.timer.end("neighbors", st, ev);
}

int nb_read_flags(System &S, int *out4)
{
    HIP_TRY(hipMemcpyAsync(out4, S.flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    return 0;
}
