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

#include <algorithm>
#include <cmath>

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

// Bin every atom: wrap offset -floor(frac) on periodic axes
// (Geometry.f90:597-619) and the bin of the wrapped fractional coordinate.
__global__ void k_bin(int n, CellGeom g, const double4 *__restrict__ pos, int *__restrict__ key,
                      int *__restrict__ wrap, int *__restrict__ cellcount, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    double rel[3];
    frac_of(g, p.x, p.y, p.z, rel);
    int w[3], c[3];
    bool bad = false;
    for (int a = 0; a < 3; ++a) {
        // non-finite or far-away coordinates are flagged and binned at 0
        if (g.pbc[a]) {
            double fl = floor(rel[a]);
            bool ok = fabs(fl) < (double)(PYSIC_IMG_MAX / 2);
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
    if (bad) atomicOr(flags + 2, 1);
    int k = c[0] + g.dim[0] * (c[1] + g.dim[1] * c[2]);
    key[i] = k;
    wrap[i] = pack_img(w[0], w[1], w[2]);
    atomicAdd(cellcount + k, 1);
}

// --- exclusive scan of the cell counts (3 phases, 1024 elements per block) --
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

__global__ void k_scan_local(int m, const int *__restrict__ in, int *__restrict__ out,
                             int *__restrict__ blk)
{
    __shared__ int lds[SCAN_T];
    int base = blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER], s = 0;
    for (int q = 0; q < SCAN_PER; ++q) {
        v[q] = base + q < m ? in[base + q] : 0;
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

__global__ void k_scan_add(int m, int *__restrict__ out, const int *__restrict__ blk)
{
    int base = blockIdx.x * SCAN_BLK;
    int add = blk[blockIdx.x];
    for (int q = threadIdx.x; q < SCAN_BLK; q += SCAN_T)
        if (base + q < m) out[base + q] += add;
}

__global__ void k_scatter(int n, const int *__restrict__ key, const int *__restrict__ start,
                          int *__restrict__ fill, int *__restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = key[i];
    int slot = start[k] + atomicAdd(fill + k, 1);
    perm[slot] = i;
}

// Within each cell, order atoms by original index: the atomic slot claims
// above are unordered, this makes the whole permutation deterministic.
__global__ void k_cellsort(int ncell, const int *__restrict__ start, int *__restrict__ perm,
                           const int *__restrict__ id)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int a = start[c], b = start[c + 1];
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

__global__ void k_permute(int n, const int *__restrict__ perm, const double4 *__restrict__ pos,
                          const double4 *__restrict__ mom, const double4 *__restrict__ frc,
                          const double *__restrict__ q, const double *__restrict__ cut,
                          const int *__restrict__ id, const int *__restrict__ key,
                          const int *__restrict__ wrap, double4 *__restrict__ pos2,
                          double4 *__restrict__ mom2, double4 *__restrict__ frc2,
                          double *__restrict__ q2, double *__restrict__ cut2,
                          int *__restrict__ id2, int *__restrict__ key2, int *__restrict__ wrap2)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int o = perm[s];
    pos2[s] = pos[o];
    mom2[s] = mom[o];
    frc2[s] = frc[o];
    q2[s] = q[o];
    cut2[s] = cut[o];
    id2[s] = id[o];
    key2[s] = key[o];
    wrap2[s] = wrap[o];
}

// Per-atom scan of the 27 adjacent bins (Core.f90:8182-8252 as a full list).
__global__ __launch_bounds__(256) void k_neighbors(int n, CellGeom g, const double4 *__restrict__ pos,
                                                   const int *__restrict__ key,
                                                   const int *__restrict__ wrap,
                                                   const int *__restrict__ start,
                                                   const double *__restrict__ cut, int cap,
                                                   int *__restrict__ nbr, int *__restrict__ nimg,
                                                   int *__restrict__ ncnt, int *__restrict__ flags)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 ps = pos[s];
    int k = key[s];
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

}  // namespace

// Bins per axis: floor(perpendicular width / max_cut) as in
// get_optimal_splitting (Geometry.f90:679-707) without the 40 cap; the total
// is kept below ~2 bins per atom by coarsening (bins only grow, so the
// 27-bin stencil stays complete).
void nb_configure_grid(System &S, double max_cut)
{
    CellGeom &g = S.geom;
    PYSIC_REQUIRE(max_cut > 0.0, PYSIC_EINVAL, "neighbor cutoff must be positive");
    double w[3];
    for (int i = 0; i < 3; ++i) {
        const double *a = g.v + 3 * ((i + 1) % 3), *b = g.v + 3 * ((i + 2) % 3), *v = g.v + 3 * i;
        double nx = a[1] * b[2] - a[2] * b[1], ny = a[2] * b[0] - a[0] * b[2],
               nz = a[0] * b[1] - a[1] * b[0];
        w[i] = std::fabs(v[0] * nx + v[1] * ny + v[2] * nz) / std::sqrt(nx * nx + ny * ny + nz * nz);
        g.nc[i] = (int)std::floor(w[i] / max_cut);
        PYSIC_REQUIRE(g.nc[i] >= 1, PYSIC_ERANGE,
                      "cell narrower than the neighbor cutoff; the cell list needs width >= cutoff");
    }
    long limit = std::max<long>(64, 2L * S.n);
    while ((long)g.nc[0] * g.nc[1] * g.nc[2] > limit) {
        int a = 0;
        for (int i = 1; i < 3; ++i)
            if (g.nc[i] > g.nc[a]) a = i;
        g.nc[a] = std::max(1, (int)(g.nc[a] * 0.8));
    }
    for (int i = 0; i < 3; ++i) g.dim[i] = g.pbc[i] ? g.nc[i] : g.nc[i] + 2;
    g.ncell = g.dim[0] * g.dim[1] * g.dim[2];
    S.max_cut = max_cut;
}

static void swap_bufs(System &S)
{
    std::swap(S.pos, S.pos2);
    std::swap(S.mom, S.mom2);
    std::swap(S.frc, S.frc2);
    std::swap(S.q, S.q2);
    std::swap(S.cut, S.cut2);
    std::swap(S.id, S.id2);
}

void nb_sort_atoms(System &S)
{
    const int n = S.n;
    CellGeom &g = S.geom;
    hipStream_t st = S.stream;
    S.key.ensure(2 * (size_t)n);  // second half: keys after the permutation
    S.wrap.ensure(2 * (size_t)n);
    S.perm.ensure(n);
    S.cellcount.ensure(g.ncell);
    S.cellfill.ensure(g.ncell);
    S.cellstart.ensure((size_t)g.ncell + 1);
    int nb = div_up(g.ncell, SCAN_BLK);
    S.scanblk.ensure(nb);
    S.pos2.ensure(n); S.mom2.ensure(n); S.frc2.ensure(n); S.q2.ensure(n); S.cut2.ensure(n);
    S.id2.ensure(n);
    S.flags.ensure(4);

    hipEvent_t ev;
    S.timer.begin("bin_sort", st, &ev);
    HIP_TRY(hipMemsetAsync(S.cellcount.p, 0, sizeof(int) * g.ncell, st));
    HIP_TRY(hipMemsetAsync(S.cellfill.p, 0, sizeof(int) * g.ncell, st));
    int *key_old = S.key.p, *key_new = S.key.p + n;
    int *wrap_old = S.wrap.p, *wrap_new = S.wrap.p + n;
    k_bin<<<div_up(n, 256), 256, 0, st>>>(n, g, S.pos.p, key_old, wrap_old, S.cellcount.p, S.flags.p);
    k_scan_local<<<nb, SCAN_T, 0, st>>>(g.ncell, S.cellcount.p, S.cellstart.p, S.scanblk.p);
    k_scan_blocks<<<1, SCAN_T, 0, st>>>(nb, S.scanblk.p, S.cellstart.p, g.ncell);
    k_scan_add<<<nb, SCAN_T, 0, st>>>(g.ncell, S.cellstart.p, S.scanblk.p);
    k_scatter<<<div_up(n, 256), 256, 0, st>>>(n, key_old, S.cellstart.p, S.cellfill.p, S.perm.p);
    k_cellsort<<<div_up(g.ncell, 256), 256, 0, st>>>(g.ncell, S.cellstart.p, S.perm.p, S.id.p);
    k_permute<<<div_up(n, 256), 256, 0, st>>>(n, S.perm.p, S.pos.p, S.mom.p, S.frc.p, S.q.p, S.cut.p,
                                              S.id.p, key_old, wrap_old, S.pos2.p, S.mom2.p,
                                              S.frc2.p, S.q2.p, S.cut2.p, S.id2.p, key_new,
                                              wrap_new);
    HIP_TRY(hipGetLastError());
    S.timer.end("bin_sort", st, ev);
    swap_bufs(S);
    // keep the sorted keys/wraps in the first half
    HIP_TRY(hipMemcpyAsync(key_old, key_new, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(wrap_old, wrap_new, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
}

void nb_build_lists(System &S)
{
    const int n = S.n;
    hipStream_t st = S.stream;
    S.nbr.ensure((size_t)S.cap * n);
    S.nimg.ensure((size_t)S.cap * n);
    S.ncnt.ensure(n);
    hipEvent_t ev;
    S.timer.begin("neighbors", st, &ev);
    k_neighbors<<<div_up(n, 256), 256, 0, st>>>(n, S.geom, S.pos.p, S.key.p, S.wrap.p, S.cellstart.p,
                                                 S.cut.p, S.cap, S.nbr.p, S.nimg.p, S.ncnt.p,
                                                 S.flags.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("neighbors", st, ev);
}

int nb_read_flags(System &S, int *out4)
{
    HIP_TRY(hipMemcpyAsync(out4, S.flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    return 0;
}
