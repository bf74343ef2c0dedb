// common.h -- shared types of the MI355X pysic engine (host + device).
//
// Everything here is plain data: the compiled interaction model (tables
// indexed by species), the device-resident system in cell-sorted order, and
// the error/launch helpers used by every translation unit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#define PYSIC_MAXS 4      // species per system
#define PYSIC_MAXP 4      // exponential terms per ordered species pair

// ---------------------------------------------------------------------------
// interaction model, compiled from the pysic potential / bond-order entries
// ---------------------------------------------------------------------------

// One exponential pair term v(r) = eps*exp(-zeta*r) (Potentials.f90:3643-3677)
// with the create_potential soft-cutoff rule (Potentials.f90:2382-2388).
struct PairTerm {
    double eps, zeta, cut, soft;
    int smooth;   // 1 when 0 < soft < cut
    int group;    // -1 plain, >= 0 coordinated by the Tersoff group
};

// One 'tersoff' bond-order parameter set (Potentials.f90:5685-5961).
struct BondTerm {
    double beta, eta, a, c, d, h, mu, cut, soft;
    int valid;
    int a_zero;   // a == 0: exp((a dr)^mu) == 1 and its derivative vanishes
};

struct Model {
    int nspecies;
    int npair[PYSIC_MAXS][PYSIC_MAXS];
    PairTerm pair[PYSIC_MAXS][PYSIC_MAXS][PYSIC_MAXP];
    BondTerm bond[PYSIC_MAXS][PYSIC_MAXS][PYSIC_MAXS];  // [center][j][k]
    BondTerm post[PYSIC_MAXS];                           // b~ post-processing per center
    double maxpot[PYSIC_MAXS];    // max_potential_radius per species (Core.f90:1073-1125)
    double maxbond[PYSIC_MAXS];   // max_bond_radius per species (Core.f90:1034-1070)
    double rshort;                // largest cutoff any Tersoff-kernel term uses
    int group;                    // the single bond-order group (-1: none)
    int has_pairs;                // any exponential terms at all
    // charge_self singlets: sum_n eps_n q^n per species, n in {1,2,3,4}
    double self_eps[PYSIC_MAXS][5];
    int has_self;
};

// Pair potentials outside the Tersoff kernels (pair.hip): every 2-body type
// of the reference except a bond-order-coordinated exponential, and products
// of them (ProductPotential).  Type codes are the descriptor's:
// 1 exponential, 3 LJ, 4 spring, 5 Buckingham, 6 power, 7 tabulated,
// 8 shift_power, 9 step, 12 charge_exp, 13 charge_abs, 14 charge_pair;
// parameters in the reference's order.
#define PYSIC_MAXG 8      // such potentials per ordered species pair
#define PYSIC_MAXF 4      // factors of one product
#define PYSIC_MAXCF 16    // charge-dependent factors per model
struct PairFactor {
    int type;
    int toff, tlen;       // tabulated: first row in the table buffer, rows;
                          // charge-dependent types: toff indexes PairModel::cf
    double p[4];
};
// A charge-dependent pair factor of COMB (charge_exp, charge_abs,
// charge_pair): its
// parameters, the species its parameters 1 belong to (the potential's
// original first target; for a pair whose first atom is another species the
// two halves swap, Potentials.f90:3797-3829) and charge_exp's derived
// exponents and scales (Potentials.f90:3747-3766).
struct ChargeFactor {
    double p[13];
    double eta[2], beta[2];
    int orig, pad;
};
struct GenPair {
    PairFactor f[PYSIC_MAXF];  // f[0] the potential, then its multipliers
    int nf;
    int smooth;           // 1 when 0 < soft < cut
    double cut, soft;
};
// A lone LJ on an ordered species pair, with the bounds that let its force
// (which needs r^2 only) skip the square root: r^2 < lo2 means r < cut,
// r^2 >= hi2 means r >= cut, r^2 < soft_lo2 means r < soft (f_c = 1); only
// between them is r itself compared.
struct LJPair {
    double lo2, hi2, soft_lo2;
    double cut, soft, eps, sig2;
    int on, smooth;
};
// A bond_bend row, one per target permutation as the reference adds them
// (calculator.py:1247-1277): E = eps (cos^n theta - cos^n theta_0)^m on the
// angle at the vertex, acting only where the vertex is the species of the
// original second target (Potentials.f90:3196-3328).
#define PYSIC_MAXB 12
struct BendTerm {
    int e[3];             // targets in this row's order
    int vertex;           // species of the original second target
    int n, m;             // integer exponents (the reference truncates)
    int smooth, pad;
    double eps, c0;       // c0 = cos(theta_0)^n (Potentials.f90:3178-3185)
    double cut, soft;
};
struct PairModel {
    int npair[PYSIC_MAXS][PYSIC_MAXS];
    GenPair pair[PYSIC_MAXS][PYSIC_MAXS][PYSIC_MAXG];
    // squared largest cutoff of the ordered species pair, widened by 1e-12
    // relative: r^2 at or beyond it means r >= every cutoff, so the square
    // root and the exact test are skipped
    double cut2_hi[PYSIC_MAXS][PYSIC_MAXS];
    // 'constant' (E = V) and 'force' (F; E = -F . r) one-body terms per species
    double konst[PYSIC_MAXS];
    double kforce[PYSIC_MAXS][3];
    int has_pairs, has_single;
    int needs_pick;       // a 'step' or charge factor: it depends on which atom the reference takes first
    int has_charge, ncf;  // charge-dependent factors (cf[0 .. ncf-1])
    ChargeFactor cf[PYSIC_MAXCF];
    int nbend, pad_b;     // bond_bend rows (bend[0 .. nbend-1])
    double bend_cut;      // the largest bond_bend cutoff
    // per species, the largest cutoff of the potentials whose first target
    // it is (the reference's max_potential_radius, Core.f90:1034-1125)
    double bend_maxpot[PYSIC_MAXS];
    BendTerm bend[PYSIC_MAXB];
    // every ordered species pair holds at most one potential, a lone LJ
    // (lj_only), and its table
    int lj_only;
    LJPair lj[PYSIC_MAXS][PYSIC_MAXS];
};

// Pair potentials (ProductPotentials) scaled by the 'tersoff' bond-order
// factors of a coordinator (pair.hip, the bond-order pass): COMB's
// attraction.  At most one such potential per ordered species pair; each
// names its coordinator's group, whose 'tersoff' entries are keyed by the
// (center, partner, third) species triplet.  post[g][s] is the group's
// first entry with center species s, whose beta and eta turn the sum zeta
// into the factor (Core.f90:2763-2807).  maxbond[s]: the largest bond-order
// cutoff of center species s over every group (a third atom counts only
// within it, Core.f90:2518).
#define PYSIC_MAXGRP 4
struct BopModel {
    int on;
    int nbop[PYSIC_MAXS][PYSIC_MAXS];
    int group[PYSIC_MAXS][PYSIC_MAXS];
    GenPair pair[PYSIC_MAXS][PYSIC_MAXS];
    BondTerm bond[PYSIC_MAXGRP][PYSIC_MAXS][PYSIC_MAXS][PYSIC_MAXS];
    BondTerm post[PYSIC_MAXGRP][PYSIC_MAXS];
    double maxbond[PYSIC_MAXS];
    double radius;        // short-list radius: every such cutoff and bond cutoff
};

// ---------------------------------------------------------------------------
// packed periodic image offsets: three signed 10-bit fields
// ---------------------------------------------------------------------------
__host__ __device__ inline int pack_img(int a, int b, int c)
{
    return ((a & 1023)) | ((b & 1023) << 10) | ((c & 1023) << 20);
}
__host__ __device__ inline int img_x(int p) { return (p << 22) >> 22; }
__host__ __device__ inline int img_y(int p) { return (p << 12) >> 22; }
__host__ __device__ inline int img_z(int p) { return (p << 2) >> 22; }
__host__ __device__ inline int neg_img(int p)
{
    return pack_img(-img_x(p), -img_y(p), -img_z(p));
}
#define PYSIC_IMG_MAX 511
// Largest wrap offset of an atom (cells from its box) and largest scan reach
// (bins = cells of a thin axis): a pair image is a bin shift plus the
// difference of two wraps, 2 * 223 + 64 <= PYSIC_IMG_MAX.
#define PYSIC_WRAP_MAX 223
#define PYSIC_REACH_MAX 64

// separation_vector (Geometry.f90:449-466): r2 - r1 + sum_i o_i v_i in that
// order, no contraction, so distances match the reference's bits.
__device__ __forceinline__ double3 separation64(const double4 &ri, const double4 &rj, int img,
                                                const double *v)
{
#pragma clang fp contract(off)
    // same image (the common case): adding 0 * v would change no bit of a
    // nonzero component, so skip the nine multiply-adds
    if (img == 0) return make_double3(rj.x - ri.x, rj.y - ri.y, rj.z - ri.z);
    double ox = (double)img_x(img), oy = (double)img_y(img), oz = (double)img_z(img);
    double x = rj.x - ri.x, y = rj.y - ri.y, z = rj.z - ri.z;
    x = x + ox * v[0]; y = y + ox * v[1]; z = z + ox * v[2];
    x = x + oy * v[3]; y = y + oy * v[4]; z = z + oy * v[5];
    x = x + oz * v[6]; y = y + oz * v[7]; z = z + oz * v[8];
    return make_double3(x, y, z);
}
// pick (Geometry.f90:648-676) on original indices: whether the reference
// evaluates the pair from atom i1's side
__device__ __forceinline__ bool pick_dev(int i1, int i2, int img)
{
    if (i2 > i1) return true;
    if (i2 < i1) return false;
    int ox = img_x(img), oy = img_y(img), oz = img_z(img);
    return ox > 0 || (ox == 0 && (oy > 0 || (oy == 0 && oz > 0)));
}

// 1/x for a normal x: the hardware reciprocal estimate and two Newton
// steps (e = 1 - x y, y += y e), within an ulp or two of the correctly
// rounded quotient (x = 0 gives NaN; callers handle r = 0 first).
__device__ __forceinline__ double recip64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// sin and cos of (pi/2) x for 0 <= x <= 1.  The angle is moved to
// w = (pi/2) x - pi/4 in [-pi/4, pi/4], where the Taylor series of sin w
// (through w^17) and cos w (through w^18) leave a truncation error below
// 1e-19, and turned back by pi/4:
//   sin(pi/4 + w) = (cos w + sin w) / sqrt(2),
//   cos(pi/4 + w) = (cos w - sin w) / sqrt(2).
// No range reduction is needed, unlike the library sin/cos.
__device__ __forceinline__ void sincos_quarter64(double x, double *s, double *c)
{
    const double w = fma(x, 1.5707963267948966, -0.78539816339744831);
    const double w2 = w * w;
    double ps = 2.8114572543455206e-15;          // 1/17!
    ps = fma(ps, w2, -7.647163731819816e-13);    // -1/15!
    ps = fma(ps, w2, 1.6059043836821613e-10);    // 1/13!
    ps = fma(ps, w2, -2.505210838544172e-08);    // -1/11!
    ps = fma(ps, w2, 2.7557319223985893e-06);    // 1/9!
    ps = fma(ps, w2, -0.0001984126984126984);    // -1/7!
    ps = fma(ps, w2, 0.008333333333333333);      // 1/5!
    ps = fma(ps, w2, -0.16666666666666666);      // -1/3!
    const double sw = fma(ps * w2, w, w);
    double pc = -1.5619206968586225e-16;         // -1/18!
    pc = fma(pc, w2, 4.779477332387385e-14);     // 1/16!
    pc = fma(pc, w2, -1.1470745597729725e-11);   // -1/14!
    pc = fma(pc, w2, 2.08767569878681e-09);      // 1/12!
    pc = fma(pc, w2, -2.755731922398589e-07);    // -1/10!
    pc = fma(pc, w2, 2.48015873015873e-05);      // 1/8!
    pc = fma(pc, w2, -0.001388888888888889);     // -1/6!
    pc = fma(pc, w2, 4.1666666666666667e-02);    // 1/4!
    pc = fma(pc, w2, -0.5);                      // -1/2!
    const double cw = fma(pc, w2, 1.0);
    *s = (cw + sw) * 0.70710678118654752;
    *c = (cw - sw) * 0.70710678118654752;
}

// smoothening_factor and smoothening_derivative (Potentials.f90:442-509):
// f = (1 + cos(pi t)) / 2 = cos^2(pi t / 2) and
// df/dr = -(pi / (2 w)) sin(pi t) = -(pi / w) sin(pi t / 2) cos(pi t / 2),
// t = (r - soft) / w, w = hard - soft; both from one sincos_quarter64.
__device__ __forceinline__ void smooth64(double r, double hard, double soft, double *f,
                                         double *df)
{
    if (r > hard) { *f = 0.0; *df = 0.0; return; }
    if (r < soft) { *f = 1.0; *df = 0.0; return; }
    const double iw = recip64(hard - soft);
    double s, c;
    sincos_quarter64((r - soft) * iw, &s, &c);
    *f = c * c;
    *df = -M_PI * iw * s * c;
}
__device__ __forceinline__ double fc64(double r, double hard, double soft)
{
    double f, df;
    smooth64(r, hard, soft, &f, &df);
    return f;
}
__device__ __forceinline__ double dfc64(double r, double hard, double soft)
{
    double f, df;
    smooth64(r, hard, soft, &f, &df);
    return df;
}

// x^y for the bond-order post-processing, where x >= 0: exp(y log x) for
// x > 0 (a few ulp for |y log x| of order one, a third of the library pow's
// instructions); x <= 0 keeps pow's special cases.
__device__ __forceinline__ double powpos64(double x, double y)
{
    return x > 0.0 ? exp(y * log(x)) : pow(x, y);
}

// F_i = sum_j (D_ij - D_ji) over atom i's short list: D_ij = dU_i/dr_ij is
// the site kernel's output for entry (i, j, img) and D_ji the entry (j, i,
// -img) of j's list, found by scanning it (column-major: entry t of atom j
// at t n + j).  miss reports a bond with no reverse entry.
__device__ __forceinline__ double3 gather_bond_force(int i, int n, const int *__restrict__ sj,
                                                     const int *__restrict__ simg,
                                                     const int *__restrict__ scnt,
                                                     const float4 *__restrict__ dU, bool &miss)
{
    const int ns = scnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int s = 0; s < ns; ++s) {
        const size_t o = (size_t)s * n + i;
        const int j = sj[o];
        const int rim = neg_img(simg[o]);
        const float4 d = dU[o];
        const int nj = scnt[j];
        float4 dr = make_float4(0, 0, 0, 0);
        bool found = false;
        for (int t = 0; t < nj; ++t) {
            const size_t q = (size_t)t * n + j;
            if (sj[q] == i && simg[q] == rim) {
                dr = dU[q];
                found = true;
                break;
            }
        }
        miss |= !found;
        fx += (double)d.x - (double)dr.x;
        fy += (double)d.y - (double)dr.y;
        fz += (double)d.z - (double)dr.z;
    }
    return make_double3(fx, fy, fz);
}

// The same sum when every neighbor appears once in a list (no atom reaches
// two images of another within the list radius): each atom's short list
// also sits in a row srow[i * SCAP + t] (-1 past its end), so D_ji is found
// by comparing j's row against i in registers.  Four bonds go at a time,
// each stage's loads issued together -- the neighbors' rows, then the own
// and reverse gradients -- so a wave waits on three memory latencies per
// four bonds rather than several per bond.  The bonds are
// summed in the same order as gather_bond_force, so the result is the same.
template <int SCAP>
__device__ __forceinline__ double3 gather_bond_force_rows(int i, int n,
                                                          const int *__restrict__ srow,
                                                          const float4 *__restrict__ dU,
                                                          bool &miss)
{
    constexpr int B = 4;   // bonds per batch (a crystal atom has four)
    const int *row = srow + (size_t)i * SCAP;
    double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll 1
    for (int c0 = 0; c0 < SCAP; c0 += B) {
        const int4 jv = *reinterpret_cast<const int4 *>(row + c0);
        const int j[B] = {jv.x, jv.y, jv.z, jv.w};
        if (!__any(j[0] >= 0)) break;   // rows are -1 past their end
        int tr[B];
#pragma unroll
        for (int t = 0; t < B; ++t) tr[t] = -1;
#pragma unroll(SCAP == 8 ? 2 : 1)
        for (int s0 = 0; s0 < SCAP; s0 += 4) {
            int4 kv[B];
#pragma unroll
            for (int t = 0; t < B; ++t)
                kv[t] = j[t] >= 0 ? *reinterpret_cast<const int4 *>(srow + (size_t)j[t] * SCAP + s0)
                                  : make_int4(-1, -1, -1, -1);
#pragma unroll
            for (int t = 0; t < B; ++t) {
                tr[t] = kv[t].x == i ? s0 : tr[t];
                tr[t] = kv[t].y == i ? s0 + 1 : tr[t];
                tr[t] = kv[t].z == i ? s0 + 2 : tr[t];
                tr[t] = kv[t].w == i ? s0 + 3 : tr[t];
            }
        }
        float4 own[B], rev[B];
#pragma unroll
        for (int t = 0; t < B; ++t) {
            own[t] = j[t] >= 0 ? dU[(size_t)(c0 + t) * n + i] : make_float4(0, 0, 0, 0);
            rev[t] = (j[t] >= 0 && tr[t] >= 0) ? dU[(size_t)tr[t] * n + j[t]]
                                               : make_float4(0, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            if (j[t] < 0) continue;
            miss |= tr[t] < 0;
            fx += (double)own[t].x - (double)rev[t].x;
            fy += (double)own[t].y - (double)rev[t].y;
            fz += (double)own[t].z - (double)rev[t].z;
        }
    }
    return make_double3(fx, fy, fz);
}

// The gather for the short-list capacity of the site kernel that wrote the
// rows (8, 16 or 32).
__device__ __forceinline__ double3 gather_rows_any(int i, int n, int scap,
                                                   const int *__restrict__ srow,
                                                   const float4 *__restrict__ dU, bool &miss)
{
    if (scap == 8) return gather_bond_force_rows<8>(i, n, srow, dU, miss);
    if (scap == 16) return gather_bond_force_rows<16>(i, n, srow, dU, miss);
    return gather_bond_force_rows<32>(i, n, srow, dU, miss);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remap so that XCD x runs one contiguous range of logical blocks:
// spatially adjacent atoms (the cell-sorted order) then share an L2.
__device__ __forceinline__ int xcd_block(int nblocks)
{
    const int b = blockIdx.x;
    const int x = b & 7, q = b >> 3;
    const int per = nblocks >> 3, rem = nblocks & 7;
    return x * per + (x < rem ? x : rem) + q;
}

// .o. (Quaternions.f90:475-483), no contraction
__device__ __forceinline__ double dot64(const double3 &a)
{
#pragma clang fp contract(off)
    return a.x * a.x + a.y * a.y + a.z * a.z;
}

// ---------------------------------------------------------------------------
// cell geometry passed by value to kernels
// ---------------------------------------------------------------------------
struct CellGeom {
    double v[9];     // cell vectors in rows
    double inv[9];   // M^-1 row-major (fractional = inv * r)
    int pbc[3];
    int nc[3];       // cell-list bins per axis (interior)
    int dim[3];      // bins per axis incl. padding on open axes
    int ncell;       // dim[0]*dim[1]*dim[2]
    int reach[3];    // bins searched on each side: 1 unless a bin is narrower
                     // than the list radius (a cell thinner than the cutoff)
};

// Where a kernel finds the cell of an atom slot.  GeoOne: one periodic cell
// for every atom (passed by value).  GeoBatch: many independent cells, each
// owning a fixed range of atom slots (the sort reorders atoms only within
// their cell, because every cell's bins form their own key range).
struct GeoOne {
    static constexpr bool kPersistent = false;  // one workgroup per tile
    CellGeom g;
    __device__ __forceinline__ const CellGeom &of(int) const { return g; }
    __device__ __forceinline__ const double *vec(int) const { return g.v; }
    __device__ __forceinline__ int bin0(int) const { return 0; }
};
struct GeoBatch {
    static constexpr bool kPersistent = true;   // site kernels may walk many tiles
    const CellGeom *cells;   // per cell
    const int *cell_of;      // cell of each atom slot
    const int *first_bin;    // first global bin of each cell
    __device__ __forceinline__ const CellGeom &of(int s) const { return cells[cell_of[s]]; }
    __device__ __forceinline__ const double *vec(int s) const { return cells[cell_of[s]].v; }
    __device__ __forceinline__ int bin0(int s) const { return first_bin[cell_of[s]]; }
};

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
#define PYSIC_OK 0
#define PYSIC_EINVAL -1
#define PYSIC_EHIP -2
#define PYSIC_ESTATE -3
#define PYSIC_ECAP -4
#define PYSIC_ENOMEM -5
#define PYSIC_ERANGE -6

struct PysicError {
    int code;
    std::string msg;
};

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            throw PysicError{PYSIC_EHIP, std::string(#expr) + ": " +               \
                                             hipGetErrorString(_e)};               \
    } while (0)

#define PYSIC_REQUIRE(cond, code, text)                                            \
    do {                                                                           \
        if (!(cond)) throw PysicError{code, text};                                 \
    } while (0)

static inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }
