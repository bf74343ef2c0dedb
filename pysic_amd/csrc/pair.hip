// pair.hip -- pair potentials and one-body terms outside the Tersoff kernels.
//
// Restates the doublet loop of the reference for the plain (uncoordinated)
// pair potentials: core_evaluate_local_doublet_energy_B and _forces_B
// (Core.f90:4257-4385, 5371-5606) with evaluate_energy / evaluate_forces and
// their components (Potentials.f90:2640-2890) for the types LJ, spring,
// exponential (as a product factor), Buckingham, power, tabulated,
// shift_power, step and COMB's charge-dependent charge_exp, charge_abs and
// charge_pair,
// products of them (ProductPotential), and the one-body 'constant' and
// 'force' terms of core_evaluate_local_singlet (Core.f90:4095-4239).  The
// charge-dependent factors add their -dE/dq to the electronegativities
// (core_calculate_electronegativities' doublet loop, Core.f90:7113-7273,
// and evaluate_electronegativity, Potentials.f90:2580-2637).
//
// The reference visits every pair once (pick) and scatters the force to both
// atoms.  Here one thread owns one atom and walks its full neighbor list:
// each pair is evaluated from both ends, every atom sums its own force, and
// nothing is scattered (no atomics; the summation order is fixed, so runs
// are bit-for-bit repeatable).  Per atom and pair the thread adds half the
// pair energy and half of -r_ij (x) g_i to the virial, g_i being the force
// on atom i from that pair.  All arithmetic is FP64.
//
// The force on the first atom of a pair is written f1 = G s + C (1,1,1),
// s = r_2 - r_1: every type is central (C = 0) except 'step', whose
// reference force is the same scalar in all three components.  For a
// central force both ends agree on g_i = G r_ij; for 'step' the sign of C
// depends on which atom the reference takes first (pick), so that is
// evaluated only when the model has a step factor.
#include "system.h"

namespace {

constexpr int PB = 256;

struct Radial {
    double v;   // energy of the factor
    double g;   // f1 = g s + c (1,1,1)
    double c;
};

// The charges of a pair in the order the reference evaluates it (atom 1 the
// atom pick takes first) and atom 1's species.
struct QPair {
    double q1, q2;
    int s1;
};

// charge_exp (evaluate_energy_charge_exp and _electronegativity_charge_exp,
// Potentials.f90:3776-3836, 3847-3903): E = eps exp((xi1 D1 + xi2 D2) / 2),
// D_k = Rmax_k + |beta_k (Qmax_k - q_k)|^eta_k, with the parameter halves
// swapped when atom 1 is not the species parameters 1 belong to.  With x,
// x[k] = -dE/dq_k.
__device__ __noinline__ double charge_exp(const ChargeFactor &c, const QPair &qp, double *x)
{
    const int a = qp.s1 != c.orig ? 1 : 0, b = 1 - a;
    const double *p = c.p;
    const double d5 = c.beta[a] * (p[3 + 4 * a] - qp.q1);
    const double d6 = c.beta[b] * (p[3 + 4 * b] - qp.q2);
    const double d1 = p[1 + 4 * a] + pow(fabs(d5), c.eta[a]);
    const double d2 = p[1 + 4 * b] + pow(fabs(d6), c.eta[b]);
    const double xi1 = p[9 + a], xi2 = p[9 + b];
    const double e = p[0] * exp(0.5 * (xi1 * d1 + xi2 * d2));
    if (x) {
        x[1] = e * xi2 * 0.5 * c.eta[b] * pow(fabs(d6), c.eta[b] - 1.0) * copysign(1.0, d6) *
               c.beta[b];
        x[0] = e * xi1 * 0.5 * c.eta[a] * pow(fabs(d5), c.eta[a] - 1.0) * copysign(1.0, d5) *
               c.beta[a];
    }
    return e;
}

// charge_abs (evaluate_energy_charge_pair_abs and
// _electronegativity_charge_pair_abs, Potentials.f90:4571-4731):
// E = sqrt(f1 f2), f_k = max(0, a_k + b_k |q_k - Q_k|^n_k), halves swapped
// as for charge_exp.
__device__ __noinline__ double charge_abs(const ChargeFactor &c, const QPair &qp, double *x)
{
    const int a = qp.s1 != c.orig ? 4 : 0, b = 4 - a;
    const double *p = c.p;
    const double a1 = p[a], b1 = p[a + 1], Q1 = p[a + 2], n1 = p[a + 3];
    const double a2 = p[b], b2 = p[b + 1], Q2 = p[b + 2], n2 = p[b + 3];
    if (x) {
        const double abs1 = fabs(qp.q1 - Q1), abs2 = fabs(qp.q2 - Q2);
        const double pow1 = pow(abs1, n1 - 1.0), pow2 = pow(abs2, n2 - 1.0);
        double f1 = a1 + b1 * abs1 * pow1;
        double f2 = a2 + b2 * abs2 * pow2;
        const double dd1 = -n1 * pow1 * copysign(1.0, qp.q1 - Q1) * b1;
        const double dd2 = -n2 * pow2 * copysign(1.0, qp.q2 - Q2) * b2;
        double h = 0.5 / sqrt(f1 * f2);
        if (f1 < 0.0) { f1 = 0.0; h = 0.0; }
        if (f2 < 0.0) { f2 = 0.0; h = 0.0; }
        x[0] = h * dd1 * f2;
        x[1] = h * dd2 * f1;
    }
    double f1 = a1 + b1 * pow(fabs(qp.q1 - Q1), n1);
    double f2 = a2 + b2 * pow(fabs(qp.q2 - Q2), n2);
    if (f1 < 0.0) f1 = 0.0;
    if (f2 < 0.0) f2 = 0.0;
    return sqrt(f1 * f2);
}

// x^n for an integer n by repeated products (Fortran's real**integer)
__device__ __forceinline__ double ipow(double x, int n)
{
    const int m = n < 0 ? -n : n;
    double y = 1.0;
    for (int k = 0; k < m; ++k) y = y * x;
    return n < 0 ? 1.0 / y : y;
}

// charge_pair (evaluate_energy_charge_pair and
// _electronegativity_charge_pair, Potentials.f90:4415-4506):
// E = eps q1^n1 q2^n2 with integer exponents, swapped as for charge_exp.
__device__ __noinline__ double charge_pair(const ChargeFactor &c, const QPair &qp, double *x)
{
    const bool inv = qp.s1 != c.orig;
    const int n1 = (int)(inv ? c.p[2] : c.p[1]), n2 = (int)(inv ? c.p[1] : c.p[2]);
    if (x) {
        x[0] = n1 > 0 ? -c.p[0] * (n1 * ipow(qp.q1, n1 - 1) * ipow(qp.q2, n2)) : 0.0;
        x[1] = n2 > 0 ? -c.p[0] * (n2 * ipow(qp.q1, n1) * ipow(qp.q2, n2 - 1)) : 0.0;
    }
    return c.p[0] * (ipow(qp.q1, n1) * ipow(qp.q2, n2));
}

// One factor at distance r > 0 (evaluate_energy_component and
// evaluate_force_component, Potentials.f90:2717-2890; the per-type
// routines at 2971-3085, 3643-3677, 4009-4298, 4803-4967).  cut is the
// potential's cutoff (the spring's energy shift uses it).
__device__ __forceinline__ Radial factor(const PairFactor &f, double r, double cut,
                                         const double2 *__restrict__ tab, bool need_force,
                                         const ChargeFactor *cft, const QPair &qp)
{
    Radial o{0.0, 0.0, 0.0};
    const double p0 = f.p[0], p1 = f.p[1], p2 = f.p[2], p3 = f.p[3];
    switch (f.type) {
    case 1: {  // exponential: eps exp(-zeta r)
        const double ex = exp(-p1 * r);
        o.v = p0 * ex;
        if (need_force) o.g = -p1 * p0 * ex / r;
        break;
    }
    case 3: {  // LJ: eps ((sigma/r)^12 - (sigma/r)^6)
        const double ratio = p1 / r;
        const double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
        o.v = p0 * (r6 * r6 - r6);
        if (need_force) o.g = p0 * (6.0 * r6 - 12.0 * r6 * r6) / (r * r);
        break;
    }
    case 4: {  // spring: k/2 ((r - R0)^2 - (cut - R0)^2) inside |r - R0| < |cut - R0|
        const double r2 = r - p1, r6 = cut - p1;
        if (r2 * r2 < r6 * r6) {
            o.v = p0 * 0.5 * (r2 * r2 - r6 * r6);
            if (need_force) o.g = p0 * r2 / r;
        }
        break;
    }
    case 5: {  // Buckingham: A exp(-r/sigma) - C (sigma/r)^6
        const double ratio = r / p2;
        const double r6 = 1.0 / (ratio * ratio * ratio * ratio * ratio * ratio);
        const double ex = exp(-ratio);
        o.v = p0 * ex - p1 * r6;
        if (need_force) o.g = (-p0 * ratio * ex + p1 * 6.0 * r6) / (r * r);
        break;
    }
    case 6: {  // power: eps (a/r)^n
        const double pw = pow(p1 / r, p2);
        o.v = p0 * pw;
        if (need_force) o.g = p0 * (-p2 * pw) / (r * r);
        break;
    }
    case 7: {  // tabulated: cubic Hermite on (value, derivative) rows, scaled
        if (r >= p1) {
            o.v = tab[f.toff + f.tlen - 1].x;  // the reference does not scale it here
            break;
        }
        const double dr = p1 / (f.tlen - 1);
        int lo = (int)floor(r / dr) + 1;       // 1-based, as the reference
        if (lo > f.tlen - 1) lo = f.tlen - 1;  // stay inside the table
        const double t = r / dr - (lo - 1);
        const double t2 = t * t, t3 = t2 * t;
        const double2 a = tab[f.toff + lo - 1], b = tab[f.toff + lo];
        o.v = p2 * (a.x * (2 * t3 - 3 * t2 + 1) + a.y * dr / p1 * (t3 - 2 * t2 + t) +
                    b.x * (3 * t2 - 2 * t3) + b.y * dr / p1 * (t3 - t2));
        if (need_force)
            o.g = p2 *
                  (a.x / dr * 6 * (t2 - t) + a.y / p1 * (3 * t2 - 4 * t + 1) +
                   b.x / dr * 6 * (t - t2) + b.y / p1 * (3 * t2 - 2 * t)) /
                  r;
        break;
    }
    case 8: {  // shift_power: eps ((r1 - r)/(r1 - r2))^n
        const double w = p1 - p2, q = (p1 - r) / w;
        o.v = p0 * pow(q, p3);
        if (need_force) o.g = -p0 * p3 / w * pow(q, p3 - 1.0) / r;
        break;
    }
    case 9: {  // step: e1 below r1, e2 above r2, a cosine in between
        const double w = p3 - p2;
        if (r < p2) {
            o.v = p0;
        } else if (r > p3) {
            o.v = p1;
        } else {
            const double ph = M_PI * (r - p2) / w;
            o.v = (p0 + p1) * 0.5 * (1.0 + 0.5 * cos(ph));
            if (need_force) o.c = (p0 + p1) * 0.25 * M_PI / w * sin(ph);
        }
        break;
    }
    case 12:  // charge_exp: no dependence on r (its force is the cutoff's)
        o.v = charge_exp(cft[f.toff], qp, nullptr);
        break;
    case 13:  // charge_abs
        o.v = charge_abs(cft[f.toff], qp, nullptr);
        break;
    case 14:  // charge_pair
        o.v = charge_pair(cft[f.toff], qp, nullptr);
        break;
    default:
        break;
    }
    return o;
}

// A potential: the product of its factors (evaluate_energy / evaluate_forces,
// Potentials.f90:2640-2693, 2791-2814): E = prod E_i,
// f = sum_i (E / E_i) f_i, or prod_{j != i} E_j f_i where E_i = 0.
// Two passes keep every factor in registers: the first forms E, the product
// of the nonzero energies and the number of zero ones; the second
// re-evaluates each factor with its force.
__device__ __forceinline__ Radial potential(const GenPair &T, double r,
                                            const double2 *__restrict__ tab, bool need_force,
                                            const ChargeFactor *cft, const QPair &qp)
{
    const int nf = T.nf;
    if (nf == 1) return factor(T.f[0], r, T.cut, tab, need_force, cft, qp);
    double e = 1.0, enz = 1.0;
    int zeros = 0;
    for (int k = 0; k < nf; ++k) {
        const double v = factor(T.f[k], r, T.cut, tab, false, cft, qp).v;
        e = e * v;
        if (v != 0.0) enz = enz * v;
        else ++zeros;
    }
    Radial out{e, 0.0, 0.0};
    if (!need_force) return out;
    for (int k = 0; k < nf; ++k) {
        // a charge-dependent factor has no r dependence: nothing to add
        if (T.f[k].type >= 12) continue;
        const Radial q = factor(T.f[k], r, T.cut, tab, true, cft, qp);
        const double w = q.v != 0.0 ? e / q.v : (zeros == 1 ? enz : 0.0);
        out.g += w * q.g;
        out.c += w * q.c;
    }
    return out;
}

// -dE/dq of both atoms of a potential (evaluate_electronegativity,
// Potentials.f90:2580-2637): x = sum_k (E / E_k) x_k over the factors, or
// prod_{j != k} E_j x_k where E_k = 0; factors without charge dependence
// have x_k = 0.
__device__ void potential_chi(const GenPair &T, double r, const double2 *__restrict__ tab,
                              const ChargeFactor *cft, const QPair &qp, double *x)
{
    x[0] = x[1] = 0.0;
    double me[PYSIC_MAXF], mx[PYSIC_MAXF][2], e = 1.0;
    const int nf = T.nf;
    for (int k = 0; k < nf; ++k) {
        const PairFactor &f = T.f[k];
        mx[k][0] = mx[k][1] = 0.0;
        if (f.type == 12) me[k] = charge_exp(cft[f.toff], qp, mx[k]);
        else if (f.type == 13) me[k] = charge_abs(cft[f.toff], qp, mx[k]);
        else if (f.type == 14) me[k] = charge_pair(cft[f.toff], qp, mx[k]);
        else me[k] = factor(f, r, T.cut, tab, false, cft, qp).v;
    }
    if (nf == 1) {
        x[0] = mx[0][0];
        x[1] = mx[0][1];
        return;
    }
    for (int k = 0; k < nf; ++k) e = e * me[k];
    for (int k = 0; k < nf; ++k) {
        double w = 1.0;
        if (me[k] != 0.0) {
            w = e / me[k];
        } else {
            for (int j = 0; j < nf; ++j)
                if (j != k) w = w * me[j];
        }
        x[0] = x[0] + w * mx[k][0];
        x[1] = x[1] + w * mx[k][1];
    }
}

// Lennard-Jones alone, the common molecular-dynamics model: the LJPair
// table of the model (common.h).  Its force needs r^2 only; the square root
// is taken only where an exact comparison of r with the cutoff or the soft
// cutoff needs it.
// LDS of each kernel kind: 0 any pair potentials (the whole model),
// 1 LJ alone (the table per ordered species pair), 2 LJ alone with one
// species (nothing: its one table entry is read with scalar loads)
template <int KIND> struct PairLds { LJPair lj[PYSIC_MAXS][PYSIC_MAXS]; };
template <> struct PairLds<0> { PairModel pm; };
template <> struct PairLds<2> { int unused; };


// The model into LDS (a fixed copy per block; the caller synchronises).
__device__ __forceinline__ void stage(PairLds<0> &L, const PairModel *__restrict__ PM)
{
    const int *src = reinterpret_cast<const int *>(PM);
    int *dst = reinterpret_cast<int *>(&L.pm);
    for (int w = threadIdx.x; w < (int)(sizeof(PairModel) / sizeof(int)); w += PB) dst[w] = src[w];
}
__device__ __forceinline__ void stage(PairLds<1> &L, const PairModel *__restrict__ PM)
{
    const int *src = reinterpret_cast<const int *>(&PM->lj[0][0]);
    int *dst = reinterpret_cast<int *>(&L.lj[0][0]);
    for (int w = threadIdx.x; w < (int)(sizeof(L.lj) / sizeof(int)); w += PB) dst[w] = src[w];
}
__device__ __forceinline__ void stage(PairLds<2> &, const PairModel *__restrict__) {}

// The pair (i, j) at separation d = r_ij, r2 = |d|^2, for the ordered
// species pair (si, sj): energy v (both ends each add half) and the force
// on i, g = G d + C (1,1,1).  Returns false when nothing acts.
// The smoothing of an LJ pair inside its soft-cutoff shell: v f_c and
// G f_c + v f_c' / r.  Out of line: the unrolled neighbor loop keeps one
// copy of the trigonometric code instead of one per neighbor.
__device__ __noinline__ double2 smooth_lj(double cut, double soft, double r2, double r, double v,
                                          double G)
{
    if (r < 0.0) r = sqrt(r2);
    const double f = fc64(r, cut, soft);
    return make_double2(v * f, G * f + v * dfc64(r, cut, soft) / r);
}

template <bool FORCE>
__device__ __forceinline__ bool visit(const LJPair &q, double r2, double &v, double &G, double &C)
{
    if (!q.on || !(r2 < q.hi2)) return false;
    double r = -1.0;
    if (!(r2 < q.lo2)) {
        r = sqrt(r2);
        if (!(q.cut > r)) return false;
    }
    C = 0.0;
    if (FORCE && r2 == 0.0) {  // the reference's forces vanish at r == 0
        v = 0.0;
        G = 0.0;
        return true;
    }
    const double inv_r2 = recip64(r2);
    const double x = q.sig2 * inv_r2;
    const double s6 = x * x * x;
    v = q.eps * (s6 * s6 - s6);
    G = FORCE ? q.eps * (6.0 * s6 - 12.0 * s6 * s6) * inv_r2 : 0.0;
    if (q.smooth && !(r2 < q.soft_lo2)) {
        const double2 vg = smooth_lj(q.cut, q.soft, r2, r, v, G);
        v = vg.x;
        G = FORCE ? vg.y : 0.0;
    }
    return true;
}

template <bool FORCE>
__device__ __forceinline__ bool visit(const PairLds<0> &L, int si, int sj, double r2, double &v,
                                      double &G, double &C, const double2 *__restrict__ tab,
                                      const QPair &qp)
{
    const PairModel &P = L.pm;
    const int np = P.npair[si][sj];
    if (np == 0 || !(r2 < P.cut2_hi[si][sj])) return false;
    const double r = sqrt(r2);
    v = 0.0;
    G = 0.0;
    C = 0.0;
    for (int c = 0; c < np; ++c) {
        const GenPair &T = P.pair[si][sj][c];
        if (!(T.cut > r)) continue;
        if (!FORCE) {
            const Radial q = potential(T, r, tab, false, P.cf, qp);
            v += T.smooth ? q.v * fc64(r, T.cut, T.soft) : q.v;
        } else if (r > 0.0) {  // the reference's forces vanish at r == 0
            const Radial q = potential(T, r, tab, true, P.cf, qp);
            if (T.smooth) {
                // f1 fc + v grad(fc), grad(fc) = fc' s / r
                const double f = fc64(r, T.cut, T.soft), df = dfc64(r, T.cut, T.soft);
                G += q.g * f + q.v * df / r;
                C += q.c * f;
            } else {
                G += q.g;
                C += q.c;
            }
        }
    }
    return true;
}

// Neighbors are taken a batch at a time: their indices, then their
// positions, are loaded before any of them is evaluated.  8 for the LJ
// kernels (measured: 333 us per launch on 1M argon atoms vs 361 with 4 and
// 431 with 2); 2 for the general kernel, whose registers would not hold 8.
#ifndef PYSIC_PAIR_NBU
#define PYSIC_PAIR_NBU 8
#endif

// KIND 0: any pair potentials; 1: every pair potential is a lone LJ;
// 2: a lone LJ and one species.  G lanes per atom (consecutive threads):
// lane l takes the list rows l NBU .. l NBU + NBU - 1, then every G NBU-th
// batch; the lanes' forces are summed by shuffles and lane 0 writes them.
// The general kind runs 8 lanes per atom up to 2^18 atoms: its per-pair
// work (exp, pow and the products) is long enough that one thread per atom
// leaves a system of some ten thousand atoms on a fraction of the chip
// (24,576 atoms of COMB SiO2: 381 us with 1 lane, 209 with 2, 144 with 4,
// 140 with 8).  Above that, one thread per atom already fills the chip, and
// each extra block would stage the whole model into LDS once more.
constexpr int PAIR_LANES_SMALL = 8;
constexpr int PAIR_LANES_MAX_ATOMS = 1 << 18;

template <class Geo, int MODE, int KIND, int G>
__global__ __launch_bounds__(PB) void k_pairs(int n, Geo geo, const double4 *__restrict__ pos,
                                              const int *__restrict__ id,
                                              const int *__restrict__ nbr,
                                              const int *__restrict__ nimg,
                                              const int *__restrict__ ncnt,
                                              const PairModel *__restrict__ PM,
                                              const double2 *__restrict__ tab,
                                              const double *__restrict__ q,
                                              double4 *__restrict__ frc, int accumulate,
                                              double4 *__restrict__ mom, double half_dt, int kick,
                                              double *__restrict__ partial)
{
    constexpr bool FORCE = MODE != PAIR_ENERGY;
    constexpr bool VIR = MODE == PAIR_FORCES;
    constexpr int NBU = KIND == 0 ? 2 : PYSIC_PAIR_NBU;
    __shared__ PairLds<KIND> lds;
    __shared__ double red[7][PB];
    stage(lds, PM);
    __syncthreads();
    const int blk = xcd_block(gridDim.x);
    const int gid = blk * PB + threadIdx.x;
    const int i = gid / G, lane = gid % G;
    double e = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
    if (i < n) {
        const double4 pi = pos[i];
        const int si = (int)pi.w;
        if (PM->has_single && lane == 0) {
            const double kfx = PM->kforce[si][0], kfy = PM->kforce[si][1],
                         kfz = PM->kforce[si][2];
            if (FORCE) {
                fx = kfx; fy = kfy; fz = kfz;
            } else {
                // E = V and -F . r (the dot product of Quaternions.f90:475-483)
                e = PM->konst[si] + ((-kfx * pi.x + -kfy * pi.y) + -kfz * pi.z);
            }
        }
        const double *cv = geo.vec(i);
        const int nf = ncnt[i];
        // pick decides the sign of a 'step' force and which atom is atom 1
        // of a charge factor
        const bool pick = KIND == 0 && FORCE && PM->needs_pick;
        const bool charged = KIND == 0 && PM->has_charge;
        const int idi = pick || charged ? id[i] : 0;
        const double qi = charged ? q[i] : 0.0;
        LJPair one{};
        if constexpr (KIND == 2) one = PM->lj[0][0];
        // list rows t0 .. t0 + NBU - 1 of atom i (column-major: row t at t n + i)
        const size_t stride = (size_t)n;
        for (int t0 = lane * NBU; t0 < nf; t0 += G * NBU) {
            const int *pn = nbr + (size_t)t0 * stride + i, *pm = nimg + (size_t)t0 * stride + i;
            int jv[NBU], iv[NBU];
#pragma unroll
            for (int u = 0; u < NBU; ++u) {
                jv[u] = i;  // past the end: the atom itself (loaded, not evaluated)
                iv[u] = 0;
                if (t0 + u < nf) {
                    jv[u] = pn[u * stride];
                    iv[u] = pm[u * stride];
                }
            }
            double4 pv[NBU];
#pragma unroll
            for (int u = 0; u < NBU; ++u) pv[u] = pos[jv[u]];
#pragma unroll
            for (int u = 0; u < NBU; ++u) {
                if (t0 + u >= nf) break;
                const double3 d = separation64(pi, pv[u], iv[u], cv);
                const double r2 = dot64(d);
                double v, G2, C;
                bool on;
                if constexpr (KIND == 2) on = visit<FORCE>(one, r2, v, G2, C);
                else if constexpr (KIND == 1) on = visit<FORCE>(lds.lj[si][(int)pv[u].w], r2, v, G2, C);
                else {
                    QPair qp{0.0, 0.0, 0};
                    if (charged) {
                        const int j = jv[u], sj = (int)pv[u].w;
                        const double qj = q[j];
                        if (pick_dev(idi, id[j], iv[u])) qp = QPair{qi, qj, si};
                        else qp = QPair{qj, qi, sj};
                    }
                    on = visit<FORCE>(lds, si, (int)pv[u].w, r2, v, G2, C, tab, qp);
                }
                if (!on) continue;
                if (!FORCE) {
                    e += 0.5 * v;
                } else {
                    // atom i is the first atom when pick(i, j) holds: g_i = f1
                    // with s = r_ij; otherwise g_i = -f1 with s = -r_ij
                    double cs = C;
                    if (pick && C != 0.0 && !pick_dev(idi, id[jv[u]], iv[u])) cs = -C;
                    const double gx = G2 * d.x + cs, gy = G2 * d.y + cs, gz = G2 * d.z + cs;
                    fx += gx; fy += gy; fz += gz;
                    if (VIR) {
                        w0 -= 0.5 * d.x * gx;
                        w1 -= 0.5 * d.y * gy;
                        w2 -= 0.5 * d.z * gz;
                        w3 -= 0.5 * d.y * gz;
                        w4 -= 0.5 * d.x * gz;
                        w5 -= 0.5 * d.x * gy;
                    }
                }
            }
        }
    }
    if (FORCE) {
        // the lanes of one atom are consecutive threads of one wave; every
        // thread of the block reaches these shuffles
        if constexpr (G > 1) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) {
                fx += __shfl_xor(fx, o);
                fy += __shfl_xor(fy, o);
                fz += __shfl_xor(fz, o);
            }
        }
        if (i < n && lane == 0) {
            double4 f = make_double4(fx, fy, fz, 0.0);
            if (accumulate) {
                const double4 a = frc[i];
                f = make_double4(a.x + fx, a.y + fy, a.z + fz, 0.0);
            }
            frc[i] = f;
            if (kick) {
                double4 p = mom[i];
                p.x += half_dt * f.x;
                p.y += half_dt * f.y;
                p.z += half_dt * f.z;
                mom[i] = p;
            }
        }
    }
    if (MODE == PAIR_FORCES_ONLY) return;
    // block partials (fixed tree): energy, or the six virial components
    const int tid = threadIdx.x;
    red[0][tid] = e;
    red[1][tid] = w0; red[2][tid] = w1; red[3][tid] = w2;
    red[4][tid] = w3; red[5][tid] = w4; red[6][tid] = w5;
    __syncthreads();
    for (int s = PB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int c = 0; c < 7; ++c) red[c][tid] += red[c][tid + s];
        __syncthreads();
    }
    if (tid < 7) partial[(size_t)blk * 8 + tid] = red[tid][0];
}

// gtotals[c] = sum over blocks of partial[b*8 + c], fixed order (block c)
__global__ __launch_bounds__(PB) void k_pair_totals(int nblocks, const double *__restrict__ partial,
                                                    double *__restrict__ totals)
{
    __shared__ double red[PB];
    const int c = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int b = tid; b < nblocks; b += PB) s += partial[(size_t)b * 8 + c];
    red[tid] = s;
    __syncthreads();
    for (int w = PB / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) totals[c] = red[0];
}

// The charge-dependent pair terms' share of the electronegativities:
// chi[i] += sum over i's pairs of -dE/dq_i f_c(r).  G lanes per atom walk
// its full list (every G-th row each, summed by shuffles); each pair is
// oriented as the reference takes it, and each end adds its own atom's
// share (the reference adds both at once).
template <class Geo, int G>
__global__ __launch_bounds__(PB) void k_pair_chi(int n, Geo geo, const double4 *__restrict__ pos,
                                                 const int *__restrict__ id,
                                                 const int *__restrict__ nbr,
                                                 const int *__restrict__ nimg,
                                                 const int *__restrict__ ncnt,
                                                 const PairModel *__restrict__ PM,
                                                 const double2 *__restrict__ tab,
                                                 const double *__restrict__ q,
                                                 double *__restrict__ chi)
{
    __shared__ PairLds<0> lds;
    stage(lds, PM);
    __syncthreads();
    const int gid = xcd_block(gridDim.x) * PB + threadIdx.x;
    const int i = gid / G, lane = gid % G;
    double x = 0.0;
    if (i < n) {
        const PairModel &P = lds.pm;
        const double4 pi = pos[i];
        const int si = (int)pi.w, idi = id[i];
        const double qi = q[i];
        const double *cv = geo.vec(i);
        const int nf = ncnt[i];
        for (int t = lane; t < nf; t += G) {
            const int j = nbr[(size_t)t * n + i], img = nimg[(size_t)t * n + i];
            const double4 pj = pos[j];
            const int sj = (int)pj.w;
            const int np = P.npair[si][sj];
            if (np == 0) continue;
            const double3 d = separation64(pi, pj, img, cv);
            const double r2 = dot64(d);
            if (!(r2 < P.cut2_hi[si][sj])) continue;
            const double r = sqrt(r2);
            const bool first = pick_dev(idi, id[j], img);
            const double qj = q[j];
            const QPair qp = first ? QPair{qi, qj, si} : QPair{qj, qi, sj};
            for (int c = 0; c < np; ++c) {
                const GenPair &T = P.pair[si][sj][c];
                if (!(T.cut > r)) continue;
                double xs[2];
                potential_chi(T, r, tab, P.cf, qp, xs);
                const double s = first ? xs[0] : xs[1];
                x += T.smooth ? s * fc64(r, T.cut, T.soft) : s;
            }
        }
    }
    // the lanes of one atom are consecutive threads of one wave; every
    // thread reaches the shuffles
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (i < n && lane == 0) chi[i] += x;
}

template <class Geo, int KIND, int G>
void launch(System &S, const Geo &geo, int mode, bool accumulate, double half_dt, int kick)
{
    const int nb = div_up(S.n * G, PB);
    const int acc = accumulate ? 1 : 0;
    auto args = [&](auto kern) {
        kern<<<nb, PB, 0, S.stream>>>(S.n, geo, S.pos.p, S.id.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                      S.dpmodel.p, S.tables.p, S.q.p, S.frc.p, acc, S.mom.p,
                                      half_dt, kick, S.gpartial.p);
    };
    if (mode == PAIR_ENERGY) args(k_pairs<Geo, PAIR_ENERGY, KIND, G>);
    else if (mode == PAIR_FORCES) args(k_pairs<Geo, PAIR_FORCES, KIND, G>);
    else args(k_pairs<Geo, PAIR_FORCES_ONLY, KIND, G>);
    HIP_TRY(hipGetLastError());
}




// ---------------------------------------------------------------------------
// Bond-order-scaled pair potentials (COMB's attraction).  The reference
// multiplies a coordinated potential's pair energy by
// (b_i + b_j + B*_ij + B*_ji) / 2, b the per-atom (level-1) factors -- zero
// for the level-2 'tersoff' factor -- and B*_ij the pair factor from the
// sum zeta_ij over the center i's other neighbors k
// (core_calculate_pair_bond_order_factor, Core.f90:2439-2637, post-processed
// by the group's first factor of i's species, 2763-2807), and adds the
// gradients of B*_ij and B*_ji to the forces (core_add_pair_bond_order_forces,
// Core.f90:4952-5353) and the factor to the electronegativities
// (Core.f90:5688-5729).  Here that pair energy is split between the two
// atoms as site energies U_i = sum_j 1/2 f_c(r_ij) V_ij B*_ij, V_ij taken in
// the reference's orientation (pick), so that E = sum_i U_i; U_i depends
// on i's neighbor vectors r_ij and on the charges.  Pass 1 (one lane per
// partner j) forms zeta_ij, B*_ij, U_i and, per short-list slot, dU_i/dr_ij
// through V and through zeta's dependence on r_ij, and dU_i/dq_j; pass 2
// (one lane per slot t as the third atom) adds dU_i/dr_it through the
// zetas of i's other bonds.  The gathers form F_a = sum_t dU_a/dr_at -
// sum_{i: a in list(i)} dU_i/dr_ia and chi_a = -dU_a/dq_a - sum_i dU_i/dq_a,
// the reverse entries found in each neighbor's list at (a, -o).  No
// atomics, fixed order; all arithmetic is FP64.
// ---------------------------------------------------------------------------

enum { BOP_CHI = 3 };

// x^mu for the exponent of the Tersoff term: the integer exponents of the
// usual parametrisations (mu = 1 or 3, and mu - 1 = 0 or 2) by products,
// any other by the library pow
__device__ __forceinline__ double mu_pow(double x, double mu)
{
    if (mu == 1.0) return x;
    if (mu == 3.0) return x * x * x;
    if (mu == 2.0) return x * x;
    if (mu == 0.0) return 1.0;
    return pow(x, mu);
}

// The pieces of one Tersoff zeta term shared by its value and gradients:
// s1 = r_i - r_j (partner to center), s2 = r_k - r_i, r1 and r2 their norms
struct BoTerm {
    double xi, gee;    // f_c(r2) exp((a (r1 - r2))^mu) and g(theta)
    double dxi;        // d xi / d(r1 - r2) (at fixed f_c)
    double gc;         // d g / d cos, divided by r1 r2
    double dot, ex, dfc;
};

__device__ __forceinline__ BoTerm bo_parts(const double3 &s1, double r1, const double3 &s2,
                                           double r2, const BondTerm &b, bool grad)
{
    BoTerm t;
    t.dot = -s1.x * s2.x + -s1.y * s2.y + -s1.z * s2.z;
    const double ratio = recip64(r1 * r2);
    const double cosine = t.dot * ratio;
    const double cc = b.c * b.c, dd = b.d * b.d, hc = b.h - cosine;
    const double den = dd + hc * hc, iden = recip64(den);
    double decay, dfc;
    smooth64(r2, b.cut, b.soft, &decay, &dfc);
    t.dfc = dfc;
    const double x = b.a * (r1 - r2);
    t.gee = 1.0 + cc * recip64(dd) - cc * iden;
    if (grad) {
        const double e1a = mu_pow(x, b.mu - 1.0);
        t.ex = exp(x * e1a);
        t.dxi = decay * (t.ex * b.a * b.mu * e1a);
        t.gc = -cc * iden * iden * 2.0 * hc * ratio;
    } else {
        t.ex = exp(mu_pow(x, b.mu));
        t.dxi = 0.0;
        t.gc = 0.0;
    }
    t.xi = decay * t.ex;
    return t;
}

// evaluate_pair_bond_order_factor_tersoff (Potentials.f90:5744-5795)
__device__ __forceinline__ double bo_term(const double3 &s1, double r1, const double3 &s2,
                                          double r2, const BondTerm &b)
{
    const BoTerm t = bo_parts(s1, r1, s2, r2, b, false);
    return t.xi * t.gee;
}

// evaluate_pair_bond_order_gradient_tersoff (Potentials.f90:5806-5902): the
// gradient of bo_term with respect to the partner's position (gp) and the
// third atom's (gk); either may be null.  Returns bo_term's value.
__device__ __forceinline__ double bo_grad(const double3 &s1, double r1, const double3 &s2,
                                          double r2, const BondTerm &b, double3 *gp, double3 *gk)
{
    const BoTerm t = bo_parts(s1, r1, s2, r2, b, true);
    const double3 t1 = make_double3(-s1.x, -s1.y, -s1.z);
    if (gp) {
        // d r1 / d r_partner = t1 / r1; d cos / d r_partner = t3 / (r1 r2)
        const double ir1 = recip64(r1);
        const double k1 = t.dot * ir1 * ir1;
        const double3 t3 = make_double3(s2.x - k1 * t1.x, s2.y - k1 * t1.y, s2.z - k1 * t1.z);
        const double a = t.dxi * t.gee * ir1, c = t.xi * t.gc;
        *gp = make_double3(a * t1.x + c * t3.x, a * t1.y + c * t3.y, a * t1.z + c * t3.z);
    }
    if (gk) {
        // d r2 / d r_k = s2 / r2 (through f_c and through r1 - r2); d cos / d r_k = t4 / (r1 r2)
        const double ir2 = recip64(r2);
        const double k2 = t.dot * ir2 * ir2;
        const double3 t4 = make_double3(t1.x - k2 * s2.x, t1.y - k2 * s2.y, t1.z - k2 * s2.z);
        const double a = (t.dfc * t.ex - t.dxi) * t.gee * ir2, c = t.xi * t.gc;
        *gk = make_double3(a * s2.x + c * t4.x, a * s2.y + c * t4.y, a * s2.z + c * t4.z);
    }
    return t.xi * t.gee;
}

// post_process_bond_order_factor_tersoff (Potentials.f90:5910-5921) and its
// derivative (post_process_bond_order_gradient_tersoff on a unit gradient,
// 5929-5961, with the zeta = 0 branch); beta zeta >= 0
__device__ __forceinline__ double bo_post(double zeta, const BondTerm &b)
{
    return powpos64(1.0 + powpos64(b.beta * zeta, b.eta), -1.0 / (2.0 * b.eta));
}
__device__ __forceinline__ double bo_dpost(double zeta, const BondTerm &b)
{
    if (b.eta == 0.0) return 0.0;
    if (zeta == 0.0 && b.eta <= 1.0) return 0.0;
    const double inv_eta = -1.0 / (2.0 * b.eta);
    return inv_eta * powpos64(1.0 + powpos64(b.beta * zeta, b.eta), inv_eta - 1.0) * b.eta *
           powpos64(b.beta * zeta, b.eta - 1.0) * b.beta;
}

// Each atom's neighbors within the bond-order pass's radius, widened by
// 1e-10 relative so that a pair in one atom's list is in the other's.
template <class Geo>
__global__ __launch_bounds__(PB) void k_bop_list(int n, Geo geo, const double4 *__restrict__ pos,
                                                 const int *__restrict__ nbr,
                                                 const int *__restrict__ nimg,
                                                 const int *__restrict__ ncnt, double radius,
                                                 int *__restrict__ osj, int *__restrict__ osimg,
                                                 int *__restrict__ oscnt)
{
    const int v = xcd_block(gridDim.x) * PB + threadIdx.x;
    if (v >= n) return;
    const double4 pv = pos[v];
    const double *cv = geo.vec(v);
    const size_t N = (size_t)n;
    const double r2max = radius * radius * (1.0 + 1e-10);
    const int nf = ncnt[v];
    int c = 0;
    for (int t = 0; t < nf; ++t) {
        const int j = nbr[(size_t)t * N + v], img = nimg[(size_t)t * N + v];
        if (!(dot64(separation64(pv, pos[j], img, cv)) < r2max)) continue;
        osj[(size_t)c * N + v] = j;
        osimg[(size_t)c * N + v] = img;
        ++c;
    }
    oscnt[v] = c;
}

// The first BOP_STAGE entries of each atom's bond-order list, staged once per
// atom in LDS for the triplet loops of both passes: the vector r_k - r_i
// with its length, and the species of k.  Entries beyond are formed from
// global memory where they are used, by the same functions, so the values
// are the same either way.
constexpr int BOP_STAGE = 16;

template <int G>
struct BopStage {
    double4 v[BOP_STAGE][PB / G];
    int sp[BOP_STAGE][PB / G];
};

template <int G, class Geo>
__device__ __forceinline__ void bop_stage(BopStage<G> &st, int i, int n, int lane, int a,
                                          const Geo &geo, const double4 *__restrict__ pos,
                                          const int *__restrict__ osj,
                                          const int *__restrict__ osimg,
                                          const int *__restrict__ oscnt)
{
    if (i < n) {
        const size_t N = (size_t)n;
        const double4 pi = pos[i];
        const double *cv = geo.vec(i);
        const int c = oscnt[i] < BOP_STAGE ? oscnt[i] : BOP_STAGE;
        for (int u = lane; u < c; u += G) {
            const double4 pk = pos[osj[(size_t)u * N + i]];
            const double3 s2 = separation64(pi, pk, osimg[(size_t)u * N + i], cv);
            st.v[u][a] = make_double4(s2.x, s2.y, s2.z, sqrt(dot64(s2)));
            st.sp[u][a] = (int)pk.w;
        }
    }
    __syncthreads();
}

// Entry u of atom i's list: r_k - r_i, its length, the species of k.
template <int G>
__device__ __forceinline__ void bop_entry(const BopStage<G> &st, int u, int i, int a, size_t N,
                                          const double4 &pi, const double *cv,
                                          const double4 *__restrict__ pos,
                                          const int *__restrict__ osj,
                                          const int *__restrict__ osimg, double3 &s, double &r,
                                          int &sk)
{
    if (u < BOP_STAGE) {
        const double4 v = st.v[u][a];
        s = make_double3(v.x, v.y, v.z);
        r = v.w;
        sk = st.sp[u][a];
    } else {
        const double4 pk = pos[osj[(size_t)u * N + i]];
        s = separation64(pi, pk, osimg[(size_t)u * N + i], cv);
        r = sqrt(dot64(s));
        sk = (int)pk.w;
    }
}

// Pass 1: G lanes per atom, lane l takes the partner slots l, l + G, ...
template <class Geo, int MODE, int G>
__global__ __launch_bounds__(PB) void k_bop_pairs(int n, Geo geo, const double4 *__restrict__ pos,
                                                  const int *__restrict__ id,
                                                  const BopModel *__restrict__ BM,
                                                  const PairModel *__restrict__ PM,
                                                  const double2 *__restrict__ tab,
                                                  const double *__restrict__ q,
                                                  const int *__restrict__ osj,
                                                  const int *__restrict__ osimg,
                                                  const int *__restrict__ oscnt,
                                                  double4 *__restrict__ oD, double *__restrict__ oW,
                                                  double *__restrict__ oown,
                                                  double *__restrict__ partial, int blk0)
{
    constexpr bool FORCE = MODE == PAIR_FORCES || MODE == PAIR_FORCES_ONLY;
    constexpr bool CHI = MODE == BOP_CHI;
    __shared__ double red[PB];
    __shared__ BopStage<G> stg;
    const int blk = xcd_block(gridDim.x);
    const int gid = blk * PB + threadIdx.x;
    const int i = gid / G, lane = gid % G, a = threadIdx.x / G;
    const size_t N = (size_t)n;
    bop_stage<G>(stg, i, n, lane, a, geo, pos, osj, osimg, oscnt);
    double e = 0.0, own = 0.0;
    if (i < n) {
        const double4 pi = pos[i];
        const int si = (int)pi.w, idi = id[i];
        const double qi = q[i];
        const double *cv = geo.vec(i);
        const double mb = BM->maxbond[si];
        const int c = oscnt[i];
        for (int t = lane; t < c; t += G) {
            const int j = osj[(size_t)t * N + i], img = osimg[(size_t)t * N + i];
            double3 d;
            double r;
            int sj;
            bop_entry<G>(stg, t, i, a, N, pi, cv, pos, osj, osimg, d, r, sj);
            double4 D = make_double4(0.0, 0.0, 0.0, 0.0);
            double wk = 0.0;
            if (BM->nbop[si][sj] && BM->pair[si][sj].cut > r && r > 0.0) {
                const GenPair &T = BM->pair[si][sj];
                const int g = BM->group[si][sj];
                const double3 s1 = make_double3(-d.x, -d.y, -d.z);
                double zeta = 0.0;
                double3 gp = make_double3(0.0, 0.0, 0.0);
                for (int u = 0; u < c; ++u) {
                    double3 s2;
                    double r2;
                    int sk;
                    bop_entry<G>(stg, u, i, a, N, pi, cv, pos, osj, osimg, s2, r2, sk);
                    const BondTerm &b = BM->bond[g][si][sj][sk];
                    if (!b.valid) continue;
                    // the partner itself (the same vector) is not a third atom
                    if (!(s1.x + s2.x != 0.0 || s1.y + s2.y != 0.0 || s1.z + s2.z != 0.0))
                        continue;
                    if (!(r2 < mb)) continue;
                    if (FORCE) {
                        double3 gpk;
                        zeta += bo_grad(s1, r, s2, r2, b, &gpk, nullptr);
                        gp.x += gpk.x;
                        gp.y += gpk.y;
                        gp.z += gpk.z;
                    } else {
                        zeta += bo_term(s1, r, s2, r2, b);
                    }
                }
                const BondTerm &pp = BM->post[g][si];
                const double bt = pp.valid ? bo_post(zeta, pp) : zeta;
                const bool first = pick_dev(idi, id[j], img);
                const double qj = q[j];
                const QPair qp = first ? QPair{qi, qj, si} : QPair{qj, qi, sj};
                const double fc = T.smooth ? fc64(r, T.cut, T.soft) : 1.0;
                if (CHI) {
                    // potential_chi gives -dV/dq in the pair's orientation
                    double xs[2];
                    potential_chi(T, r, tab, PM->cf, qp, xs);
                    own += -0.5 * fc * bt * (first ? xs[0] : xs[1]);
                    D.w = -0.5 * fc * bt * (first ? xs[1] : xs[0]);
                } else {
                    const Radial pv = potential(T, r, tab, FORCE, PM->cf, qp);
                    e += 0.5 * fc * pv.v * bt;
                    if (FORCE) {
                        // d(f_c V)/dr_ij = (G f_c + V f_c' / r) d: central (a
                        // 'step' factor is refused for these potentials)
                        double gs = pv.g * fc;
                        if (T.smooth) gs += pv.v * dfc64(r, T.cut, T.soft) / r;
                        const double dbt = pp.valid ? bo_dpost(zeta, pp) : 1.0;
                        wk = 0.5 * fc * pv.v * dbt;
                        D.x = 0.5 * bt * gs * d.x + wk * gp.x;
                        D.y = 0.5 * bt * gs * d.y + wk * gp.y;
                        D.z = 0.5 * bt * gs * d.z + wk * gp.z;
                    }
                }
            }
            if (FORCE || CHI) oD[(size_t)t * N + i] = D;
            if (FORCE) oW[(size_t)t * N + i] = wk;
        }
    }
    if (CHI) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) own += __shfl_xor(own, o);
        if (i < n && lane == 0) oown[i] = own;
        return;
    }
    if (MODE == PAIR_FORCES_ONLY) return;
    const int tid = threadIdx.x;
    red[tid] = e;
    __syncthreads();
    for (int w = PB / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid < 8) partial[(size_t)(blk0 + blk) * 8 + tid] = tid == 0 ? red[0] : 0.0;
}

// Pass 2: G lanes per atom, lane l takes the slots t = l, l + G, ... as the
// third atom of every other bond of the atom; then the virial
// -sum_t r_it (x) dU_i/dr_it.
template <class Geo, int MODE, int G>
__global__ __launch_bounds__(PB) void k_bop_third(int n, Geo geo, const double4 *__restrict__ pos,
                                                  const BopModel *__restrict__ BM,
                                                  const int *__restrict__ osj,
                                                  const int *__restrict__ osimg,
                                                  const int *__restrict__ oscnt,
                                                  double4 *__restrict__ oD,
                                                  const double *__restrict__ oW,
                                                  double *__restrict__ partial, int blk0)
{
    constexpr bool VIR = MODE == PAIR_FORCES;
    __shared__ double red[6][PB];
    __shared__ BopStage<G> stg;
    const int blk = xcd_block(gridDim.x);
    const int gid = blk * PB + threadIdx.x;
    const int i = gid / G, lane = gid % G, a = threadIdx.x / G;
    const size_t N = (size_t)n;
    bop_stage<G>(stg, i, n, lane, a, geo, pos, osj, osimg, oscnt);
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
    if (i < n) {
        const double4 pi = pos[i];
        const int si = (int)pi.w;
        const double *cv = geo.vec(i);
        const double mb = BM->maxbond[si];
        const int c = oscnt[i];
        for (int t = lane; t < c; t += G) {
            double3 s2;
            double r2;
            int sk;
            bop_entry<G>(stg, t, i, a, N, pi, cv, pos, osj, osimg, s2, r2, sk);
            double4 D = oD[(size_t)t * N + i];
            if (r2 < mb) {
                for (int u = 0; u < c; ++u) {
                    const double wj = oW[(size_t)u * N + i];
                    if (wj == 0.0) continue;
                    double3 d;
                    double r1;
                    int sj;
                    bop_entry<G>(stg, u, i, a, N, pi, cv, pos, osj, osimg, d, r1, sj);
                    const BondTerm &b = BM->bond[BM->group[si][sj]][si][sj][sk];
                    if (!b.valid) continue;
                    const double3 s1 = make_double3(-d.x, -d.y, -d.z);
                    if (!(s1.x + s2.x != 0.0 || s1.y + s2.y != 0.0 || s1.z + s2.z != 0.0))
                        continue;
                    double3 gk;
                    bo_grad(s1, r1, s2, r2, b, nullptr, &gk);
                    D.x += wj * gk.x;
                    D.y += wj * gk.y;
                    D.z += wj * gk.z;
                }
                oD[(size_t)t * N + i] = D;
            }
            if (VIR) {
                w0 -= s2.x * D.x;
                w1 -= s2.y * D.y;
                w2 -= s2.z * D.z;
                w3 -= s2.y * D.z;
                w4 -= s2.x * D.z;
                w5 -= s2.x * D.y;
            }
        }
    }
    if (MODE == PAIR_FORCES_ONLY) return;
    const int tid = threadIdx.x;
    red[0][tid] = w0; red[1][tid] = w1; red[2][tid] = w2;
    red[3][tid] = w3; red[4][tid] = w4; red[5][tid] = w5;
    __syncthreads();
    for (int s = PB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int q = 0; q < 6; ++q) red[q][tid] += red[q][tid + s];
        __syncthreads();
    }
    if (tid < 8) partial[(size_t)(blk0 + blk) * 8 + tid] = (tid >= 1 && tid <= 6) ? red[tid - 1][0] : 0.0;
}

// F_a += sum_t dU_a/dr_at - sum of the reverse entries (F_a as in the
// header); with CHI, chi_a += -dU_a/dq_a - sum of the reverse entries' w.
// G lanes per atom split a's slots; shuffles sum them.
template <int G, bool CHI>
__global__ __launch_bounds__(PB) void k_bop_gather(int n, const int *__restrict__ osj,
                                                   const int *__restrict__ osimg,
                                                   const int *__restrict__ oscnt,
                                                   const double4 *__restrict__ oD,
                                                   const double *__restrict__ oown,
                                                   double4 *__restrict__ frc,
                                                   double4 *__restrict__ mom, double half_dt,
                                                   int kick, double *__restrict__ chi)
{
    const int gid = xcd_block(gridDim.x) * PB + threadIdx.x;
    const int a = gid / G, lane = gid % G;
    const size_t N = (size_t)n;
    double fx = 0.0, fy = 0.0, fz = 0.0, x = 0.0;
    if (a < n) {
        const int c = oscnt[a];
        for (int t = lane; t < c; t += G) {
            if (!CHI) {
                const double4 own = oD[(size_t)t * N + a];
                fx += own.x;
                fy += own.y;
                fz += own.z;
            }
            const int v = osj[(size_t)t * N + a];
            const int want = neg_img(osimg[(size_t)t * N + a]);
            const int cv = oscnt[v];
            for (int u = 0; u < cv; ++u) {
                if (osj[(size_t)u * N + v] == a && osimg[(size_t)u * N + v] == want) {
                    const double4 g = oD[(size_t)u * N + v];
                    fx -= g.x;
                    fy -= g.y;
                    fz -= g.z;
                    x -= g.w;
                    break;
                }
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o);
        fy += __shfl_xor(fy, o);
        fz += __shfl_xor(fz, o);
        x += __shfl_xor(x, o);
    }
    if (a >= n || lane != 0) return;
    if (CHI) {
        chi[a] += x - oown[a];
        return;
    }
    double4 f = frc[a];
    f.x += fx;
    f.y += fy;
    f.z += fz;
    frc[a] = f;
    if (kick) {
        double4 p = mom[a];
        p.x += half_dt * f.x;
        p.y += half_dt * f.y;
        p.z += half_dt * f.z;
        mom[a] = p;
    }
}

}  // namespace

void pr_eval(System &S, int mode, bool accumulate, double half_dt, int kick)
{
    PYSIC_REQUIRE(S.dpmodel.p && S.tables.p, PYSIC_ESTATE, "pair model not compiled");
    PYSIC_REQUIRE(S.n > 0, PYSIC_ESTATE, "no atoms");
    const int kind = S.pmodel.lj_only ? (S.model.nspecies == 1 ? 2 : 1) : 0;
    const bool lanes = kind == 0 && S.n <= PAIR_LANES_MAX_ATOMS;
    const int nb = div_up(S.n * (lanes ? PAIR_LANES_SMALL : 1), PB);
    // bond_bend angles, then the bond-order-scaled pairs, add their
    // partials after the pair blocks and their forces after the pair forces
    // (the last of them the half-kick, if asked)
    const int bb = bend_blocks(S);
    const int ob = bop_blocks(S, mode);
    S.gpartial.ensure((size_t)(nb + bb + ob) * 8);
    S.gtotals.ensure(8);
    S.gpartial_blocks = nb + bb + ob;
    hipEvent_t ev{};
    const char *name = mode == PAIR_ENERGY ? "pair_energy" : "pair_forces";
    S.timer.begin(name, S.stream, &ev);
    const int pkick = (bb || ob) ? 0 : kick;
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        if (kind == 2) launch<Geo, 2, 1>(S, geo, mode, accumulate, half_dt, pkick);
        else if (kind == 1) launch<Geo, 1, 1>(S, geo, mode, accumulate, half_dt, pkick);
        else if (lanes) launch<Geo, 0, PAIR_LANES_SMALL>(S, geo, mode, accumulate, half_dt, pkick);
        else launch<Geo, 0, 1>(S, geo, mode, accumulate, half_dt, pkick);
    });
    S.timer.end(name, S.stream, ev);
    if (bb) bend_eval(S, mode, nb, half_dt, ob ? 0 : kick);
    if (ob) bop_eval(S, mode, nb + bb, half_dt, kick);
    if (mode != PAIR_FORCES_ONLY) {
        k_pair_totals<<<7, PB, 0, S.stream>>>(nb + bb + ob, S.gpartial.p, S.gtotals.p);
        HIP_TRY(hipGetLastError());
    }
}

void pr_chi(System &S, double *chi)
{
    if (!S.pmodel.has_charge) return;
    PYSIC_REQUIRE(S.dpmodel.p && S.tables.p, PYSIC_ESTATE, "pair model not compiled");
    constexpr int G = PAIR_LANES_SMALL;
    const int nb = div_up(S.n * G, PB);
    with_geo(S, [&](auto geo) {
        k_pair_chi<decltype(geo), G><<<nb, PB, 0, S.stream>>>(S.n, geo, S.pos.p, S.id.p, S.nbr.p,
                                                             S.nimg.p, S.ncnt.p, S.dpmodel.p,
                                                             S.tables.p, S.q.p, chi);
    });
    HIP_TRY(hipGetLastError());
    bop_chi(S, chi);
}

// lanes per atom of the bond-order passes: an oxide's few neighbors within
// 3 A, one or two slots per lane
constexpr int BOP_LANES = 8;

int bop_blocks(const System &S, int mode)
{
    if (!S.bmodel.on) return 0;
    const int nb = div_up(S.n * BOP_LANES, PB);
    return mode == PAIR_ENERGY ? nb : 2 * nb;
}

static void bop_prepare(System &S)
{
    const int n = S.n;
    const size_t cap = (size_t)S.cap * n;
    S.osj.ensure(cap);
    S.osimg.ensure(cap);
    S.oscnt.ensure(n);
    S.oD.ensure(cap);
    S.oW.ensure(cap);
    S.oown.ensure(n);
    with_geo(S, [&](auto geo) {
        k_bop_list<decltype(geo)><<<div_up(n, PB), PB, 0, S.stream>>>(
            n, geo, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p, S.bmodel.radius, S.osj.p, S.osimg.p,
            S.oscnt.p);
    });
    HIP_TRY(hipGetLastError());
}

void bop_eval(System &S, int mode, int blk0, double half_dt, int kick)
{
    if (!S.bmodel.on) return;
    const int n = S.n, nb = div_up(n * BOP_LANES, PB);
    hipEvent_t ev{};
    S.timer.begin("bond_order_pairs", S.stream, &ev);
    bop_prepare(S);
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        auto one = [&](auto kern) {
            kern<<<nb, PB, 0, S.stream>>>(n, geo, S.pos.p, S.id.p, S.dbmodel.p, S.dpmodel.p,
                                          S.tables.p, S.q.p, S.osj.p, S.osimg.p, S.oscnt.p,
                                          S.oD.p, S.oW.p, S.oown.p, S.gpartial.p, blk0);
        };
        auto two = [&](auto kern) {
            kern<<<nb, PB, 0, S.stream>>>(n, geo, S.pos.p, S.dbmodel.p, S.osj.p, S.osimg.p,
                                          S.oscnt.p, S.oD.p, S.oW.p, S.gpartial.p, blk0 + nb);
        };
        if (mode == PAIR_ENERGY) {
            one(k_bop_pairs<Geo, PAIR_ENERGY, BOP_LANES>);
        } else if (mode == PAIR_FORCES) {
            one(k_bop_pairs<Geo, PAIR_FORCES, BOP_LANES>);
            two(k_bop_third<Geo, PAIR_FORCES, BOP_LANES>);
        } else {
            one(k_bop_pairs<Geo, PAIR_FORCES_ONLY, BOP_LANES>);
            two(k_bop_third<Geo, PAIR_FORCES_ONLY, BOP_LANES>);
        }
    });
    HIP_TRY(hipGetLastError());
    if (mode != PAIR_ENERGY) {
        k_bop_gather<BOP_LANES, false><<<div_up(n * BOP_LANES, PB), PB, 0, S.stream>>>(
            n, S.osj.p, S.osimg.p, S.oscnt.p, S.oD.p, S.oown.p, S.frc.p, S.mom.p, half_dt, kick,
            nullptr);
        HIP_TRY(hipGetLastError());
    }
    S.timer.end("bond_order_pairs", S.stream, ev);
}

void bop_chi(System &S, double *chi)
{
    if (!S.bmodel.on) return;
    const int n = S.n, nb = div_up(n * BOP_LANES, PB);
    bop_prepare(S);
    with_geo(S, [&](auto geo) {
        k_bop_pairs<decltype(geo), BOP_CHI, BOP_LANES><<<nb, PB, 0, S.stream>>>(
            n, geo, S.pos.p, S.id.p, S.dbmodel.p, S.dpmodel.p, S.tables.p, S.q.p, S.osj.p,
            S.osimg.p, S.oscnt.p, S.oD.p, S.oW.p, S.oown.p, nullptr, 0);
    });
    HIP_TRY(hipGetLastError());
    k_bop_gather<BOP_LANES, true><<<nb, PB, 0, S.stream>>>(n, S.osj.p, S.osimg.p, S.oscnt.p,
                                                           S.oD.p, S.oown.p, nullptr, nullptr,
                                                           0.0, 0, chi);
    HIP_TRY(hipGetLastError());
}
