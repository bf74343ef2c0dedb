// ewald.hip -- Ewald summation and EEM charge relaxation on the MI355X.
//
// Restates calculate_ewald_energy / _forces / _electronegativities
// (Potentials.f90:1428-2249) plus the inline real-space terms of the
// interaction loop (Core.f90:3416-3460) and the charge_self singlets
// (Core.f90:4095-4239, Potentials.f90:4341-4372), all in FP64.
//
//  * real space: every atom gathers over its full neighbor list (each pair
//    is evaluated from both ends, no scatter); unscaled charges as in the
//    reference's main loop;
//  * reciprocal space: the half-space k-vectors with |k| < k_c are enumerated
//    once on the host in the reference's order and grouped into segments of
//    up to 16 consecutive k3 at fixed (k1, k2).  With the fractional
//    coordinates u_a (k . r = 2 pi (k1 u1 + k2 u2 + k3 u3)), one phase per
//    (segment, atom) gives a segment's first factor and the per-atom row
//    exp(2 pi i l u3) the rest.  Structure factors S(k) = sum_i q_i s_i
//    exp(i k.r_i): a (segment x atom) by (atom x 16) complex product on the
//    FP64 matrix cores, deterministic partial sums over atom chunks.  Per-atom
//    consumers (forces, electronegativities): per segment, the polynomial
//    sum_l z(l) exp(2 pi i l u3) of each atom, again a complex product on
//    the matrix cores (16 atoms x 16 powers by 16 powers x 16 segments);
//  * self and charged-background terms are per-atom closed forms.
// The reciprocal "stress" uses absolute positions (Potentials.f90:1940-1945)
// and the background electronegativity uses a zero-initialised charge sum
// (the reference leaves tmp_qsum uninitialised, Potentials.f90:2002).
#include "ewald.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int EB = 256;

// Real-space Ewald + charge_self singlets + self/background chi, per atom.
// G lanes share an atom (G divides 64): lane g takes list entries g, g + G,
// ...; the per-atom sums are reduced over the G lanes with shuffles, so
// small systems still fill the GPU.
template <int G>
__global__ __launch_bounds__(EB) void k_real(int n, CellGeom g, const double4 *__restrict__ pos,
                                             const double *__restrict__ q,
                                             const double *__restrict__ sc,
                                             const int *__restrict__ nbr,
                                             const int *__restrict__ nimg,
                                             const int *__restrict__ ncnt,
                                             const Model *__restrict__ M, int ewald_on, double rc,
                                             double inv_eps_4pi, double inv_sig_sqrt2,
                                             double inv_sig_sqrt2pi, double inv_sig_sq2,
                                             double bg_chi, const double *__restrict__ bg_dev,
                                             int mode, int self_on, double *__restrict__ out1,
                                             double4 *__restrict__ frc, double *__restrict__ part)
{
    __shared__ double red[EB][7];
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gid / G, sub = gid % G;
    double e = 0.0, chi = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 pi = pos[i];
        int si = (int)pi.w;
        double qi = q[i];
        // charge_self singlets: E = eps q^n, chi = -n eps q^(n-1)
        if (self_on && M->has_self && sub == 0) {
            double qp = 1.0;  // q^(nn-1)
            e += M->self_eps[si][0];
            for (int nn = 1; nn <= 4; ++nn) {
                double eps = M->self_eps[si][nn];
                chi += -nn * eps * qp;
                qp *= qi;
                e += eps * qp;
            }
        }
        if (ewald_on) {
            if (sub == 0) {
                // self (Potentials.f90:2126) and background (2166) electronegativity
                const double qs = qi * sc[i];
                const double bg = bg_dev ? bg_dev[0] : bg_chi;
                chi += 2.0 * qs * inv_eps_4pi * inv_sig_sqrt2pi + sc[i] * bg;
            }
            const double inv2perpi = 2.0 * inv_sig_sqrt2pi;
            int nf = ncnt[i];
            for (int t = sub; t < nf; t += G) {
                int j = nbr[(size_t)t * n + i];
                int im = nimg[(size_t)t * n + i];
                double3 sp3 = separation64(pi, pos[j], im, g.v);
                double x = sp3.x, y = sp3.y, z = sp3.z;
                double d = sqrt(dot64(sp3));
                if (!(d < rc)) continue;
                double qj = q[j];
                double erfcv = erfc(d * inv_sig_sqrt2);
                if (mode == 1) {
                    e += 0.5 * qi * qj * erfcv / d * inv_eps_4pi;
                } else if (mode == 2) {
                    double id = 1.0 / d;
                    double pre = -inv_eps_4pi * qi * qj *
                                 (erfcv * id * id + inv2perpi * exp(-d * d * inv_sig_sq2) * id) * id;
                    double tx = pre * x, ty = pre * y, tz = pre * z;
                    fx += tx; fy += ty; fz += tz;
                    w[0] -= 0.5 * x * tx;
                    w[1] -= 0.5 * y * ty;
                    w[2] -= 0.5 * z * tz;
                    w[3] -= 0.5 * y * tz;
                    w[4] -= 0.5 * x * tz;
                    w[5] -= 0.5 * x * ty;
                } else {
                    chi += -inv_eps_4pi / d * erfcv * qj;
                }
            }
        }
    }
    // per-atom sums over the G lanes of an atom (all lanes of a wave take
    // part: G divides 64 and the lanes past n carry zeros)
    if (G > 1) {
#pragma unroll
        for (int m = 1; m < G; m <<= 1) {
            if (mode == 2) {
                fx += __shfl_xor(fx, m, G);
                fy += __shfl_xor(fy, m, G);
                fz += __shfl_xor(fz, m, G);
            }
            if (mode == 3) chi += __shfl_xor(chi, m, G);
        }
    }
    if (i < n && sub == 0) {
        if (mode == 2) {
            double4 f = frc[i];
            frc[i] = make_double4(f.x + fx, f.y + fy, f.z + fz, 0.0);
        }
        if (mode == 3) out1[i] = chi;
    }
    // block partials: energy (mode 1) or the 6 virial components (mode 2),
    // summed over every lane
    red[threadIdx.x][0] = (i < n) ? e : 0.0;
    for (int c = 0; c < 6; ++c) red[threadIdx.x][1 + c] = (i < n) ? w[c] : 0.0;
    __syncthreads();
    for (int s = EB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 7; ++c) red[threadIdx.x][c] += red[threadIdx.x + s][c];
        __syncthreads();
    }
    if (threadIdx.x < 7) part[(size_t)blockIdx.x * 8 + threadIdx.x] = red[0][threadIdx.x];
}

constexpr int LSEG = 16;   // k-vectors per segment

struct Recip {
    double b[9];   // reciprocal vectors in rows: k = k1 b1 + k2 b2 + k3 b3
    double f[9];   // the same over 2 pi: rows of the inverse cell matrix
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(2 pi i u) for a phase u in turns.  The reduction is exact: removing
// the integer part leaves r in [-1/2, 1/2] (|u| is far below 2^52), and
// removing the nearest quarter turn q/4 leaves f in [-1/8, 1/8] (Sterbenz).
// sin and cos of x = 2 pi f, |x| <= pi/4, are their Taylor series through
// x^17 and x^16 (the first dropped terms are below 1e-17), and the quarter
// turns rotate the result: exp(2 pi i u) = i^q exp(i x).
__device__ __forceinline__ double2 turn(double u)
{
    const double r = u - rint(u);
    const double qd = rint(4.0 * r);
    const double x = 6.283185307179586 * fma(-0.25, qd, r);
    const double x2 = x * x;
    // Horner in x^2 with the coefficients (-1)^k / (2k+1)! and (-1)^k / (2k)!
    double ps = -2.8114572543455206e-15;              // -1/17!
    ps = fma(ps, x2, 7.647163731819816e-13);          //  1/15!
    ps = fma(ps, x2, -1.6059043836821613e-10);        // -1/13!
    ps = fma(ps, x2, 2.505210838544172e-08);          //  1/11!
    ps = fma(ps, x2, -2.7557319223985893e-06);        // -1/9!
    ps = fma(ps, x2, 1.984126984126984e-04);          //  1/7!
    ps = fma(ps, x2, -8.333333333333333e-03);         // -1/5!
    ps = fma(ps, x2, 1.6666666666666666e-01);         //  1/3!
    const double sn = fma(-x * x2, ps, x);
    double pc = 4.779477332387385e-14;                //  1/16!
    pc = fma(pc, x2, -1.1470745597729725e-11);        // -1/14!
    pc = fma(pc, x2, 2.08767569878681e-09);           //  1/12!
    pc = fma(pc, x2, -2.755731922398589e-07);         // -1/10!
    pc = fma(pc, x2, 2.48015873015873e-05);           //  1/8!
    pc = fma(pc, x2, -1.388888888888889e-03);         // -1/6!
    pc = fma(pc, x2, 4.1666666666666664e-02);         //  1/4!
    pc = fma(pc, x2, -0.5);                           // -1/2!
    const double cs = fma(pc, x2, 1.0);
    const int q = (int)qd & 3;   // i^q: 1, i, -1, -i
    const double c = (q & 1) ? sn : cs;
    const double s = (q & 1) ? cs : sn;
    return make_double2((q == 1 || q == 2) ? -c : c, (q >= 2) ? -s : s);
}

// Per-atom phase table: th[i] = (u1, u2, u3, q s), u_a = f_a . r the
// fractional coordinates (k . r = 2 pi (k1 u1 + k2 u2 + k3 u3)), the
// structure-factor row pw[i][l] = q s exp(2 pi i l u3), l < LSEG, and the
// step between consecutive segments of a row, wl[i] = exp(2 pi i LSEG u3).
__global__ __launch_bounds__(EB) void k_phase_table(int n, Recip R, const double4 *__restrict__ pos,
                                                    const double *__restrict__ q,
                                                    const double *__restrict__ sc,
                                                    double4 *__restrict__ th,
                                                    double2 *__restrict__ pw,
                                                    double2 *__restrict__ wl)
{
    const int i = blockIdx.x * EB + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    const double u1 = R.f[0] * p.x + R.f[1] * p.y + R.f[2] * p.z;
    const double u2 = R.f[3] * p.x + R.f[4] * p.y + R.f[5] * p.z;
    const double u3 = R.f[6] * p.x + R.f[7] * p.y + R.f[8] * p.z;
    const double qs = q[i] * sc[i];
    th[i] = make_double4(u1, u2, u3, qs);
    wl[i] = turn(LSEG * u3);
    double2 *row = pw + (size_t)i * LSEG;
    for (int l = 0; l < LSEG; ++l) {
        const double2 e = turn(l * u3);
        row[l] = make_double2(qs * e.x, qs * e.y);
    }
}

// S(k) partials on the matrix cores.  Over an atom chunk, for a segment
// (k1, k2, l0) and its entries l < LSEG,
//   S_seg(l) = sum_a exp(2 pi i (k1 u1 + k2 u2 + l0 u3)_a) pw_a(l),
// a complex product of the (segment x atom) phase matrix, built on the fly,
// and the (atom x entry) table pw: FP64 MFMA 16x16x4, four real products
// per complex one.  A wave owns SK_MT tiles of 16 segments (the work plan
// below) and steps through the atoms 4 at a time; lane L forms the phase of
// the segment in row L % 16 and atom L / 16 of each tile (the A operand)
// and loads entry L % 16 of that atom's row (the B operand, coalesced).
// When tile t holds the segments that follow tile t - 1's in the same
// (k1, k2) rows, its phases are tile t - 1's times exp(2 pi i LSEG u3): one
// complex multiply instead of a phase.  Operand layout of
// v_mfma_f64_16x16x4f64: A (16 x 4) lane i + 16 k holds A[i][k]; B (4 x 16)
// lane j + 16 k holds B[k][j]; D (16 x 16) lane j + 16 (i % 4), register
// i / 4, holds D[i][j].
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
constexpr int SK_MT = 4;   // 16-segment tiles per wave


__global__ __launch_bounds__(EB) void k_structure(int nwaves, int nk, int n, int chunk,
                                                  const int *__restrict__ plan,
                                                  const int *__restrict__ pflags,
                                                  const int4 *__restrict__ segs,
                                                  const double4 *__restrict__ th,
                                                  const double2 *__restrict__ wl,
                                                  const double2 *__restrict__ pw,
                                                  double2 *__restrict__ skpart)
{
    const int wave = blockIdx.x * (EB / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave >= nwaves) return;   // whole waves only; the kernel has no barrier
    const int lane = threadIdx.x & 63, row = lane & 15, kq = lane >> 4;
    const int *wplan = plan + (size_t)wave * (16 * SK_MT);
    const int flags = pflags[wave];   // bit t: tile t continues tile t - 1
    const int a0 = blockIdx.y * chunk, a1 = min(n, a0 + chunk);
    // each tile's (k1, k2, l0) packed in one register: 11 + 11 signed bits
    // and 10 bits (|k1|, |k2|, l0 < 1024, checked on the host)
    int kp[SK_MT];
    mfma_d4 dre[SK_MT], dim[SK_MT];
#pragma unroll
    for (int t = 0; t < SK_MT; ++t) {
        const int sidx = wplan[t * 16 + row];
        const int4 sg = sidx >= 0 ? segs[sidx] : make_int4(0, 0, 0, 0);
        kp[t] = (sg.x & 0x7ff) | ((sg.y & 0x7ff) << 11) | ((sg.z & 0x3ff) << 22);
        dre[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        dim[t] = dre[t];
    }
    // software pipeline: the products of step a are issued before the
    // phases of step a + 4 are formed from operands loaded one step
    // earlier; step a + 8 is being loaded.  Atoms past the chunk carry a
    // zero row and add nothing.
    const double4 zero4 = make_double4(0, 0, 0, 0);
    const double2 zero2 = make_double2(0, 0);
    double2 v[SK_MT];
    auto phases = [&](const double4 &u, const double2 &w) {
#pragma unroll
        for (int t = 0; t < SK_MT; ++t)
            v[t] = (t > 0 && ((flags >> t) & 1))
                       ? cmul(v[t > 0 ? t - 1 : 0], w)
                       : turn((double)((kp[t] << 21) >> 21) * u.x +
                              (double)((kp[t] << 10) >> 21) * u.y +
                              (double)((unsigned)kp[t] >> 22) * u.z);
    };
    int at = a0 + kq;
    double2 pa = at < a1 ? pw[(size_t)at * LSEG + row] : zero2;
    phases(at < a1 ? th[at] : zero4, at < a1 ? wl[at] : zero2);
    at += 4;
    double4 tb = at < a1 ? th[at] : zero4;
    double2 wb = at < a1 ? wl[at] : zero2;
    double2 pb = at < a1 ? pw[(size_t)at * LSEG + row] : zero2;
    for (int a = a0; a < a1; a += 4) {
        at = a + 8 + kq;
        const double4 tc = at < a1 ? th[at] : zero4;
        const double2 wc = at < a1 ? wl[at] : zero2;
        const double2 pc = at < a1 ? pw[(size_t)at * LSEG + row] : zero2;
        // the two products into one accumulator are 2 SK_MT products apart,
        // so neither waits for the other's result
#pragma unroll
        for (int t = 0; t < SK_MT; ++t) {
            dre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[t].x, pa.x, dre[t], 0, 0, 0);
            dim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[t].x, pa.y, dim[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < SK_MT; ++t) {
            dre[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-v[t].y, pa.y, dre[t], 0, 0, 0);
            dim[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[t].y, pa.x, dim[t], 0, 0, 0);
        }
        phases(tb, wb);
        pa = pb;
        tb = tc;
        wb = wc;
        pb = pc;
    }
    // lane holds D[i][j] for row i = kq + 4 r and entry j = row
    double2 *out = skpart + (size_t)blockIdx.y * nk;
#pragma unroll
    for (int t = 0; t < SK_MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int sidx = wplan[t * 16 + kq + 4 * r];
            if (sidx < 0) continue;
            const int4 sg = segs[sidx];
            if (row < (sg.z >> 16)) out[sg.w + row] = make_double2(dre[t][r], dim[t][r]);
        }
}

__global__ void k_structure_sum(int nk, int nchunk, const double2 *__restrict__ skpart,
                                double2 *__restrict__ sk)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    double c = 0.0, s = 0.0;
    for (int b = 0; b < nchunk; ++b) {
        double2 v = skpart[(size_t)b * nk + k];
        c += v.x;
        s += v.y;
    }
    sk[k] = make_double2(c, s);
}

// E_rec = sum_k c_k |S(k)|^2 (Potentials.f90:1610-1639), block partials
__global__ void k_recip_energy(int nk, const double4 *__restrict__ kvec,
                               const double2 *__restrict__ sk, double *__restrict__ part)
{
    __shared__ double red[EB];
    int k = blockIdx.x * EB + threadIdx.x;
    double v = 0.0;
    if (k < nk) {
        double2 s = sk[k];
        v = kvec[k].w * (s.x * s.x + s.y * s.y);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Coefficients of the per-atom sums in the order of the structure-factor
// work plan, for k_recip_mfma.  Forces (mode 2): z = 2 c_k (S_sin + i S_cos),
// the force on atom i being -q_i s_i sum_k Re(z exp(i k.r_i)) k
// (Potentials.f90:1895-1931).  Electronegativities (mode 3):
// z = 2 c_k (S_cos - i S_sin), chi_i -= sum_k Re(z exp(i k.r_i))
// (Potentials.f90:2196-2229).  zp[(T * 16 + l) * 16 + c] = (z, l z) for
// entry l of the segment in column c of plan tile T (zero past its length
// or for an empty column); kpk[T * 16 + c] is its (k1, k2, l0) packed as in
// k_structure.  Thread per (tile, entry, column).
__global__ void k_recip_coef_plan(int ntiles, int mode, const int *__restrict__ plan,
                                  const int4 *__restrict__ segs, const double4 *__restrict__ kvec,
                                  const double2 *__restrict__ sk, double4 *__restrict__ zp,
                                  int *__restrict__ kpk)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntiles * 256) return;
    const int T = e >> 8, l = (e >> 4) & 15, c = e & 15;
    const int sidx = plan[T * 16 + c];
    const int4 sg = sidx >= 0 ? segs[sidx] : make_int4(0, 0, 0, 0);
    double4 o = make_double4(0, 0, 0, 0);
    if (l < (sg.z >> 16)) {
        const int k = sg.w + l;
        const double c2 = 2.0 * kvec[k].w;
        const double2 S = sk[k];
        const double zr = mode == 3 ? c2 * S.x : c2 * S.y;
        const double zi = mode == 3 ? -(c2 * S.y) : c2 * S.x;
        o = make_double4(zr, zi, l * zr, l * zi);
    }
    zp[e] = o;
    if (l == 0) kpk[T * 16 + c] = (sg.x & 0x7ff) | ((sg.y & 0x7ff) << 11) | ((sg.z & 0x3ff) << 22);
}

// Per-atom reciprocal forces (MODE 2) or electronegativities (MODE 3) on
// the matrix cores.  For atom i and segment j with coefficients z_j(l),
//   P[i][j] = sum_l w_i^l z_j(l),   Q[i][j] = sum_l w_i^l l z_j(l),
// w_i = exp(2 pi i u3_i): a (16 atoms x 16 powers) by (16 x 16 segments)
// complex product per tile, FP64 MFMA 16x16x4 over 4 steps of l, four real
// products per complex one.  With v_ij the phase of segment j's first k at
// atom i,
//   chi_i -= sum_j Re(v P),
//   F_i = -q_i s_i sum_j [Re(v P) (k1 b1 + k2 b2 + l0 b3) + Re(v Q) b3].
// A wave owns 16 atoms and walks a chunk of the structure-factor work plan
// (waves of 4 tiles of 16 segments); in a continuing tile v is the previous
// tile's times exp(2 pi i 16 u3).  Lane L: A operand w^(4 ks + L / 16) of
// atom L % 16; B operand coefficient 4 ks + L / 16 of the tile's segment
// L % 16; D holds (atom L / 16 + 4 r, segment L % 16), r < 4, and the
// per-atom sums are reduced over the 16 lanes of a row at the end.
template <int MODE>
__global__ __launch_bounds__(EB) void k_recip_mfma(int n, int nwaves, int wchunk, Recip R,
                                                   const int *__restrict__ pflags,
                                                   const double4 *__restrict__ zp,
                                                   const int *__restrict__ kpk,
                                                   const double4 *__restrict__ th,
                                                   const double2 *__restrict__ wl,
                                                   double4 *__restrict__ part)
{
    const int lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
    const int ab = (blockIdx.x * (EB / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * 16;
    if (ab >= n) return;   // whole waves only; the kernel has no barrier
    const int T0 = blockIdx.y * wchunk * 4, T1 = min(nwaves, (blockIdx.y + 1) * wchunk) * 4;
    const double4 zero4 = make_double4(0, 0, 0, 0);
    // A operands: powers of w of atom ab + col (rows past n are dropped)
    double2 pa[4];
    {
        const int i = ab + col;
        const double u = i < n ? th[i].z : 0.0;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) pa[ks] = turn((4 * ks + kq) * u);
    }
    // electronegativities (P only) take three real products per complex
    // one: at 2 waves per SIMD either way, it is the faster form there
    // (forces, with Q as well, would drop to 1 wave per SIMD)
    constexpr bool GAUSS = MODE == 3;
    double asum[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) asum[ks] = pa[ks].x + pa[ks].y;
    // the atoms of this lane's D rows
    double u1[4], u2[4], u3[4], g1[4], g2[4], g3[4];
    double2 stp[4], v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = ab + kq + 4 * r;
        const double4 t = i < n ? th[i] : zero4;
        u1[r] = t.x;
        u2[r] = t.y;
        u3[r] = t.z;
        stp[r] = i < n ? wl[i] : make_double2(1.0, 0.0);
        v[r] = make_double2(1.0, 0.0);
        g1[r] = g2[r] = g3[r] = 0.0;
    }
    // the next tile's operands are loaded while this tile's products run
    double4 cn[4];
    int kpn = 0;
    if (T0 < T1) {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) cn[ks] = zp[((size_t)T0 * 16 + 4 * ks + kq) * 16 + col];
        kpn = kpk[T0 * 16 + col];
    }
#pragma unroll 1
    for (int T = T0; T < T1; ++T) {
        double4 c[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) c[ks] = cn[ks];
        const int kp = kpn;
        if (T + 1 < T1) {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                cn[ks] = zp[((size_t)(T + 1) * 16 + 4 * ks + kq) * 16 + col];
            kpn = kpk[(T + 1) * 16 + col];
        }
        mfma_d4 pr = {0.0, 0.0, 0.0, 0.0}, pi = pr, qr = pr, qi = pr;
        if (GAUSS) {
            // pr, pi, ps: Re A Re B, Im A Im B, (Re A + Im A)(Re B + Im B)
            mfma_d4 ps = pr;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                pr = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].x, c[ks].x, pr, 0, 0, 0);
                pi = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].y, c[ks].y, pi, 0, 0, 0);
                ps = __builtin_amdgcn_mfma_f64_16x16x4f64(asum[ks], c[ks].x + c[ks].y, ps, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double a = pr[r], b = pi[r];
                pr[r] = a - b;
                pi[r] = ps[r] - a - b;
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                pr = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].x, c[ks].x, pr, 0, 0, 0);
                pi = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].x, c[ks].y, pi, 0, 0, 0);
                if (MODE == 2) {
                    qr = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].x, c[ks].z, qr, 0, 0, 0);
                    qi = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].x, c[ks].w, qi, 0, 0, 0);
                }
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                pr = __builtin_amdgcn_mfma_f64_16x16x4f64(-pa[ks].y, c[ks].y, pr, 0, 0, 0);
                pi = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].y, c[ks].x, pi, 0, 0, 0);
                if (MODE == 2) {
                    qr = __builtin_amdgcn_mfma_f64_16x16x4f64(-pa[ks].y, c[ks].w, qr, 0, 0, 0);
                    qi = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks].y, c[ks].z, qi, 0, 0, 0);
                }
            }
        }
        const double k1 = (kp << 21) >> 21, k2 = (kp << 10) >> 21, l0 = (unsigned)kp >> 22;
        const bool cont = (pflags[T >> 2] >> (T & 3)) & 1;   // never set for tile 0 of a wave
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            v[r] = cont ? cmul(v[r], stp[r]) : turn(k1 * u1[r] + k2 * u2[r] + l0 * u3[r]);
            const double r0 = v[r].x * pr[r] - v[r].y * pi[r];
            if (MODE == 2) {
                const double r1 = v[r].x * qr[r] - v[r].y * qi[r];
                g1[r] = fma(k1, r0, g1[r]);
                g2[r] = fma(k2, r0, g2[r]);
                g3[r] = fma(l0, r0, g3[r]) + r1;
            } else {
                g1[r] += r0;
            }
        }
    }
    // sum over the 16 lanes (segment columns) of each atom row
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
            g1[r] += __shfl_xor(g1[r], m, 16);
            if (MODE == 2) {
                g2[r] += __shfl_xor(g2[r], m, 16);
                g3[r] += __shfl_xor(g3[r], m, 16);
            }
        }
    if (col != 0) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = ab + kq + 4 * r;
        if (i >= n) continue;
        double4 o = make_double4(-g1[r], 0.0, 0.0, 0.0);
        if (MODE == 2) {
            const double m = -th[i].w;
            o.x = m * (g1[r] * R.b[0] + g2[r] * R.b[3] + g3[r] * R.b[6]);
            o.y = m * (g1[r] * R.b[1] + g2[r] * R.b[4] + g3[r] * R.b[7]);
            o.z = m * (g1[r] * R.b[2] + g2[r] * R.b[5] + g3[r] * R.b[8]);
        }
        part[(size_t)blockIdx.y * n + i] = o;
    }
}

// Sum the k-chunk partials per atom; mode 2 adds forces and forms the
// absolute-position virial partials, mode 3 adds to chi.
__global__ __launch_bounds__(EB) void k_recip_finish(int n, int nchunk, int mode,
                                                     const double4 *__restrict__ part,
                                                     const double4 *__restrict__ pos,
                                                     double4 *__restrict__ frc,
                                                     double *__restrict__ chi,
                                                     double *__restrict__ bpart)
{
    __shared__ double red[EB][6];
    int i = blockIdx.x * EB + threadIdx.x;
    double w[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int b = 0; b < nchunk; ++b) {
            double4 v = part[(size_t)b * n + i];
            ax += v.x;
            ay += v.y;
            az += v.z;
        }
        if (mode == 3) {
            chi[i] += ax;
        } else {
            double4 f = frc[i];
            frc[i] = make_double4(f.x + ax, f.y + ay, f.z + az, 0.0);
            double4 r = pos[i];
            w[0] = r.x * ax; w[1] = r.y * ay; w[2] = r.z * az;
            w[3] = r.y * az; w[4] = r.x * az; w[5] = r.x * ay;
        }
    }
    for (int c = 0; c < 6; ++c) red[threadIdx.x][c] = w[c];
    __syncthreads();
    for (int s = EB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 6; ++c) red[threadIdx.x][c] += red[threadIdx.x + s][c];
        __syncthreads();
    }
    if (threadIdx.x < 6) bpart[(size_t)blockIdx.x * 8 + threadIdx.x] = red[0][threadIdx.x];
}

// Fingerprint of what S(k) and the phase table depend on besides the cell
// and the k-vectors: positions, charges and scalers in the current atom
// order.  Each atom's words and index go through the splitmix64 finaliser;
// the atoms' values are summed (block partials, then k_hash_sum).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(EB) void k_state_hash(int n, const double4 *__restrict__ pos,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ sc,
                                                   unsigned long long *__restrict__ part)
{
    __shared__ unsigned long long red[EB];
    const int i = blockIdx.x * EB + threadIdx.x;
    unsigned long long h = 0;
    if (i < n) {
        const double4 p = pos[i];
        h = mix64(((unsigned long long)i * 0x9e3779b97f4a7c15ULL) ^ __double_as_longlong(p.x));
        h = mix64(h ^ __double_as_longlong(p.y));
        h = mix64(h ^ __double_as_longlong(p.z));
        h = mix64(h ^ __double_as_longlong(q[i]));
        h = mix64(h ^ __double_as_longlong(sc[i]));
    }
    red[threadIdx.x] = h;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(EB) void k_hash_sum(int count, const unsigned long long *__restrict__ part,
                                                 unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long red[EB];
    unsigned long long h = 0;
    for (int b = threadIdx.x; b < count; b += EB) h += part[b];
    red[threadIdx.x] = h;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// sum of q*s and of (q*s)^2 (block partials)
__global__ void k_charge_sums(int n, const double *__restrict__ q, const double *__restrict__ sc,
                              double *__restrict__ part)
{
    __shared__ double r1[EB], r2[EB];
    int i = blockIdx.x * EB + threadIdx.x;
    double v = i < n ? q[i] * sc[i] : 0.0;
    r1[threadIdx.x] = v;
    r2[threadIdx.x] = v * v;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[(size_t)blockIdx.x * 8] = r1[0];
        part[(size_t)blockIdx.x * 8 + 1] = r2[0];
    }
}

// block partials of the dipole sum_i q_i s_i r_i (positions as stored:
// absolute, unwrapped), components 0..2 of each 8-double row
__global__ void k_dipole_sums(int n, const double4 *__restrict__ pos,
                              const double *__restrict__ q, const double *__restrict__ sc,
                              double *__restrict__ part)
{
    __shared__ double r[3][EB];
    int i = blockIdx.x * EB + threadIdx.x;
    double v = i < n ? q[i] * sc[i] : 0.0;
    double4 p = i < n ? pos[i] : make_double4(0, 0, 0, 0);
    r[0][threadIdx.x] = v * p.x;
    r[1][threadIdx.x] = v * p.y;
    r[2][threadIdx.x] = v * p.z;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 3; ++c) r[c][threadIdx.x] += r[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[(size_t)blockIdx.x * 8 + threadIdx.x] = r[threadIdx.x][0];
}

// background electronegativity from the reduced charge sum (same operations
// as the host form - inv_eps_2v * (-sigma^2 Q), Potentials.f90:2166-2167)
__global__ void k_background(const double *__restrict__ qsum, double *__restrict__ out,
                             double inv_eps_2v, double sigma)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = -inv_eps_2v * (-sigma * sigma * qsum[0]);
}

// Deterministic sum of `count` strided partials (component c of stride 8).
__global__ void k_sum_partials(int count, int stride, int ncomp, const double *__restrict__ part,
                               double *__restrict__ out)
{
    __shared__ double lds[EB];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < count; b += EB) s += part[(size_t)b * stride + c];
        lds[threadIdx.x] = s;
        __syncthreads();
        for (int w = EB / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = lds[0];
        __syncthreads();
    }
}

__global__ void k_gather_scaler(int n, const int *__restrict__ id, const double *__restrict__ src,
                                double *__restrict__ dst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[s] = src[id[s]];
}

// ---- damped charge dynamics (relaxation.py:426-448) -------------------------
__global__ void k_relax_update(int n, double *__restrict__ q, double *__restrict__ rate,
                               const double *__restrict__ dchi, double dt, double dt2,
                               double inv_mq, double friction)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = rate[i], f = dchi[i];
    q[i] += r * dt + (f - friction * r) * inv_mq * dt2;
    rate[i] = (1.0 - 0.5 * friction * dt) * r + f * 0.5 * inv_mq * dt;
}

// dchi = chi - mean; optionally finish the rate update; block max of dchi
__global__ void k_relax_finish(int n, const double *__restrict__ chi,
                               const double *__restrict__ sum, double *__restrict__ dchi,
                               double *__restrict__ rate, int update, double future,
                               double half_dt_inv_mq, double *__restrict__ part)
{
    __shared__ double red[EB];
    int i = blockIdx.x * EB + threadIdx.x;
    double mean = sum[0] / n;
    double v = -INFINITY;
    if (i < n) {
        double f = chi[i] - mean;
        dchi[i] = f;
        if (update) rate[i] = future * (rate[i] + f * half_dt_inv_mq);
        v = f;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_max_partials(int count, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double lds[EB];
    double m = -INFINITY;
    for (int b = threadIdx.x; b < count; b += EB) m = fmax(m, part[b]);
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = lds[0];
}

__global__ void k_sum_chi(int n, const double *__restrict__ chi, double *__restrict__ part)
{
    __shared__ double red[EB];
    int i = blockIdx.x * EB + threadIdx.x;
    red[threadIdx.x] = i < n ? chi[i] : 0.0;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8] = red[0];
}

// ---- host side ----------------------------------------------------------------

// Work plan of k_structure.  Chains of up to SK_MT consecutive segments of
// one (k1, k2) row (the segments of a row are consecutive in the natural
// order, all but the last 16 long), in groups of 16 chains of one length,
// placed first-fit (longest first) into waves of SK_MT tiles.
// plan[w][t][i] is the segment in row i of tile t of wave w (-1: none);
// flags[w] bit t says tile t continues tile t - 1.
void build_plan(const std::vector<int4> &segs, std::vector<int> &plan, std::vector<int> &flags)
{
    std::vector<int> first[SK_MT + 1];   // first segment of each chain, by length
    for (size_t s = 0; s < segs.size();) {
        size_t e = s + 1;
        while (e < segs.size() && segs[e].x == segs[s].x && segs[e].y == segs[s].y) ++e;
        for (size_t o = s; o < e; o += SK_MT)
            first[std::min<size_t>(SK_MT, e - o)].push_back((int)o);
        s = e;
    }
    plan.clear();
    flags.clear();
    std::vector<int> used;   // tiles taken in each wave
    size_t open = 0;         // waves before this one are full
    for (int len = SK_MT; len >= 1; --len) {
        const std::vector<int> &c = first[len];
        for (size_t g = 0; g < c.size(); g += 16) {
            while (open < used.size() && used[open] == SK_MT) ++open;
            size_t w = open;
            while (w < used.size() && used[w] + len > SK_MT) ++w;
            if (w == used.size()) {
                used.push_back(0);
                flags.push_back(0);
                plan.resize(plan.size() + 16 * SK_MT, -1);
            }
            for (int j = 0; j < len; ++j) {
                const int t = used[w] + j;
                for (size_t i = 0; i < 16 && g + i < c.size(); ++i)
                    plan[w * 16 * SK_MT + t * 16 + i] = c[g + i] + j;
                if (j > 0) flags[w] |= 1 << t;
            }
            used[w] += len;
        }
    }
}

struct Consts {
    double inv_eps_4pi, inv_eps_2v, inv_sig_sqrt2, inv_sig_sqrt2pi, inv_sig_sq2;
};

Consts consts_of(const EwaldParams &P, double volume)
{
    const double pi = 4.0 * std::atan(1.0);
    Consts c{};
    if (!P.on) return c;
    c.inv_eps_4pi = 1.0 / (4.0 * pi * P.eps);
    c.inv_eps_2v = 1.0 / (2.0 * volume * P.eps);
    c.inv_sig_sqrt2 = 1.0 / (std::sqrt(2.0) * P.sigma);
    c.inv_sig_sqrt2pi = 1.0 / (std::sqrt(2.0 * pi) * P.sigma);
    c.inv_sig_sq2 = c.inv_sig_sqrt2 * c.inv_sig_sqrt2;
    return c;
}

void prepare(System &S, const EwaldParams &P, EwaldState &E)
{
    const int n = S.n;
    const double *v = S.geom.v;
    double cr[3] = {v[1] * v[5] - v[2] * v[4], v[2] * v[3] - v[0] * v[5], v[0] * v[4] - v[1] * v[3]};
    E.volume = std::fabs(cr[0] * v[6] + cr[1] * v[7] + cr[2] * v[8]);
    E.scaler.ensure(n);
    E.chi.ensure(n);
    E.red.ensure(16);
    if (!P.on) return;
    if (E.dirty) {
        E.sk_valid = false;
        // k-vectors in the reference's loop order (Potentials.f90:1526-1548)
        const double pi = 4.0 * std::atan(1.0);
        double rcl[9];
        for (int q = 0; q < 9; ++q) rcl[q] = 2 * pi * S.geom.inv[q];
        Consts c = consts_of(P, E.volume);
        double kc2 = P.kc * P.kc;
        std::vector<double4> kv;
        std::vector<int4> segs;   // {k1, k2, k3 start | len << 16, first k index}
        for (int k1 = -P.K[0]; k1 <= P.K[0]; ++k1)
            for (int k2 = -P.K[1]; k2 <= P.K[1]; ++k2)
                for (int k3 = 0; k3 <= P.K[2]; ++k3) {
                    if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                    double x = k1 * rcl[0] + k2 * rcl[3] + k3 * rcl[6];
                    double y = k1 * rcl[1] + k2 * rcl[4] + k3 * rcl[7];
                    double z = k1 * rcl[2] + k2 * rcl[5] + k3 * rcl[8];
                    double k2n = x * x + y * y + z * z;
                    if (!(k2n < kc2)) continue;
                    double w = (k3 == 0) ? 1.0 : 2.0;
                    double coef = c.inv_eps_2v * w * std::exp(-P.sigma * P.sigma * k2n * 0.5) / k2n;
                    int idx = (int)kv.size();
                    kv.push_back(make_double4(x, y, z, coef));
                    bool extend = false;
                    if (!segs.empty()) {
                        int4 &g = segs.back();
                        int gl = g.z >> 16, g0 = g.z & 0xffff;
                        extend = g.x == k1 && g.y == k2 && g0 + gl == k3 && gl < LSEG;
                        if (extend) g.z = g0 | ((gl + 1) << 16);
                    }
                    if (!extend) segs.push_back(make_int4(k1, k2, k3 | (1 << 16), idx));
                }
        PYSIC_REQUIRE(P.K[0] < 1024 && P.K[1] < 1024 && P.K[2] < 1024, PYSIC_EINVAL,
                      "at most 1023 reciprocal cells along each axis");
        for (int q = 0; q < 9; ++q) {
            E.b[q] = rcl[q];
            E.f[q] = S.geom.inv[q];
        }
        E.nseg = (int)segs.size();
        std::vector<int> plan, flags;
        build_plan(segs, plan, flags);
        E.nwaves = (int)flags.size();
        E.seg.ensure(segs.size());
        E.plan.ensure(plan.size());
        E.pflags.ensure(flags.size());
        if (!segs.empty()) {
            HIP_TRY(hipMemcpyAsync(E.seg.p, segs.data(), sizeof(int4) * segs.size(),
                                   hipMemcpyHostToDevice, S.stream));
            HIP_TRY(hipMemcpyAsync(E.plan.p, plan.data(), sizeof(int) * plan.size(),
                                   hipMemcpyHostToDevice, S.stream));
            HIP_TRY(hipMemcpyAsync(E.pflags.p, flags.data(), sizeof(int) * flags.size(),
                                   hipMemcpyHostToDevice, S.stream));
        }
        E.nk = (int)kv.size();
        E.kvec.ensure(kv.size());
        E.sk.ensure(kv.size());
        if (!kv.empty())
            HIP_TRY(hipMemcpyAsync(E.kvec.p, kv.data(), sizeof(double4) * kv.size(),
                                   hipMemcpyHostToDevice, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));   // before the host vectors go
        E.dirty = false;
    }
    // scaler in the current atom order
    S.host_stage.ensure(n);
    std::vector<double> sc = P.scaler;
    sc.resize(n, 1.0);
    HIP_TRY(hipMemcpyAsync(S.host_stage.p, sc.data(), sizeof(double) * n, hipMemcpyHostToDevice, S.stream));
    k_gather_scaler<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.host_stage.p, E.scaler.p);
    HIP_TRY(hipStreamSynchronize(S.stream));
}

void ensure_scaler_identity(System &S, EwaldState &E)
{
    // used when Ewald is off: chi needs a scaler array anyway
    std::vector<double> one(S.n, 1.0);
    E.scaler.ensure(S.n);
    HIP_TRY(hipMemcpyAsync(E.scaler.p, one.data(), sizeof(double) * S.n, hipMemcpyHostToDevice, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
}

Recip recip_of(const EwaldState &E)
{
    Recip R;
    for (int q = 0; q < 9; ++q) {
        R.b[q] = E.b[q];
        R.f[q] = E.f[q];
    }
    return R;
}

// S(k) into E.sk; also leaves the per-atom phase table E.th that
// recip_atoms reads.
void structure_factors(System &S, EwaldState &E)
{
    const int n = S.n, nk = E.nk;
    if (nk == 0) return;
    int stiles = div_up(E.nwaves, EB / 64);
    int nchunk = std::max(1, std::min(div_up(n, EB), div_up(2048, stiles)));
    int chunk = div_up(n, nchunk);
    nchunk = div_up(n, chunk);
    E.skpart.ensure((size_t)nchunk * nk);
    E.th.ensure(n);
    E.pw.ensure((size_t)n * LSEG);
    E.wl.ensure(n);
    hipEvent_t ev;
    S.timer.begin("ewald_sk", S.stream, &ev);
    k_phase_table<<<div_up(n, EB), EB, 0, S.stream>>>(n, recip_of(E), S.pos.p, S.q.p, E.scaler.p,
                                                      E.th.p, E.pw.p, E.wl.p);
    k_structure<<<dim3(stiles, nchunk), EB, 0, S.stream>>>(E.nwaves, nk, n, chunk, E.plan.p,
                                                          E.pflags.p, E.seg.p, E.th.p, E.wl.p,
                                                          E.pw.p, E.skpart.p);
    k_structure_sum<<<div_up(nk, 256), 256, 0, S.stream>>>(nk, nchunk, E.skpart.p, E.sk.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("ewald_sk", S.stream, ev);
}

// mode 2: add forces + return stress; mode 3: add into chi.  Needs the
// structure factors and phase table of the current positions.
void recip_atoms(System &S, EwaldState &E, int mode, double *stress6)
{
    const int n = S.n, nk = E.nk;
    const int atiles = div_up(n, 16 * (EB / 64));
    int nchunk = std::max(1, std::min(E.nwaves, div_up(4096, atiles)));
    const int wchunk = div_up(E.nwaves, nchunk);
    nchunk = nk > 0 ? div_up(E.nwaves, wchunk) : 0;
    if (nchunk == 0) {
        if (stress6) for (int c = 0; c < 6; ++c) stress6[c] = 0.0;
        return;
    }
    const int ftiles = div_up(n, EB);   // k_recip_finish: one atom per thread
    DBuf<double4> &part = E.rpart;
    part.ensure((size_t)nchunk * n);
    E.part.ensure((size_t)ftiles * 8);
    hipEvent_t ev;
    S.timer.begin(mode == 3 ? "ewald_chi" : "ewald_force", S.stream, &ev);
    const int ntiles = E.nwaves * 4;
    E.zw.ensure((size_t)ntiles * 256);
    E.kpk.ensure((size_t)ntiles * 16);
    k_recip_coef_plan<<<div_up(ntiles * 256, EB), EB, 0, S.stream>>>(
        ntiles, mode, E.plan.p, E.seg.p, E.kvec.p, E.sk.p, E.zw.p, E.kpk.p);
    if (mode == 3)
        k_recip_mfma<3><<<dim3(atiles, nchunk), EB, 0, S.stream>>>(
            n, E.nwaves, wchunk, recip_of(E), E.pflags.p, E.zw.p, E.kpk.p, E.th.p, E.wl.p, part.p);
    else
        k_recip_mfma<2><<<dim3(atiles, nchunk), EB, 0, S.stream>>>(
            n, E.nwaves, wchunk, recip_of(E), E.pflags.p, E.zw.p, E.kpk.p, E.th.p, E.wl.p, part.p);
    k_recip_finish<<<ftiles, EB, 0, S.stream>>>(n, nchunk, mode, part.p, S.pos.p, S.frc.p, E.chi.p,
                                                E.part.p);
    HIP_TRY(hipGetLastError());
    S.timer.end(mode == 3 ? "ewald_chi" : "ewald_force", S.stream, ev);
    if (mode == 2) {
        k_sum_partials<<<1, EB, 0, S.stream>>>(ftiles, 8, 6, E.part.p, E.red.p);
        double h[6];
        HIP_TRY(hipMemcpyAsync(h, E.red.p, sizeof(double) * 6, hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
        for (int c = 0; c < 6; ++c) stress6[c] = h[c];
    }
}

unsigned long long state_hash(System &S, EwaldState &E)
{
    const int nb = div_up(S.n, EB);
    E.hpart.ensure(nb + 1);
    k_state_hash<<<nb, EB, 0, S.stream>>>(S.n, S.pos.p, S.q.p, E.scaler.p, E.hpart.p);
    k_hash_sum<<<1, EB, 0, S.stream>>>(nb, E.hpart.p, E.hpart.p + nb);
    HIP_TRY(hipGetLastError());
    unsigned long long h;
    HIP_TRY(hipMemcpyAsync(&h, E.hpart.p + nb, sizeof(h), hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    return h;
}

// S(k) and the phase table of the current state.  An energy call followed by
// a forces call on the same configuration (the usual calculator sequence)
// computes them once: they are kept while the fingerprint of positions,
// charges and scalers matches and neither the cell nor the k-vectors changed.
void current_structure_factors(System &S, EwaldState &E)
{
    const unsigned long long h = state_hash(S, E);
    if (E.sk_valid && E.sk_hash == h) return;
    structure_factors(S, E);
    E.sk_hash = h;
    E.sk_valid = true;
}

// q sums: out[0] = sum q s, out[1] = sum (q s)^2
void charge_sums(System &S, EwaldState &E, double *out)
{
    int nb = div_up(S.n, EB);
    E.part.ensure((size_t)nb * 8);
    k_charge_sums<<<nb, EB, 0, S.stream>>>(S.n, S.q.p, E.scaler.p, E.part.p);
    k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 8, 2, E.part.p, E.red.p);
    HIP_TRY(hipMemcpyAsync(out, E.red.p, sizeof(double) * 2, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
}

// real-space + singlet pass; returns the energy (mode 1) or stress (mode 2)
void real_pass(System &S, const EwaldParams &P, EwaldState &E, int mode, double bg_chi,
               double *out7, const double *bg_dev = nullptr, bool with_self = true)
{
    const int n = S.n;
    Consts c = consts_of(P, E.volume);
    // lanes per atom: enough atoms x lanes for about 2 waves per SIMD
    const int lanes = n >= 131072 ? 1 : n >= 65536 ? 2 : n >= 32768 ? 4 : n >= 16384 ? 8 : 16;
    int nb = div_up(n * lanes, EB);
    E.part.ensure((size_t)nb * 8);
    hipEvent_t ev;
    S.timer.begin("ewald_real", S.stream, &ev);
    auto launch = [&](auto kern) {
        kern<<<nb, EB, 0, S.stream>>>(n, S.geom, S.pos.p, S.q.p, E.scaler.p, S.nbr.p, S.nimg.p,
                                      S.ncnt.p, S.dmodel.p, P.on, P.rc, c.inv_eps_4pi,
                                      c.inv_sig_sqrt2, c.inv_sig_sqrt2pi, c.inv_sig_sq2, bg_chi,
                                      bg_dev, mode, with_self ? 1 : 0, E.chi.p, S.frc.p,
                                      E.part.p);
    };
    switch (lanes) {
    case 1: launch(k_real<1>); break;
    case 2: launch(k_real<2>); break;
    case 4: launch(k_real<4>); break;
    case 8: launch(k_real<8>); break;
    default: launch(k_real<16>); break;
    }
    HIP_TRY(hipGetLastError());
    S.timer.end("ewald_real", S.stream, ev);
    if (out7) {
        k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 8, 7, E.part.p, E.red.p);
        HIP_TRY(hipMemcpyAsync(out7, E.red.p, sizeof(double) * 7, hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
    }
}

// Electronegativities into E.chi without a host round trip: the charge sum
// and the background term stay on the device.
void chi_eval(System &S, const EwaldParams &P, EwaldState &E)
{
    const double *bg_dev = nullptr;
    if (P.on) {
        int nb = div_up(S.n, EB);
        E.part.ensure((size_t)nb * 8);
        k_charge_sums<<<nb, EB, 0, S.stream>>>(S.n, S.q.p, E.scaler.p, E.part.p);
        k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 8, 2, E.part.p, E.red.p);
        Consts c = consts_of(P, E.volume);
        // - s inv_eps_2v (-sigma^2 Q)  (Potentials.f90:2166-2167)
        k_background<<<1, 64, 0, S.stream>>>(E.red.p, E.red.p + 12, c.inv_eps_2v, P.sigma);
        HIP_TRY(hipGetLastError());
        bg_dev = E.red.p + 12;
    }
    real_pass(S, P, E, 3, 0.0, nullptr, bg_dev);
    if (P.on && E.nk > 0) {
        // the relaxation changes the charges every step: no fingerprint
        structure_factors(S, E);
        E.sk_valid = false;
        recip_atoms(S, E, 3, nullptr);
    }
    pr_chi(S, E.chi.p);
}

}  // namespace

void ewald_release(EwaldState &E)
{
    E.kvec.release();
    E.sk.release();
    E.scaler.release();
    E.chi.release();
    E.part.release();
    E.red.release();
    E.skpart.release();
    E.rpart.release();
    E.seg.release();
    E.plan.release();
    E.pflags.release();
    E.wl.release();
    E.th.release();
    E.pw.release();
    E.zw.release();
    E.kpk.release();
    E.hpart.release();
    E.sk_valid = false;
}

namespace {
// The energy terms: [charge_self singlets +] Ewald real, self, background and
// reciprocal parts [+ the dipole term].
double energy_terms(System &S, const EwaldParams &P, EwaldState &E, bool with_self,
                    bool with_dipole)
{
    prepare(S, P, E);
    if (!P.on) ensure_scaler_identity(S, E);
    double r7[7];
    real_pass(S, P, E, 1, 0.0, r7, nullptr, with_self);
    double e = r7[0];
    if (P.on && with_dipole) {
        // |sum q s r|^2 inv_eps_2v / 3 (Potentials.f90:1563-1565, 1647)
        Consts c = consts_of(P, E.volume);
        int nb = div_up(S.n, EB);
        E.part.ensure((size_t)nb * 8);
        k_dipole_sums<<<nb, EB, 0, S.stream>>>(S.n, S.pos.p, S.q.p, E.scaler.p, E.part.p);
        k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 8, 3, E.part.p, E.red.p);
        double d[3];
        HIP_TRY(hipMemcpyAsync(d, E.red.p, sizeof(double) * 3, hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
        e += (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * c.inv_eps_2v / 3.0;
    }
    if (P.on) {
        Consts c = consts_of(P, E.volume);
        double qs[2];
        charge_sums(S, E, qs);
        e += -qs[1] * c.inv_eps_4pi * c.inv_sig_sqrt2pi;                         // self
        e += -qs[0] * qs[0] * 0.5 * c.inv_eps_2v * P.sigma * P.sigma;            // background
        if (E.nk > 0) {
            current_structure_factors(S, E);
            int nb = div_up(E.nk, EB);
            E.part.ensure((size_t)nb);
            k_recip_energy<<<nb, EB, 0, S.stream>>>(E.nk, E.kvec.p, E.sk.p, E.part.p);
            k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 1, 1, E.part.p, E.red.p);
            double er;
            HIP_TRY(hipMemcpyAsync(&er, E.red.p, sizeof(double), hipMemcpyDeviceToHost, S.stream));
            HIP_TRY(hipStreamSynchronize(S.stream));
            e += er;
        }
    }
    return e;
}
}  // namespace

double ewald_eem_energy(System &S, const EwaldParams &P, EwaldState &E)
{
    return energy_terms(S, P, E, true, false);
}

double ewald_standalone_energy(System &S, const EwaldParams &P, EwaldState &E)
{
    PYSIC_REQUIRE(P.on, PYSIC_EINVAL, "get_ewald_energy: no Ewald parameters");
    return energy_terms(S, P, E, false, true);
}

void ewald_add_forces(System &S, const EwaldParams &P, EwaldState &E, double *stress)
{
    prepare(S, P, E);
    double r7[7];
    real_pass(S, P, E, 2, 0.0, r7);
    for (int c = 0; c < 6; ++c) stress[c] = r7[1 + c];
    if (P.on && E.nk > 0) {
        current_structure_factors(S, E);
        double st[6];
        recip_atoms(S, E, 2, st);
        for (int c = 0; c < 6; ++c) stress[c] += st[c];
    }
}

DBuf<double> &ewald_eem_chi(System &S, const EwaldParams &P, EwaldState &E)
{
    prepare(S, P, E);
    if (!P.on) ensure_scaler_identity(S, E);
    chi_eval(S, P, E);
    HIP_TRY(hipStreamSynchronize(S.stream));
    return E.chi;
}

void ewald_relax(System &S, const EwaldParams &P, EwaldState &E, double *rates, int n_steps,
                 double dt, double inertia, double friction, double tolerance, int *steps,
                 double *error)
{
    const int n = S.n;
    prepare(S, P, E);
    if (!P.on) ensure_scaler_identity(S, E);
    DBuf<double> dchi, mx;
    dchi.ensure(n);
    int nb = div_up(n, EB);
    mx.ensure(nb);
    const double dt2 = dt * dt, inv_mq = 1.0 / inertia;
    const double future = 1.0 / (1.0 + 0.5 * friction * dt);
    auto dchi_of_q = [&](int update) {
        chi_eval(S, P, E);
        E.part.ensure((size_t)nb * 8);
        k_sum_chi<<<nb, EB, 0, S.stream>>>(n, E.chi.p, E.part.p);
        k_sum_partials<<<1, EB, 0, S.stream>>>(nb, 8, 1, E.part.p, E.red.p);
        k_relax_finish<<<nb, EB, 0, S.stream>>>(n, E.chi.p, E.red.p, dchi.p, rates, update, future,
                                                0.5 * inv_mq * dt, mx.p);
        k_max_partials<<<1, EB, 0, S.stream>>>(nb, mx.p, E.red.p + 8);
        double err;
        HIP_TRY(hipMemcpyAsync(&err, E.red.p + 8, sizeof(double), hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
        return err;
    };
    double err = dchi_of_q(0);
    int step = 0;
    // the reference tests the signed max(dchi) (relaxation.py:449); a
    // non-finite value means the dynamics ran away (time step too long for
    // the stiffest charge mode, or an indefinite charge Hessian)
    auto check = [&]() {
        if (!std::isfinite(err)) {
            dchi.release();
            mx.release();
            throw PysicError{PYSIC_ERANGE, "charge relaxation diverged at step " +
                                               std::to_string(step) +
                                               " (reduce the time step or raise the inertia)"};
        }
    };
    check();
    err = 2.0 * tolerance;
    while (err > tolerance && step < n_steps) {
        k_relax_update<<<nb, EB, 0, S.stream>>>(n, S.q.p, rates, dchi.p, dt, dt2, inv_mq, friction);
        err = dchi_of_q(1);
        ++step;
        check();
    }
    *steps = step;
    *error = err;
    dchi.release();
    mx.release();
}
