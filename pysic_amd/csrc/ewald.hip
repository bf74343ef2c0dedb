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
//    up to 16 consecutive k3 at fixed (k1, k2).  With theta_a = b_a . r
//    (b_a the reciprocal vectors), exp(i k.r) = exp(i (k1 th1 + k2 th2 +
//    k3 th3)): one sincos per (segment, atom) gives the first phase, and
//    the rest of the segment follows by the recurrence v <- v exp(i th3)
//    (a complex multiply instead of a sincos per k and atom).  Structure
//    factors S(k) = sum_i q_i s_i exp(i k.r_i): thread per segment over atom
//    chunks staged in LDS (2-D grid, deterministic partial sums); per-atom
//    consumers (forces, electronegativities): thread per atom over segment
//    tiles whose k-vectors and S(k) are staged in LDS;
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
                                             int mode, double *__restrict__ out1,
                                             double4 *__restrict__ frc, double *__restrict__ part)
{
    __shared__ double red[EB][7];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, chi = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 pi = pos[i];
        int si = (int)pi.w;
        double qi = q[i];
        // charge_self singlets: E = eps q^n, chi = -n eps q^(n-1)
        if (M->has_self) {
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
            double qs = qi * sc[i];
            // self (Potentials.f90:2126) and background (2166) electronegativity
            const double bg = bg_dev ? bg_dev[0] : bg_chi;
            chi += 2.0 * qs * inv_eps_4pi * inv_sig_sqrt2pi + sc[i] * bg;
            const double inv2perpi = 2.0 * inv_sig_sqrt2pi;
            int nf = ncnt[i];
            for (int t = 0; t < nf; ++t) {
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
        if (mode == 2) {
            double4 f = frc[i];
            frc[i] = make_double4(f.x + fx, f.y + fy, f.z + fz, 0.0);
        }
        if (mode == 3) out1[i] = chi;
    }
    // block partials: energy (mode 1) or the 6 virial components (mode 2)
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
constexpr int STILE = 32;  // segments per LDS tile of the per-atom kernel

struct Recip {
    double b[9];   // reciprocal vectors in rows: k = k1 b1 + k2 b2 + k3 b3
};

__device__ __forceinline__ void phases(const Recip &R, const double4 &p, double &t1, double &t2,
                                       double &t3)
{
    t1 = R.b[0] * p.x + R.b[1] * p.y + R.b[2] * p.z;
    t2 = R.b[3] * p.x + R.b[4] * p.y + R.b[5] * p.z;
    t3 = R.b[6] * p.x + R.b[7] * p.y + R.b[8] * p.z;
}

// S(k) partial sums: block (segment tile, atom chunk); thread = segment.
// Atoms are staged SK_TA at a time with the table q s exp(i l theta3),
// l < LSEG, so each k of a segment costs one complex multiply-add with a
// wave-wide (broadcast) LDS read; one sincos per (segment, atom) gives the
// segment's first phase.
constexpr int SK_TA = 64;

__global__ __launch_bounds__(EB) void k_structure(int nseg, int nk, int n, int chunk, Recip R,
                                                  const int4 *__restrict__ segs,
                                                  const double4 *__restrict__ pos,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ sc,
                                                  double2 *__restrict__ skpart)
{
    __shared__ double4 at[SK_TA];          // theta1, theta2, theta3, q s
    __shared__ double2 pw[SK_TA][LSEG];    // q s exp(i l theta3)
    const int sidx = blockIdx.x * EB + threadIdx.x;
    const int a0 = blockIdx.y * chunk, a1 = min(n, a0 + chunk);
    int4 sg = sidx < nseg ? segs[sidx] : make_int4(0, 0, 0, 0);
    const double k1 = sg.x, k2 = sg.y, l0 = sg.z & 0xffff;
    const int len = sg.z >> 16;  // 0 for lanes past the last segment
    double cr[LSEG], ci[LSEG];
#pragma unroll
    for (int l = 0; l < LSEG; ++l) cr[l] = ci[l] = 0.0;
    for (int base = a0; base < a1; base += SK_TA) {
        if (threadIdx.x < SK_TA) {
            const int a = base + threadIdx.x;
            double t1 = 0.0, t2 = 0.0, t3 = 0.0, qs = 0.0;
            if (a < a1) {
                phases(R, pos[a], t1, t2, t3);
                qs = q[a] * sc[a];
            }
            at[threadIdx.x] = make_double4(t1, t2, t3, qs);
            double s3, c3;
            sincos(t3, &s3, &c3);
            double vr = qs, vi = 0.0;
#pragma unroll
            for (int l = 0; l < LSEG; ++l) {
                pw[threadIdx.x][l] = make_double2(vr, vi);
                const double nr = vr * c3 - vi * s3;
                vi = vr * s3 + vi * c3;
                vr = nr;
            }
        }
        __syncthreads();
        const int m = min(SK_TA, a1 - base);
        for (int t = 0; t < m; ++t) {
            const double4 th = at[t];
            if (th.w == 0.0) continue;
            double vs, vc;
            sincos(k1 * th.x + k2 * th.y + l0 * th.z, &vs, &vc);
#pragma unroll
            for (int l = 0; l < LSEG; ++l) {
                if (l < len) {
                    const double2 p = pw[t][l];
                    cr[l] += vc * p.x - vs * p.y;
                    ci[l] += vc * p.y + vs * p.x;
                }
            }
        }
        __syncthreads();
    }
    double2 *out = skpart + (size_t)blockIdx.y * nk + sg.w;
#pragma unroll
    for (int l = 0; l < LSEG; ++l)
        if (l < len) out[l] = make_double2(cr[l], ci[l]);
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

// Per-atom reciprocal forces (mode 2) or electronegativities (mode 3) over
// one chunk of segments (natural order, so each tile of STILE segments covers
// one contiguous k range); block (atom tile, segment chunk).  Each thread
// carries RAT atoms (i, i + EB, ...) so that every k-vector and S(k) read
// from LDS serves RAT atoms.  Measured on the 24,576-atom SiO2 cell: RAT = 1
// 1.20 ms, 2: 0.89 ms, 4: 1.32 ms (occupancy 3 waves/SIMD); scalar-cache
// reads instead of LDS: 0.96 ms.
constexpr int RAT = 2;

__global__ __launch_bounds__(EB) void k_recip_atom(int n, int nseg, int schunk, int mode, Recip R,
                                                   const int4 *__restrict__ segs,
                                                   const double4 *__restrict__ kvec,
                                                   const double2 *__restrict__ sk,
                                                   const double4 *__restrict__ pos,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ sc,
                                                   double4 *__restrict__ part)
{
    __shared__ int4 sl[STILE];
    __shared__ double4 kl[STILE * LSEG];
    __shared__ double2 sv[STILE * LSEG];
    const int i0 = blockIdx.x * (EB * RAT) + threadIdx.x;
    const int s0 = blockIdx.y * schunk, s1 = min(nseg, s0 + schunk);
    double t1[RAT], t2[RAT], t3[RAT], wc[RAT], ws[RAT], qs[RAT];
    double ax[RAT], ay[RAT], az[RAT];
    double vs[RAT], vc[RAT];   // the running phase exp(i k.r) of each atom
#pragma unroll
    for (int a = 0; a < RAT; ++a) {
        const int i = i0 + a * EB;
        const double4 p = i < n ? pos[i] : make_double4(0, 0, 0, 0);
        qs[a] = i < n ? q[i] * sc[i] : 0.0;
        phases(R, p, t1[a], t2[a], t3[a]);
        sincos(t3[a], &ws[a], &wc[a]);
        ax[a] = ay[a] = az[a] = 0.0;
        vs[a] = 0.0;
        vc[a] = 1.0;
    }
    // (k1, k2, next k3) that the running phase stands at; a segment that
    // continues the previous one's row starts there without a sincos
    int rk1 = 0x7fffffff, rk2 = 0, rl = -1;
    for (int base = s0; base < s1; base += STILE) {
        const int m = min(STILE, s1 - base);
        const int kb = segs[base].w;
        const int4 last = segs[base + m - 1];
        const int nkt = last.w + (last.z >> 16) - kb;   // <= STILE * LSEG
        if (threadIdx.x < m) sl[threadIdx.x] = segs[base + threadIdx.x];
        for (int k = threadIdx.x; k < nkt; k += EB) {
            kl[k] = kvec[kb + k];
            sv[k] = sk[kb + k];
        }
        __syncthreads();
        for (int t = 0; t < m; ++t) {
            const int4 sg = sl[t];
            const int len = sg.z >> 16;
            const int o = sg.w - kb;
            const int il0 = sg.z & 0xffff;
            if (!(sg.x == rk1 && sg.y == rk2 && il0 == rl)) {
                const double k1 = sg.x, k2 = sg.y, l0 = il0;
#pragma unroll
                for (int a = 0; a < RAT; ++a)
                    sincos(k1 * t1[a] + k2 * t2[a] + l0 * t3[a], &vs[a], &vc[a]);
            }
            rk1 = sg.x;
            rk2 = sg.y;
            rl = il0 + len;
            for (int l = 0; l < len; ++l) {
                const double4 kv = kl[o + l];
                const double2 S = sv[o + l];
                const double c2 = 2.0 * kv.w;
#pragma unroll
                for (int a = 0; a < RAT; ++a) {
                    if (mode == 3) {
                        // chi -= 2 c_k (C cos + S sin)   (Potentials.f90:2196-2229)
                        ax[a] -= c2 * (S.x * vc[a] + S.y * vs[a]);
                    } else {
                        // f -= 2 c_k q s (S cos - C sin) k   (Potentials.f90:1895-1931)
                        const double pre = c2 * qs[a] * (S.y * vc[a] - S.x * vs[a]);
                        ax[a] -= pre * kv.x;
                        ay[a] -= pre * kv.y;
                        az[a] -= pre * kv.z;
                    }
                    const double nc = vc[a] * wc[a] - vs[a] * ws[a];
                    vs[a] = vc[a] * ws[a] + vs[a] * wc[a];
                    vc[a] = nc;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < RAT; ++a) {
        const int i = i0 + a * EB;
        if (i < n) part[(size_t)blockIdx.y * n + i] = make_double4(ax[a], ay[a], az[a], 0.0);
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
        PYSIC_REQUIRE(P.K[2] < 0xffff, PYSIC_EINVAL, "too many reciprocal cells along k3");
        for (int q = 0; q < 9; ++q) E.b[q] = rcl[q];
        E.nseg = (int)segs.size();
        std::vector<int4> sorted = segs;
        std::stable_sort(sorted.begin(), sorted.end(),
                         [](const int4 &a, const int4 &b) { return (a.z >> 16) > (b.z >> 16); });
        E.seg.ensure(segs.size());
        E.seg_sorted.ensure(segs.size());
        if (!segs.empty()) {
            HIP_TRY(hipMemcpyAsync(E.seg.p, segs.data(), sizeof(int4) * segs.size(),
                                   hipMemcpyHostToDevice, S.stream));
            HIP_TRY(hipMemcpyAsync(E.seg_sorted.p, sorted.data(), sizeof(int4) * sorted.size(),
                                   hipMemcpyHostToDevice, S.stream));
        }
        E.nk = (int)kv.size();
        E.kvec.ensure(kv.size());
        E.sk.ensure(kv.size());
        if (!kv.empty())
            HIP_TRY(hipMemcpyAsync(E.kvec.p, kv.data(), sizeof(double4) * kv.size(),
                                   hipMemcpyHostToDevice, S.stream));
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
    for (int q = 0; q < 9; ++q) R.b[q] = E.b[q];
    return R;
}

void structure_factors(System &S, EwaldState &E)
{
    const int n = S.n, nk = E.nk;
    if (nk == 0) return;
    int stiles = div_up(E.nseg, EB);
    int nchunk = std::max(1, std::min(div_up(n, EB), div_up(2048, stiles)));
    int chunk = div_up(div_up(n, nchunk), EB) * EB;
    nchunk = div_up(n, chunk);
    E.skpart.ensure((size_t)nchunk * nk);
    hipEvent_t ev;
    S.timer.begin("ewald_sk", S.stream, &ev);
    k_structure<<<dim3(stiles, nchunk), EB, 0, S.stream>>>(E.nseg, nk, n, chunk, recip_of(E),
                                                          E.seg_sorted.p, S.pos.p, S.q.p,
                                                          E.scaler.p, E.skpart.p);
    k_structure_sum<<<div_up(nk, 256), 256, 0, S.stream>>>(nk, nchunk, E.skpart.p, E.sk.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("ewald_sk", S.stream, ev);
}

// mode 2: add forces + return stress; mode 3: add into chi
void recip_atoms(System &S, EwaldState &E, int mode, double *stress6)
{
    const int n = S.n, nk = E.nk;
    int atiles = div_up(n, EB * RAT);
    int nchunk = std::max(1, std::min(div_up(E.nseg, STILE), div_up(2048, atiles)));
    int schunk = div_up(div_up(E.nseg, nchunk), STILE) * STILE;
    nchunk = nk > 0 ? div_up(E.nseg, schunk) : 0;
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
    k_recip_atom<<<dim3(atiles, nchunk), EB, 0, S.stream>>>(n, E.nseg, schunk, mode, recip_of(E),
                                                           E.seg.p, E.kvec.p, E.sk.p, S.pos.p,
                                                           S.q.p, E.scaler.p, part.p);
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
               double *out7, const double *bg_dev = nullptr)
{
    const int n = S.n;
    Consts c = consts_of(P, E.volume);
    int nb = div_up(n, EB);
    E.part.ensure((size_t)nb * 8);
    hipEvent_t ev;
    S.timer.begin("ewald_real", S.stream, &ev);
    k_real<<<nb, EB, 0, S.stream>>>(n, S.geom, S.pos.p, S.q.p, E.scaler.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                    S.dmodel.p, P.on, P.rc, c.inv_eps_4pi, c.inv_sig_sqrt2,
                                    c.inv_sig_sqrt2pi, c.inv_sig_sq2, bg_chi, bg_dev, mode, E.chi.p,
                                    S.frc.p,
                                    E.part.p);
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
        structure_factors(S, E);
        recip_atoms(S, E, 3, nullptr);
    }
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
    E.seg_sorted.release();
}

double ewald_eem_energy(System &S, const EwaldParams &P, EwaldState &E)
{
    prepare(S, P, E);
    if (!P.on) ensure_scaler_identity(S, E);
    double r7[7];
    real_pass(S, P, E, 1, 0.0, r7);
    double e = r7[0];
    if (P.on) {
        Consts c = consts_of(P, E.volume);
        double qs[2];
        charge_sums(S, E, qs);
        e += -qs[1] * c.inv_eps_4pi * c.inv_sig_sqrt2pi;                         // self
        e += -qs[0] * qs[0] * 0.5 * c.inv_eps_2v * P.sigma * P.sigma;            // background
        if (E.nk > 0) {
            structure_factors(S, E);
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

void ewald_add_forces(System &S, const EwaldParams &P, EwaldState &E, double *stress)
{
    prepare(S, P, E);
    double r7[7];
    real_pass(S, P, E, 2, 0.0, r7);
    for (int c = 0; c < 6; ++c) stress[c] = r7[1 + c];
    if (P.on && E.nk > 0) {
        structure_factors(S, E);
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
