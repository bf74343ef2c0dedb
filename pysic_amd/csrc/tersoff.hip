// tersoff.hip -- Tersoff bond-order energy, forces and virial on the MI355X.
//
// The reference evaluates, for every pair i<j picked from the neighbor list
// (Core.f90:3337-3953), the pair terms and the bond-order factor
// b_ij = (b~(zeta_ij) + b~(zeta_ji))/2 (Core.f90:4359-4368, 2439-2637), and
// scatters the three-body gradient -1/2 v f_c b~'(zeta) grad(zeta) onto
// i, j and every k (Core.f90:4952-5353).  Here the same energy is written as
// a sum of site energies
//
//   U_i = sum_j 1/2 f_c(r_ij) [ v_R(r_ij) + v_A(r_ij) b~(zeta_ij) ],
//   zeta_ij = sum_{k != j} f_c^b(r_ik) exp((a (r_ij - r_ik))^mu) g(theta_jik)
//
// (identical to the reference's pair sum once the pair entries are symmetric,
// which the permutation expansion of calculator.py:1166-1192 guarantees),
// each depending only on the vectors r_ij = r_j - r_i of atom i's own
// neighbors.  Kernel 1 (site) computes U_i and D_ij = dU_i/dr_ij for every
// neighbor j of i, using only i's neighborhood (staged in LDS).  Kernel 2
// (gather) forms F_i = sum_j (D_ij - D_ji): every atom owns its force, no
// atomics, deterministic summation order.  The virial is
// W = -sum_i sum_j r_ij (x) D_ij, which equals the reference's
// sum of r (x) f over pair and triplet terms.
//
// Precision: separations are formed in FP64 from FP64 positions; zeta, b~,
// b~' and the pair energies are FP64 (the energy contract of 1e-7 eV/atom);
// the three-body gradient arithmetic is FP32; D, forces, energy and virial
// accumulate in FP64.
#include "system.h"

#include <cmath>

namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ float fc32(float r, float hard, float soft)
{
    if (r > hard) return 0.0f;
    if (r < soft) return 1.0f;
    return 0.5f * (1.0f + __cosf(3.14159265358979f * (r - soft) / (hard - soft)));
}
__device__ __forceinline__ float dfc32(float r, float hard, float soft)
{
    if (r > hard || r < soft) return 0.0f;
    float w = hard - soft;
    return -0.5f * 3.14159265358979f / w * __sinf(3.14159265358979f * (r - soft) / w);
}

// One zeta term in FP64 (evaluate_pair_bond_order_factor_tersoff,
// Potentials.f90:5744-5795), with g written without cancellation:
// 1 + c^2/d^2 - c^2/(d^2+t^2) = 1 + c^2 t^2 / (d^2 (d^2 + t^2)).
__device__ __forceinline__ double zeta_term64(const double4 &u, const double4 &v, const BondTerm &b)
{
    double cosine = (u.x * v.x + u.y * v.y + u.z * v.z) / (u.w * v.w);
    double t = b.h - cosine;
    double dd = b.d * b.d, cc = b.c * b.c;
    double gee = 1.0 + cc * t * t / (dd * (dd + t * t));
    double xi = fc64(v.w, b.cut, b.soft);
    if (!b.a_zero) xi *= exp(pow(b.a * (u.w - v.w), b.mu));
    return xi * gee;
}

// b~(zeta) and db~/dzeta (Potentials.f90:5910-5961)
__device__ __forceinline__ void bt64(double zeta, const BondTerm &p, double *bt, double *dbt)
{
    double be = p.beta * zeta;
    double pw = powpos64(be, p.eta);
    double inv_eta = -1.0 / (2.0 * p.eta);
    *bt = powpos64(1.0 + pw, inv_eta);
    if (p.eta == 0.0 || (zeta == 0.0 && p.eta <= 1.0))
        *dbt = 0.0;
    else
        *dbt = inv_eta * powpos64(1.0 + pw, inv_eta - 1.0) * p.eta * powpos64(be, p.eta - 1.0) *
               p.beta;
}

template <int G, int SCAP>
struct SiteShared {
    static constexpr int APB = BLOCK / G;
    // [slot][atom]: the atoms of a wave sit on consecutive LDS banks
    double4 sep[SCAP][APB];   // r_ij (FP64) and |r_ij|
    int jdx[SCAP][APB];
    int img[SCAP][APB];
    int spc[SCAP][APB];
    float P[SCAP][APB];       // 1/2 v_A f_c db~/dzeta for the bond (i,j)
    double dpair[SCAP][APB];  // d/dr of 1/2 f_c (v_R + v_A b~) at fixed b~
    double red[APB][8];
};

template <int G, int SCAP, class Geo>
__device__ __forceinline__ void site_tile(int blk, int n, const Geo &geo,
                                          const double4 *__restrict__ pos,
                                          const int *__restrict__ id,
                                          const int *__restrict__ nbr,
                                          const int *__restrict__ nimg,
                                          const int *__restrict__ ncnt,
                                          const Model *__restrict__ M,
                                          int *__restrict__ sj, int *__restrict__ simg,
                                          int *__restrict__ scnt, int *__restrict__ srow,
                                          int unique, float4 *__restrict__ dU,
                                          double *__restrict__ partial,
                                          double *__restrict__ atom_ev,
                                          int *__restrict__ flags)
{
    using SH = SiteShared<G, SCAP>;
    constexpr int APB = SH::APB;
    __shared__ SH sh;
    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int i = blk * APB + a;
    const bool valid = i < n;
    const double *cv = geo.vec(valid ? i : 0);  // vectors of this atom's cell
    const int wl = threadIdx.x & 63;
    const int gbase = (wl / G) * G;
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    const int si = (int)pi.w;
    const double rshort = M->rshort;
    const double maxbond_i = valid ? M->maxbond[si] : 0.0;

    // ---- 1. compact the full list to the short list (r < rshort) ----------
    int nf = valid ? ncnt[i] : 0;
    int cnt = 0;
    for (int base = 0; base < nf; base += G) {
        int e = base + l;
        bool keep = false;
        double4 d = make_double4(0, 0, 0, 0);
        int j = 0, im = 0, sp = 0;
        if (e < nf) {
            j = nbr[(size_t)e * n + i];
            im = nimg[(size_t)e * n + i];
            double4 pj = pos[j];
            double3 sp3 = separation64(pi, pj, im, cv);
            double r = sqrt(dot64(sp3));
            d = make_double4(sp3.x, sp3.y, sp3.z, r);
            sp = (int)pj.w;
            keep = r < rshort;
        }
        unsigned long long b = __ballot(keep);
        unsigned long long gb = (b >> gbase) & gmask;
        int rank = __popcll(gb & ((1ull << l) - 1ull));
        int slot = cnt + rank;
        if (keep && slot < SCAP) {
            sh.sep[slot][a] = d;
            sh.jdx[slot][a] = j;
            sh.img[slot][a] = im;
            sh.spc[slot][a] = sp;
        }
        cnt += __popcll(gb);
    }
    if (cnt > SCAP && l == 0) atomicMax(flags + 1, cnt);
    const int ns = cnt < SCAP ? cnt : SCAP;
    __syncthreads();

    // ---- 2. zeta, b~, pair terms (FP64) --------------------------------------
    double e_lane = 0.0;
    const int grp = M->group;
    for (int jj = l; jj < ns; jj += G) {
        double4 u = sh.sep[jj][a];
        int sjp = sh.spc[jj][a];
        double zeta = 0.0;
        if (grp >= 0) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                double4 v = sh.sep[k][a];
                if (!(v.w < maxbond_i)) continue;
                const BondTerm &bt = M->bond[si][sjp][sh.spc[k][a]];
                if (bt.valid) zeta += zeta_term64(u, v, bt);
            }
        }
        double bt = 1.0, dbt = 0.0;
        if (grp >= 0) {
            const BondTerm &pp = M->post[si];
            if (pp.valid) bt64(zeta, pp, &bt, &dbt);
            else { bt = zeta; dbt = 1.0; }  // no post-processing: b = raw sum (Core.f90:2800-2805)
        }
        // pair cutoff tests of Core.f90:3462 and 4297: r < max_potential_radius of
        // the picked atom1, and r < the term's own cutoff
        double mp = M->maxpot[si];
        if (M->maxpot[sjp] != mp) {
            int idi = id[i], idj = id[sh.jdx[jj][a]];
            if (!pick_dev(idi, idj, sh.img[jj][a])) mp = M->maxpot[sjp];
        }
        double eh = 0.0, dp = 0.0, P = 0.0;
        if (u.w < mp) {
            int np = M->npair[si][sjp];
            for (int t = 0; t < np; ++t) {
                const PairTerm &pt = M->pair[si][sjp][t];
                if (!(pt.cut > u.w)) continue;
                double v = pt.eps * exp(-pt.zeta * u.w);
                double dv = -pt.zeta * v;
                double f = 1.0, df = 0.0;
                if (pt.smooth) smooth64(u.w, pt.cut, pt.soft, &f, &df);
                if (pt.group >= 0) {
                    eh += 0.5 * v * f * bt;
                    dp += 0.5 * (dv * f + v * df) * bt;
                    P += 0.5 * v * f * dbt;
                } else {
                    eh += 0.5 * v * f;
                    dp += 0.5 * (dv * f + v * df);
                }
            }
        }
        e_lane += eh;
        sh.P[jj][a] = (float)P;
        sh.dpair[jj][a] = dp;
    }
    __syncthreads();

    // ---- 3. D_ij = dU_i/dr_ij (FP32 three-body arithmetic) -------------------
    double w6[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = l; jj < ns; jj += G) {
        double4 ud = sh.sep[jj][a];
        float ux = (float)ud.x, uy = (float)ud.y, uz = (float)ud.z, r1 = (float)ud.w;
        float inv_r1 = 1.0f / r1;
        float Pj = sh.P[jj][a];
        int sjp = sh.spc[jj][a];
        float cu = 0.f;  // three-body coefficient of u
        float dcx = 0.f, dcy = 0.f, dcz = 0.f;  // accumulated coefficient * dcos/du
        if (grp >= 0) {
            const float mb = (float)maxbond_i;
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                double4 vd = sh.sep[k][a];
                float vx = (float)vd.x, vy = (float)vd.y, vz = (float)vd.z, r2 = (float)vd.w;
                float Pk = sh.P[k][a];
                int skp = sh.spc[k][a];
                float inv_r12 = inv_r1 / r2;
                float cs = (ux * vx + uy * vy + uz * vz) * inv_r12;
                float coef_u = 0.f, coef_c = 0.f;
                // term of zeta_ij with k third (derivative w.r.t. r_ij)
                const BondTerm &bjk = M->bond[si][sjp][skp];
                if (Pj != 0.f && bjk.valid && vd.w < maxbond_i) {
                    float cc = (float)(bjk.c * bjk.c), dd = (float)(bjk.d * bjk.d);
                    float t = (float)bjk.h - cs;
                    float den = dd + t * t;
                    float gee = 1.0f + cc * t * t / (dd * den);
                    float gp = -2.0f * cc * t / (den * den);
                    float fk = fc32(r2, (float)bjk.cut, (float)bjk.soft);
                    float ex = 1.f, exd = 0.f;
                    if (!bjk.a_zero) {
                        float av = (float)bjk.a, mu = (float)bjk.mu;
                        float x = av * (r1 - r2);
                        float pm1 = powf(x, mu - 1.0f);
                        ex = __expf(x * pm1);
                        exd = ex * av * mu * pm1;
                    }
                    coef_u += Pj * gee * fk * exd;
                    coef_c += Pj * fk * ex * gp;
                }
                // term of zeta_ik with j third (derivative w.r.t. its 2nd vector)
                const BondTerm &bkj = M->bond[si][skp][sjp];
                if (Pk != 0.f && bkj.valid && r1 < mb) {
                    float cc = (float)(bkj.c * bkj.c), dd = (float)(bkj.d * bkj.d);
                    float t = (float)bkj.h - cs;
                    float den = dd + t * t;
                    float gee = 1.0f + cc * t * t / (dd * den);
                    float gp = -2.0f * cc * t / (den * den);
                    float fj = fc32(r1, (float)bkj.cut, (float)bkj.soft);
                    float dfj = dfc32(r1, (float)bkj.cut, (float)bkj.soft);
                    float ex = 1.f, exd = 0.f;
                    if (!bkj.a_zero) {
                        float av = (float)bkj.a, mu = (float)bkj.mu;
                        float x = av * (r2 - r1);
                        float pm1 = powf(x, mu - 1.0f);
                        ex = __expf(x * pm1);
                        exd = ex * av * mu * pm1;
                    }
                    coef_u += Pk * gee * (dfj * ex - fj * exd);
                    coef_c += Pk * fj * ex * gp;
                }
                // dcos/dr_ij = v/(r1 r2) - cos u / r1^2
                float cu2 = cs * inv_r1 * inv_r1;
                dcx += coef_c * (vx * inv_r12 - cu2 * ux);
                dcy += coef_c * (vy * inv_r12 - cu2 * uy);
                dcz += coef_c * (vz * inv_r12 - cu2 * uz);
                cu += coef_u * inv_r1;
            }
        }
        // pair part in FP64, three-body part from FP32
        double cp = sh.dpair[jj][a] / ud.w;
        double Dx = cp * ud.x + (double)(cu * ux + dcx);
        double Dy = cp * ud.y + (double)(cu * uy + dcy);
        double Dz = cp * ud.z + (double)(cu * uz + dcz);
        size_t o = (size_t)jj * n + i;
        dU[o] = make_float4((float)Dx, (float)Dy, (float)Dz, 0.f);
        if (!unique) {
            sj[o] = sh.jdx[jj][a];
            simg[o] = sh.img[jj][a];
        }
        srow[(size_t)i * SCAP + jj] = sh.jdx[jj][a];
        // virial W = -sum r_ij (x) D_ij ; order xx yy zz yz xz xy
        w6[0] -= ud.x * Dx;
        w6[1] -= ud.y * Dy;
        w6[2] -= ud.z * Dz;
        w6[3] -= ud.y * Dz;
        w6[4] -= ud.x * Dz;
        w6[5] -= ud.x * Dy;
    }
    if (valid) {
        if (l == 0) scnt[i] = ns;
        for (int t = ns + l; t < SCAP; t += G) srow[(size_t)i * SCAP + t] = -1;
    }

    // ---- 4. deterministic reduction: group lanes, then atoms of the block ----
    double vals[7] = {e_lane, w6[0], w6[1], w6[2], w6[3], w6[4], w6[5]};
    for (int q = 0; q < 7; ++q) {
        double v = vals[q];
        for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
        vals[q] = v;
    }
    if (l == 0)
        for (int q = 0; q < 7; ++q) sh.red[a][q] = valid ? vals[q] : 0.0;
    if (atom_ev && valid && l == 0)
        for (int q = 0; q < 7; ++q) atom_ev[(size_t)q * n + i] = vals[q];
    __syncthreads();
    if (threadIdx.x < 7) {
        double s = 0.0;
        for (int t = 0; t < APB; ++t) s += sh.red[t][threadIdx.x];
        partial[(size_t)blk * 8 + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Single-species fast path.  Every parameter is uniform (scalar registers)
// and each neighbor's cutoff functions and pair terms are evaluated once by
// the lane that loads it, not once per triplet.  Two kernels split the work
// the way callers need it:
//   k_site1e  energy only: zeta, b~ and the pair energies in FP64 (the
//             1e-7 eV/atom contract); no gradients;
//   k_site1f  forces + virial: zeta, b~, b~' and the three-body gradient in
//             FP32 (the force path's precision form), pair terms, D and the
//             virial in FP64.  A velocity-Verlet step needs exactly this, as
//             calculate_forces returns forces and stress (PyInterface.f90:668).
// ---------------------------------------------------------------------------
//
// k_site1e with SHORT reads the short list a k_site1f pass on the same
// positions left (scnt, srow and the images in simg) instead of compacting
// the full list again: the same slots in the same order, so the same sums.
template <int G, int SCAP, class Geo, bool SHORT>
__device__ __forceinline__ void site1e_tile(int blk, int n, const Geo &geo,
                                            const double4 *__restrict__ pos,
                                            const int *__restrict__ nbr,
                                            const int *__restrict__ nimg,
                                            const int *__restrict__ ncnt,
                                            const int *__restrict__ scnt,
                                            const int *__restrict__ srow,
                                            const int *__restrict__ simg,
                                            const Model *__restrict__ M,
                                            double *__restrict__ partial,
                                            double *__restrict__ atom_ev,
                                            int *__restrict__ flags)
{
    constexpr int APB = BLOCK / G;
    constexpr int PER = (SCAP + G - 1) / G;
    // [slot][atom]: the atoms of a wave sit on consecutive LDS banks
    __shared__ double4 s_sep[SCAP][APB];      // x, y, z, r
    __shared__ double2 s_fi[SCAP][APB];       // bond f_c(r) (0 at r >= max_bond_radius), 1/r
    __shared__ double s_red[APB];

    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int i = blk * APB + a;
    const bool valid = i < n;
    const double *cv = geo.vec(valid ? i : 0);  // vectors of this atom's cell
    const int gbase = ((threadIdx.x & 63) / G) * G;
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    const BondTerm B = M->bond[0][0][0];
    const bool bond_on = M->group >= 0 && B.valid;
    const double maxbond = M->maxbond[0], maxpot = M->maxpot[0], rshort = M->rshort;
    const int npair = M->npair[0][0];
    const double dd = B.d * B.d, cd2 = B.c * B.c / dd;

    double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    int ns;
    if constexpr (SHORT) {
        ns = valid ? scnt[i] : 0;
        for (int e = l; e < ns; e += G) {
            const int j = srow[(size_t)i * SCAP + e];
            const double3 sp3 = separation64(pi, pos[j], simg[(size_t)e * n + i], cv);
            const double r = sqrt(dot64(sp3));
            const bool inb = bond_on && r < maxbond;
            s_sep[e][a] = make_double4(sp3.x, sp3.y, sp3.z, r);
            s_fi[e][a] = make_double2(inb ? fc64(r, B.cut, B.soft) : 0.0, recip64(r));
        }
    } else {
        int nf = valid ? ncnt[i] : 0;
        int cnt = 0;
        for (int base = 0; base < nf; base += G) {
            int e = base + l;
            bool keep = false;
            double4 d = make_double4(0, 0, 0, 0);
            if (e < nf) {
                int j = nbr[(size_t)e * n + i];
                int im = nimg[(size_t)e * n + i];
                double3 sp3 = separation64(pi, pos[j], im, cv);
                double r = sqrt(dot64(sp3));
                d = make_double4(sp3.x, sp3.y, sp3.z, r);
                keep = r < rshort;
            }
            unsigned long long b = __ballot(keep);
            unsigned long long gb = (b >> gbase) & gmask;
            int slot = cnt + __popcll(gb & ((1ull << l) - 1ull));
            if (keep && slot < SCAP) {
                bool inb = bond_on && d.w < maxbond;
                s_sep[slot][a] = d;
                s_fi[slot][a] = make_double2(inb ? fc64(d.w, B.cut, B.soft) : 0.0, recip64(d.w));
            }
            cnt += __popcll(gb);
        }
        if (cnt > SCAP && l == 0) atomicMax(flags + 1, cnt);
        ns = cnt < SCAP ? cnt : SCAP;
    }
    __syncthreads();

    double e_lane = 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int jj = l + q * G;
        if (jj >= ns) continue;
        double4 u = s_sep[jj][a];
        double ir1 = s_fi[jj][a].y;
        double zeta = 0.0;
        if (bond_on) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                double2 fk = s_fi[k][a];
                if (fk.x == 0.0) continue;
                double4 v = s_sep[k][a];
                double cosine = (u.x * v.x + u.y * v.y + u.z * v.z) * ir1 * fk.y;
                double t = B.h - cosine;
                double t2 = t * t;
                double xi = fk.x;
                if (!B.a_zero) xi *= exp(pow(B.a * (u.w - v.w), B.mu));
                zeta += xi * (1.0 + cd2 * (t2 * recip64(dd + t2)));
            }
        }
        double bt = 1.0;
        if (bond_on) bt = powpos64(1.0 + powpos64(B.beta * zeta, B.eta), -1.0 / (2.0 * B.eta));
        if (u.w < maxpot) {
            for (int t = 0; t < npair; ++t) {
                const PairTerm pt = M->pair[0][0][t];
                if (!(pt.cut > u.w)) continue;
                double v = pt.eps * exp(-pt.zeta * u.w);
                double f = 1.0;
                if (pt.smooth) {
                    // a term with the bond factor's cutoffs (Tersoff's usual
                    // form) takes the staged bond smoothing, the same value
                    if (bond_on && pt.cut == B.cut && pt.soft == B.soft && u.w < maxbond)
                        f = s_fi[jj][a].x;
                    else
                        f = fc64(u.w, pt.cut, pt.soft);
                }
                e_lane += 0.5 * v * f * (pt.group >= 0 ? bt : 1.0);
            }
        }
    }
    for (int off = G / 2; off > 0; off >>= 1) e_lane += __shfl_xor(e_lane, off, G);
    if (l == 0) s_red[a] = valid ? e_lane : 0.0;
    if (atom_ev && valid && l == 0) atom_ev[i] = e_lane;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sacc = 0.0;
        for (int t = 0; t < APB; ++t) sacc += s_red[t];
        partial[(size_t)blk * 8] = sacc;
        for (int c = 1; c < 7; ++c) partial[(size_t)blk * 8 + c] = 0.0;
    }
}

// e^x for the pair exponentials of the MD force pass, x = -zeta r formed in
// FP64: x log2(e) is split in FP64 into an integer and a fraction |f| <= 1/2,
// and only 2^f is taken in FP32 (v_exp_f32), so the result carries FP32
// rounding (~1e-7 relative) without the 3e-7 that rounding r or x to FP32
// first would add.
__device__ __forceinline__ float exp_split(double x)
{
    const double y = x * 1.4426950408889634;
    const double nd = rint(y);
    return ldexpf(__builtin_amdgcn_exp2f((float)(y - nd)), (int)nd);
}

// The force pass of a velocity-Verlet step (forces only, no virial), one
// species.  Same structure as site1f_tile, with the pair arithmetic in FP32
// as the three-body arithmetic already is: separations and distances are
// formed in FP64 (the short-list and cutoff decisions stay exact), then
// everything per bond is FP32 -- the exponentials through exp_split, the
// smoothing functions from precomputed 1/(cut - soft), the reciprocal
// distance from v_rcp_f32 and one Newton step -- and D_ij is accumulated and
// stored in FP32.  The energy pass (site1e_tile) stays FP64.
template <int G, int SCAP, class Geo>
__device__ __forceinline__ void site1m_tile(int blk, int n, const Geo &geo,
                                            const double4 *__restrict__ pos,
                                            const int *__restrict__ nbr,
                                            const int *__restrict__ nimg,
                                            const int *__restrict__ ncnt,
                                            const Model *__restrict__ M,
                                            int *__restrict__ sj, int *__restrict__ simg,
                                            int *__restrict__ scnt, int *__restrict__ srow,
                                            int unique, float4 *__restrict__ dU,
                                            int *__restrict__ flags)
{
    constexpr int APB = BLOCK / G;
    constexpr int PER = (SCAP + G - 1) / G;
    // [slot][atom]: the atoms of a wave sit on consecutive LDS banks
    __shared__ float4 s_u[SCAP][APB];      // x, y, z, r
    __shared__ float4 s_aux[SCAP][APB];    // 1/r, f_c(r), f_c'(r), A (then P)
    __shared__ float2 s_pair[SCAP][APB];   // 1/2 d(f v_R)/dr / r, 1/2 d(f v_A)/dr / r

    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int i = blk * APB + a;
    const bool valid = i < n;
    const double *cv = geo.vec(valid ? i : 0);  // vectors of this atom's cell
    const int gbase = ((threadIdx.x & 63) / G) * G;
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    const BondTerm B = M->bond[0][0][0];
    const bool bond_on = M->group >= 0 && B.valid;
    const double maxbond = M->maxbond[0], maxpot = M->maxpot[0], rshort = M->rshort;
    const int npair = M->npair[0][0];
    const float hf = (float)B.h, ddf = (float)(B.d * B.d), ccf = (float)(B.c * B.c);
    const float cd2f = (float)(B.c * B.c / (B.d * B.d));
    const float betaf = (float)B.beta, etaf = (float)B.eta;
    const float inv_eta = (float)(-0.5 / B.eta);
    // bond cutoff: f_c = (1 + cos(pi t)) / 2, t = (r - soft) / (cut - soft);
    // v_cos_f32 / v_sin_f32 take revolutions, so pi t = 2 pi (t / 2)
    const bool bsmooth = B.soft > 0.0 && B.soft < B.cut;
    const float bsoft = (float)B.soft, bhalfw = bsmooth ? (float)(0.5 / (B.cut - B.soft)) : 0.f;
    const float bdk = bsmooth ? (float)(-0.5 * M_PI / (B.cut - B.soft)) : 0.f;

    double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);

    // ---- 1. short list; per neighbor: cutoff functions and pair terms -------
    int nf = valid ? ncnt[i] : 0;
    int cnt = 0;
    for (int base = 0; base < nf; base += G) {
        int e = base + l;
        bool keep = false;
        double3 sp3 = make_double3(0, 0, 0);
        double r = 0.0;
        int j = 0, im = 0;
        if (e < nf) {
            j = nbr[(size_t)e * n + i];
            im = nimg[(size_t)e * n + i];
            sp3 = separation64(pi, pos[j], im, cv);
            r = sqrt(dot64(sp3));
            keep = r < rshort;
        }
        unsigned long long b = __ballot(keep);
        unsigned long long gb = (b >> gbase) & gmask;
        int slot = cnt + __popcll(gb & ((1ull << l) - 1ull));
        if (keep && slot < SCAP) {
            const float rf = (float)r;
            float dR = 0.f, dA = 0.f, A = 0.f;
            if (r < maxpot) {
                for (int t = 0; t < npair; ++t) {
                    const PairTerm &pt = M->pair[0][0][t];
                    if (!(pt.cut > r)) continue;
                    const float v = (float)pt.eps * exp_split(-pt.zeta * r);
                    float f = 1.f, df = 0.f;
                    if (pt.smooth && r > pt.soft) {
                        const float hw = (float)(0.5 / (pt.cut - pt.soft));
                        const float x = (rf - (float)pt.soft) * hw;
                        f = 0.5f + 0.5f * __builtin_amdgcn_cosf(x);
                        df = -(float)M_PI * hw * __builtin_amdgcn_sinf(x);
                    }
                    const float dv = 0.5f * v * (df - (float)pt.zeta * f);
                    if (pt.group >= 0) { dA += dv; A += 0.5f * v * f; }
                    else dR += dv;
                }
            }
            float ir = __builtin_amdgcn_rcpf(rf);
            ir = ir * (2.0f - rf * ir);
            float fb = 0.f, dfb = 0.f;
            if (bond_on && r < maxbond) {
                fb = 1.f;
                if (bsmooth && r > B.soft) {
                    const float x = (rf - bsoft) * bhalfw;
                    fb = 0.5f + 0.5f * __builtin_amdgcn_cosf(x);
                    dfb = bdk * __builtin_amdgcn_sinf(x);
                }
            }
            s_u[slot][a] = make_float4((float)sp3.x, (float)sp3.y, (float)sp3.z, rf);
            s_aux[slot][a] = make_float4(ir, fb, dfb, A);
            s_pair[slot][a] = make_float2(dR * ir, dA * ir);  // divided by r once
            if (!unique) {
                size_t o = (size_t)slot * n + i;
                sj[o] = j;
                simg[o] = im;
            }
            srow[(size_t)i * SCAP + slot] = j;
        }
        cnt += __popcll(gb);
    }
    if (cnt > SCAP && l == 0) atomicMax(flags + 1, cnt);
    const int ns = cnt < SCAP ? cnt : SCAP;
    __syncthreads();

    // ---- 2. zeta, b~, b~' ---------------------------------------------------
    float cpair[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        cpair[q] = 0.f;
        int jj = l + q * G;
        if (jj >= ns) continue;
        const float4 u = s_u[jj][a];
        const float4 au = s_aux[jj][a];
        float zeta = 0.f;
        if (bond_on) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                const float4 ak = s_aux[k][a];
                if (ak.y == 0.f) continue;
                const float4 v = s_u[k][a];
                const float cs = (u.x * v.x + u.y * v.y + u.z * v.z) * au.x * ak.x;
                const float t = hf - cs;
                const float t2 = t * t;
                float xi = ak.y;
                if (!B.a_zero) xi *= __expf(powf((float)B.a * (u.w - v.w), (float)B.mu));
                zeta += xi * (1.0f + cd2f * t2 * __builtin_amdgcn_rcpf(ddf + t2));
            }
        }
        float bt = 1.f, dbt = 0.f;
        if (bond_on) {
            // b~ = (1 + (beta zeta)^eta)^(-1/(2 eta)); b~' = b~ / (1 + pw) *
            // (-1/2) (beta zeta)^(eta - 1) beta
            const float be = betaf * zeta;
            const float pw = be > 0.f ? __powf(be, etaf) : 0.f;
            bt = __powf(1.0f + pw, inv_eta);
            if (etaf == 0.f)
                dbt = 0.f;  // Potentials.f90:5940-5941
            else if (be > 0.f)
                dbt = inv_eta * etaf * betaf * bt * pw *
                      __builtin_amdgcn_rcpf((1.0f + pw) * be);
            else if (etaf == 1.f)
                dbt = inv_eta * betaf;  // (beta zeta)^0 = 1 at zeta = 0
            // eta < 1 at zeta = 0: the reference's b~' = 0 (Potentials.f90:5946-5953)
        }
        const float2 pp = s_pair[jj][a];
        cpair[q] = pp.x + pp.y * bt;   // (d/dr of 1/2 f (v_R + v_A b~) at fixed b~) / r
        s_aux[jj][a].w = au.w * dbt;   // P = 1/2 f v_A db~/dzeta
    }
    __syncthreads();

    // ---- 3. D_ij ------------------------------------------------------------
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int jj = l + q * G;
        if (jj >= ns) continue;
        const float4 u = s_u[jj][a];
        const float4 aj = s_aux[jj][a];
        const float ir1 = aj.x, fj = aj.y, dfj = aj.z, Pj = aj.w;
        float cu = 0.f, dcx = 0.f, dcy = 0.f, dcz = 0.f;
        if (bond_on) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                const float4 ak = s_aux[k][a];
                const float fk = ak.y, Pk = ak.w;
                if (fk == 0.f && fj == 0.f) continue;
                const float4 v = s_u[k][a];
                const float ir12 = ir1 * ak.x;
                const float cs = (u.x * v.x + u.y * v.y + u.z * v.z) * ir12;
                const float t = hf - cs;
                const float iden = __builtin_amdgcn_rcpf(ddf + t * t);
                const float gee = 1.0f + cd2f * t * t * iden;
                const float gp = -2.0f * ccf * t * iden * iden;
                float coef_u, coef_c;
                if (B.a_zero) {
                    coef_u = Pk * gee * dfj;
                    coef_c = gp * (Pj * fk + Pk * fj);
                } else {
                    const float av = (float)B.a, mu = (float)B.mu;
                    const float x1 = av * (u.w - v.w), x2 = -x1;
                    const float p1 = powf(x1, mu - 1.0f), p2 = powf(x2, mu - 1.0f);
                    const float e1 = __expf(x1 * p1), e2 = __expf(x2 * p2);
                    const float d1 = e1 * av * mu * p1, d2 = e2 * av * mu * p2;
                    coef_u = Pj * gee * fk * d1 + Pk * gee * (dfj * e2 - fj * d2);
                    coef_c = gp * (Pj * fk * e1 + Pk * fj * e2);
                }
                const float cu2 = cs * ir1 * ir1;
                dcx += coef_c * (v.x * ir12 - cu2 * u.x);
                dcy += coef_c * (v.y * ir12 - cu2 * u.y);
                dcz += coef_c * (v.z * ir12 - cu2 * u.z);
                cu += coef_u * ir1;
            }
        }
        const float c = cpair[q] + cu;
        dU[(size_t)jj * n + i] = make_float4(c * u.x + dcx, c * u.y + dcy, c * u.z + dcz, 0.f);
    }
    if (valid) {
        if (l == 0) scnt[i] = ns;
        for (int t = ns + l; t < SCAP; t += G) srow[(size_t)i * SCAP + t] = -1;
    }
}

template <int G, int SCAP, class Geo, bool VIR>
__device__ __forceinline__ void site1f_tile(int blk, int n, const Geo &geo,
                                            const double4 *__restrict__ pos,
                                            const int *__restrict__ nbr,
                                            const int *__restrict__ nimg,
                                            const int *__restrict__ ncnt,
                                            const Model *__restrict__ M,
                                            int *__restrict__ sj, int *__restrict__ simg,
                                            int *__restrict__ scnt, int *__restrict__ srow,
                                            int unique, float4 *__restrict__ dU,
                                            double *__restrict__ partial,
                                            double *__restrict__ atom_ev,
                                            int *__restrict__ flags)
{
    constexpr int APB = BLOCK / G;
    constexpr int PER = (SCAP + G - 1) / G;
    // [slot][atom]: the atoms of a wave sit on consecutive LDS banks
    __shared__ double4 s_sep[SCAP][APB];      // x, y, z, r (FP64)
    __shared__ float4 s_aux[SCAP][APB];       // 1/r, f_c(r), f_c'(r), A (then P)
    __shared__ double2 s_pair[SCAP][APB];     // 1/2 d(f v_R)/dr / r, 1/2 d(f v_A)/dr / r
    __shared__ double s_red[APB][8];

    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int i = blk * APB + a;
    const bool valid = i < n;
    const double *cv = geo.vec(valid ? i : 0);  // vectors of this atom's cell
    const int gbase = ((threadIdx.x & 63) / G) * G;
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    const BondTerm B = M->bond[0][0][0];
    const bool bond_on = M->group >= 0 && B.valid;
    const double maxbond = M->maxbond[0], maxpot = M->maxpot[0], rshort = M->rshort;
    const int npair = M->npair[0][0];
    const float hf = (float)B.h, ddf = (float)(B.d * B.d), ccf = (float)(B.c * B.c);
    const float cd2f = (float)(B.c * B.c / (B.d * B.d));
    const float betaf = (float)B.beta, etaf = (float)B.eta;

    double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);

    // ---- 1. short list; per neighbor: cutoff functions and pair terms -------
    int nf = valid ? ncnt[i] : 0;
    int cnt = 0;
    for (int base = 0; base < nf; base += G) {
        int e = base + l;
        bool keep = false;
        double3 sp3 = make_double3(0, 0, 0);
        double r = 0.0;
        int j = 0, im = 0;
        if (e < nf) {
            j = nbr[(size_t)e * n + i];
            im = nimg[(size_t)e * n + i];
            sp3 = separation64(pi, pos[j], im, cv);
            r = sqrt(dot64(sp3));
            keep = r < rshort;
        }
        unsigned long long b = __ballot(keep);
        unsigned long long gb = (b >> gbase) & gmask;
        int slot = cnt + __popcll(gb & ((1ull << l) - 1ull));
        if (keep && slot < SCAP) {
            bool inb = bond_on && r < maxbond;
            double fb = 0.0, dfb = 0.0;
            if (inb) smooth64(r, B.cut, B.soft, &fb, &dfb);
            double dR = 0.0, dA = 0.0, A = 0.0;
            if (r < maxpot) {
                for (int t = 0; t < npair; ++t) {
                    const PairTerm pt = M->pair[0][0][t];
                    if (!(pt.cut > r)) continue;
                    double v = pt.eps * exp(-pt.zeta * r);
                    double f = 1.0, df = 0.0;
                    if (pt.smooth) {
                        // the bond factor's cutoffs: its smoothing, the same values
                        if (inb && pt.cut == B.cut && pt.soft == B.soft) { f = fb; df = dfb; }
                        else smooth64(r, pt.cut, pt.soft, &f, &df);
                    }
                    double dv = 0.5 * (-pt.zeta * v * f + v * df);
                    if (pt.group >= 0) { dA += dv; A += 0.5 * v * f; }
                    else dR += dv;
                }
            }
            const double ir = recip64(r);  // r > 0: the list holds no atom at its own position
            s_sep[slot][a] = make_double4(sp3.x, sp3.y, sp3.z, r);
            s_aux[slot][a] = make_float4((float)ir, (float)fb, (float)dfb, (float)A);
            s_pair[slot][a] = make_double2(dR * ir, dA * ir);  // divided by r once
            // the images also with unique images: an energy pass reads them
            const size_t o = (size_t)slot * n + i;
            if (!unique) sj[o] = j;
            simg[o] = im;
            srow[(size_t)i * SCAP + slot] = j;
        }
        cnt += __popcll(gb);
    }
    if (cnt > SCAP && l == 0) atomicMax(flags + 1, cnt);
    const int ns = cnt < SCAP ? cnt : SCAP;
    __syncthreads();

    // ---- 2. zeta, b~, b~' in FP32 -------------------------------------------------
    double cpair[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        cpair[q] = 0.0;
        int jj = l + q * G;
        if (jj >= ns) continue;
        double4 ud = s_sep[jj][a];
        float ux = (float)ud.x, uy = (float)ud.y, uz = (float)ud.z, r1 = (float)ud.w;
        float4 au = s_aux[jj][a];
        float zeta = 0.f;
        if (bond_on) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                float4 ak = s_aux[k][a];
                if (ak.y == 0.f) continue;
                double4 vd = s_sep[k][a];
                float cs = (ux * (float)vd.x + uy * (float)vd.y + uz * (float)vd.z) * au.x * ak.x;
                float t = hf - cs;
                float t2 = t * t;
                float xi = ak.y;
                if (!B.a_zero) xi *= __expf(powf((float)B.a * (r1 - (float)vd.w), (float)B.mu));
                zeta += xi * (1.0f + cd2f * t2 * __builtin_amdgcn_rcpf(ddf + t2));
            }
        }
        float bt = 1.f, dbt = 0.f;
        if (bond_on) {
            float be = betaf * zeta;
            float pw = __powf(be, etaf);
            float inv_eta = -0.5f / etaf;
            bt = __powf(1.0f + pw, inv_eta);
            if (!(etaf == 0.f || (zeta == 0.f && etaf <= 1.f)))
                dbt = inv_eta * (bt / (1.0f + pw)) * etaf * (pw / be) * betaf;
        }
        double2 pp = s_pair[jj][a];
        cpair[q] = pp.x + pp.y * (double)bt;  // (d/dr of 1/2 f (v_R + v_A b~) at fixed b~) / r
        s_aux[jj][a].w = au.w * dbt;           // P = 1/2 f v_A db~/dzeta
    }
    __syncthreads();

    // ---- 3. D_ij: FP64 pair part + FP32 three-body part; virial in FP64 ------
    double w6[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int jj = l + q * G;
        if (jj >= ns) continue;
        double4 ud = s_sep[jj][a];
        float ux = (float)ud.x, uy = (float)ud.y, uz = (float)ud.z, r1 = (float)ud.w;
        float4 aj = s_aux[jj][a];
        float ir1 = aj.x, fj = aj.y, dfj = aj.z, Pj = aj.w;
        float cu = 0.f, dcx = 0.f, dcy = 0.f, dcz = 0.f;
        if (bond_on) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                float4 ak = s_aux[k][a];
                float fk = ak.y, Pk = ak.w;
                if (fk == 0.f && fj == 0.f) continue;
                double4 vd = s_sep[k][a];
                float vx = (float)vd.x, vy = (float)vd.y, vz = (float)vd.z;
                float ir12 = ir1 * ak.x;
                float cs = (ux * vx + uy * vy + uz * vz) * ir12;
                float t = hf - cs;
                float iden = __builtin_amdgcn_rcpf(ddf + t * t);
                float gee = 1.0f + cd2f * t * t * iden;
                float gp = -2.0f * ccf * t * iden * iden;
                float coef_u, coef_c;
                if (B.a_zero) {
                    coef_u = Pk * gee * dfj;
                    coef_c = gp * (Pj * fk + Pk * fj);
                } else {
                    float av = (float)B.a, mu = (float)B.mu, r2 = (float)vd.w;
                    float x1 = av * (r1 - r2), x2 = -x1;
                    float p1 = powf(x1, mu - 1.0f), p2 = powf(x2, mu - 1.0f);
                    float e1 = __expf(x1 * p1), e2 = __expf(x2 * p2);
                    float d1 = e1 * av * mu * p1, d2 = e2 * av * mu * p2;
                    coef_u = Pj * gee * fk * d1 + Pk * gee * (dfj * e2 - fj * d2);
                    coef_c = gp * (Pj * fk * e1 + Pk * fj * e2);
                }
                float cu2 = cs * ir1 * ir1;
                dcx += coef_c * (vx * ir12 - cu2 * ux);
                dcy += coef_c * (vy * ir12 - cu2 * uy);
                dcz += coef_c * (vz * ir12 - cu2 * uz);
                cu += coef_u * ir1;
            }
        }
        double cp = cpair[q];
        double Dx = cp * ud.x + (double)(cu * ux + dcx);
        double Dy = cp * ud.y + (double)(cu * uy + dcy);
        double Dz = cp * ud.z + (double)(cu * uz + dcz);
        dU[(size_t)jj * n + i] = make_float4((float)Dx, (float)Dy, (float)Dz, 0.f);
        if constexpr (VIR) {
            w6[0] -= ud.x * Dx;
            w6[1] -= ud.y * Dy;
            w6[2] -= ud.z * Dz;
            w6[3] -= ud.y * Dz;
            w6[4] -= ud.x * Dz;
            w6[5] -= ud.x * Dy;
        }
    }
    if (valid) {
        if (l == 0) scnt[i] = ns;
        for (int t = ns + l; t < SCAP; t += G) srow[(size_t)i * SCAP + t] = -1;
    }

    if constexpr (VIR) {
#pragma unroll
        for (int q = 0; q < 6; ++q)
            for (int off = G / 2; off > 0; off >>= 1) w6[q] += __shfl_xor(w6[q], off, G);
        if (l == 0)
            for (int q = 0; q < 6; ++q) s_red[a][1 + q] = valid ? w6[q] : 0.0;
        if (atom_ev && valid && l == 0)
            for (int q = 0; q < 6; ++q) atom_ev[(size_t)(1 + q) * n + i] = w6[q];
        __syncthreads();
        if (threadIdx.x < 7) {
            double sacc = 0.0;
            if (threadIdx.x > 0)
                for (int t = 0; t < APB; ++t) sacc += s_red[t][threadIdx.x];
            partial[(size_t)blk * 8 + threadIdx.x] = sacc;  // energy slot 0 left at 0
        }
    }
}

// Kernels over tiles of BLOCK / G atoms.  With one block per tile each block
// takes its XCD-remapped tile; with fewer blocks (a persistent launch) block
// b walks the contiguous tile range [b T / grid, (b + 1) T / grid), so one
// launch covers any number of cells and each XCD still sees a contiguous run
// of atoms.
template <class Geo>
__device__ __forceinline__ void tile_range(int ntiles, int &t0, int &t1)
{
    const long pb = xcd_block(gridDim.x);
    if constexpr (Geo::kPersistent) {
        t0 = (int)(pb * ntiles / gridDim.x);
        t1 = (int)((pb + 1) * ntiles / gridDim.x);
    } else {
        t0 = (int)pb;   // the launch has exactly one block per tile
        t1 = t0 + 1;
    }
}

template <int G, int SCAP, class Geo>
__global__ __launch_bounds__(BLOCK) void k_site(int n, int ntiles, Geo geo,
                                                const double4 *__restrict__ pos,
                                                const int *__restrict__ id,
                                                const int *__restrict__ nbr,
                                                const int *__restrict__ nimg,
                                                const int *__restrict__ ncnt,
                                                const Model *__restrict__ M,
                                                int *__restrict__ sj, int *__restrict__ simg,
                                                int *__restrict__ scnt, int *__restrict__ srow,
                                                int unique, float4 *__restrict__ dU,
                                                double *__restrict__ partial,
                                                double *__restrict__ atom_ev,
                                                int *__restrict__ flags)
{
    int t0, t1;
    tile_range<Geo>(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        site_tile<G, SCAP>(t, n, geo, pos, id, nbr, nimg, ncnt, M, sj, simg, scnt, srow, unique,
                           dU, partial, atom_ev, flags);
        __syncthreads();
    }
}

template <int G, int SCAP, class Geo, bool SHORT>
__global__ __launch_bounds__(BLOCK) void k_site1e(int n, int ntiles, Geo geo,
                                                  const double4 *__restrict__ pos,
                                                  const int *__restrict__ nbr,
                                                  const int *__restrict__ nimg,
                                                  const int *__restrict__ ncnt,
                                                  const int *__restrict__ scnt,
                                                  const int *__restrict__ srow,
                                                  const int *__restrict__ simg,
                                                  const Model *__restrict__ M,
                                                  double *__restrict__ partial,
                                                  double *__restrict__ atom_ev,
                                                  int *__restrict__ flags)
{
    int t0, t1;
    tile_range<Geo>(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        site1e_tile<G, SCAP, Geo, SHORT>(t, n, geo, pos, nbr, nimg, ncnt, scnt, srow, simg, M,
                                         partial, atom_ev, flags);
        __syncthreads();
    }
}

template <int G, int SCAP, class Geo, bool VIR>
__global__ __launch_bounds__(BLOCK) void k_site1f(int n, int ntiles, Geo geo,
                                                  const double4 *__restrict__ pos,
                                                  const int *__restrict__ nbr,
                                                  const int *__restrict__ nimg,
                                                  const int *__restrict__ ncnt,
                                                  const Model *__restrict__ M,
                                                  int *__restrict__ sj, int *__restrict__ simg,
                                                  int *__restrict__ scnt, int *__restrict__ srow,
                                                  int unique, float4 *__restrict__ dU,
                                                  double *__restrict__ partial,
                                                  double *__restrict__ atom_ev,
                                                  int *__restrict__ flags)
{
    int t0, t1;
    tile_range<Geo>(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) {
        if constexpr (VIR)
            site1f_tile<G, SCAP, Geo, true>(t, n, geo, pos, nbr, nimg, ncnt, M, sj, simg, scnt,
                                            srow, unique, dU, partial, atom_ev, flags);
        else
            site1m_tile<G, SCAP, Geo>(t, n, geo, pos, nbr, nimg, ncnt, M, sj, simg, scnt, srow,
                                      unique, dU, flags);
        __syncthreads();
    }
}

// Per-cell sums of the per-atom energies and virials: one block per cell,
// one wave per component.  Each lane keeps 8 partial sums over a stride of
// 512 atoms (8 loads in flight), adds them in a fixed order, and a butterfly
// sums the wave: a fixed order throughout, no barriers.
constexpr int CELL_SUM_COMPONENTS = 7;
__global__ __launch_bounds__(64 * CELL_SUM_COMPONENTS) void k_cell_sums(
    int n, const int *__restrict__ atom0, const double *__restrict__ atom_ev,
    double *__restrict__ cell_ev)
{
    constexpr int U = 8;
    const int c = blockIdx.x;
    const int q = threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    const int a0 = atom0[c], a1 = atom0[c + 1];
    const double *v = atom_ev + (size_t)q * n;
    double p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) p[u] = 0.0;
    for (int b = a0 + lane; b < a1; b += 64 * U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = b + 64 * u;
            p[u] += i < a1 ? v[i] : 0.0;
        }
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) s += p[u];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) cell_ev[(size_t)c * 8 + q] = s;
}

// F_i = sum_j (D_ij - D_ji); D_ji found in j's short list at (i, -o).
__global__ __launch_bounds__(256) void k_gather(int n, const int *__restrict__ sj,
                                                const int *__restrict__ simg,
                                                const int *__restrict__ scnt,
                                                const int *__restrict__ srow, int scap,
                                                int unique, const float4 *__restrict__ dU,
                                                double4 *__restrict__ frc,
                                                double4 *__restrict__ mom, double half_dt,
                                                int kick, int *__restrict__ flags)
{
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool miss = false;
    const double3 f = unique ? gather_rows_any(i, n, scap, srow, dU, miss)
                             : gather_bond_force(i, n, sj, simg, scnt, dU, miss);
    // flags[3]: bit 0 a bond without its reverse, bit 1 a non-finite force
    const int fbad = (miss ? 1 : 0) |
                     ((isfinite(f.x) && isfinite(f.y) && isfinite(f.z)) ? 0 : 2);
    if (fbad) atomicOr(flags + 3, fbad);
    frc[i] = make_double4(f.x, f.y, f.z, 0.0);
    if (kick) {
        // the second half-kick of velocity Verlet (fma: the fused
        // gather-drift kernel of the MD loop rounds it the same way)
        double4 p = mom[i];
        p.x = fma(half_dt, f.x, p.x);
        p.y = fma(half_dt, f.y, p.y);
        p.z = fma(half_dt, f.z, p.z);
        mom[i] = p;
    }
}

// one block per component: block c sums partial[b*8 + c] over b (fixed order)
__global__ __launch_bounds__(1024) void k_reduce_partials(int nb, const double *__restrict__ partial,
                                                          double *__restrict__ totals)
{
    __shared__ double lds[1024];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 1024) s += partial[(size_t)b * 8 + q];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[q] = lds[0];
}

struct Variant {
    const char *name;
    int G, SCAP;
};
const Variant kVariants[] = {{"g4_s8", 4, 8},   {"g4_s16", 4, 16}, {"g8_s8", 8, 8},
                             {"g8_s16", 8, 16}, {"g2_s8", 2, 8},   {"g16_s32", 16, 32}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

template <int G, int SCAP>
void launch_site(System &S, int mode)
{
    constexpr int APB = BLOCK / G;
    const int ntiles = div_up(S.n, APB);
    // persistent launch (batches of cells): at most site_grid blocks walk
    // all tiles; one cell: one block per tile
    const int grid =
        (S.ncells > 0 && S.site_grid > 0 && S.site_grid < ntiles) ? S.site_grid : ntiles;
    S.partial.ensure((size_t)ntiles * 8);
    // rows only: the column-major short list is not written (nor read)
    const int uq = S.unique_images ? 1 : 0;
    double *aev = nullptr;
    if (S.ncells > 0) {
        S.atom_ev.ensure((size_t)7 * S.n);
        aev = S.atom_ev.p;
    }
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        if (S.model.nspecies == 1 && mode == SITE_ENERGY)
            k_site1e<G, SCAP, Geo, false><<<grid, BLOCK, 0, S.stream>>>(
                S.n, ntiles, geo, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p, nullptr, nullptr, nullptr,
                S.dmodel.p, S.partial.p, aev, S.flags.p);
        else if (S.model.nspecies == 1 && mode == SITE_ENERGY_SHORT)
            k_site1e<G, SCAP, Geo, true><<<grid, BLOCK, 0, S.stream>>>(
                S.n, ntiles, geo, S.pos.p, nullptr, nullptr, nullptr, S.scnt.p, S.srow.p,
                S.simg.p, S.dmodel.p, S.partial.p, aev, S.flags.p);
        else if (S.model.nspecies == 1 && mode == SITE_FORCES)
            k_site1f<G, SCAP, Geo, true><<<grid, BLOCK, 0, S.stream>>>(
                S.n, ntiles, geo, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p, S.dmodel.p, S.sj.p,
                S.simg.p, S.scnt.p, S.srow.p, uq, S.dU.p, S.partial.p, aev, S.flags.p);
        else if (S.model.nspecies == 1 && mode == SITE_FORCES_ONLY)
            k_site1f<G, SCAP, Geo, false><<<grid, BLOCK, 0, S.stream>>>(
                S.n, ntiles, geo, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p, S.dmodel.p, S.sj.p,
                S.simg.p, S.scnt.p, S.srow.p, uq, S.dU.p, S.partial.p, aev, S.flags.p);
        else
            k_site<G, SCAP, Geo><<<grid, BLOCK, 0, S.stream>>>(
                S.n, ntiles, geo, S.pos.p, S.id.p, S.nbr.p, S.nimg.p, S.ncnt.p, S.dmodel.p, S.sj.p,
                S.simg.p, S.scnt.p, S.srow.p, uq, S.dU.p, S.partial.p, aev, S.flags.p);
    });
}

}  // namespace

int ts_variant_count() { return kNumVariants; }
const char *ts_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kVariants[v].name : ""; }
int ts_variant_capacity(int v) { return (v >= 0 && v < kNumVariants) ? kVariants[v].SCAP : 0; }
// smallest-G variant whose short-list capacity holds `need` neighbors
int ts_variant_for_capacity(int need)
{
    int best = -1;
    for (int v = 0; v < kNumVariants; ++v)
        if (kVariants[v].SCAP >= need && (best < 0 || kVariants[v].SCAP < kVariants[best].SCAP))
            best = v;
    return best;
}

void ts_site(System &S, int variant, int mode)
{
    PYSIC_REQUIRE(variant >= 0 && variant < kNumVariants, PYSIC_EINVAL, "bad Tersoff kernel variant");
    S.scap = kVariants[variant].SCAP;
    S.sj.ensure((size_t)S.scap * S.n);
    S.simg.ensure((size_t)S.scap * S.n);
    S.scnt.ensure(S.n);
    S.dU.ensure((size_t)S.scap * S.n);
    S.srow.ensure((size_t)S.scap * S.n);
    hipEvent_t ev;
    const char *name =
        (mode == SITE_ENERGY || mode == SITE_ENERGY_SHORT) ? "tersoff_energy" : "tersoff_site";
    S.timer.begin(name, S.stream, &ev);
    switch (variant) {
    case 0: launch_site<4, 8>(S, mode); break;
    case 1: launch_site<4, 16>(S, mode); break;
    case 2: launch_site<8, 8>(S, mode); break;
    case 3: launch_site<8, 16>(S, mode); break;
    case 4: launch_site<2, 8>(S, mode); break;
    case 5: launch_site<16, 32>(S, mode); break;
    }
    HIP_TRY(hipGetLastError());
    S.timer.end(name, S.stream, ev);
    S.partial_blocks = div_up(S.n, BLOCK / kVariants[variant].G);
}

// Site-kernel tiles of the single-species FP64 energy pass for this system
// and variant: the number of per-block partial rows it writes.
int ts_energy_tiles(const System &S, int variant)
{
    PYSIC_REQUIRE(variant >= 0 && variant < kNumVariants, PYSIC_EINVAL, "bad Tersoff kernel variant");
    return div_up(S.n, BLOCK / kVariants[variant].G);
}

namespace {
template <int G, int SCAP>
void launch_energy_into(System &S, double *partial)
{
    const int ntiles = div_up(S.n, BLOCK / G);
    k_site1e<G, SCAP, GeoOne, false><<<ntiles, BLOCK, 0, S.stream>>>(
        S.n, ntiles, GeoOne{S.geom}, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p, nullptr, nullptr,
        nullptr, S.dmodel.p, partial, nullptr, S.flags.p);
}
}  // namespace

// The single-species FP64 energy pass of one cell with caller-provided
// scratch: partial[ts_energy_tiles * 8] receives the per-block sums and
// totals[0] the energy.  Allocates nothing and records no timing events;
// the flags are cleared by a kernel first (short-list overflow lands in
// flags[1], as in ts_site).
void ts_energy_into(System &S, int variant, double *partial, double *totals)
{
    PYSIC_REQUIRE(S.ncells == 0 && S.model.nspecies == 1, PYSIC_EINVAL,
                  "the scratch energy pass takes one cell of one species");
    nb_clear_flags(S);
    switch (variant) {
    case 0: launch_energy_into<4, 8>(S, partial); break;
    case 1: launch_energy_into<4, 16>(S, partial); break;
    case 2: launch_energy_into<8, 8>(S, partial); break;
    case 3: launch_energy_into<8, 16>(S, partial); break;
    case 4: launch_energy_into<2, 8>(S, partial); break;
    case 5: launch_energy_into<16, 32>(S, partial); break;
    default: throw PysicError{PYSIC_EINVAL, "bad Tersoff kernel variant"};
    }
    HIP_TRY(hipGetLastError());
    k_reduce_partials<<<1, 1024, 0, S.stream>>>(ts_energy_tiles(S, variant), partial, totals);
    HIP_TRY(hipGetLastError());
}

void ts_gather(System &S, double half_dt, int kick)
{
    hipEvent_t ev;
    S.timer.begin("tersoff_gather", S.stream, &ev);
    k_gather<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, S.sj.p, S.simg.p, S.scnt.p, S.srow.p,
                                                     S.scap, S.unique_images ? 1 : 0, S.dU.p,
                                                     S.frc.p, S.mom.p, half_dt, kick, S.flags.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("tersoff_gather", S.stream, ev);
}

void ts_cell_sums(System &S)
{
    PYSIC_REQUIRE(S.ncells > 0, PYSIC_ESTATE, "not a batch of cells");
    S.cell_ev.ensure((size_t)8 * S.ncells);
    // the energy kernel writes component 0, the force kernels 1..6; the
    // generic kernel writes all seven
    k_cell_sums<<<S.ncells, 64 * CELL_SUM_COMPONENTS, 0, S.stream>>>(S.n, S.atom0.p, S.atom_ev.p,
                                                                    S.cell_ev.p);
    HIP_TRY(hipGetLastError());
}

void ts_reduce(System &S)
{
    S.totals.ensure(8);
    k_reduce_partials<<<7, 1024, 0, S.stream>>>(S.partial_blocks, S.partial.p, S.totals.p);
    HIP_TRY(hipGetLastError());
}
