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

__device__ __forceinline__ double fc64(double r, double hard, double soft)
{
    // smoothening_factor (Potentials.f90:442-458)
    if (r > hard) return 0.0;
    if (r < soft) return 1.0;
    return 0.5 * (1.0 + cos(M_PI * (r - soft) / (hard - soft)));
}
__device__ __forceinline__ double dfc64(double r, double hard, double soft)
{
    // smoothening_derivative (Potentials.f90:485-501)
    if (r > hard || r < soft) return 0.0;
    return -0.5 * M_PI / (hard - soft) * sin(M_PI * (r - soft) / (hard - soft));
}
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
    double pw = pow(be, p.eta);
    double inv_eta = -1.0 / (2.0 * p.eta);
    *bt = pow(1.0 + pw, inv_eta);
    if (p.eta == 0.0 || (zeta == 0.0 && p.eta <= 1.0))
        *dbt = 0.0;
    else
        *dbt = inv_eta * pow(1.0 + pw, inv_eta - 1.0) * p.eta * pow(be, p.eta - 1.0) * p.beta;
}

// pick (Geometry.f90:648-676) on original indices
__device__ __forceinline__ bool pick_dev(int i1, int i2, int img)
{
    if (i2 > i1) return true;
    if (i2 < i1) return false;
    int ox = img_x(img), oy = img_y(img), oz = img_z(img);
    return ox > 0 || (ox == 0 && (oy > 0 || (oy == 0 && oz > 0)));
}

template <int G, int SCAP>
struct SiteShared {
    static constexpr int APB = BLOCK / G;
    double4 sep[APB][SCAP];   // r_ij (FP64) and |r_ij|
    int jdx[APB][SCAP];
    int img[APB][SCAP];
    int spc[APB][SCAP];
    float P[APB][SCAP];       // 1/2 v_A f_c db~/dzeta for the bond (i,j)
    double dpair[APB][SCAP];  // d/dr of 1/2 f_c (v_R + v_A b~) at fixed b~
    double red[APB][8];
};

template <int G, int SCAP>
__global__ __launch_bounds__(BLOCK) void k_site(int n, CellGeom g, const double4 *__restrict__ pos,
                                                const int *__restrict__ id,
                                                const int *__restrict__ nbr,
                                                const int *__restrict__ nimg,
                                                const int *__restrict__ ncnt,
                                                const Model *__restrict__ M,
                                                int *__restrict__ sj, int *__restrict__ simg,
                                                int *__restrict__ scnt, float4 *__restrict__ dU,
                                                double *__restrict__ partial,
                                                int *__restrict__ flags)
{
    using SH = SiteShared<G, SCAP>;
    constexpr int APB = SH::APB;
    __shared__ SH sh;
    const int a = threadIdx.x / G;
    const int l = threadIdx.x % G;
    const int i = blockIdx.x * APB + a;
    const bool valid = i < n;
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
            double3 sp3 = separation64(pi, pj, im, g.v);
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
            sh.sep[a][slot] = d;
            sh.jdx[a][slot] = j;
            sh.img[a][slot] = im;
            sh.spc[a][slot] = sp;
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
        double4 u = sh.sep[a][jj];
        int sjp = sh.spc[a][jj];
        double zeta = 0.0;
        if (grp >= 0) {
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                double4 v = sh.sep[a][k];
                if (!(v.w < maxbond_i)) continue;
                const BondTerm &bt = M->bond[si][sjp][sh.spc[a][k]];
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
            int idi = id[i], idj = id[sh.jdx[a][jj]];
            if (!pick_dev(idi, idj, sh.img[a][jj])) mp = M->maxpot[sjp];
        }
        double eh = 0.0, dp = 0.0, P = 0.0;
        if (u.w < mp) {
            int np = M->npair[si][sjp];
            for (int t = 0; t < np; ++t) {
                const PairTerm &pt = M->pair[si][sjp][t];
                if (!(pt.cut > u.w)) continue;
                double v = pt.eps * exp(-pt.zeta * u.w);
                double dv = -pt.zeta * v;
                double f = pt.smooth ? fc64(u.w, pt.cut, pt.soft) : 1.0;
                double df = pt.smooth ? dfc64(u.w, pt.cut, pt.soft) : 0.0;
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
        sh.P[a][jj] = (float)P;
        sh.dpair[a][jj] = dp;
    }
    __syncthreads();

    // ---- 3. D_ij = dU_i/dr_ij (FP32 three-body arithmetic) -------------------
    double w6[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = l; jj < ns; jj += G) {
        double4 ud = sh.sep[a][jj];
        float ux = (float)ud.x, uy = (float)ud.y, uz = (float)ud.z, r1 = (float)ud.w;
        float inv_r1 = 1.0f / r1;
        float Pj = sh.P[a][jj];
        int sjp = sh.spc[a][jj];
        float cu = 0.f;  // three-body coefficient of u
        float dcx = 0.f, dcy = 0.f, dcz = 0.f;  // accumulated coefficient * dcos/du
        if (grp >= 0) {
            const float mb = (float)maxbond_i;
            for (int k = 0; k < ns; ++k) {
                if (k == jj) continue;
                double4 vd = sh.sep[a][k];
                float vx = (float)vd.x, vy = (float)vd.y, vz = (float)vd.z, r2 = (float)vd.w;
                float Pk = sh.P[a][k];
                int skp = sh.spc[a][k];
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
        double cp = sh.dpair[a][jj] / ud.w;
        double Dx = cp * ud.x + (double)(cu * ux + dcx);
        double Dy = cp * ud.y + (double)(cu * uy + dcy);
        double Dz = cp * ud.z + (double)(cu * uz + dcz);
        size_t o = (size_t)jj * n + i;
        dU[o] = make_float4((float)Dx, (float)Dy, (float)Dz, 0.f);
        sj[o] = sh.jdx[a][jj];
        simg[o] = sh.img[a][jj];
        // virial W = -sum r_ij (x) D_ij ; order xx yy zz yz xz xy
        w6[0] -= ud.x * Dx;
        w6[1] -= ud.y * Dy;
        w6[2] -= ud.z * Dz;
        w6[3] -= ud.y * Dz;
        w6[4] -= ud.x * Dz;
        w6[5] -= ud.x * Dy;
    }
    if (valid && l == 0) scnt[i] = ns;

    // ---- 4. deterministic reduction: group lanes, then atoms of the block ----
    double vals[7] = {e_lane, w6[0], w6[1], w6[2], w6[3], w6[4], w6[5]};
    for (int q = 0; q < 7; ++q) {
        double v = vals[q];
        for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
        vals[q] = v;
    }
    if (l == 0)
        for (int q = 0; q < 7; ++q) sh.red[a][q] = valid ? vals[q] : 0.0;
    __syncthreads();
    if (threadIdx.x < 7) {
        double s = 0.0;
        for (int t = 0; t < APB; ++t) s += sh.red[t][threadIdx.x];
        partial[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
    }
}

// F_i = sum_j (D_ij - D_ji); D_ji found in j's short list at (i, -o).
__global__ __launch_bounds__(256) void k_gather(int n, const int *__restrict__ sj,
                                                const int *__restrict__ simg,
                                                const int *__restrict__ scnt,
                                                const float4 *__restrict__ dU,
                                                double4 *__restrict__ frc,
                                                double4 *__restrict__ mom, double half_dt,
                                                int kick, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ns = scnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    bool miss = false;
    for (int s = 0; s < ns; ++s) {
        size_t o = (size_t)s * n + i;
        int j = sj[o];
        int rim = neg_img(simg[o]);
        float4 d = dU[o];
        int nj = scnt[j];
        float4 dr = make_float4(0, 0, 0, 0);
        bool found = false;
        for (int t = 0; t < nj; ++t) {
            size_t q = (size_t)t * n + j;
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
    if (miss) atomicOr(flags + 3, 1);
    frc[i] = make_double4(fx, fy, fz, 0.0);
    if (kick) {
        double4 p = mom[i];
        p.x += half_dt * fx;
        p.y += half_dt * fy;
        p.z += half_dt * fz;
        mom[i] = p;
    }
}

__global__ void k_reduce_partials(int nb, const double *__restrict__ partial,
                                  double *__restrict__ totals)
{
    __shared__ double lds[256];
    for (int q = 0; q < 7; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) s += partial[(size_t)b * 8 + q];
        lds[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[q] = lds[0];
        __syncthreads();
    }
}

struct Variant {
    const char *name;
    int G, SCAP;
};
const Variant kVariants[] = {{"g4_s16", 4, 16}, {"g8_s16", 8, 16}, {"g16_s32", 16, 32},
                             {"g2_s16", 2, 16}, {"g8_s32", 8, 32}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

template <int G, int SCAP>
void launch_site(System &S)
{
    constexpr int APB = BLOCK / G;
    int nb = div_up(S.n, APB);
    S.partial.ensure((size_t)nb * 8);
    k_site<G, SCAP><<<nb, BLOCK, 0, S.stream>>>(S.n, S.geom, S.pos.p, S.id.p, S.nbr.p, S.nimg.p,
                                               S.ncnt.p, S.dmodel.p, S.sj.p, S.simg.p, S.scnt.p,
                                               S.dU.p, S.partial.p, S.flags.p);
}

}  // namespace

int ts_variant_count() { return kNumVariants; }
const char *ts_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kVariants[v].name : ""; }

void ts_site(System &S, int variant)
{
    PYSIC_REQUIRE(variant >= 0 && variant < kNumVariants, PYSIC_EINVAL, "bad Tersoff kernel variant");
    S.scap = kVariants[variant].SCAP;
    S.sj.ensure((size_t)S.scap * S.n);
    S.simg.ensure((size_t)S.scap * S.n);
    S.scnt.ensure(S.n);
    S.dU.ensure((size_t)S.scap * S.n);
    hipEvent_t ev;
    S.timer.begin("tersoff_site", S.stream, &ev);
    switch (variant) {
    case 0: launch_site<4, 16>(S); break;
    case 1: launch_site<8, 16>(S); break;
    case 2: launch_site<16, 32>(S); break;
    case 3: launch_site<2, 16>(S); break;
    case 4: launch_site<8, 32>(S); break;
    }
    HIP_TRY(hipGetLastError());
    S.timer.end("tersoff_site", S.stream, ev);
    S.partial_blocks = div_up(S.n, BLOCK / kVariants[variant].G);
}

void ts_gather(System &S, double half_dt, int kick)
{
    hipEvent_t ev;
    S.timer.begin("tersoff_gather", S.stream, &ev);
    k_gather<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, S.sj.p, S.simg.p, S.scnt.p, S.dU.p,
                                                     S.frc.p, S.mom.p, half_dt, kick, S.flags.p);
    HIP_TRY(hipGetLastError());
    S.timer.end("tersoff_gather", S.stream, ev);
}

void ts_reduce(System &S)
{
    S.totals.ensure(8);
    k_reduce_partials<<<1, 256, 0, S.stream>>>(S.partial_blocks, S.partial.p, S.totals.p);
    HIP_TRY(hipGetLastError());
}
