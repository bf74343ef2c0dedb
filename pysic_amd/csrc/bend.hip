// bend.hip -- COMB's bond-bending angle term (bond_bend).
//
// The reference finds angles while it walks the pairs (Core.f90:3567-3610,
// 3760-3800): for each pair (atom1, atom2) taken once by pick, first the
// triplets (atom2, atom1, atom3) over atom1's neighbors with
// pick(atom2, atom3), then (atom1, atom2, atom3) over atom2's neighbors with
// pick(atom1, atom3); each is evaluated by the many-body potentials whose
// targets match the atoms in that order (core_evaluate_local_triplet_B,
// Core.f90:6656-7028).  So every angle a-v-b at a vertex v is visited once,
// oriented so that pick(a, b) holds: from the pair (v, a) with v as atom 1
// when pick(v, a) holds (targets matched as (v, a, b)), else from (a, v)
// (targets matched as (a, v, b)).
//
// Here one thread owns one vertex v.  Its neighbors within the largest bend
// cutoff form a short list; every unordered pair of the list is oriented and
// matched against the rows as the reference does it.  The forces an angle
// puts on its two outer atoms are stored per short-list slot and summed by
// those atoms' own threads afterwards (no atomics, fixed order, bit-for-bit
// repeatable).  All arithmetic is FP64.
#include "system.h"

namespace {

constexpr int BB = 256;

// x^n for an integer n by repeated products (Fortran's real**integer)
__device__ __forceinline__ double bend_ipow(double x, int n)
{
    const int m = n < 0 ? -n : n;
    double y = 1.0;
    for (int k = 0; k < m; ++k) y = y * x;
    return n < 0 ? 1.0 / y : y;
}

// One neighbor of the vertex: index, image, species, original index,
// r_j - r_v and its length.
struct Leg {
    int j, img, s, id;
    double3 d;
    double r;
};

__device__ __forceinline__ Leg load_leg(const double4 &pv, int j, int img,
                                        const double4 *__restrict__ pos,
                                        const int *__restrict__ id, const double *cv)
{
    Leg L;
    const double4 pj = pos[j];
    L.j = j;
    L.img = img;
    L.s = (int)pj.w;
    L.id = id[j];
    L.d = separation64(pv, pj, img, cv);
    L.r = sqrt(dot64(L.d));
    return L;
}

// The angle a-v-b of one row (evaluate_energy_bond_bending and
// evaluate_force_bond_bending, Potentials.f90:3196-3328), with
// s1 = r_v - r_a, s2 = r_b - r_v: E = eps (cos^n - c0)^m and, with FORCE,
// the forces f[0..2] on a, v, b.
template <bool FORCE>
__device__ __forceinline__ double bend_term(const BendTerm &P, const double3 &s1, double r1,
                                            const double3 &s2, double r2, double3 *f)
{
    const double3 t1 = make_double3(-s1.x, -s1.y, -s1.z);
    const double r3 = t1.x * s2.x + t1.y * s2.y + t1.z * s2.z;
    const double e = P.eps * bend_ipow(bend_ipow(r3 / (r1 * r2), P.n) - P.c0, P.m);
    if (FORCE) {
        const double k1 = r3 / (r1 * r1), k2 = r3 / (r2 * r2);
        const double3 t3 = make_double3(s2.x - k1 * t1.x, s2.y - k1 * t1.y, s2.z - k1 * t1.z);
        const double3 t4 = make_double3(t1.x - k2 * s2.x, t1.y - k2 * s2.y, t1.z - k2 * s2.z);
        const double ratio = 1.0 / (r1 * r2);
        const double cosplain = r3 * ratio;
        const double cospower = bend_ipow(cosplain, P.n - 1);
        const double r6 = P.eps * (double)(P.m * P.n) *
                          bend_ipow(cospower * cosplain - P.c0, P.m - 1) * cospower;
        f[0] = make_double3(-t3.x * ratio * r6, -t3.y * ratio * r6, -t3.z * ratio * r6);
        f[1] = make_double3((t3.x + t4.x) * ratio * r6, (t3.y + t4.y) * ratio * r6,
                            (t3.z + t4.z) * ratio * r6);
        f[2] = make_double3(-t4.x * ratio * r6, -t4.y * ratio * r6, -t4.z * ratio * r6);
    }
    return e;
}

__device__ __forceinline__ void add3(double4 *p, const double3 &f)
{
    double4 g = *p;
    g.x += f.x;
    g.y += f.y;
    g.z += f.z;
    *p = g;
}

template <class Geo, int MODE>
__global__ __launch_bounds__(BB) void k_bend_site(int n, Geo geo, const double4 *__restrict__ pos,
                                                  const int *__restrict__ id,
                                                  const int *__restrict__ nbr,
                                                  const int *__restrict__ nimg,
                                                  const int *__restrict__ ncnt,
                                                  const PairModel *__restrict__ PM,
                                                  int *__restrict__ bsj, int *__restrict__ bsimg,
                                                  int *__restrict__ bscnt,
                                                  double4 *__restrict__ bG,
                                                  double4 *__restrict__ bF,
                                                  double *__restrict__ partial, int blk0)
{
    constexpr bool FORCE = MODE != PAIR_ENERGY;
    constexpr bool VIR = MODE == PAIR_FORCES;
    __shared__ double red[7][BB];
    const int blk = xcd_block(gridDim.x);
    const int v = blk * BB + threadIdx.x;
    double e = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
    if (v < n) {
        const double4 pv = pos[v];
        const int sv = (int)pv.w, idv = id[v];
        const double *cv = geo.vec(v);
        const size_t N = (size_t)n;
        // the short list; its radius is widened by 1e-10 relative so that a
        // pair inside a cutoff seen from either end is in both atoms' lists
        const double bc2 = PM->bend_cut * PM->bend_cut * (1.0 + 1e-10);
        const double vmax = PM->bend_maxpot[sv];
        const int nf = ncnt[v];
        int c = 0;
        for (int t = 0; t < nf; ++t) {
            const int j = nbr[(size_t)t * N + v], img = nimg[(size_t)t * N + v];
            const double3 d = separation64(pv, pos[j], img, cv);
            if (!(dot64(d) < bc2)) continue;
            bsj[(size_t)c * N + v] = j;
            bsimg[(size_t)c * N + v] = img;
            if (FORCE) bG[(size_t)c * N + v] = make_double4(0.0, 0.0, 0.0, 0.0);
            ++c;
        }
        bscnt[v] = c;
        double3 fv = make_double3(0.0, 0.0, 0.0);
        const int nbend = PM->nbend;
        for (int ta = 0; ta < c; ++ta) {
            const Leg A = load_leg(pv, bsj[(size_t)ta * N + v], bsimg[(size_t)ta * N + v], pos, id,
                                   cv);
            for (int tb = ta + 1; tb < c; ++tb) {
                const Leg B = load_leg(pv, bsj[(size_t)tb * N + v], bsimg[(size_t)tb * N + v],
                                       pos, id, cv);
                // the reference's orientation: first leg a with pick(a, b)
                const int dimg = pack_img(img_x(B.img) - img_x(A.img), img_y(B.img) - img_y(A.img),
                                          img_z(B.img) - img_z(A.img));
                const bool ab = pick_dev(A.id, B.id, dimg);
                const Leg &Fl = ab ? A : B;
                const Leg &Ll = ab ? B : A;
                const int tf_slot = ab ? ta : tb, tl_slot = ab ? tb : ta;
                // from the pair (v, a) [v atom 1] or (a, v) [a atom 1]
                const bool vfirst = pick_dev(idv, Fl.id, Fl.img);
                const int e0 = vfirst ? sv : Fl.s, e1 = vfirst ? Fl.s : sv;
                // the second leg is searched within atom 1's radius, which
                // for (a, v) is the vertex's own largest potential cutoff
                if (!vfirst && !(Ll.r < vmax)) continue;
                const double3 s1 = make_double3(-Fl.d.x, -Fl.d.y, -Fl.d.z), s2 = Ll.d;
                const double r1 = Fl.r, r2 = Ll.r;
                for (int k = 0; k < nbend; ++k) {
                    const BendTerm &P = PM->bend[k];
                    if (P.e[0] != e0 || P.e[1] != e1 || P.e[2] != Ll.s) continue;
                    if (!(P.cut > r1) || !(P.cut > r2) || P.vertex != sv) continue;
                    double fc1 = 1.0, fc2 = 1.0;
                    if (P.smooth) {
                        fc1 = fc64(r1, P.cut, P.soft);
                        fc2 = fc64(r2, P.cut, P.soft);
                    }
                    if (!FORCE) {
                        const double ev = bend_term<false>(P, s1, r1, s2, r2, nullptr);
                        e += ev * fc1 * fc2 * (1.0 + 1.0 + 1.0) / 3.0;
                        continue;
                    }
                    double3 f[3];
                    const double ev = bend_term<true>(P, s1, r1, s2, r2, f);
                    double3 g1 = make_double3(0.0, 0.0, 0.0), g2 = g1;
                    double es = 0.0;
                    if (P.smooth) {
                        es = ev;
                        const double d1 = dfc64(r1, P.cut, P.soft) / r1;
                        const double d2 = r2 == 0.0 ? 0.0 : dfc64(r2, P.cut, P.soft) / r2;
                        g1 = make_double3(s1.x * d1, s1.y * d1, s1.z * d1);
                        g2 = make_double3(s2.x * d2, s2.y * d2, s2.z * d2);
                    }
                    const double ff = fc1 * fc2;
                    const double3 t0 = make_double3(f[0].x * ff + g1.x * fc2 * es,
                                                    f[0].y * ff + g1.y * fc2 * es,
                                                    f[0].z * ff + g1.z * fc2 * es);
                    const double3 t1 = make_double3(
                        f[1].x * ff + (-g1.x * fc2 + g2.x * fc1) * es,
                        f[1].y * ff + (-g1.y * fc2 + g2.y * fc1) * es,
                        f[1].z * ff + (-g1.z * fc2 + g2.z * fc1) * es);
                    const double3 t2 = make_double3(f[2].x * ff - g2.x * fc1 * es,
                                                    f[2].y * ff - g2.y * fc1 * es,
                                                    f[2].z * ff - g2.z * fc1 * es);
                    fv.x += t1.x;
                    fv.y += t1.y;
                    fv.z += t1.z;
                    add3(bG + (size_t)tf_slot * N + v, t0);
                    add3(bG + (size_t)tl_slot * N + v, t2);
                    if (VIR) {
                        // sum of r (x) f over the three atoms, relative to a
                        const double3 s12 = make_double3(s1.x + s2.x, s1.y + s2.y, s1.z + s2.z);
                        w0 += s1.x * t1.x + s12.x * t2.x;
                        w1 += s1.y * t1.y + s12.y * t2.y;
                        w2 += s1.z * t1.z + s12.z * t2.z;
                        w3 += s1.y * t1.z + s12.y * t2.z;
                        w4 += s1.x * t1.z + s12.x * t2.z;
                        w5 += s1.x * t1.y + s12.x * t2.y;
                    }
                }
            }
        }
        if (FORCE) bF[v] = make_double4(fv.x, fv.y, fv.z, 0.0);
    }
    if (MODE == PAIR_FORCES_ONLY) return;
    const int tid = threadIdx.x;
    red[0][tid] = e;
    red[1][tid] = w0; red[2][tid] = w1; red[3][tid] = w2;
    red[4][tid] = w3; red[5][tid] = w4; red[6][tid] = w5;
    __syncthreads();
    for (int s = BB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int q = 0; q < 7; ++q) red[q][tid] += red[q][tid + s];
        __syncthreads();
    }
    if (tid < 7) partial[(size_t)(blk0 + blk) * 8 + tid] = red[tid][0];
}

// F_a += (force of a's own angles on a) + sum over a's short-list neighbors
// v of the force v's angles put on a, found in v's short list at (a, -o).
__global__ __launch_bounds__(BB) void k_bend_gather(int n, const int *__restrict__ bsj,
                                                    const int *__restrict__ bsimg,
                                                    const int *__restrict__ bscnt,
                                                    const double4 *__restrict__ bG,
                                                    const double4 *__restrict__ bF,
                                                    double4 *__restrict__ frc,
                                                    double4 *__restrict__ mom, double half_dt,
                                                    int kick)
{
    const int a = xcd_block(gridDim.x) * BB + threadIdx.x;
    if (a >= n) return;
    const size_t N = (size_t)n;
    const double4 own = bF[a];
    double fx = own.x, fy = own.y, fz = own.z;
    const int c = bscnt[a];
    for (int t = 0; t < c; ++t) {
        const int v = bsj[(size_t)t * N + a];
        const int want = neg_img(bsimg[(size_t)t * N + a]);
        const int cv = bscnt[v];
        for (int u = 0; u < cv; ++u) {
            if (bsj[(size_t)u * N + v] == a && bsimg[(size_t)u * N + v] == want) {
                const double4 g = bG[(size_t)u * N + v];
                fx += g.x;
                fy += g.y;
                fz += g.z;
                break;
            }
        }
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

int bend_blocks(const System &S) { return S.pmodel.nbend > 0 ? div_up(S.n, BB) : 0; }

void bend_eval(System &S, int mode, int blk0, double half_dt, int kick)
{
    if (S.pmodel.nbend == 0) return;
    const int n = S.n, nb = div_up(n, BB);
    const size_t cap = (size_t)S.cap * n;
    S.bsj.ensure(cap);
    S.bsimg.ensure(cap);
    S.bscnt.ensure(n);
    S.bG.ensure(cap);
    S.bF.ensure(n);
    hipEvent_t ev{};
    S.timer.begin("bend", S.stream, &ev);
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        auto go = [&](auto kern) {
            kern<<<nb, BB, 0, S.stream>>>(n, geo, S.pos.p, S.id.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                          S.dpmodel.p, S.bsj.p, S.bsimg.p, S.bscnt.p, S.bG.p,
                                          S.bF.p, S.gpartial.p, blk0);
        };
        if (mode == PAIR_ENERGY) go(k_bend_site<Geo, PAIR_ENERGY>);
        else if (mode == PAIR_FORCES) go(k_bend_site<Geo, PAIR_FORCES>);
        else go(k_bend_site<Geo, PAIR_FORCES_ONLY>);
    });
    HIP_TRY(hipGetLastError());
    if (mode != PAIR_ENERGY) {
        k_bend_gather<<<nb, BB, 0, S.stream>>>(n, S.bsj.p, S.bsimg.p, S.bscnt.p, S.bG.p, S.bF.p,
                                               S.frc.p, S.mom.p, half_dt, kick);
        HIP_TRY(hipGetLastError());
    }
    S.timer.end("bend", S.stream, ev);
}
