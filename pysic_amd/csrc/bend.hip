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
// Here each vertex v gets a short list of its neighbors within the largest
// bend cutoff, then G lanes: each lane takes some of the legs and, for each,
// every angle it forms with another leg, oriented and matched against the
// rows as the reference does it.  The force on a leg's atom is stored per
// short-list slot and summed by that atom's own thread afterwards, with the
// vertex's own (minus the sum over its slots): no atomics, fixed order,
// bit-for-bit repeatable.  All arithmetic is FP64.
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

// The short list of each vertex: its neighbors within the largest bend
// cutoff, widened by 1e-10 relative so that a pair inside a cutoff seen from
// either end is in both atoms' lists.
template <class Geo>
__global__ __launch_bounds__(BB) void k_bend_list(int n, Geo geo, const double4 *__restrict__ pos,
                                                  const int *__restrict__ nbr,
                                                  const int *__restrict__ nimg,
                                                  const int *__restrict__ ncnt,
                                                  const PairModel *__restrict__ PM,
                                                  int *__restrict__ bsj, int *__restrict__ bsimg,
                                                  int *__restrict__ bscnt)
{
    const int v = xcd_block(gridDim.x) * BB + threadIdx.x;
    if (v >= n) return;
    const double4 pv = pos[v];
    const double *cv = geo.vec(v);
    const size_t N = (size_t)n;
    const double bc2 = PM->bend_cut * PM->bend_cut * (1.0 + 1e-10);
    const int nf = ncnt[v];
    int c = 0;
    for (int t = 0; t < nf; ++t) {
        const int j = nbr[(size_t)t * N + v], img = nimg[(size_t)t * N + v];
        const double3 d = separation64(pv, pos[j], img, cv);
        if (!(dot64(d) < bc2)) continue;
        bsj[(size_t)c * N + v] = j;
        bsimg[(size_t)c * N + v] = img;
        ++c;
    }
    bscnt[v] = c;
}

// G lanes per vertex v; lane l takes the short-list legs ta = l, l + G, ...
// and every angle (ta, tb), tb != ta, so each angle is evaluated by the two
// lanes of its legs.  A lane keeps only its own leg's share: the force on
// that atom (stored per slot), half the energy, and d_ta (x) F_ta of the
// virial (the three forces of an angle sum to zero, so sum r (x) F over its
// atoms is d_a (x) F_a + d_b (x) F_b, taking r_v as the origin).
template <class Geo, int MODE, int G, int H>
__global__ __launch_bounds__(BB) void k_bend_angles(int n, Geo geo, const double4 *__restrict__ pos,
                                                    const int *__restrict__ id,
                                                    const PairModel *__restrict__ PM,
                                                    const int *__restrict__ bsj,
                                                    const int *__restrict__ bsimg,
                                                    const int *__restrict__ bscnt,
                                                    double4 *__restrict__ bG,
                                                    double *__restrict__ partial, int blk0)
{
    constexpr bool FORCE = MODE != PAIR_ENERGY;
    constexpr bool VIR = MODE == PAIR_FORCES;
    __shared__ double red[7][BB];
    const int blk = xcd_block(gridDim.x);
    const int gid = blk * BB + threadIdx.x;
    // G * H consecutive lanes per vertex: lane l + G h takes the legs
    // l, l + G, ... and, of each, the partner legs h, h + H, ...
    const int v = gid / (G * H), lane = gid % G, part = (gid / G) % H;
    double e = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
    if (v < n) {
        const double4 pv = pos[v];
        const int sv = (int)pv.w, idv = id[v];
        const double *cv = geo.vec(v);
        const size_t N = (size_t)n;
        const double vmax = PM->bend_maxpot[sv];
        const int c = bscnt[v];
        const int nbend = PM->nbend;
        for (int ta = lane; ta < c; ta += G) {
            const Leg A = load_leg(pv, bsj[(size_t)ta * N + v], bsimg[(size_t)ta * N + v], pos, id,
                                   cv);
            double3 fa = make_double3(0.0, 0.0, 0.0);
            for (int tb = part; tb < c; tb += H) {
                if (tb == ta) continue;
                const Leg B = load_leg(pv, bsj[(size_t)tb * N + v], bsimg[(size_t)tb * N + v],
                                       pos, id, cv);
                // the reference's orientation: first leg with pick(first, last)
                const bool ab = pick_dev(A.id, B.id,
                                         pack_img(img_x(B.img) - img_x(A.img),
                                                  img_y(B.img) - img_y(A.img),
                                                  img_z(B.img) - img_z(A.img)));
                const Leg &Fl = ab ? A : B;
                const Leg &Ll = ab ? B : A;
                // from the pair (v, first) [v atom 1] or (first, v)
                const bool vfirst = pick_dev(idv, Fl.id, Fl.img);
                const int e0 = vfirst ? sv : Fl.s, e1 = vfirst ? Fl.s : sv;
                // the second leg is searched within atom 1's radius, which
                // for (first, v) is the vertex's own largest potential cutoff
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
                        e += 0.5 * (ev * fc1 * fc2 * (1.0 + 1.0 + 1.0) / 3.0);
                        continue;
                    }
                    double3 f[3];
                    const double ev = bend_term<true>(P, s1, r1, s2, r2, f);
                    const double ff = fc1 * fc2;
                    double3 t;
                    if (ab) {  // this lane's leg is the first: f[0] and fc(r1)'s gradient
                        double3 g1 = make_double3(0.0, 0.0, 0.0);
                        if (P.smooth) {
                            const double d1 = dfc64(r1, P.cut, P.soft) / r1;
                            g1 = make_double3(s1.x * d1 * fc2 * ev, s1.y * d1 * fc2 * ev,
                                              s1.z * d1 * fc2 * ev);
                        }
                        t = make_double3(f[0].x * ff + g1.x, f[0].y * ff + g1.y,
                                         f[0].z * ff + g1.z);
                    } else {  // the last: f[2] and fc(r2)'s gradient
                        double3 g2 = make_double3(0.0, 0.0, 0.0);
                        if (P.smooth && r2 != 0.0) {
                            const double d2 = dfc64(r2, P.cut, P.soft) / r2;
                            g2 = make_double3(s2.x * d2 * fc1 * ev, s2.y * d2 * fc1 * ev,
                                              s2.z * d2 * fc1 * ev);
                        }
                        t = make_double3(f[2].x * ff - g2.x, f[2].y * ff - g2.y,
                                         f[2].z * ff - g2.z);
                    }
                    fa.x += t.x;
                    fa.y += t.y;
                    fa.z += t.z;
                    if (VIR) {
                        const double3 d = A.d;
                        w0 += d.x * t.x;
                        w1 += d.y * t.y;
                        w2 += d.z * t.z;
                        w3 += d.y * t.z;
                        w4 += d.x * t.z;
                        w5 += d.x * t.y;
                    }
                }
            }
            if (FORCE) {
                // the H parts of one leg run the same ta loop, so they are
                // active together here
#pragma unroll
                for (int o = G; o < G * H; o <<= 1) {
                    fa.x += __shfl_xor(fa.x, o);
                    fa.y += __shfl_xor(fa.y, o);
                    fa.z += __shfl_xor(fa.z, o);
                }
                if (part == 0) bG[(size_t)ta * N + v] = make_double4(fa.x, fa.y, fa.z, 0.0);
            }
        }
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

// F_a += -(the forces of a's own angles on their outer atoms) + the forces
// a's short-list neighbors' angles put on a (found in each neighbor's short
// list at (a, -o)).  G lanes per atom split a's slots; shuffles sum them.
template <int G>
__global__ __launch_bounds__(BB) void k_bend_gather(int n, const int *__restrict__ bsj,
                                                    const int *__restrict__ bsimg,
                                                    const int *__restrict__ bscnt,
                                                    const double4 *__restrict__ bG,
                                                    double4 *__restrict__ frc,
                                                    double4 *__restrict__ mom, double half_dt,
                                                    int kick)
{
    const int gid = xcd_block(gridDim.x) * BB + threadIdx.x;
    const int a = gid / G, lane = gid % G;
    const size_t N = (size_t)n;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (a < n) {
        const int c = bscnt[a];
        for (int t = lane; t < c; t += G) {
            const double4 own = bG[(size_t)t * N + a];
            fx -= own.x;
            fy -= own.y;
            fz -= own.z;
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
    }
    // the lanes of one atom are consecutive threads of one wave; every
    // thread reaches the shuffles
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o);
        fy += __shfl_xor(fy, o);
        fz += __shfl_xor(fz, o);
    }
    if (a >= n || lane != 0) return;
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

// lanes per vertex: 16 covers the legs of a 3 A cutoff in oxides with one
// leg per lane (24,576 atoms of COMB SiO2: angles + gather 337 us with 4
// lanes, 247 with 8, 196 with 16)
constexpr int BG_LANES = 16;
// and each leg's partners split over BG_PARTS lanes
constexpr int BG_PARTS = 2;
// lanes per atom in the gather: its slots split over 8 lanes
constexpr int BG_GATHER_LANES = 8;

int bend_blocks(const System &S)
{
    return S.pmodel.nbend > 0 ? div_up(S.n * BG_LANES * BG_PARTS, BB) : 0;
}

void bend_eval(System &S, int mode, int blk0, double half_dt, int kick)
{
    if (S.pmodel.nbend == 0) return;
    const int n = S.n, nb = div_up(n, BB), na = bend_blocks(S);
    const size_t cap = (size_t)S.cap * n;
    S.bsj.ensure(cap);
    S.bsimg.ensure(cap);
    S.bscnt.ensure(n);
    S.bG.ensure(cap);
    hipEvent_t ev{};
    S.timer.begin("bend", S.stream, &ev);
    with_geo(S, [&](auto geo) {
        using Geo = decltype(geo);
        k_bend_list<Geo><<<nb, BB, 0, S.stream>>>(n, geo, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                                  S.dpmodel.p, S.bsj.p, S.bsimg.p, S.bscnt.p);
        auto go = [&](auto kern) {
            kern<<<na, BB, 0, S.stream>>>(n, geo, S.pos.p, S.id.p, S.dpmodel.p, S.bsj.p,
                                          S.bsimg.p, S.bscnt.p, S.bG.p, S.gpartial.p, blk0);
        };
        if (mode == PAIR_ENERGY) go(k_bend_angles<Geo, PAIR_ENERGY, BG_LANES, BG_PARTS>);
        else if (mode == PAIR_FORCES) go(k_bend_angles<Geo, PAIR_FORCES, BG_LANES, BG_PARTS>);
        else go(k_bend_angles<Geo, PAIR_FORCES_ONLY, BG_LANES, BG_PARTS>);
    });
    HIP_TRY(hipGetLastError());
    if (mode != PAIR_ENERGY) {
        k_bend_gather<BG_GATHER_LANES><<<div_up(n * BG_GATHER_LANES, BB), BB, 0, S.stream>>>(
            n, S.bsj.p, S.bsimg.p, S.bscnt.p, S.bG.p, S.frc.p, S.mom.p, half_dt, kick);
        HIP_TRY(hipGetLastError());
    }
    S.timer.end("bend", S.stream, ev);
}
