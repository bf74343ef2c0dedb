// capi.hip -- the C-ABI of libpysic_hip.so (declared in include/pysic_hip.h).
//
// One pysic_ctx owns one System (device buffers in cell-sorted order) and the
// host-side copy of the interaction model.  Calls mirror the f2py entry
// points of fortran/PyInterface.f90; errors are returned as codes with a
// message, never by exiting the process (the reference stops or exits,
// Core.f90:4909-4911, MPI.f90:58-70).
#include "system.h"
#include "ewald.h"
#include "learned.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/pysic_hip.h"
#include "../../include/pysic_learned.h"
#include "../../include/pysic_batch.h"

#include <mutex>

static_assert(PYSIC_CORRECTOR_NCOEF == CORR_NC, "corrector coefficient count");
static_assert(PYSIC_CORRECTOR_MAX_STAGES == CORR_MAX_STAGES, "corrector stage limit");

// ---------------------------------------------------------------------------
// Fixed device scratch of pysic_calculate_energy_noalloc: the per-block
// partial sums of the energy pass (8 doubles per site-kernel tile) and the
// totals row.  Each GPU holds its own copy for the life of the process; host
// code uses the address hipGetSymbolAddress returns on that device (the
// array's name is only a host-side stand-in), looked up once per device when
// the first handle on it is created.
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoallocTiles = PYSIC_NOALLOC_MAX_TILES;
constexpr size_t kScratchDoubles = (size_t)kNoallocTiles * 8 + 8;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
// with `device` current
void scratch_setup(int device)
{
    PYSIC_REQUIRE(device < kMaxDevices, PYSIC_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    PYSIC_REQUIRE(rc == hipSuccess && p, PYSIC_EHIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    g_scratch[device] = (double *)p;
}

double *scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}
}  // namespace

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------
void KernelTimer::begin(const char *name, hipStream_t s, hipEvent_t *a)
{
    if (!on) return;
    HIP_TRY(hipEventCreate(a));
    HIP_TRY(hipEventRecord(*a, s));
    (void)name;
}
void KernelTimer::end(const char *name, hipStream_t s, hipEvent_t a)
{
    if (!on) return;
    hipEvent_t b;
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(b, s));
    pending.push_back({name, a, b});
    if (pending.size() > 4096) collect();
}
void KernelTimer::collect()
{
    for (auto &r : pending) {
        HIP_TRY(hipEventSynchronize(r.b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        bool found = false;
        for (size_t k = 0; k < total.size(); ++k)
            if (total[k].first == r.name) {
                total[k].second += ms;
                calls[k].second += 1;
                found = true;
            }
        if (!found) {
            total.push_back({r.name, ms});
            calls.push_back({r.name, 1});
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pending.clear();
}
void KernelTimer::reset()
{
    collect();
    total.clear();
    calls.clear();
}

void System::release_all()
{
    DBuf<double4> *b4[] = {&pos, &pos2, &mom, &mom2, &frc, &frc2, &out4};
    for (auto *b : b4) b->release();
    DBuf<double> *bd[] = {&q, &q2, &cut, &cut2, &partial, &totals, &host_stage, &gpartial,
                          &gtotals, &atom_ev, &cell_ev};
    for (auto *b : bd) b->release();
    DBuf<int> *bi[] = {&id, &id2, &key, &key2, &wrap, &wrap2, &perm, &cellcount, &cellstart, &cellfill, &scanblk,
                       &prevstart, &dirty, &scanloc, &srow,
                       &nbr, &nimg, &ncnt, &flags, &sj, &simg, &scnt, &cell_of, &first_bin, &atom0,
                       &bsj, &bsimg, &bscnt, &osj, &osimg, &oscnt, &qc_check};
    for (auto *b : bi) b->release();
    bG.release();
    oD.release();
    oW.release();
    oown.release();
    dbmodel.release();
    dU.release();
    wpos.release();
    dmodel.release();
    dpmodel.release();
    tables.release();
    cells.release();
    order_valid = false;
    bins_clean = false;
}

// ---------------------------------------------------------------------------
// order-restoring helpers
// ---------------------------------------------------------------------------
namespace {
__global__ void k_gather4(int n, const int *__restrict__ id, const double4 *__restrict__ src,
                          double4 *__restrict__ dst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[s] = src[id[s]];
}
__global__ void k_gather1(int n, const int *__restrict__ id, const double *__restrict__ src,
                          double *__restrict__ dst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[s] = src[id[s]];
}
__global__ void k_scatter1(int n, const int *__restrict__ id, const double *__restrict__ src,
                           double *__restrict__ dst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[id[s]] = src[s];
}
__global__ void k_iota(int n, int *__restrict__ id)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) id[s] = s;
}
__global__ void k_zero4(int n, double4 *__restrict__ a)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) a[s] = make_double4(0, 0, 0, 0);
}
__global__ void k_sum_energy(const double *__restrict__ t1, const double *__restrict__ t2,
                             double *__restrict__ dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = (t1 ? t1[0] : 0.0) + (t2 ? t2[0] : 0.0);
}
}  // namespace

struct pysic_ctx {
    int device = 0;
    System S;
    std::string err;
    std::vector<int> species;     // original order
    std::vector<double> masses;   // original order
    std::vector<double> potdesc, bonddesc;
    std::map<int, std::vector<double>> tables;  // 'tabulated' tables by id: rows of (value, derivative)
    int variant = 0;
    bool atoms_ready = false, cell_ready = false;
    bool lists_valid = false;
    bool forces_valid = false;
    EwaldParams ewald;
    EwaldState ew;
    double min_cut = 0.0;         // smallest per-atom list radius of the current lists
    DBuf<double> rates;           // charge rates of relax_charges
    // captured device evaluations of a batch of cells, one per position
    // buffer (the sort swaps buffers, so consecutive evaluations alternate)
    struct Captured {
        hipGraphExec_t exec = nullptr;
        std::vector<const void *> sig;   // buffers and settings it was captured with
    } graph[2], mdgraph[2];
};

namespace {

int fail(pysic_ctx *h, const PysicError &e)
{
    if (h) h->err = e.msg;
    return e.code;
}

#define API_BEGIN(h)                                                                         \
    if (!(h)) return PYSIC_EINVAL;                                                           \
    try {                                                                                    \
        HIP_TRY(hipSetDevice((h)->device));
#define API_END(h)                                                                           \
    }                                                                                        \
    catch (const PysicError &e) { return fail(h, e); }                                       \
    catch (const std::exception &e) { return fail(h, PysicError{PYSIC_ENOMEM, e.what()}); }  \
    return PYSIC_OK;

constexpr int POT_STRIDE = 12;  // doubles per potential descriptor row

void drop_graphs(pysic_ctx *h);

// Compile the descriptor rows into species tables (Core.f90:1034-1125 builds
// the per-atom lists the same way: position-1 element match).
void compile_model(pysic_ctx *h)
{
    Model M;
    std::memset(&M, 0, sizeof(M));
    int ns = 1;
    for (int s : h->species) ns = std::max(ns, s + 1);
    PYSIC_REQUIRE(ns <= PYSIC_MAXS, PYSIC_EINVAL, "at most 4 species per system");
    M.nspecies = ns;
    M.group = -1;
    BopModel B;
    std::memset(&B, 0, sizeof(B));
    auto use_group = [&](int g) {
        if (g < 0) return;
        PYSIC_REQUIRE(M.group < 0 || M.group == g, PYSIC_EINVAL,
                      "one bond-order group (coordinator) per system is supported");
        M.group = g;
    };
    PairModel P;
    std::memset(&P, 0, sizeof(P));
    std::vector<double2> htab;
    std::map<int, int> tab_off;
    const size_t np = h->potdesc.size() / POT_STRIDE;
    auto row = [&](size_t k) { return h->potdesc.data() + POT_STRIDE * k; };
    for (size_t k = 0; k < np; ++k) {
        const double *d = row(k);
        int type = (int)d[0], nt = (int)d[1], e1 = (int)d[2], e2 = (int)d[3];
        PYSIC_REQUIRE(d[11] == 0.0, PYSIC_EINVAL,
                      "a multiplier row must follow the potential it multiplies");
        // the factor rows of this potential (link 0, then its ProductPotential
        // multipliers, link 1), each charge factor followed by its
        // continuation row (link 2)
        size_t frow[PYSIC_MAXF + 1], crow[PYSIC_MAXF + 1];
        size_t nfac = 0, next = k;
        while (next < np && (next == k || row(next)[11] == 1.0)) {
            PYSIC_REQUIRE(nfac < PYSIC_MAXF, PYSIC_EINVAL, "at most 4 factors in a product");
            frow[nfac] = next;
            crow[nfac] = SIZE_MAX;
            ++next;
            if (next < np && row(next)[11] == 2.0) crow[nfac] = next++;
            const int ft = (int)row(frow[nfac])[0];
            PYSIC_REQUIRE((ft >= 12 && ft <= 15) == (crow[nfac] != SIZE_MAX) &&
                              (crow[nfac] == SIZE_MAX || (int)row(crow[nfac])[0] == ft),
                          PYSIC_EINVAL,
                          "a charge-dependent or bond_bend row needs its continuation row, and "
                          "only it");
            ++nfac;
        }
        PYSIC_REQUIRE(next >= np || row(next)[11] == 0.0, PYSIC_EINVAL,
                      "a continuation row must follow a charge-dependent or bond_bend row");
        const size_t nmul = nfac - 1;  // the ProductPotential multipliers
        PYSIC_REQUIRE(e1 >= 0 && e1 < PYSIC_MAXS && (nt < 2 || (e2 >= 0 && e2 < PYSIC_MAXS)),
                      PYSIC_EINVAL, "potential species out of range");
        PYSIC_REQUIRE(type >= 1 && type <= 15, PYSIC_EINVAL, "unsupported potential type");
        const bool single = type == 2 || type == 10 || type == 11;
        PYSIC_REQUIRE(nt == (single ? 1 : type == 15 ? 3 : 2), PYSIC_EINVAL,
                      single ? "one-body potential needs 1 target"
                             : type == 15 ? "bond_bend needs 3 targets"
                                          : "pair potential needs 2 targets");
        PYSIC_REQUIRE(!(single && nmul), PYSIC_EINVAL,
                      "products of one-body potentials are not supported");
        if (e1 < ns) M.maxpot[e1] = std::max(M.maxpot[e1], d[6]);
        const int group = (int)d[8];
        if (type == 1 && nmul == 0) {
            // the Tersoff kernels' pair term (plain or bond-order scaled)
            int &c = M.npair[e1][e2];
            PYSIC_REQUIRE(c < PYSIC_MAXP, PYSIC_EINVAL, "too many exponential terms for one pair");
            PairTerm &t = M.pair[e1][e2][c++];
            t.eps = d[4];
            t.zeta = d[5];
            t.cut = d[6];
            // soft-cutoff rule of create_potential (Potentials.f90:2382-2388)
            if (d[7] > 0.0 && d[7] < d[6]) { t.soft = d[7]; t.smooth = 1; }
            else { t.soft = d[6]; t.smooth = 0; }
            t.group = group;
            use_group(t.group);
            M.has_pairs = 1;
        } else if (type == 2) {
            int nexp = (int)d[5];
            PYSIC_REQUIRE(nexp >= 0 && nexp <= 4, PYSIC_EINVAL, "charge_self exponent must be 0..4");
            M.self_eps[e1][nexp] += d[4];
            M.has_self = 1;
        } else if (type == 15) {
            PYSIC_REQUIRE(nmul == 0 && group < 0, PYSIC_EINVAL,
                          "bond_bend takes no multipliers and no bond-order scaling");
            PYSIC_REQUIRE(P.nbend < PYSIC_MAXB, PYSIC_EINVAL,
                          "too many bond_bend rows (at most 12)");
            const double *c = row(crow[0]);
            const int e3 = (int)c[1];
            PYSIC_REQUIRE(e3 >= 0 && e3 < PYSIC_MAXS, PYSIC_EINVAL, "potential species out of range");
            BendTerm &B = P.bend[P.nbend++];
            B.e[0] = e1; B.e[1] = e2; B.e[2] = e3;
            B.vertex = (int)c[3];
            B.eps = d[4];
            B.n = (int)d[9];
            B.m = (int)d[10];
            // calculate_derived_parameters_bond_bending (Potentials.f90:3178-3185)
            B.c0 = std::pow(std::cos(d[5]), d[9]);
            B.cut = d[6];
            if (d[7] > 0.0 && d[7] < d[6]) { B.soft = d[7]; B.smooth = 1; }
            else { B.soft = d[6]; B.smooth = 0; }
            P.bend_cut = std::max(P.bend_cut, d[6]);
            P.has_pairs = 1;
        } else if (type == 10 || type == 11) {
            PYSIC_REQUIRE(group < 0, PYSIC_EINVAL,
                          "bond-order scaling of one-body potentials is not supported");
            if (type == 10) P.konst[e1] += d[4];
            else { P.kforce[e1][0] += d[4]; P.kforce[e1][1] += d[5]; P.kforce[e1][2] += d[9]; }
            P.has_single = 1;
        } else {
            PYSIC_REQUIRE(nmul + 1 <= PYSIC_MAXF, PYSIC_EINVAL, "at most 4 factors in a product");
            // a bond-order-scaled potential goes to the bond-order pass
            // (one per ordered species pair), any other to the pair kernel
            const bool scaled = group >= 0;
            if (scaled) {
                PYSIC_REQUIRE(group < PYSIC_MAXGRP, PYSIC_EINVAL,
                              "bond-order-scaled pair potentials take coordinator groups 0..3");
                PYSIC_REQUIRE(B.nbop[e1][e2] == 0, PYSIC_EINVAL,
                              "at most one bond-order-scaled potential per species pair");
            } else {
                PYSIC_REQUIRE(P.npair[e1][e2] < PYSIC_MAXG, PYSIC_EINVAL,
                              "too many pair potentials for one species pair (at most 8)");
            }
            GenPair &T = scaled ? B.pair[e1][e2] : P.pair[e1][e2][P.npair[e1][e2]++];
            if (scaled) {
                B.nbop[e1][e2] = 1;
                B.group[e1][e2] = group;
                B.on = 1;
                B.radius = std::max(B.radius, d[6]);
            }
            T.cut = d[6];
            if (d[7] > 0.0 && d[7] < d[6]) { T.soft = d[7]; T.smooth = 1; }
            else { T.soft = d[6]; T.smooth = 0; }
            T.nf = (int)nfac;
            for (int f = 0; f < T.nf; ++f) {
                const double *r = row(frow[f]);
                PairFactor &F = T.f[f];
                F.type = (int)r[0];
                const bool charged = F.type >= 12 && F.type <= 14;
                PYSIC_REQUIRE(F.type == 1 || (F.type >= 3 && F.type <= 9) || charged,
                              PYSIC_EINVAL, "a product takes pair potentials only");
                F.p[0] = r[4]; F.p[1] = r[5]; F.p[2] = r[9]; F.p[3] = r[10];
                if (charged) {
                    const double *c = row(crow[f]);
                    ChargeFactor Q;
                    std::memset(&Q, 0, sizeof(Q));
                    for (int a = 0; a < 4; ++a) Q.p[a] = F.p[a];
                    for (int a = 0; a < 9; ++a) Q.p[4 + a] = c[2 + a];
                    Q.orig = (int)c[1];
                    if (F.type == 12) {
                        // calculate_derived_parameters_charge_exp (Potentials.f90:3747-3766)
                        const double *q = Q.p;
                        Q.eta[0] = std::log(q[1] / (q[1] - q[2])) / std::log(q[3] / (q[3] - q[4]));
                        Q.eta[1] = std::log(q[5] / (q[5] - q[6])) / std::log(q[7] / (q[7] - q[8]));
                        Q.beta[0] = std::pow(q[2] - q[1], 1.0 / Q.eta[0]) / (q[3] - q[4]);
                        Q.beta[1] = std::pow(q[6] - q[5], 1.0 / Q.eta[1]) / (q[7] - q[8]);
                    }
                    // the permuted copies of a potential share one entry
                    int slot = 0;
                    while (slot < P.ncf && std::memcmp(&P.cf[slot], &Q, sizeof(Q)) != 0) ++slot;
                    if (slot == P.ncf) {
                        PYSIC_REQUIRE(P.ncf < PYSIC_MAXCF, PYSIC_EINVAL,
                                      "too many distinct charge-dependent pair factors "
                                      "(at most 16)");
                        P.cf[P.ncf++] = Q;
                    }
                    F.toff = slot;
                    P.has_charge = 1;
                    // the factor's value depends on which atom comes first
                    P.needs_pick = 1;
                }
                if (F.type == 7) {
                    const int id = (int)r[4];
                    auto it = h->tables.find(id);
                    PYSIC_REQUIRE(it != h->tables.end(), PYSIC_EINVAL,
                                  "no table " + std::to_string(id) +
                                      " for a 'tabulated' potential (pysic_set_potential_table)");
                    auto off = tab_off.find(id);
                    if (off == tab_off.end()) {
                        off = tab_off.emplace(id, (int)htab.size()).first;
                        const std::vector<double> &v = it->second;
                        for (size_t q = 0; q + 1 < v.size(); q += 2)
                            htab.push_back(make_double2(v[q], v[q + 1]));
                    }
                    F.toff = off->second;
                    F.tlen = (int)(it->second.size() / 2);
                }
                if (F.type == 9) P.needs_pick = 1;
                // the bond-order pass takes central pair forces only
                PYSIC_REQUIRE(!(scaled && F.type == 9), PYSIC_EINVAL,
                              "a bond-order-scaled potential cannot hold a 'step' factor");
            }
            if (!scaled) {
                const double c2 = T.cut * T.cut * (1.0 + 1e-12);
                P.cut2_hi[e1][e2] = std::max(P.cut2_hi[e1][e2], c2);
            }
            P.has_pairs = 1;
        }
        k = next - 1;
    }
    size_t nb = h->bonddesc.size() / 14;
    for (size_t k = 0; k < nb; ++k) {
        const double *d = h->bonddesc.data() + 14 * k;
        int e1 = (int)d[0], e2 = (int)d[1], e3 = (int)d[2];
        PYSIC_REQUIRE(e1 >= 0 && e1 < PYSIC_MAXS && e2 >= 0 && e2 < PYSIC_MAXS && e3 >= 0 &&
                          e3 < PYSIC_MAXS,
                      PYSIC_EINVAL, "bond-order species out of range");
        BondTerm b;
        std::memset(&b, 0, sizeof(b));
        b.beta = d[3]; b.eta = d[4]; b.a = d[5]; b.c = d[6]; b.d = d[7]; b.h = d[8];
        b.mu = d[9]; b.cut = d[10]; b.soft = d[11];
        b.valid = 1;
        b.a_zero = (d[5] == 0.0 && d[9] > 0.0) ? 1 : 0;
        const int g = (int)d[12];
        // a third atom counts within the largest bond cutoff of the center
        // species over every factor (atom%max_bond_radius, Core.f90:1034-1070)
        M.maxbond[e1] = std::max(M.maxbond[e1], d[10]);
        // the Tersoff kernels' group: that of the bond-order-scaled lone
        // exponentials (with none, the first group named)
        if (M.group < 0 && !M.has_pairs) use_group(g);
        if (g == M.group) {
            PYSIC_REQUIRE(!M.bond[e1][e2][e3].valid, PYSIC_EINVAL,
                          "duplicate tersoff entry for one species triplet");
            M.bond[e1][e2][e3] = b;
            if (!M.post[e1].valid) M.post[e1] = b;  // first factor of the center species
        }
        if (g >= 0 && g < PYSIC_MAXGRP) {
            PYSIC_REQUIRE(!B.bond[g][e1][e2][e3].valid, PYSIC_EINVAL,
                          "duplicate tersoff entry for one species triplet in a group");
            B.bond[g][e1][e2][e3] = b;
            if (!B.post[g][e1].valid) B.post[g][e1] = b;  // the group's first of the center
        }
    }
    double rs = 0.0;
    for (int s = 0; s < PYSIC_MAXS; ++s) {
        double mpair = 0.0;
        for (int t = 0; t < PYSIC_MAXS; ++t)
            for (int c = 0; c < M.npair[s][t]; ++c) mpair = std::max(mpair, M.pair[s][t][c].cut);
        rs = std::max(rs, std::min(mpair, M.maxpot[s]));
        rs = std::max(rs, M.maxbond[s]);
    }
    M.rshort = rs;
    h->S.model = M;
    h->S.dmodel.ensure(1);
    HIP_TRY(hipMemcpyAsync(h->S.dmodel.p, &h->S.model, sizeof(Model), hipMemcpyHostToDevice,
                           h->S.stream));
    for (int a = 0; a < PYSIC_MAXS; ++a) P.bend_maxpot[a] = M.maxpot[a];
    // the LJ table of a model whose pair potentials are lone LJs
    P.lj_only = 1;
    for (int a = 0; a < PYSIC_MAXS; ++a)
        for (int b = 0; b < PYSIC_MAXS; ++b) {
            const int np = P.npair[a][b];
            if (np > 1 || (np == 1 && (P.pair[a][b][0].nf != 1 || P.pair[a][b][0].f[0].type != 3)))
                P.lj_only = 0;
            LJPair &q = P.lj[a][b];
            q.on = np > 0;
            if (!q.on) continue;
            const GenPair &T = P.pair[a][b][0];
            q.cut = T.cut;
            q.soft = T.soft;
            q.smooth = T.smooth;
            q.lo2 = T.cut * T.cut * (1.0 - 1e-12);
            q.hi2 = P.cut2_hi[a][b];
            q.soft_lo2 = T.soft * T.soft * (1.0 - 1e-12);
            q.eps = T.f[0].p[0];
            q.sig2 = T.f[0].p[1] * T.f[0].p[1];
        }
    h->S.pmodel = P;
    h->S.dpmodel.ensure(1);
    HIP_TRY(hipMemcpyAsync(h->S.dpmodel.p, &h->S.pmodel, sizeof(PairModel),
                           hipMemcpyHostToDevice, h->S.stream));
    for (int a = 0; a < PYSIC_MAXS; ++a) {
        B.maxbond[a] = M.maxbond[a];
        if (B.on) B.radius = std::max(B.radius, M.maxbond[a]);
    }
    h->S.bmodel = B;
    h->S.dbmodel.ensure(1);
    HIP_TRY(hipMemcpyAsync(h->S.dbmodel.p, &h->S.bmodel, sizeof(BopModel),
                           hipMemcpyHostToDevice, h->S.stream));
    drop_graphs(h);  // captured sequences depend on which kernel families the model needs
    h->S.tables.ensure(htab.empty() ? 1 : htab.size());
    if (!htab.empty())
        HIP_TRY(hipMemcpyAsync(h->S.tables.p, htab.data(), sizeof(double2) * htab.size(),
                               hipMemcpyHostToDevice, h->S.stream));
    HIP_TRY(hipStreamSynchronize(h->S.stream));
    h->S.model_dirty = false;
}

void upload_state(pysic_ctx *h, const double *positions, const double *momenta, bool sorted_by_id)
{
    System &S = h->S;
    int n = S.n;
    std::vector<double4> hp(n), hm(n);
    for (int i = 0; i < n; ++i) {
        hp[i] = make_double4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2],
                             (double)h->species[i]);
        double px = momenta ? momenta[3 * i] : 0.0, py = momenta ? momenta[3 * i + 1] : 0.0,
               pz = momenta ? momenta[3 * i + 2] : 0.0;
        hm[i] = make_double4(px, py, pz, h->masses[i]);
    }
    if (!sorted_by_id) {
        HIP_TRY(hipMemcpyAsync(S.pos.p, hp.data(), sizeof(double4) * n, hipMemcpyHostToDevice, S.stream));
        HIP_TRY(hipMemcpyAsync(S.mom.p, hm.data(), sizeof(double4) * n, hipMemcpyHostToDevice, S.stream));
    } else {
        S.out4.ensure(n);
        HIP_TRY(hipMemcpyAsync(S.out4.p, hp.data(), sizeof(double4) * n, hipMemcpyHostToDevice, S.stream));
        k_gather4<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.out4.p, S.pos.p);
        if (momenta) {   // an update without momenta keeps them
            HIP_TRY(hipMemcpyAsync(S.out4.p, hm.data(), sizeof(double4) * n, hipMemcpyHostToDevice,
                                   S.stream));
            k_gather4<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.out4.p, S.mom.p);
        }
    }
    HIP_TRY(hipStreamSynchronize(S.stream));
}

void upload_scalar(System &S, const double *host, DBuf<double> &dst)
{
    int n = S.n;
    S.host_stage.ensure(n);
    HIP_TRY(hipMemcpyAsync(S.host_stage.p, host, sizeof(double) * n, hipMemcpyHostToDevice, S.stream));
    k_gather1<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.host_stage.p, dst.p);
    HIP_TRY(hipStreamSynchronize(S.stream));
}

void download_scalar(System &S, const double *src, double *host)
{
    int n = S.n;
    S.host_stage.ensure(n);
    k_scatter1<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, src, S.host_stage.p);
    HIP_TRY(hipMemcpyAsync(host, S.host_stage.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
}

void download_vec3(System &S, const double4 *src, double *host)
{
    int n = S.n;
    S.out4.ensure(n);
    sys_scatter_original(S, src, S.out4.p);
    std::vector<double4> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), S.out4.p, sizeof(double4) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    for (int i = 0; i < n; ++i) {
        host[3 * i] = tmp[i].x;
        host[3 * i + 1] = tmp[i].y;
        host[3 * i + 2] = tmp[i].z;
    }
}

void require_ready(pysic_ctx *h)
{
    PYSIC_REQUIRE(h->atoms_ready, PYSIC_ESTATE, "atoms not created");
    PYSIC_REQUIRE(h->cell_ready, PYSIC_ESTATE, "cell not created");
    PYSIC_REQUIRE(h->lists_valid, PYSIC_ESTATE, "neighbor lists not generated");
    if (h->S.model_dirty) compile_model(h);
}

// flags[2]: bit 0 an atom lies more than 223 cells from its box; bit 1 the
// cell sort found an inconsistent permutation (its guards kept every access
// in bounds; the state is not to be trusted)
void require_range_flag(int f2)
{
    PYSIC_REQUIRE(!(f2 & 2), PYSIC_ESTATE, "internal error: inconsistent cell-list sort");
    PYSIC_REQUIRE(f2 == 0, PYSIC_ERANGE, "an atom lies more than 223 cells from its box");
}

// flags[3]: bit 0 a Tersoff bond without its reverse in the short lists,
// bit 1 a non-finite force from the gather
void require_force_flags(int f3)
{
    PYSIC_REQUIRE(!(f3 & 1), PYSIC_EINVAL, "neighbor list is not symmetric");
    PYSIC_REQUIRE(!(f3 & 2), PYSIC_ERANGE,
                  "non-finite force (atoms on top of each other, or a time step too long for "
                  "the dynamics)");
}

void check_flags(pysic_ctx *h, bool lists_only)
{
    int f[4];
    nb_read_flags(h->S, f);
    require_range_flag(f[2]);
    if (!lists_only) {
        PYSIC_REQUIRE(f[1] == 0, PYSIC_ECAP,
                      "Tersoff short list overflow: " + std::to_string(f[1]) + " neighbors");
        require_force_flags(f[3]);
    }
}

// Neighbor lists for the current positions; grows the capacity on overflow.
void build_lists(pysic_ctx *h)
{
    System &S = h->S;
    for (int attempt = 0; attempt < 4; ++attempt) {
        HIP_TRY(hipMemsetAsync(S.flags.p, 0, sizeof(int) * 4, S.stream));
        nb_sort_atoms(S);
        nb_build_lists(S);
        int f[4];
        nb_read_flags(S, f);
        require_range_flag(f[2]);
        if (f[0] <= S.cap) {
            h->lists_valid = true;
            return;
        }
        S.cap = f[0] + 8;
    }
    throw PysicError{PYSIC_ECAP, "neighbor list capacity could not be established"};
}

// the model has terms for pair.hip (pair potentials or one-body terms)
bool has_gen(const System &S) { return S.pmodel.has_pairs || S.pmodel.has_single; }

// forces (+ optional momentum kick) and the virial into totals[1..6]
void tersoff_forces(pysic_ctx *h, bool kick, double half_dt)
{
    System &S = h->S;
    ts_site(S, h->variant, SITE_FORCES);
    ts_gather(S, half_dt, kick ? 1 : 0);
    ts_reduce(S);
}

// energy into totals[0]
void tersoff_energy(pysic_ctx *h)
{
    System &S = h->S;
    ts_site(S, h->variant, SITE_ENERGY);
    ts_reduce(S);
}

// One evaluation for the API path: if an atom had more short neighbors than
// the variant holds, switch to a variant that holds them and run again.
void tersoff_checked(pysic_ctx *h, int mode)
{
    System &S = h->S;
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIP_TRY(hipMemsetAsync(S.flags.p, 0, sizeof(int) * 4, S.stream));
        if (mode == SITE_ENERGY) tersoff_energy(h);
        else tersoff_forces(h, false, 0.0);
        int f[4];
        nb_read_flags(S, f);
        if (f[1] == 0) break;
        int v = ts_variant_for_capacity(f[1]);
        PYSIC_REQUIRE(v >= 0, PYSIC_ECAP,
                      "more short-range neighbors (" + std::to_string(f[1]) + ") than any kernel holds");
        h->variant = v;
    }
}

// Forces of every term except Ewald into S.frc; virials into totals[1..6]
// (Tersoff kernels) and gtotals[1..6] (pair.hip).
void local_forces(pysic_ctx *h)
{
    System &S = h->S;
    const bool gen = has_gen(S);
    if (S.model.has_pairs) tersoff_checked(h, SITE_FORCES);
    else if (!gen) k_zero4<<<div_up(S.n, 256), 256, 0, S.stream>>>(S.n, S.frc.p);
    if (gen) pr_eval(S, PAIR_FORCES, S.model.has_pairs != 0);
}

// The forces of a velocity-Verlet step (no virial), with the second
// half-kick fused into the last kernel.  Kernel launches only (capturable).
void md_forces(pysic_ctx *h, double half)
{
    System &S = h->S;
    const bool gen = has_gen(S);
    if (S.model.has_pairs) {
        ts_site(S, h->variant, SITE_FORCES_ONLY);
        ts_gather(S, half, gen ? 0 : 1);
    }
    if (gen) pr_eval(S, PAIR_FORCES_ONLY, S.model.has_pairs != 0, half, 1);
}

// Potential energy (all terms but Ewald) into *dst on the device.
void md_energy(pysic_ctx *h, double *dst)
{
    System &S = h->S;
    const bool gen = has_gen(S);
    if (S.model.has_pairs) tersoff_energy(h);
    if (gen) pr_eval(S, PAIR_ENERGY, false);
    k_sum_energy<<<1, 64, 0, S.stream>>>(S.model.has_pairs ? S.totals.p : nullptr,
                                         gen ? S.gtotals.p : nullptr, dst);
}

// Put the atoms back into the caller's order (slot i = atom i).
void restore_original_order(System &S)
{
    const int n = S.n;
    S.pos2.ensure(n); S.mom2.ensure(n); S.q2.ensure(n); S.cut2.ensure(n);
    sys_scatter_original(S, S.pos.p, S.pos2.p);
    sys_scatter_original(S, S.mom.p, S.mom2.p);
    k_scatter1<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.q.p, S.q2.p);
    k_scatter1<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p, S.cut.p, S.cut2.p);
    std::swap(S.pos, S.pos2); std::swap(S.mom, S.mom2); std::swap(S.q, S.q2);
    std::swap(S.cut, S.cut2);
    k_iota<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(S.stream));
}

void drop_graphs(pysic_ctx *h)
{
    for (auto *set : {h->graph, h->mdgraph})
        for (int k = 0; k < 2; ++k) {
            if (set[k].exec) (void)hipGraphExecDestroy(set[k].exec);
            set[k].exec = nullptr;
            set[k].sig.clear();
        }
}

// The whole device evaluation of a batch: list rebuild (bin, sort, 27-bin
// scan), Tersoff forces and virials, FP64 site energies, per-cell sums.
// Only kernel launches (no memset, no host synchronisation), so that it can
// be captured into a graph.
void batch_sequence(pysic_ctx *h, bool rebuild)
{
    System &S = h->S;
    nb_clear_flags(S);
    if (rebuild) {
        nb_sort_atoms(S);
        nb_build_lists(S);
    }
    ts_site(S, h->variant, SITE_FORCES);
    ts_gather(S, 0.0, 0);
    ts_site(S, h->variant, SITE_ENERGY_SHORT);  // over the force pass's short list
    ts_cell_sums(S);
}

std::vector<const void *> graph_signature(pysic_ctx *h, bool rebuild)
{
    System &S = h->S;
    return {S.pos.p, S.pos2.p, S.mom.p, S.q.p, S.cut.p, S.id.p, S.key.p, S.wrap.p, S.perm.p,
            S.wpos.p, S.cellcount.p, S.cellfill.p, S.cellstart.p, S.scanblk.p, S.prevstart.p,
            S.dirty.p, S.scanloc.p, (const void *)(long)(S.order_valid ? 1 : 0), S.nbr.p, S.nimg.p,
            S.ncnt.p, S.sj.p, S.simg.p, S.scnt.p, S.srow.p,
            (const void *)(long)(S.unique_images ? 1 : 0), S.dU.p, S.partial.p, S.atom_ev.p,
            S.cell_ev.p,
            S.frc.p, S.flags.p, S.dmodel.p, S.cells.p, S.first_bin.p, S.cell_of.p, S.atom0.p,
            S.dpmodel.p, S.tables.p, S.gpartial.p, S.gtotals.p, S.totals.p, S.bsj.p,
            S.bsimg.p, S.bscnt.p, S.bG.p, S.dbmodel.p, S.osj.p, S.osimg.p, S.oscnt.p, S.oD.p,
            S.oW.p, S.oown.p,
            (const void *)(long)S.cap, (const void *)(long)h->variant,
            (const void *)(long)S.site_grid, (const void *)(long)S.n, (const void *)(long)S.nbins,
            (const void *)(long)S.nb_method, (const void *)(long)(rebuild ? 1 : 0)};
}

void swap_state(System &S) { nb_swap_state(S); }

// Capture batch_sequence for the current buffer state into slot k.  Nothing
// runs; the host-side buffer swap that the sort performs during capture is
// undone, so host and device state stay as they were.
template <class Seq>
void capture_seq(pysic_ctx *h, pysic_ctx::Captured &slot, std::vector<const void *> sig,
                 bool swaps, Seq &&seq)
{
    System &S = h->S;
    if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
    slot.exec = nullptr;
    slot.sig = std::move(sig);
    const bool timing = S.timer.on;
    S.timer.on = false;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    try {
        seq();
    } catch (...) {
        (void)hipStreamEndCapture(S.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        S.timer.on = timing;
        slot.sig.clear();
        throw;
    }
    S.timer.on = timing;
    HIP_TRY(hipStreamEndCapture(S.stream, &graph));
    if (swaps) swap_state(S);
    hipError_t e = hipGraphInstantiate(&slot.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        slot.exec = nullptr;
        slot.sig.clear();
    }
    HIP_TRY(e);
}

void capture_into(pysic_ctx *h, int k, bool rebuild)
{
    capture_seq(h, h->graph[k], graph_signature(h, rebuild), rebuild,
                [&]() { batch_sequence(h, rebuild); });
}

// One velocity-Verlet step with a list rebuild: drift fused with binning,
// counting sort, neighbor scan, Tersoff forces, gather with the second
// half-kick.  Kernel launches only (capturable).
void md_step(pysic_ctx *h, double dt, double half)
{
    System &S = h->S;
    nb_sort_atoms(S, true, dt, half);
    nb_build_lists(S);
    md_forces(h, half);
}

// A model whose forces all come from the Tersoff kernels: its MD steps can
// fuse the force gather into the next step's drift (md_step_fused).
bool fused_md(pysic_ctx *h) { return h->S.model.has_pairs && !has_gen(h->S); }

// The step boundary and the next step up to its per-bond gradients: gather
// of the previous step's forces with its second half-kick, first half-kick,
// drift, bin and sort (one pass, forces never stored), neighbor scan,
// Tersoff site kernel.  Kernel launches only (capturable).
void md_step_fused(pysic_ctx *h, double dt, double half)
{
    System &S = h->S;
    nb_gather_drift_sort(S, dt, half);
    nb_build_lists(S);
    ts_site(S, h->variant, SITE_FORCES_ONLY);
}

std::vector<const void *> md_signature(pysic_ctx *h, double dt)
{
    auto sig = graph_signature(h, true);
    double half = 0.5 * dt;
    const void *d1, *d2;
    std::memcpy(&d1, &dt, sizeof(double));
    std::memcpy(&d2, &half, sizeof(double));
    sig.push_back(d1);
    sig.push_back(d2);
    sig.push_back((const void *)(long)(fused_md(h) ? 1 : 0));
    sig.push_back((const void *)(long)(h->S.qc_uniform ? 1 : 0));
    return sig;
}

// nsteps MD steps replayed from graphs captured for both buffer states (the
// sort alternates them).  The first step runs eagerly, which allocates
// whatever is missing, so nothing is allocated under capture.
void md_run_graph(pysic_ctx *h, int nsteps, double dt)
{
    System &S = h->S;
    const double half = 0.5 * dt;
    const bool fused = fused_md(h);
    // the first step runs eagerly, which allocates whatever is missing; with
    // the fused loop it stops at its per-bond gradients, whose gather is the
    // first thing the next step does
    if (fused) {
        nb_sort_atoms(S, true, dt, half);
        nb_build_lists(S);
        ts_site(S, h->variant, SITE_FORCES_ONLY);
    } else {
        md_step(h, dt, half);
    }
    if (nsteps > 1) {
        for (int k = 0; k < 2; ++k) {
            auto sig = md_signature(h, dt);
            pysic_ctx::Captured &slot = h->mdgraph[k];
            if (!(slot.exec && slot.sig == sig))
                capture_seq(h, slot, sig, true, [&]() {
                    if (fused) md_step_fused(h, dt, half);
                    else md_step(h, dt, half);
                });
            swap_state(S);   // the state of the other parity
        }
        for (int step = 1; step < nsteps; ++step) {
            auto sig = md_signature(h, dt);
            pysic_ctx::Captured *g = nullptr;
            for (auto &c : h->mdgraph)
                if (c.exec && c.sig == sig) g = &c;
            PYSIC_REQUIRE(g, PYSIC_ESTATE, "md graph lost its buffers");
            HIP_TRY(hipGraphLaunch(g->exec, S.stream));
            swap_state(S);
        }
    }
    // the last step's forces and its second half-kick
    if (fused) ts_gather(S, half, 1);
}

// One device evaluation.  With use_graph: replay the graph captured for the
// current buffer state.  On a miss the evaluation runs eagerly (which also
// allocates whatever is missing, so no allocation happens under capture),
// then graphs are captured for both buffer states (the sort alternates
// them), and later calls only replay.
void batch_evaluate(pysic_ctx *h, bool use_graph, bool rebuild)
{
    System &S = h->S;
    if (!use_graph) {
        batch_sequence(h, rebuild);
        return;
    }
    auto sig = graph_signature(h, rebuild);
    for (auto &g : h->graph)
        if (g.exec && g.sig == sig) {
            HIP_TRY(hipGraphLaunch(g.exec, S.stream));
            if (rebuild) swap_state(S);
            return;
        }
    drop_graphs(h);
    batch_sequence(h, rebuild);   // eager; buffers now in the next state
    capture_into(h, 0, rebuild);  // graph for this state
    if (rebuild) {
        swap_state(S);            // back to the state before the eager run
        capture_into(h, 1, rebuild);
        swap_state(S);            // and forward again: host matches device
    }
}

// Host arrays that become device state are checked before anything changes:
// a NaN position would bin as a valid one and drop out of every cutoff test.
void require_finite(const double *x, size_t count, const std::string &what)
{
    if (!x) return;
    for (size_t k = 0; k < count; ++k)
        PYSIC_REQUIRE(std::isfinite(x[k]), PYSIC_EINVAL,
                      "non-finite " + what + " (entry " + std::to_string(k) + ")");
}

// Three cell vectors (rows of v) must be finite and span a volume: the bin
// grid divides by each axis' width.
void require_cell_vectors(const double *v, const std::string &who)
{
    for (int k = 0; k < 9; ++k)
        PYSIC_REQUIRE(std::isfinite(v[k]), PYSIC_EINVAL, who + ": non-finite cell vector");
    const double det = v[0] * (v[4] * v[8] - v[5] * v[7]) - v[1] * (v[3] * v[8] - v[5] * v[6]) +
                       v[2] * (v[3] * v[7] - v[4] * v[6]);
    double len = 1.0;
    for (int i = 0; i < 3; ++i)
        len *= std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                         v[3 * i + 2] * v[3 * i + 2]);
    PYSIC_REQUIRE(std::fabs(det) > 1e-9 * len, PYSIC_EINVAL,
                  who + ": the cell vectors span no volume");
}

// A setter has just put its new rows in `now` (the previous ones in
// `old`): compile them if the atoms are there; if they are refused, put the
// previous rows back, so that the handle keeps the model it had.
void recompile_or_restore(pysic_ctx *h, std::vector<double> &now, std::vector<double> &old)
{
    h->S.model_dirty = true;
    h->forces_valid = false;
    if (!h->atoms_ready) return;
    try {
        compile_model(h);
    } catch (...) {
        now.swap(old);
        h->S.model_dirty = true;   // recompiled from the previous rows when next used
        throw;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int pysic_create(pysic_ctx **out, int device)
{
    if (!out) return PYSIC_EINVAL;
    pysic_ctx *h = new pysic_ctx();
    h->device = device;
    try {
        int count = 0;
        HIP_TRY(hipGetDeviceCount(&count));
        PYSIC_REQUIRE(device >= 0 && device < count, PYSIC_EINVAL, "no such HIP device");
        HIP_TRY(hipSetDevice(device));
        scratch_setup(device);
        HIP_TRY(hipStreamCreate(&h->S.stream));
        h->S.flags.ensure(4);
        HIP_TRY(hipMemset(h->S.flags.p, 0, sizeof(int) * 4));
        h->S.cap = 16;
    } catch (const PysicError &e) {
        delete h;
        *out = nullptr;
        return e.code;
    }
    *out = h;
    return PYSIC_OK;
}

void pysic_destroy(pysic_ctx *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->S.stream) (void)hipStreamSynchronize(h->S.stream);
    h->S.timer.reset();
    h->S.release_all();
    ewald_release(h->ew);
    drop_graphs(h);
    h->rates.release();
    if (h->S.stream) (void)hipStreamDestroy(h->S.stream);
    delete h;
}

const char *pysic_last_error(pysic_ctx *h) { return h ? h->err.c_str() : "null handle"; }

int pysic_create_atoms(pysic_ctx *h, int n, const double *masses, const double *charges,
                       const double *positions, const double *momenta, const int *species)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(n > 0 && positions && species, PYSIC_EINVAL, "create_atoms: bad arguments");
    for (int i = 0; i < n; ++i)
        PYSIC_REQUIRE(species[i] >= 0 && species[i] < PYSIC_MAXS, PYSIC_EINVAL,
                      "species must be 0..3");
    require_finite(positions, 3 * (size_t)n, "position");
    require_finite(momenta, 3 * (size_t)n, "momentum");
    require_finite(charges, (size_t)n, "charge");
    if (masses)
        for (int i = 0; i < n; ++i)
            PYSIC_REQUIRE(masses[i] > 0.0 && std::isfinite(masses[i]), PYSIC_EINVAL,
                          "masses must be positive and finite");
    // arguments checked: from here the handle holds no usable system until
    // the call completes, so a failure part-way (an allocation) leaves
    // evaluations refused rather than running on buffers of another size
    h->atoms_ready = false;
    h->lists_valid = false;
    h->forces_valid = false;
    System &S = h->S;
    S.n = n;
    h->species.assign(species, species + n);
    h->masses.assign(n, 1.0);
    if (masses) h->masses.assign(masses, masses + n);
    S.ncells = 0;
    S.order_valid = false;  // atoms in original order: no previous sort to follow
    drop_graphs(h);
    S.pos.ensure(n); S.mom.ensure(n); S.frc.ensure(n); S.q.ensure(n); S.cut.ensure(n); S.id.ensure(n);
    k_iota<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.id.p);
    k_zero4<<<div_up(n, 256), 256, 0, S.stream>>>(n, S.frc.p);
    upload_state(h, positions, momenta, false);
    std::vector<double> q(n, 0.0);
    if (charges) q.assign(charges, charges + n);
    HIP_TRY(hipMemcpyAsync(S.q.p, q.data(), sizeof(double) * n, hipMemcpyHostToDevice, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    h->atoms_ready = true;
    h->lists_valid = false;
    h->forces_valid = false;
    S.model_dirty = true;
    API_END(h)
}

int pysic_create_cell(pysic_ctx *h, const double *vectors, const double *inverse, const int *pbc)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(vectors && inverse && pbc, PYSIC_EINVAL, "create_cell: bad arguments");
    require_cell_vectors(vectors, "create_cell");
    std::memcpy(h->S.geom.v, vectors, sizeof(double) * 9);
    std::memcpy(h->S.geom.inv, inverse, sizeof(double) * 9);
    for (int a = 0; a < 3; ++a) h->S.geom.pbc[a] = pbc[a] ? 1 : 0;
    h->S.ncells = 0;
    drop_graphs(h);
    h->ew.dirty = true;   // the k-vectors follow the cell
    h->cell_ready = true;
    h->lists_valid = false;
    h->forces_valid = false;
    API_END(h)
}

int pysic_update_atom_coordinates(pysic_ctx *h, int n, const double *positions, const double *momenta)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->atoms_ready && n == h->S.n && positions, PYSIC_EINVAL,
                  "update_atom_coordinates: atom count does not match");
    require_finite(positions, 3 * (size_t)n, "position");
    require_finite(momenta, 3 * (size_t)n, "momentum");
    upload_state(h, positions, momenta, true);
    h->forces_valid = false;
    API_END(h)
}

int pysic_update_atom_charges(pysic_ctx *h, int n, const double *charges)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->atoms_ready && n == h->S.n && charges, PYSIC_EINVAL,
                  "update_atom_charges: atom count does not match");
    require_finite(charges, (size_t)n, "charge");
    upload_scalar(h->S, charges, h->S.q);
    h->forces_valid = false;
    API_END(h)
}

int pysic_get_number_of_atoms(pysic_ctx *h, int *n)
{
    if (!h || !n) return PYSIC_EINVAL;
    *n = h->S.n;
    return PYSIC_OK;
}

int pysic_set_potentials(pysic_ctx *h, int npot, const double *desc)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(npot >= 0 && (npot == 0 || desc), PYSIC_EINVAL, "set_potentials: bad arguments");
    std::vector<double> old(desc, desc + POT_STRIDE * (size_t)npot);
    h->potdesc.swap(old);
    recompile_or_restore(h, h->potdesc, old);
    API_END(h)
}

int pysic_set_potential_table(pysic_ctx *h, int id, int nrows, const double *rows)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(nrows >= 2 && rows, PYSIC_EINVAL,
                  "set_potential_table: a table needs at least 2 rows");
    const bool had = h->tables.count(id) > 0;
    std::vector<double> old(rows, rows + 2 * (size_t)nrows);
    h->tables[id].swap(old);
    try {
        recompile_or_restore(h, h->tables[id], old);
    } catch (...) {
        if (!had) h->tables.erase(id);   // as before the call: no such table
        throw;
    }
    API_END(h)
}

int pysic_set_bond_order_factors(pysic_ctx *h, int nbond, const double *desc)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(nbond >= 0 && (nbond == 0 || desc), PYSIC_EINVAL, "set_bond_order_factors: bad arguments");
    std::vector<double> old(desc, desc + 14 * (size_t)nbond);
    h->bonddesc.swap(old);
    recompile_or_restore(h, h->bonddesc, old);
    API_END(h)
}

int pysic_set_ewald(pysic_ctx *h, int on, double real_cutoff, double k_cutoff, const int *kcells,
                    double sigma, double epsilon, const double *scaler)
{
    API_BEGIN(h)
    EwaldParams &p = h->ewald;
    if (on) {   // checked before anything changes
        PYSIC_REQUIRE(kcells && sigma > 0 && epsilon > 0 && real_cutoff > 0 && k_cutoff > 0,
                      PYSIC_EINVAL, "set_ewald: bad parameters");
        for (int a = 0; a < 3; ++a)
            PYSIC_REQUIRE(kcells[a] >= 0 && kcells[a] < 1024, PYSIC_EINVAL,
                          "set_ewald: at most 1023 reciprocal cells per axis");
    }
    p.on = on ? 1 : 0;
    if (p.on) {
        p.rc = real_cutoff;
        p.kc = k_cutoff;
        for (int a = 0; a < 3; ++a) p.K[a] = kcells[a];
        p.sigma = sigma;
        p.eps = epsilon;
        p.has_scaler = scaler != nullptr;
        p.scaler.assign(h->S.n, 1.0);
        if (scaler) p.scaler.assign(scaler, scaler + h->S.n);
    }
    h->ew.dirty = true;
    h->forces_valid = false;
    API_END(h)
}

int pysic_generate_neighbor_lists(pysic_ctx *h, int n, const double *cutoffs)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->atoms_ready && h->cell_ready, PYSIC_ESTATE, "atoms and cell must be created first");
    PYSIC_REQUIRE(n == h->S.n && cutoffs, PYSIC_EINVAL, "generate_neighbor_lists: bad arguments");
    double mc = 0.0, lo = cutoffs[0];
    for (int i = 0; i < n; ++i) {
        mc = std::max(mc, cutoffs[i]);
        lo = std::min(lo, cutoffs[i]);
    }
    PYSIC_REQUIRE(mc > 0.0, PYSIC_EINVAL, "neighbor cutoff must be positive");
    // the old lists no longer match the cutoffs: refused until this build ends
    h->lists_valid = false;
    h->forces_valid = false;
    h->min_cut = lo;
    upload_scalar(h->S, cutoffs, h->S.cut);
    nb_configure_grid(h->S, mc);
    build_lists(h);
    h->forces_valid = false;
    API_END(h)
}

int pysic_get_neighbor_counts(pysic_ctx *h, int n, int *counts)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->lists_valid && n == h->S.n && counts, PYSIC_ESTATE, "no neighbor lists");
    System &S = h->S;
    std::vector<int> cnt(n), id(n);
    HIP_TRY(hipMemcpy(cnt.data(), S.ncnt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(id.data(), S.id.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int s = 0; s < n; ++s) counts[id[s]] = cnt[s];
    API_END(h)
}

int pysic_get_neighbor_lists(pysic_ctx *h, long nnz, int *idx, int *offsets)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->lists_valid && idx && offsets, PYSIC_ESTATE, "no neighbor lists");
    System &S = h->S;
    int n = S.n;
    std::vector<int> cnt(n), id(n), nbr((size_t)S.cap * n), img((size_t)S.cap * n);
    HIP_TRY(hipMemcpy(cnt.data(), S.ncnt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(id.data(), S.id.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nbr.data(), S.nbr.p, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(img.data(), S.nimg.p, sizeof(int) * img.size(), hipMemcpyDeviceToHost));
    std::vector<int> slot(n);
    for (int s = 0; s < n; ++s) slot[id[s]] = s;
    long total = 0;
    for (int s = 0; s < n; ++s) total += cnt[s];
    PYSIC_REQUIRE(total == nnz, PYSIC_EINVAL, "get_neighbor_lists: nnz does not match the counts");
    long e = 0;
    for (int a = 0; a < n; ++a) {
        int s = slot[a];
        for (int k = 0; k < cnt[s]; ++k, ++e) {
            size_t o = (size_t)k * n + s;
            idx[e] = id[nbr[o]];
            offsets[3 * e] = img_x(img[o]);
            offsets[3 * e + 1] = img_y(img[o]);
            offsets[3 * e + 2] = img_z(img[o]);
        }
    }
    API_END(h)
}

int pysic_calculate_energy(pysic_ctx *h, double *energy)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(energy, PYSIC_EINVAL, "calculate_energy: null output");
    require_ready(h);
    System &S = h->S;
    double e = 0.0;
    if (S.model.has_pairs) {
        tersoff_checked(h, SITE_ENERGY);
        double t[8];
        HIP_TRY(hipMemcpyAsync(t, S.totals.p, sizeof(double) * 7, hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
        check_flags(h, false);
        e += t[0];
    }
    if (has_gen(S)) {
        pr_eval(S, PAIR_ENERGY, false);
        double t = 0.0;
        HIP_TRY(hipMemcpyAsync(&t, S.gtotals.p, sizeof(double), hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));
        e += t;
    }
    if (S.model.has_self || h->ewald.on) e += ewald_eem_energy(S, h->ewald, h->ew);
    *energy = e;
    API_END(h)
}

// Bond-order factor queries.  Every factor type this engine takes ('tersoff')
// is a per-bond factor (level 2); the per-atom factors and their gradients
// that the reference returns for a group are sums over level-1 factors only
// (Core.f90:2015-2404, the n_level == 1 filter), so for these groups they
// are zero, as the reference's are.
int pysic_calculate_bond_order_factors(pysic_ctx *h, int n, int group, double *factors)
{
    API_BEGIN(h)
    require_ready(h);
    PYSIC_REQUIRE(n == h->S.n && factors && group >= 0, PYSIC_EINVAL,
                  "calculate_bond_order_factors: bad arguments");
    std::fill(factors, factors + n, 0.0);
    API_END(h)
}

int pysic_calculate_bond_order_gradients(pysic_ctx *h, int n, int group, int atom,
                                         double *gradients)
{
    API_BEGIN(h)
    require_ready(h);
    PYSIC_REQUIRE(n == h->S.n && gradients && group >= 0 && atom >= 0 && atom < n, PYSIC_EINVAL,
                  "calculate_bond_order_gradients: bad arguments");
    std::fill(gradients, gradients + 3 * (size_t)n, 0.0);
    API_END(h)
}

int pysic_calculate_bond_order_gradients_of_factor(pysic_ctx *h, int n, int group, int atom,
                                                   double *gradients)
{
    API_BEGIN(h)
    require_ready(h);
    PYSIC_REQUIRE(n == h->S.n && gradients && group >= 0 && atom >= 0 && atom < n, PYSIC_EINVAL,
                  "calculate_bond_order_gradients_of_factor: bad arguments");
    std::fill(gradients, gradients + 3 * (size_t)n, 0.0);
    API_END(h)
}

int pysic_get_ewald_energy(pysic_ctx *h, double real_cut, double k_cut, const int *kcells,
                           double sigma, double epsilon, double *energy)
{
    API_BEGIN(h)
    require_ready(h);
    PYSIC_REQUIRE(kcells && energy && sigma > 0.0 && epsilon > 0.0 && real_cut >= 0.0,
                  PYSIC_EINVAL, "get_ewald_energy: bad arguments");
    for (int a = 0; a < 3; ++a)
        PYSIC_REQUIRE(kcells[a] >= 0 && kcells[a] < 1024, PYSIC_EINVAL,
                      "get_ewald_energy: reciprocal cells per axis must lie in 0..1023");
    EwaldParams P;
    P.on = 1;
    P.rc = real_cut;
    P.kc = k_cut;
    P.sigma = sigma;
    P.eps = epsilon;
    for (int a = 0; a < 3; ++a) P.K[a] = kcells[a];
    // unit scalers (Core.f90:8084) and a state of its own: the model's
    // Ewald parameters and cached structure factors are left as they are
    EwaldState E;
    try {
        *energy = ewald_standalone_energy(h->S, P, E);
    } catch (...) {
        ewald_release(E);
        throw;
    }
    ewald_release(E);
    API_END(h)
}

int pysic_calculate_energy_noalloc(pysic_ctx *h, double *energy)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(energy, PYSIC_EINVAL, "calculate_energy_noalloc: null output");
    PYSIC_REQUIRE(h->atoms_ready && h->cell_ready && h->lists_valid, PYSIC_ESTATE,
                  "calculate_energy_noalloc: atoms, cell and neighbor lists must be set up first");
    System &S = h->S;
    PYSIC_REQUIRE(!S.model_dirty && S.dmodel.p, PYSIC_ESTATE,
                  "calculate_energy_noalloc: the potentials changed after the last evaluation; "
                  "evaluate once with pysic_calculate_energy first");
    PYSIC_REQUIRE(S.ncells == 0 && S.model.nspecies == 1 && S.model.has_pairs && !has_gen(S) &&
                      !S.model.has_self && !h->ewald.on,
                  PYSIC_EINVAL,
                  "calculate_energy_noalloc takes one cell of one species with Tersoff and "
                  "exponential terms only");
    PYSIC_REQUIRE(S.flags.p && S.flags.cap >= 4, PYSIC_ESTATE, "calculate_energy_noalloc: no flags");
    double *ptr = scratch_of(h->device);
    PYSIC_REQUIRE(ptr, PYSIC_ESTATE, "calculate_energy_noalloc: scratch not set up on this device");
    const int tiles = ts_energy_tiles(S, h->variant);
    PYSIC_REQUIRE(tiles <= kNoallocTiles, PYSIC_ERANGE,
                  "calculate_energy_noalloc: " + std::to_string(S.n) + " atoms need " +
                      std::to_string(tiles) + " site-kernel tiles of kernel variant " +
                      ts_variant_name(h->variant) + "; the fixed scratch holds " +
                      std::to_string(kNoallocTiles));
    double *partial = ptr, *totals = ptr + (size_t)kNoallocTiles * 8;
    ts_energy_into(S, h->variant, partial, totals);
    double e = 0.0;
    int f[4];
    HIP_TRY(hipMemcpyAsync(&e, totals, sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipMemcpyAsync(f, S.flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    require_range_flag(f[2]);
    PYSIC_REQUIRE(f[1] == 0, PYSIC_ECAP,
                  "calculate_energy_noalloc: " + std::to_string(f[1]) +
                      " short-range neighbors exceed kernel variant " +
                      ts_variant_name(h->variant) + "; pick a larger one (pysic_calculate_energy "
                      "switches by itself)");
    *energy = e;
    API_END(h)
}

int pysic_calculate_forces(pysic_ctx *h, int n, double *forces, double *stress)
{
    API_BEGIN(h)
    require_ready(h);
    System &S = h->S;
    PYSIC_REQUIRE(n == S.n && forces && stress, PYSIC_EINVAL, "calculate_forces: bad arguments");
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0}, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    local_forces(h);
    if (S.model.has_pairs)
        HIP_TRY(hipMemcpyAsync(t, S.totals.p, sizeof(double) * 7, hipMemcpyDeviceToHost, S.stream));
    if (has_gen(S))
        HIP_TRY(hipMemcpyAsync(g, S.gtotals.p, sizeof(double) * 7, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    if (S.model.has_pairs) check_flags(h, false);
    for (int q = 0; q < 6; ++q) stress[q] = t[1 + q] + g[1 + q];
    if (h->ewald.on) {
        double st[6];
        ewald_add_forces(S, h->ewald, h->ew, st);  // adds into S.frc
        for (int q = 0; q < 6; ++q) stress[q] += st[q];
    }
    download_vec3(S, S.frc.p, forces);
    h->forces_valid = true;
    API_END(h)
}

int pysic_calculate_electronegativities(pysic_ctx *h, int n, double *enegs)
{
    API_BEGIN(h)
    require_ready(h);
    PYSIC_REQUIRE(n == h->S.n && enegs, PYSIC_EINVAL, "calculate_electronegativities: bad arguments");
    DBuf<double> &chi = ewald_eem_chi(h->S, h->ewald, h->ew);
    download_scalar(h->S, chi.p, enegs);
    API_END(h)
}

int pysic_relax_charges(pysic_ctx *h, int n, double *charges, double *rates, int n_steps,
                        double timestep, double inertia, double friction, double tolerance,
                        int *steps, double *error)
{
    API_BEGIN(h)
    require_ready(h);
    System &S = h->S;
    PYSIC_REQUIRE(n == S.n && charges && rates && steps && error, PYSIC_EINVAL,
                  "relax_charges: bad arguments");
    require_finite(charges, (size_t)n, "charge");
    require_finite(rates, (size_t)n, "charge rate");
    upload_scalar(S, charges, S.q);
    h->rates.ensure(n);
    upload_scalar(S, rates, h->rates);
    ewald_relax(S, h->ewald, h->ew, h->rates.p, n_steps, timestep, inertia, friction, tolerance,
                steps, error);
    download_scalar(S, S.q.p, charges);
    download_scalar(S, h->rates.p, rates);
    h->forces_valid = false;
    API_END(h)
}

int pysic_md_run(pysic_ctx *h, int nsteps, double dt, int rebuild, double *energies)
{
    API_BEGIN(h)
    require_ready(h);
    System &S = h->S;
    PYSIC_REQUIRE(nsteps >= 0 && rebuild >= 0 && std::isfinite(dt), PYSIC_EINVAL,
                  "md_run: bad arguments");
    PYSIC_REQUIRE(!h->ewald.on && (S.model.has_pairs || has_gen(S)), PYSIC_EINVAL,
                  "md_run drives the Tersoff and pair potentials (no Ewald)");
    HIP_TRY(hipMemsetAsync(S.flags.p, 0, sizeof(int) * 4, S.stream));
    DBuf<double> ebuf;
    if (energies) ebuf.ensure(nsteps > 0 ? nsteps : 1);
    if (!h->forces_valid) local_forces(h);
    HIP_TRY(hipMemsetAsync(S.flags.p, 0, sizeof(int) * 4, S.stream));
    const double half = 0.5 * dt;
    if (rebuild == 1 && !energies && !S.timer.on && nsteps >= 3) {
        // the launch-bound step loop from graphs (per-kernel timing wants
        // the eager path, where each launch is bracketed by events).  Nothing
        // in the loop writes charges or cutoffs: when each is one value for
        // every atom, the sorts leave them in place.
        nb_begin_uniform_qc(S);
        try {
            md_run_graph(h, nsteps, dt);
        } catch (...) {
            S.qc_uniform = false;
            throw;
        }
        S.qc_uniform = false;
    } else {
        for (int step = 0; step < nsteps; ++step) {
            if (rebuild > 0 && (step + 1) % rebuild == 0) {
                md_step(h, dt, half);
            } else {
                md_drift(S, dt, half);
                md_forces(h, half);
            }
            if (energies) md_energy(h, ebuf.p + step);  // an observer, not part of the step
        }
    }
    HIP_TRY(hipGetLastError());
    if (energies && nsteps > 0)
        HIP_TRY(hipMemcpyAsync(energies, ebuf.p, sizeof(double) * nsteps, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    ebuf.release();
    int f[4];
    nb_read_flags(S, f);
    PYSIC_REQUIRE(f[0] <= S.cap, PYSIC_ECAP,
                  "neighbor list overflow during MD (" + std::to_string(f[0]) +
                      "); raise the capacity and rerun");
    check_flags(h, false);
    h->forces_valid = true;
    API_END(h)
}

int pysic_get_state(pysic_ctx *h, int n, double *positions, double *momenta)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(h->atoms_ready && n == h->S.n, PYSIC_EINVAL, "get_state: bad arguments");
    if (positions) download_vec3(h->S, h->S.pos.p, positions);
    if (momenta) download_vec3(h->S, h->S.mom.p, momenta);
    API_END(h)
}

int pysic_set_kernel_variant(pysic_ctx *h, int variant)
{
    if (!h) return PYSIC_EINVAL;
    if (variant < 0 || variant >= ts_variant_count()) {
        h->err = "no such kernel variant";
        return PYSIC_EINVAL;
    }
    h->variant = variant;
    return PYSIC_OK;
}

int pysic_kernel_variant_count(void) { return ts_variant_count(); }
const char *pysic_kernel_variant_name(int v) { return ts_variant_name(v); }

int pysic_set_neighbor_capacity(pysic_ctx *h, int cap)
{
    if (!h || cap < 1) return PYSIC_EINVAL;
    h->S.cap = cap;
    h->lists_valid = false;
    return PYSIC_OK;
}

int pysic_set_neighbor_method(pysic_ctx *h, int method)
{
    if (!h || method < 0 || method > 5) return PYSIC_EINVAL;
    h->S.nb_method = method;
    h->lists_valid = false;
    return PYSIC_OK;
}

int pysic_timing(pysic_ctx *h, int on)
{
    API_BEGIN(h)
    h->S.timer.reset();
    h->S.timer.on = on != 0;
    API_END(h)
}

int pysic_timing_read(pysic_ctx *h, int max, char *names, int name_len, double *ms, long *calls)
{
    if (!h) return PYSIC_EINVAL;
    try {
        HIP_TRY(hipSetDevice(h->device));
        h->S.timer.collect();
    } catch (const PysicError &e) {
        return fail(h, e);
    }
    int k = 0;
    for (; k < (int)h->S.timer.total.size() && k < max; ++k) {
        std::strncpy(names + (size_t)k * name_len, h->S.timer.total[k].first.c_str(), name_len - 1);
        names[(size_t)k * name_len + name_len - 1] = 0;
        ms[k] = h->S.timer.total[k].second;
        calls[k] = h->S.timer.calls[k].second;
    }
    return k;
}

// ---------------------------------------------------------------------------
// batches of independent cells (include/pysic_batch.h)
// ---------------------------------------------------------------------------
int pysic_set_cells(pysic_ctx *h, int ncells, const int *natoms, const double *vectors,
                    const double *inverses, const int *pbc)
{
    API_BEGIN(h)
    System &S = h->S;
    PYSIC_REQUIRE(h->atoms_ready, PYSIC_ESTATE, "set_cells: create the atoms first");
    PYSIC_REQUIRE(ncells > 0 && natoms && vectors && inverses && pbc, PYSIC_EINVAL,
                  "set_cells: bad arguments");
    std::vector<int> a0(ncells + 1, 0);
    for (int c = 0; c < ncells; ++c) {
        PYSIC_REQUIRE(natoms[c] > 0, PYSIC_EINVAL, "set_cells: every cell needs atoms");
        a0[c + 1] = a0[c] + natoms[c];
    }
    PYSIC_REQUIRE(a0[ncells] == S.n, PYSIC_EINVAL,
                  "set_cells: the cells' atom counts must add up to the atoms created");
    for (int c = 0; c < ncells; ++c)
        require_cell_vectors(vectors + 9 * c, "set_cells: cell " + std::to_string(c));
    // the atoms go back to their input order below: the lists no longer apply
    h->lists_valid = false;
    h->forces_valid = false;
    // atoms keep their cell's slot range: start from the input order
    restore_original_order(S);
    S.hcells.assign(ncells, CellGeom{});
    for (int c = 0; c < ncells; ++c) {
        CellGeom &g = S.hcells[c];
        std::memcpy(g.v, vectors + 9 * c, sizeof(double) * 9);
        std::memcpy(g.inv, inverses + 9 * c, sizeof(double) * 9);
        for (int a = 0; a < 3; ++a) g.pbc[a] = pbc[3 * c + a] ? 1 : 0;
    }
    S.hatom0 = a0;
    std::vector<int> cell_of(S.n);
    for (int c = 0; c < ncells; ++c)
        for (int i = a0[c]; i < a0[c + 1]; ++i) cell_of[i] = c;
    S.cell_of.ensure(S.n);
    S.atom0.ensure(ncells + 1);
    HIP_TRY(hipMemcpyAsync(S.cell_of.p, cell_of.data(), sizeof(int) * S.n, hipMemcpyHostToDevice,
                           S.stream));
    HIP_TRY(hipMemcpyAsync(S.atom0.p, a0.data(), sizeof(int) * (ncells + 1), hipMemcpyHostToDevice,
                           S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    S.ncells = ncells;
    S.order_valid = false;
    drop_graphs(h);
    h->cell_ready = true;
    h->lists_valid = false;
    h->forces_valid = false;
    API_END(h)
}

int pysic_get_number_of_cells(pysic_ctx *h, int *ncells)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(ncells, PYSIC_EINVAL, "get_number_of_cells: null");
    *ncells = h->S.ncells;
    API_END(h)
}

int pysic_set_site_grid(pysic_ctx *h, int blocks)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(blocks >= 0, PYSIC_EINVAL, "set_site_grid: blocks must be >= 0");
    h->S.site_grid = blocks;
    API_END(h)
}

int pysic_evaluate_cells(pysic_ctx *h, int use_graph, int rebuild)
{
    API_BEGIN(h)
    require_ready(h);
    System &S = h->S;
    PYSIC_REQUIRE(S.ncells > 0, PYSIC_ESTATE, "evaluate_cells: no batch of cells set");
    PYSIC_REQUIRE(S.model.has_pairs && !h->ewald.on && !has_gen(S), PYSIC_EINVAL,
                  "evaluate_cells: a batch evaluates Tersoff / exponential terms only");
    S.atom_ev.ensure((size_t)7 * S.n);
    S.cell_ev.ensure((size_t)8 * S.ncells);
    batch_evaluate(h, use_graph != 0, rebuild != 0);
    h->forces_valid = true;
    API_END(h)
}

int pysic_check_cells(pysic_ctx *h, int *flags4)
{
    API_BEGIN(h)
    PYSIC_REQUIRE(flags4, PYSIC_EINVAL, "check_cells: null");
    nb_read_flags(h->S, flags4);
    API_END(h)
}

int pysic_get_cell_results(pysic_ctx *h, int ncells, double *energies, double *virials, int n,
                           double *forces)
{
    API_BEGIN(h)
    System &S = h->S;
    PYSIC_REQUIRE(S.ncells > 0 && ncells == S.ncells && n == S.n, PYSIC_EINVAL,
                  "get_cell_results: bad arguments");
    PYSIC_REQUIRE(h->forces_valid, PYSIC_ESTATE, "get_cell_results: nothing evaluated");
    int f[4];
    nb_read_flags(S, f);
    PYSIC_REQUIRE(f[0] <= S.cap, PYSIC_ECAP,
                  "neighbor list overflow (" + std::to_string(f[0]) + " > capacity " +
                      std::to_string(S.cap) + "); raise the capacity and evaluate again");
    PYSIC_REQUIRE(f[1] == 0, PYSIC_ECAP,
                  "Tersoff short list overflow: " + std::to_string(f[1]) + " neighbors");
    require_range_flag(f[2]);
    require_force_flags(f[3]);
    std::vector<double> ev((size_t)8 * ncells);
    HIP_TRY(hipMemcpyAsync(ev.data(), S.cell_ev.p, sizeof(double) * ev.size(),
                           hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    for (int c = 0; c < ncells; ++c) {
        if (energies) energies[c] = ev[8 * c];
        if (virials)
            for (int q = 0; q < 6; ++q) virials[6 * c + q] = ev[8 * c + 1 + q];
    }
    if (forces) download_vec3(S, S.frc.p, forces);
    API_END(h)
}

int pysic_synchronize(pysic_ctx *h)
{
    API_BEGIN(h)
    HIP_TRY(hipStreamSynchronize(h->S.stream));
    API_END(h)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned charge warm start (include/pysic_learned.h)
// ---------------------------------------------------------------------------
struct pysic_mlp {
    MlpState M;
    std::string err;
};

namespace {

int fail_m(pysic_mlp *m, const PysicError &e)
{
    if (m) m->err = e.msg;
    return e.code;
}

#define MLP_BEGIN(m)                                                                         \
    if (!(m)) return PYSIC_EINVAL;                                                           \
    try {                                                                                    \
        HIP_TRY(hipSetDevice((m)->M.device));
#define MLP_END(m)                                                                           \
    }                                                                                        \
    catch (const PysicError &e) { return fail_m(m, e); }                                     \
    catch (const std::exception &e) { return fail_m(m, PysicError{PYSIC_ENOMEM, e.what()}); } \
    return PYSIC_OK;

void mlp_require_params(const MlpState &M)
{
    PYSIC_REQUIRE(M.configured, PYSIC_ESTATE, "network not configured");
    PYSIC_REQUIRE(M.has_params, PYSIC_ESTATE, "network parameters not set");
}

// sample indices are used as addresses on the device: check every one
void mlp_upload_samples(MlpState &M, long count, const int *samples)
{
    PYSIC_REQUIRE(samples && count > 0, PYSIC_EINVAL, "empty sample list");
    PYSIC_REQUIRE(M.ntrain > 0, PYSIC_ESTATE, "no training data");
    for (long k = 0; k < count; ++k)
        PYSIC_REQUIRE(samples[k] >= 0 && samples[k] < M.ntrain, PYSIC_ERANGE,
                      "sample index out of range: " + std::to_string(samples[k]));
    M.order.ensure(count);
    HIP_TRY(hipMemcpyAsync(M.order.p, samples, sizeof(int) * count, hipMemcpyHostToDevice,
                           M.stream));
}

// descriptors of h's system into M.sfeat / M.sspec (current order), on h's stream
void mlp_system_descriptors(MlpState &M, pysic_ctx *h)
{
    PYSIC_REQUIRE(M.configured, PYSIC_ESTATE, "network not configured");
    PYSIC_REQUIRE(h->device == M.device, PYSIC_EINVAL, "network and system live on different GPUs");
    require_ready(h);
    PYSIC_REQUIRE(h->min_cut >= M.dp.rc, PYSIC_EINVAL,
                  "neighbor lists are shorter than the descriptor cutoff");
    for (int s : h->species)
        PYSIC_REQUIRE(s >= 0 && s < M.nspecies, PYSIC_EINVAL,
                      "system has more species than the network was configured for");
    System &S = h->S;
    M.sfeat.ensure((size_t)S.n * MLP_F);
    M.sspec.ensure(S.n);
    mlp_descriptors(S, M, S.stream);
}

}  // namespace

extern "C" {

int pysic_mlp_param_count(void) { return MLP_NP; }

int pysic_mlp_create(pysic_mlp **out, int device)
{
    if (!out) return PYSIC_EINVAL;
    *out = nullptr;
    pysic_mlp *m = new pysic_mlp();
    m->M.device = device;
    try {
        int ndev = 0;
        HIP_TRY(hipGetDeviceCount(&ndev));
        PYSIC_REQUIRE(device >= 0 && device < ndev, PYSIC_EINVAL, "no such GPU");
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreate(&m->M.stream));
        m->M.P.ensure(MLP_NP);
        m->M.W.ensure(MLP_NPK);
        m->M.m1.ensure(MLP_NP);
        m->M.m2.ensure(MLP_NP);
        m->M.grad.ensure(MLP_NP);
        m->M.norm.ensure(2 * MLP_F);
    } catch (const PysicError &e) {
        m->M.release_all();
        delete m;
        return e.code;
    }
    *out = m;
    return PYSIC_OK;
}

void pysic_mlp_destroy(pysic_mlp *m)
{
    if (!m) return;
    (void)hipSetDevice(m->M.device);
    if (m->M.stream) (void)hipStreamSynchronize(m->M.stream);
    m->M.release_all();
    if (m->M.stream) (void)hipStreamDestroy(m->M.stream);
    delete m;
}

const char *pysic_mlp_last_error(pysic_mlp *m) { return m ? m->err.c_str() : "null handle"; }

int pysic_mlp_configure(pysic_mlp *m, int nspecies, double rc, const double *mu, const double *eta,
                        const float *mean, const float *inv_scale)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(nspecies == 1 || nspecies == 2, PYSIC_EINVAL, "descriptors take 1 or 2 species");
    PYSIC_REQUIRE(rc > 0.0 && mu && eta && mean && inv_scale, PYSIC_EINVAL,
                  "configure: bad arguments");
    MlpState &M = m->M;
    M.nspecies = nspecies;
    M.dp.rc = (float)rc;
    for (int g = 0; g < MLP_G; ++g) {
        M.dp.mu[g] = (float)mu[g];
        M.dp.eta[g] = (float)eta[g];
    }
    float nb[2 * MLP_F];
    std::memcpy(nb, mean, sizeof(float) * MLP_F);
    std::memcpy(nb + MLP_F, inv_scale, sizeof(float) * MLP_F);
    HIP_TRY(hipMemcpyAsync(M.norm.p, nb, sizeof(nb), hipMemcpyHostToDevice, M.stream));
    HIP_TRY(hipStreamSynchronize(M.stream));
    M.configured = true;
    MLP_END(m)
}

int pysic_mlp_set_params(pysic_mlp *m, const float *params)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(params, PYSIC_EINVAL, "set_params: null");
    MlpState &M = m->M;
    HIP_TRY(hipMemcpyAsync(M.P.p, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice, M.stream));
    HIP_TRY(hipMemsetAsync(M.m1.p, 0, sizeof(float) * MLP_NP, M.stream));
    HIP_TRY(hipMemsetAsync(M.m2.p, 0, sizeof(float) * MLP_NP, M.stream));
    M.adam_t = 0;
    mlp_pack(M);
    HIP_TRY(hipStreamSynchronize(M.stream));
    M.has_params = true;
    MLP_END(m)
}

int pysic_mlp_get_params(pysic_mlp *m, float *params)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(params, PYSIC_EINVAL, "get_params: null");
    mlp_require_params(m->M);
    HIP_TRY(hipMemcpyAsync(params, m->M.P.p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost,
                           m->M.stream));
    HIP_TRY(hipStreamSynchronize(m->M.stream));
    MLP_END(m)
}

int pysic_mlp_descriptors(pysic_mlp *m, pysic_ctx *h, int n, float *feat, int *species)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(h && feat && species && n == h->S.n, PYSIC_EINVAL, "descriptors: bad arguments");
    MlpState &M = m->M;
    try {
        mlp_system_descriptors(M, h);
    } catch (const PysicError &e) {
        h->err = e.msg;
        throw;
    }
    System &S = h->S;
    std::vector<float> f((size_t)n * MLP_F);
    std::vector<int> sp(n), id(n);
    HIP_TRY(hipMemcpyAsync(f.data(), M.sfeat.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipMemcpyAsync(sp.data(), M.sspec.p, sizeof(int) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipMemcpyAsync(id.data(), S.id.p, sizeof(int) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    for (int s = 0; s < n; ++s) {
        std::memcpy(feat + (size_t)id[s] * MLP_F, f.data() + (size_t)s * MLP_F, sizeof(float) * MLP_F);
        species[id[s]] = sp[s];
    }
    MLP_END(m)
}

int pysic_mlp_predict_charges(pysic_mlp *m, pysic_ctx *h, int n, int neutralize, double *charges)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(h && charges && n == h->S.n, PYSIC_EINVAL, "predict_charges: bad arguments");
    MlpState &M = m->M;
    mlp_require_params(M);
    try {
        mlp_system_descriptors(M, h);
    } catch (const PysicError &e) {
        h->err = e.msg;
        throw;
    }
    System &S = h->S;
    M.y.ensure(n);
    M.q.ensure(n);
    mlp_forward(M, n, nullptr, M.sfeat.p, M.sspec.p, M.y.p, S.stream);
    mlp_scatter_charges(S, M.y.p, M.q.p, S.stream);
    HIP_TRY(hipMemcpyAsync(charges, M.q.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    if (neutralize) {
        double mean = 0.0;
        for (int i = 0; i < n; ++i) mean += charges[i];
        mean /= n;
        for (int i = 0; i < n; ++i) charges[i] -= mean;
    }
    MLP_END(m)
}

int pysic_mlp_set_corrector(pysic_mlp *m, int n_stages, const double *coef)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(n_stages >= 0 && n_stages <= CORR_MAX_STAGES && (coef || n_stages == 0),
                  PYSIC_EINVAL, "set_corrector: bad arguments");
    MlpState &M = m->M;
    const size_t count = (size_t)n_stages * CORR_NC;
    for (size_t k = 0; k < count; ++k)
        PYSIC_REQUIRE(std::isfinite(coef[k]), PYSIC_EINVAL,
                      "set_corrector: non-finite coefficient");
    M.corr_stages = 0;
    if (n_stages > 0) {
        M.ccoef.ensure(count);
        HIP_TRY(hipMemcpyAsync(M.ccoef.p, coef, sizeof(double) * count, hipMemcpyHostToDevice,
                               M.stream));
        HIP_TRY(hipStreamSynchronize(M.stream));
    }
    M.corr_stages = n_stages;
    MLP_END(m)
}

int pysic_mlp_corrector_stages(pysic_mlp *m)
{
    return m ? m->M.corr_stages : PYSIC_EINVAL;
}

int pysic_mlp_corrector_features(pysic_mlp *m, pysic_ctx *h, int n, const double *charges,
                                 const double *previous, double *features, double *enegs)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(h && charges && previous && features && n == h->S.n, PYSIC_EINVAL,
                  "corrector_features: bad arguments");
    MlpState &M = m->M;
    try {
        // the same checks as the descriptors: configured, lists reach rc
        mlp_system_descriptors(M, h);
    } catch (const PysicError &e) {
        h->err = e.msg;
        throw;
    }
    System &S = h->S;
    upload_scalar(S, charges, S.q);
    M.cprev.ensure(n);
    upload_scalar(S, previous, M.cprev);
    h->forces_valid = false;
    DBuf<double> &chi = ewald_eem_chi(S, h->ewald, h->ew);
    M.cfeat.ensure((size_t)n * CORR_NC);
    mlp_corrector_features(S, M, chi.p, M.cprev.p, S.stream);
    std::vector<double> f((size_t)n * CORR_NC);
    std::vector<int> id(n);
    HIP_TRY(hipMemcpyAsync(f.data(), M.cfeat.p, sizeof(double) * f.size(), hipMemcpyDeviceToHost,
                           S.stream));
    HIP_TRY(hipMemcpyAsync(id.data(), S.id.p, sizeof(int) * n, hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    for (int s = 0; s < n; ++s)
        std::memcpy(features + (size_t)id[s] * CORR_NC, f.data() + (size_t)s * CORR_NC,
                    sizeof(double) * CORR_NC);
    if (enegs) download_scalar(S, chi.p, enegs);
    MLP_END(m)
}

int pysic_mlp_warm_start(pysic_mlp *m, pysic_ctx *h, int n, int n_stages, double *charges)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(h && charges && n == h->S.n, PYSIC_EINVAL, "warm_start: bad arguments");
    MlpState &M = m->M;
    mlp_require_params(M);
    PYSIC_REQUIRE(n_stages >= 0 && n_stages <= M.corr_stages, PYSIC_EINVAL,
                  "warm_start: more stages asked for than the corrector holds");
    try {
        mlp_system_descriptors(M, h);
    } catch (const PysicError &e) {
        h->err = e.msg;
        throw;
    }
    System &S = h->S;
    M.y.ensure(n);
    M.cprev.ensure(n);
    M.cnext.ensure(n);
    mlp_forward(M, n, nullptr, M.sfeat.p, M.sspec.p, M.y.p, S.stream);
    mlp_neutral_charges(S, M, M.y.p, M.cprev.p, S.stream);
    h->forces_valid = false;
    double *prev = M.cprev.p, *next = M.cnext.p;
    for (int k = 0; k < n_stages; ++k) {
        DBuf<double> &chi = ewald_eem_chi(S, h->ewald, h->ew);
        mlp_corrector_stage(S, M, M.ccoef.p + (size_t)k * CORR_NC, chi.p, prev, next, S.stream);
        std::swap(prev, next);
    }
    download_scalar(S, S.q.p, charges);
    MLP_END(m)
}

int pysic_mlp_set_training_data(pysic_mlp *m, int n, const float *feat, const int *species,
                                const float *targets)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(n > 0 && feat && species && targets, PYSIC_EINVAL, "training data: bad arguments");
    for (int i = 0; i < n; ++i)
        PYSIC_REQUIRE(species[i] >= 0 && species[i] < MLP_S, PYSIC_EINVAL,
                      "training species out of range");
    MlpState &M = m->M;
    M.feat.ensure((size_t)n * MLP_F);
    M.spec.ensure(n);
    M.targ.ensure(n);
    HIP_TRY(hipMemcpyAsync(M.feat.p, feat, sizeof(float) * (size_t)n * MLP_F, hipMemcpyHostToDevice, M.stream));
    HIP_TRY(hipMemcpyAsync(M.spec.p, species, sizeof(int) * n, hipMemcpyHostToDevice, M.stream));
    HIP_TRY(hipMemcpyAsync(M.targ.p, targets, sizeof(float) * n, hipMemcpyHostToDevice, M.stream));
    HIP_TRY(hipStreamSynchronize(M.stream));
    M.ntrain = n;
    MLP_END(m)
}

int pysic_mlp_evaluate(pysic_mlp *m, int count, const int *samples, float *y)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(y, PYSIC_EINVAL, "evaluate: null output");
    MlpState &M = m->M;
    mlp_require_params(M);
    mlp_upload_samples(M, count, samples);
    M.y.ensure(count);
    mlp_forward(M, count, M.order.p, M.feat.p, M.spec.p, M.y.p, M.stream);
    HIP_TRY(hipMemcpyAsync(y, M.y.p, sizeof(float) * count, hipMemcpyDeviceToHost, M.stream));
    HIP_TRY(hipStreamSynchronize(M.stream));
    MLP_END(m)
}

int pysic_mlp_gradients(pysic_mlp *m, int count, const int *samples, float *grad, float *loss)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(grad && loss, PYSIC_EINVAL, "gradients: null output");
    PYSIC_REQUIRE(count <= (1 << 18), PYSIC_EINVAL, "gradients: at most 262144 samples per call");
    MlpState &M = m->M;
    mlp_require_params(M);
    mlp_upload_samples(M, count, samples);
    M.losses.ensure(1);
    mlp_gradient(M, count, M.order.p, M.losses.p);
    HIP_TRY(hipMemcpyAsync(grad, M.grad.p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, M.stream));
    HIP_TRY(hipMemcpyAsync(loss, M.losses.p, sizeof(float), hipMemcpyDeviceToHost, M.stream));
    HIP_TRY(hipStreamSynchronize(M.stream));
    MLP_END(m)
}

int pysic_mlp_train(pysic_mlp *m, int nsteps, int batch, const int *order, float lr, float beta1,
                    float beta2, float eps, float *losses)
{
    MLP_BEGIN(m)
    PYSIC_REQUIRE(nsteps > 0 && batch > 0 && batch <= (1 << 18) && losses, PYSIC_EINVAL,
                  "train: bad arguments");
    MlpState &M = m->M;
    mlp_require_params(M);
    mlp_upload_samples(M, (long)nsteps * batch, order);
    M.losses.ensure(nsteps);
    for (int k = 0; k < nsteps; ++k) {
        mlp_gradient(M, batch, M.order.p + (size_t)k * batch, M.losses.p + k);
        mlp_adam(M, lr, beta1, beta2, eps);
    }
    HIP_TRY(hipMemcpyAsync(losses, M.losses.p, sizeof(float) * nsteps, hipMemcpyDeviceToHost, M.stream));
    HIP_TRY(hipStreamSynchronize(M.stream));
    MLP_END(m)
}

}  // extern "C"
