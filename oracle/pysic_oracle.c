/*
 * pysic_oracle.c -- FP64 CPU restatement of the pysic force core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product path (pysic_amd) never does.
 *
 * Every routine restates one routine of the reference Fortran core and cites
 * the reference file:line it follows (paths relative to the reference root).
 * The restatement keeps the reference's loop structure and floating-point
 * operation order where that order decides a result bit (the neighbor
 * distance test) and otherwise follows the same formulas.
 *
 * Parity pins (what checks this file, besides the reference build):
 *   neighbor lists   brute force O(N^2 * images) enumeration, symmetry
 *                    (i,j,o) <-> (j,i,-o), Fortran-reference fixtures
 *   Tersoff E/F/W    closed form of the perfect diamond lattice
 *                    (E/atom = -4.629593127667 eV), zero force by symmetry,
 *                    central finite differences of E (F = -dE/dr), stress
 *                    as dE/d(strain), Fortran-reference fixtures
 *   Ewald E/F/chi    Madelung energy of rocksalt, sigma invariance, finite
 *                    differences (F = -dE/dr, chi = -dE/dq), direct sums,
 *                    Fortran-reference fixtures
 *   EEM relaxation   analytic solution without Coulomb, damped-oscillator
 *                    reduction
 *
 * Conventions: positions are row-major (N,3); the cell is given as its three
 * vectors in rows (cell[3*a+c] = component c of vector a) and inv is the
 * inverse of the matrix whose COLUMNS are the vectors, stored row-major, so
 * that fractional = inv * r (Geometry.f90:531 relative_coordinates).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ENOMEM -1
#define ORC_ECAP -2
#define ORC_EINVAL -3

/* pi = 4 atan(1)  (Utility.f90:11) */
static double orc_pi(void) { return 4.0 * atan(1.0); }

/* ------------------------------------------------------------------------ */
/* geometry (Geometry.f90)                                                   */
/* ------------------------------------------------------------------------ */

/* separation_vector: r2 - r1 + sum_i offset_i * v_i, accumulated in that
 * order (Geometry.f90:449-466). */
static void orc_separation(const double *r1, const double *r2, const int *off,
                           const double *cell, double *sep)
{
    sep[0] = r2[0] - r1[0];
    sep[1] = r2[1] - r1[1];
    sep[2] = r2[2] - r1[2];
    for (int i = 0; i < 3; ++i) {
        double o = (double)off[i];
        sep[0] = sep[0] + o * cell[3 * i + 0];
        sep[1] = sep[1] + o * cell[3 * i + 1];
        sep[2] = sep[2] + o * cell[3 * i + 2];
    }
}

/* .o. dot product (Quaternions.f90:475-483) */
static double orc_dot(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* relative_coordinates: inv * r (Geometry.f90:527-534) */
static void orc_relative(const double *inv, const double *r, double *rel)
{
    for (int i = 0; i < 3; ++i)
        rel[i] = inv[3 * i + 0] * r[0] + inv[3 * i + 1] * r[1] + inv[3 * i + 2] * r[2];
}

/* wrapped_coordinates offset: -floor(relative) on periodic axes
 * (Geometry.f90:597-619).  Also returns the wrapped fractional coordinates
 * (relative + offset), which find_subcell_for_atom bins on
 * (Geometry.f90:884-898). */
static void orc_wrap(const double *inv, const int *pbc, const double *r, int *off,
                     double *wrapped_frac)
{
    double rel[3];
    orc_relative(inv, r, rel);
    for (int i = 0; i < 3; ++i) {
        off[i] = 0;
        if (pbc[i]) {
            off[i] = -(int)floor(rel[i]);
            rel[i] = rel[i] + off[i];
        }
        wrapped_frac[i] = rel[i];
    }
}

/* pick: true for exactly one of (i,j,o) and (j,i,-o)  (Geometry.f90:648-676) */
int orc_pick(int i1, int i2, const int *off)
{
    if (i2 > i1) return 1;
    if (i2 == i1) {
        if (off[0] > 0) return 1;
        if (off[0] == 0) {
            if (off[1] > 0) return 1;
            if (off[1] == 0 && off[2] > 0) return 1;
        }
    }
    return 0;
}

/* get_optimal_splitting (Geometry.f90:679-707) */
static void orc_splitting(const double *cell, double max_cut, int *splits)
{
    for (int i = 0; i < 3; ++i) {
        int v1 = (i + 1) % 3, v2 = (i + 2) % 3;
        const double *a = cell + 3 * v1, *b = cell + 3 * v2, *v = cell + 3 * i;
        double nrm[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]};
        double length = fabs(orc_dot(v, nrm)) / sqrt(orc_dot(nrm, nrm));
        if (max_cut < 0.1e-4)
            splits[i] = (int)floor(10.0 * length);
        else
            splits[i] = (int)floor(length / max_cut);
        if (splits[i] > 40) splits[i] = 40;
    }
}

/*
 * Neighbor lists: core_create_space_partitioning + core_build_neighbor_lists
 * (Core.f90:8130-8280, with divide_cell Geometry.f90:751-865 and
 * find_subcell_for_atom Geometry.f90:884-924).
 *
 * Produces both-ways lists in the reference's insertion order: per atom a
 * list of (j, offset) with |r_j + o.V - r_i|^2 < cutoff_i^2.  Output is CSR:
 * counts[N], and for atom i its entries start at start[i] = sum counts[<i];
 * idx[] holds 0-based atom indices, off[3*e..3*e+2] the integer image offsets.
 * Returns the total entry count, or ORC_ECAP with *needed set if cap is too
 * small (no output written in that case), or another negative code.
 */
long orc_neighbor_list(int n, const double *pos, const double *cell, const double *inv,
                       const int *pbc, const double *cutoffs, long cap, int *counts,
                       int *idx, int *off, long *needed)
{
    double max_cut = 0.0;
    for (int i = 0; i < n; ++i)
        if (i == 0 || cutoffs[i] > max_cut) max_cut = cutoffs[i];
    int sp[3];
    orc_splitting(cell, max_cut, sp);
    for (int a = 0; a < 3; ++a)
        if (sp[a] < 1) return ORC_EINVAL; /* the reference falls back to ASE lists here */

    /* subcells with padding layers 0 and split+1 (divide_cell) */
    int dim[3] = {sp[0] + 2, sp[1] + 2, sp[2] + 2};
    long ncell = (long)dim[0] * dim[1] * dim[2];
    int *sc_count = calloc(ncell, sizeof(int));
    int *sc_start = malloc((ncell + 1) * sizeof(int));
    int *sc_atoms = malloc((n > 0 ? n : 1) * sizeof(int));
    int *atom_sc = malloc(3 * (n > 0 ? n : 1) * sizeof(int));
    int *wrapoff = malloc(3 * (n > 0 ? n : 1) * sizeof(int));
    if (!sc_count || !sc_start || !sc_atoms || !atom_sc || !wrapoff) {
        free(sc_count); free(sc_start); free(sc_atoms); free(atom_sc); free(wrapoff);
        return ORC_ENOMEM;
    }
    for (int i = 0; i < n; ++i) {
        double wf[3];
        int o[3];
        orc_wrap(inv, pbc, pos + 3 * i, o, wf);
        for (int a = 0; a < 3; ++a) {
            wrapoff[3 * i + a] = o[a];
            int ind;
            if (pbc[a]) {
                ind = ((int)floor(wf[a] * sp[a])) % sp[a] + 1;
                /* Fortran mod keeps the sign of the dividend; floor of a
                 * wrapped coordinate is >= 0 so C % agrees */
            } else {
                ind = (int)floor(wf[a] * sp[a]) + 1;
                if (ind < 0) ind = 0;
                if (ind > sp[a] + 1) ind = sp[a] + 1;
            }
            atom_sc[3 * i + a] = ind;
        }
        long c = atom_sc[3 * i] + (long)dim[0] * (atom_sc[3 * i + 1] + (long)dim[1] * atom_sc[3 * i + 2]);
        sc_count[c]++;
    }
    sc_start[0] = 0;
    for (long c = 0; c < ncell; ++c) sc_start[c + 1] = sc_start[c] + sc_count[c];
    memset(sc_count, 0, ncell * sizeof(int));
    for (int i = 0; i < n; ++i) { /* atoms enter their subcell in index order */
        long c = atom_sc[3 * i] + (long)dim[0] * (atom_sc[3 * i + 1] + (long)dim[1] * atom_sc[3 * i + 2]);
        sc_atoms[sc_start[c] + sc_count[c]++] = i;
    }

    /* two passes: count, then fill (the reference appends to growing arrays) */
    int *cnt = calloc(n > 0 ? n : 1, sizeof(int));
    long *fill = NULL;
    long total = 0;
    if (!cnt) { free(sc_count); free(sc_start); free(sc_atoms); free(atom_sc); free(wrapoff); return ORC_ENOMEM; }
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            if (total > cap) {
                if (needed) *needed = total;
                total = ORC_ECAP;
                break;
            }
            fill = malloc((n > 0 ? n : 1) * sizeof(long));
            long s = 0;
            for (int i = 0; i < n; ++i) { fill[i] = s; s += cnt[i]; counts[i] = cnt[i]; }
        }
        for (int i1 = 0; i1 < n; ++i1) {
            const int *ci = atom_sc + 3 * i1;
            const int *w1 = wrapoff + 3 * i1;
            for (int kn = -1; kn <= 1; ++kn)
                for (int jn = -1; jn <= 1; ++jn)
                    for (int in = -1; in <= 1; ++in) {
                        int nb[3] = {ci[0] + in, ci[1] + jn, ci[2] + kn};
                        int noff[3] = {0, 0, 0};
                        int include = 1;
                        for (int a = 0; a < 3; ++a) {
                            if (pbc[a]) {
                                while (nb[a] > sp[a]) { nb[a] -= sp[a]; noff[a] += 1; }
                                while (nb[a] < 1) { nb[a] += sp[a]; noff[a] -= 1; }
                            } else if (nb[a] < 0 || nb[a] > sp[a] + 1) {
                                include = 0;
                                nb[a] = 0;
                            }
                        }
                        if (!include) continue;
                        long c = nb[0] + (long)dim[0] * (nb[1] + (long)dim[1] * nb[2]);
                        for (int t = sc_start[c]; t < sc_start[c + 1]; ++t) {
                            int i2 = sc_atoms[t];
                            if (!orc_pick(i1, i2, noff)) continue;
                            const int *w2 = wrapoff + 3 * i2;
                            int o[3] = {noff[0] - w1[0] + w2[0], noff[1] - w1[1] + w2[1],
                                        noff[2] - w1[2] + w2[2]};
                            double sep[3];
                            orc_separation(pos + 3 * i1, pos + 3 * i2, o, cell, sep);
                            double d2 = orc_dot(sep, sep);
                            if (d2 < cutoffs[i1] * cutoffs[i1]) {
                                if (pass == 0) { cnt[i1]++; total++; }
                                else {
                                    long e = fill[i1]++;
                                    idx[e] = i2;
                                    off[3 * e] = o[0]; off[3 * e + 1] = o[1]; off[3 * e + 2] = o[2];
                                }
                            }
                            if (d2 < cutoffs[i2] * cutoffs[i2]) {
                                if (pass == 0) { cnt[i2]++; total++; }
                                else {
                                    long e = fill[i2]++;
                                    idx[e] = i1;
                                    off[3 * e] = -o[0]; off[3 * e + 1] = -o[1]; off[3 * e + 2] = -o[2];
                                }
                            }
                        }
                    }
        }
    }
    free(fill); free(cnt);
    free(sc_count); free(sc_start); free(sc_atoms); free(atom_sc); free(wrapoff);
    return total;
}

/* ------------------------------------------------------------------------ */
/* interaction model                                                         */
/* ------------------------------------------------------------------------ */

/* Potential descriptor, 12 doubles:
 *   type, n_targets, e1, e2 (species ids in target order, like
 *   apply_elements), p1, p2, cutoff, soft cutoff (cutoff - margin),
 *   group (-1: no bond-order coordinator), p3, p4, link.
 * type: 1 exponential, 2 charge_self, 3 LJ, 4 spring, 5 Buckingham, 6 power,
 *   7 tabulated, 8 shift_power, 9 step, 10 constant, 11 force; parameters
 *   p1..p4 in the order of the reference's characterizers
 *   (Potentials.f90:2909-4990).
 * link 0 starts a potential; link 1 makes the row a multiplier of the
 *   nearest preceding potential (a ProductPotential factor, which the
 *   reference adds with is_multiplier = .true. before its first factor,
 *   calculator.py:1197-1369, Potentials.f90:2353-2472): only the row's type
 *   and parameters count; targets, cutoffs and group are the potential's.
 * type 12 charge_exp (11 parameters) and 13 charge_abs (8): the row is
 *   followed by a continuation row, link 2: type, the species of the
 *   original first target (original_elements(1)), parameters 5..13. */
#define ORC_POT_STRIDE 12
#define ORC_MAXF 4
/* Bond-order descriptor ('tersoff' only), 14 doubles: e1 e2 e3 (center i,
 * j, k as original_elements), beta, eta, a, c, d, h, mu, cutoff, soft
 * cutoff, group, unused. */
#define ORC_BOND_STRIDE 14

/* one factor of a (product) potential */
typedef struct {
    int type;
    double p[13];
    const double *tab; /* tabulated: n rows of (value, derivative) */
    int tlen;
    int orig;          /* charge-dependent types: species of the original first target */
    double c0;         /* bond_bend: cos(theta_0)^n (Potentials.f90:3178-3185) */
    double eta[2], beta[2]; /* charge_exp derived parameters (Potentials.f90:3747-3766) */
} fac_t;
/* what a charge-dependent factor reads of its pair: the first atom's
 * species and both charges */
typedef struct {
    int sp1;
    double q1, q2;
} qpair_t;
typedef struct {
    int nt, e[3], group, smooth, nf;
    int o[3];          /* bond_bend: the original targets (o[1] the vertex) */
    double cut, soft;
    fac_t f[ORC_MAXF]; /* f[0] the potential, then its multipliers */
} pot_t;
typedef struct {
    int e[3], group;
    double beta, eta, a, c, d, h, mu, cut, soft;
} bond_t;

/* smoothening_factor / _derivative (Potentials.f90:442-509) */
static double orc_fc(double r, double hard, double soft)
{
    if (r > hard) return 0.0;
    if (r < soft) return 1.0;
    return 0.5 * (1.0 + cos(orc_pi() * (r - soft) / (hard - soft)));
}
static double orc_dfc(double r, double hard, double soft)
{
    if (r > hard) return 0.0;
    if (r < soft) return 0.0;
    return -0.5 * orc_pi() / (hard - soft) * sin(orc_pi() * (r - soft) / (hard - soft));
}

/* Tables of the 'tabulated' potential, by id.  The reference reads them from
 * table_<id>.txt when the potential is created (Potentials.f90:2462-2465,
 * read_table Utility.f90:384-433): column 1 the value, column 2 the
 * derivative with respect to r / range. */
#define ORC_MAXTAB 64
static struct {
    int id, n;
    double *v;
} orc_tables[ORC_MAXTAB];
static int orc_ntables;

int orc_set_table(int id, int n, const double *values)
{
    if (n < 2 || !values) return ORC_EINVAL;
    double *v = malloc(2 * (size_t)n * sizeof(double));
    if (!v) return ORC_ENOMEM;
    memcpy(v, values, 2 * (size_t)n * sizeof(double));
    for (int k = 0; k < orc_ntables; ++k)
        if (orc_tables[k].id == id) {
            free(orc_tables[k].v);
            orc_tables[k].v = v;
            orc_tables[k].n = n;
            return ORC_OK;
        }
    if (orc_ntables == ORC_MAXTAB) { free(v); return ORC_ENOMEM; }
    orc_tables[orc_ntables].id = id;
    orc_tables[orc_ntables].n = n;
    orc_tables[orc_ntables].v = v;
    orc_ntables++;
    return ORC_OK;
}

/* Interval of the tabulation for r < range: the reference's 1-based
 * lower_index = floor(r/dr) + 1 (Potentials.f90:4245, 4286), returned 0-based
 * and kept inside the table (the reference would read past its end if r/dr
 * rounded up to n - 1). */
static int orc_table_interval(const fac_t *f, double r, double *dr, double *t)
{
    *dr = f->p[1] / (f->tlen - 1);
    int lower = (int)floor(r / *dr) + 1;
    if (lower > f->tlen - 1) lower = f->tlen - 1;
    *t = r / *dr - (lower - 1);
    return lower - 1;
}

/* evaluate_energy_component for one factor of a pair potential at distance r
 * (Potentials.f90:2838-2890); cut is the potential's cutoff (the spring's
 * energy shift uses it). */
/* The per-atom parameters of a charge-dependent pair term as the reference
 * picks them: atom 1 takes parameter set 1 when its species is the
 * potential's original first target, set 2 otherwise (inverse_params,
 * Potentials.f90:3797-3829, 4580-4626). */
static int orc_charge_inverse(const fac_t *f, const qpair_t *qp)
{
    return qp->sp1 != f->orig;
}

/* evaluate_energy_charge_exp (Potentials.f90:3847-3903) and
 * _electronegativity_charge_exp (3776-3836); x1, x2 = -dE/dq1, -dE/dq2 */
static double orc_charge_exp(const fac_t *f, const qpair_t *qp, double *x1, double *x2)
{
    const double *p = f->p;
    double r2, r3, r4, r5, r6, r7, r8, r9, d3, d4;
    if (orc_charge_inverse(f, qp)) {
        r2 = p[5]; r3 = p[1]; r4 = p[7]; r5 = p[3];
        r6 = f->eta[1]; r7 = f->eta[0]; r8 = f->beta[1]; r9 = f->beta[0];
        d3 = p[10]; d4 = p[9];
    } else {
        r2 = p[1]; r3 = p[5]; r4 = p[3]; r5 = p[7];
        r6 = f->eta[0]; r7 = f->eta[1]; r8 = f->beta[0]; r9 = f->beta[1];
        d3 = p[9]; d4 = p[10];
    }
    double d5 = r8 * (r4 - qp->q1);
    double d6 = r9 * (r5 - qp->q2);
    double d1 = r2 + pow(fabs(d5), r6);
    double d2 = r3 + pow(fabs(d6), r7);
    double e = p[0] * exp(0.5 * (d3 * d1 + d4 * d2));
    if (x1) {
        *x2 = e * d4 * 0.5 * r7 * pow(fabs(d6), r7 - 1.0) * copysign(1.0, d6) * r9;
        *x1 = e * d3 * 0.5 * r6 * pow(fabs(d5), r6 - 1.0) * copysign(1.0, d5) * r8;
    }
    return e;
}

/* evaluate_energy_charge_pair_abs (Potentials.f90:4680-4731) and
 * _electronegativity_charge_pair_abs (4571-4670) */
static double orc_charge_abs(const fac_t *f, const qpair_t *qp, double *x1, double *x2)
{
    const double *p = f->p;
    int inv = orc_charge_inverse(f, qp);
    const double *s1 = inv ? p + 4 : p, *s2 = inv ? p : p + 4;
    double a1 = s1[0], b1 = s1[1], Q1 = s1[2], n1 = s1[3];
    double a2 = s2[0], b2 = s2[1], Q2 = s2[2], n2 = s2[3];
    if (x1) {
        double abs1 = fabs(qp->q1 - Q1), abs2 = fabs(qp->q2 - Q2);
        double pow1 = pow(abs1, n1 - 1.0), pow2 = pow(abs2, n2 - 1.0);
        double f1 = a1 + b1 * abs1 * pow1;
        double f2 = a2 + b2 * abs2 * pow2;
        double dd1 = -n1 * pow1 * copysign(1.0, qp->q1 - Q1) * b1;
        double dd2 = -n2 * pow2 * copysign(1.0, qp->q2 - Q2) * b2;
        double inv_sqrt = 0.5 / sqrt(f1 * f2);
        if (f1 < 0.0) { f1 = 0.0; inv_sqrt = 0.0; }
        if (f2 < 0.0) { f2 = 0.0; inv_sqrt = 0.0; }
        *x1 = inv_sqrt * dd1 * f2;
        *x2 = inv_sqrt * dd2 * f1;
    }
    double f1 = a1 + b1 * pow(fabs(qp->q1 - Q1), n1);
    double f2 = a2 + b2 * pow(fabs(qp->q2 - Q2), n2);
    if (f1 < 0.0) f1 = 0.0;
    if (f2 < 0.0) f2 = 0.0;
    return sqrt(f1 * f2);
}

/* x**n for an integer n, as Fortran's real**integer (n < 0: 1 / x**-n) */
static double orc_ipow(double x, int n)
{
    int m = n < 0 ? -n : n;
    double y = 1.0;
    for (int k = 0; k < m; ++k) y = y * x;
    return n < 0 ? 1.0 / y : y;
}

/* evaluate_energy_charge_pair (Potentials.f90:4469-4506) and
 * _electronegativity_charge_pair (4415-4462): E = eps q1^n1 q2^n2 with
 * integer exponents int(n1), int(n2), swapped as for charge_exp */
static double orc_charge_pair(const fac_t *f, const qpair_t *qp, double *x1, double *x2)
{
    const double *p = f->p;
    int inv = orc_charge_inverse(f, qp);
    int n1 = (int)(inv ? p[2] : p[1]), n2 = (int)(inv ? p[1] : p[2]);
    if (x1) {
        *x1 = n1 > 0 ? -p[0] * (n1 * orc_ipow(qp->q1, n1 - 1) * orc_ipow(qp->q2, n2)) : 0.0;
        *x2 = n2 > 0 ? -p[0] * (n2 * orc_ipow(qp->q1, n1) * orc_ipow(qp->q2, n2 - 1)) : 0.0;
    }
    return p[0] * (orc_ipow(qp->q1, n1) * orc_ipow(qp->q2, n2));
}

static int orc_is_charge(int type)
{
    return type >= 12 && type <= 14;
}

static double orc_pair_energy(const fac_t *f, double r, double cut, const qpair_t *qp)
{
    const double *p = f->p;
    switch (f->type) {
    case 1: /* evaluate_energy_exp, 3666-3677 */
        return p[0] * exp(-p[1] * r);
    case 3: { /* evaluate_energy_LJ, 2993-3005 */
        double ratio = p[1] / r;
        double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
        return p[0] * (r6 * r6 - r6);
    }
    case 4: { /* evaluate_energy_spring, 3069-3085 */
        double r2 = r - p[1], r6 = cut - p[1];
        if (r2 * r2 < r6 * r6) return p[0] * 0.5 * (r2 * r2 - r6 * r6);
        return 0.0;
    }
    case 5: { /* evaluate_energy_buckingham, 4033-4045 */
        double ratio = r / p[2];
        double r6 = 1.0 / (ratio * ratio * ratio * ratio * ratio * ratio);
        return p[0] * exp(-ratio) - p[1] * r6;
    }
    case 6: /* evaluate_energy_power, 4141-4152 */
        return p[0] * pow(p[1] / r, p[2]);
    case 7: { /* evaluate_energy_table, 4267-4298; constant (unscaled) past the range */
        if (r >= p[1]) return f->tab[2 * (f->tlen - 1)];
        double dr, t;
        int lo = orc_table_interval(f, r, &dr, &t);
        double t2 = t * t, t3 = t2 * t, mr = p[1];
        const double *a = f->tab + 2 * lo, *b = a + 2;
        return p[2] * (a[0] * (2 * t3 - 3 * t2 + 1) + a[1] * dr / mr * (t3 - 2 * t2 + t) +
                       b[0] * (3 * t2 - 2 * t3) + b[1] * dr / mr * (t3 - t2));
    }
    case 8: /* evaluate_energy_shift, 4827-4842 */
        return p[0] * pow((p[1] - r) / (p[1] - p[2]), p[3]);
    case 12:
        return orc_charge_exp(f, qp, NULL, NULL);
    case 13:
        return orc_charge_abs(f, qp, NULL, NULL);
    case 14:
        return orc_charge_pair(f, qp, NULL, NULL);
    case 9: /* evaluate_energy_step, 4946-4967 */
        if (r < p[2]) return p[0];
        if (r > p[3]) return p[1];
        return (p[0] + p[1]) * 0.5 * (1.0 + 0.5 * cos(orc_pi() * (r - p[2]) / (p[3] - p[2])));
    }
    return 0.0;
}

/* evaluate_force_component for one factor: the force on the first atom,
 * f1 (the second gets -f1); zero at r == 0 (Potentials.f90:2717-2730). */
static void orc_pair_force(const fac_t *f, const double *sep, double r, double cut, double *f1)
{
    const double *p = f->p;
    f1[0] = f1[1] = f1[2] = 0.0;
    if (r == 0.0) return;
    switch (f->type) {
    case 1: /* evaluate_force_exp, 3643-3655 */
        for (int a = 0; a < 3; ++a) f1[a] = -p[1] * p[0] * exp(-p[1] * r) * sep[a] / r;
        return;
    case 3: { /* evaluate_force_LJ, 2971-2984 */
        double ratio = p[1] / r;
        double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
        for (int a = 0; a < 3; ++a) f1[a] = p[0] * (6.0 * r6 - 12.0 * r6 * r6) * sep[a] / (r * r);
        return;
    }
    case 4: { /* evaluate_force_spring, 3042-3060 */
        double r2 = r - p[1], r6 = cut - p[1];
        if (r2 * r2 < r6 * r6)
            for (int a = 0; a < 3; ++a) f1[a] = p[0] * r2 * sep[a] / r;
        return;
    }
    case 5: { /* evaluate_force_buckingham, 4009-4023 */
        double ratio = r / p[2];
        double r6 = 1.0 / (ratio * ratio * ratio * ratio * ratio * ratio);
        for (int a = 0; a < 3; ++a)
            f1[a] = (-p[0] * ratio * exp(-ratio) + p[1] * 6.0 * r6) * sep[a] / (r * r);
        return;
    }
    case 6: { /* evaluate_force_power, 4119-4132 */
        double n = p[2], ratio = p[1] / r;
        for (int a = 0; a < 3; ++a) f1[a] = p[0] * (-n * pow(ratio, n)) * sep[a] / (r * r);
        return;
    }
    case 7: { /* evaluate_force_table, 4227-4258; zero past the range */
        if (r >= p[1]) return;
        double dr, t;
        int lo = orc_table_interval(f, r, &dr, &t);
        double t2 = t * t, mr = p[1];
        const double *a = f->tab + 2 * lo, *b = a + 2;
        double s = a[0] / dr * 6 * (t2 - t) + a[1] / mr * (3 * t2 - 4 * t + 1) +
                   b[0] / dr * 6 * (t - t2) + b[1] / mr * (3 * t2 - 2 * t);
        for (int c = 0; c < 3; ++c) f1[c] = p[2] * s / r * sep[c];
        return;
    }
    case 8: { /* evaluate_force_shift, 4803-4818 */
        double r1 = p[1], r2 = p[2], n = p[3];
        for (int a = 0; a < 3; ++a)
            f1[a] = -p[0] * n / (r1 - r2) * pow((r1 - r) / (r1 - r2), n - 1) * sep[a] / r;
        return;
    }
    case 9: /* evaluate_force_step, 4917-4937: the same value in every
             * component (the reference does not project onto r) */
        if (r < p[2] || r > p[3]) return;
        for (int a = 0; a < 3; ++a)
            f1[a] = (p[0] + p[1]) * 0.25 * orc_pi() / (p[3] - p[2]) *
                    sin(orc_pi() * (r - p[2]) / (p[3] - p[2]));
        return;
    }
}

/* evaluate_energy / evaluate_forces over the factors of a product
 * (Potentials.f90:2640-2693, 2791-2814): E = prod E_i and
 * F = sum_i (E / E_i) F_i, or prod_{j != i} E_j F_i where E_i = 0. */
static double orc_product_energy(const pot_t *P, double r, const qpair_t *qp)
{
    double e = 1.0;
    for (int i = 0; i < P->nf; ++i) e = e * orc_pair_energy(P->f + i, r, P->cut, qp);
    return e;
}
static void orc_product_force(const pot_t *P, const double *sep, double r, const qpair_t *qp,
                              double *f1)
{
    if (P->nf == 1) {
        orc_pair_force(P->f, sep, r, P->cut, f1);
        return;
    }
    double me[ORC_MAXF], mf[ORC_MAXF][3], e = 1.0;
    for (int i = 0; i < P->nf; ++i) {
        me[i] = orc_pair_energy(P->f + i, r, P->cut, qp);
        orc_pair_force(P->f + i, sep, r, P->cut, mf[i]);
    }
    for (int i = 0; i < P->nf; ++i) e = e * me[i];
    f1[0] = f1[1] = f1[2] = 0.0;
    for (int i = 0; i < P->nf; ++i) {
        if (me[i] != 0.0) {
            for (int a = 0; a < 3; ++a) f1[a] = f1[a] + e / me[i] * mf[i][a];
        } else {
            double t = 1.0;
            for (int j = 0; j < P->nf; ++j)
                if (j != i) t = t * me[j];
            for (int a = 0; a < 3; ++a) f1[a] = f1[a] + t * mf[i][a];
        }
    }
}

/* evaluate_electronegativity over the factors of a product (Potentials.f90:
 * 2580-2637): factors without charge dependence contribute nothing;
 * X = sum_i (E / E_i) X_i, or prod_{j != i} E_j X_i where E_i = 0. */
static void orc_pair_eneg(const fac_t *f, const qpair_t *qp, double *x)
{
    x[0] = x[1] = 0.0;
    if (f->type == 12) orc_charge_exp(f, qp, x, x + 1);
    if (f->type == 13) orc_charge_abs(f, qp, x, x + 1);
    if (f->type == 14) orc_charge_pair(f, qp, x, x + 1);
}
static void orc_product_eneg(const pot_t *P, double r, const qpair_t *qp, double *x)
{
    if (P->nf == 1) {
        orc_pair_eneg(P->f, qp, x);
        return;
    }
    double me[ORC_MAXF], mx[ORC_MAXF][2], e = 1.0;
    for (int i = 0; i < P->nf; ++i) {
        me[i] = orc_pair_energy(P->f + i, r, P->cut, qp);
        orc_pair_eneg(P->f + i, qp, mx[i]);
    }
    for (int i = 0; i < P->nf; ++i) e = e * me[i];
    x[0] = x[1] = 0.0;
    for (int i = 0; i < P->nf; ++i) {
        if (me[i] != 0.0) {
            for (int a = 0; a < 2; ++a) x[a] = x[a] + e / me[i] * mx[i][a];
        } else {
            double t = 1.0;
            for (int j = 0; j < P->nf; ++j)
                if (j != i) t = t * me[j];
            for (int a = 0; a < 2; ++a) x[a] = x[a] + t * mx[i][a];
        }
    }
}

/* post_process_bond_order_factor_tersoff (Potentials.f90:5910-5921) */
static double orc_tersoff_bt(double zeta, const bond_t *b)
{
    return pow(1.0 + pow(b->beta * zeta, b->eta), -1.0 / (2.0 * b->eta));
}
/* post_process_bond_order_gradient_tersoff applied to a unit gradient
 * (Potentials.f90:5929-5961): d(bt)/d(zeta) with the zeta == 0 branch */
static double orc_tersoff_dbt(double zeta, const bond_t *b)
{
    double beta = b->beta, eta = b->eta;
    if (eta == 0.0) return 0.0;
    if (zeta == 0.0 && eta <= 1.0) return 0.0;
    double inv_eta = -1.0 / (2.0 * eta);
    return inv_eta * pow(1.0 + pow(beta * zeta, eta), inv_eta - 1.0) * eta *
           pow(beta * zeta, eta - 1.0) * beta;
}

/* evaluate_pair_bond_order_factor_tersoff (Potentials.f90:5744-5795).
 * s1 = separations(:,1) (from the pair partner to the center), s2 = center
 * to k; r1, r2 their norms. */
static double orc_tersoff_term(const double *s1, double r1, const double *s2, double r2,
                               const bond_t *b)
{
    double t1[3] = {-s1[0], -s1[1], -s1[2]};
    double cosine = orc_dot(t1, s2) / (r1 * r2);
    double cc = b->c * b->c, dd = b->d * b->d, h = b->h;
    double decay = orc_fc(r2, b->cut, b->soft);
    double xi = decay * exp(pow(b->a * (r1 - r2), b->mu));
    double gee = 1 + cc / dd - cc / (dd + (h - cosine) * (h - cosine));
    return xi * gee;
}

/* evaluate_pair_bond_order_gradient_tersoff (Potentials.f90:5806-5902).
 * grad[0..2]: w.r.t. the pair partner (atoms(1)), grad[3..5]: the center
 * (atoms(2)), grad[6..8]: k (atoms(3)). */
static void orc_tersoff_grad(const double *s1, double r1, const double *s2, double r2,
                             const bond_t *b, double *grad)
{
    double t1[3] = {-s1[0], -s1[1], -s1[2]}, t2[3] = {s2[0], s2[1], s2[2]};
    double u1[3], u2[3];
    for (int c = 0; c < 3; ++c) { u1[c] = t1[c] / r1; u2[c] = t2[c] / r2; }
    double dot = orc_dot(t1, t2);
    double ratio = 1.0 / (r1 * r2);
    double cosine = dot * ratio;
    /* nablar1: partner +u1, center -u1; nablar2: center -u2, k +u2 */
    double nr1[9] = {u1[0], u1[1], u1[2], -u1[0], -u1[1], -u1[2], 0, 0, 0};
    double nr2[9] = {0, 0, 0, -u2[0], -u2[1], -u2[2], u2[0], u2[1], u2[2]};
    double t3[3], t4[3], ncos[9];
    for (int c = 0; c < 3; ++c) {
        t3[c] = t2[c] - dot / (r1 * r1) * t1[c];
        t4[c] = t1[c] - dot / (r2 * r2) * t2[c];
        ncos[c] = t3[c] * ratio;
        ncos[3 + c] = -(t3[c] + t4[c]) * ratio;
        ncos[6 + c] = t4[c] * ratio;
    }
    double mu = b->mu, a1 = b->a, cc = b->c * b->c, dd = b->d * b->d, h1 = b->h;
    double decay = orc_fc(r2, b->cut, b->soft);
    double dfc = orc_dfc(r2, b->cut, b->soft);
    double t6[3] = {u2[0] * dfc, u2[1] * dfc, u2[2] * dfc};
    double tm2[9] = {0, 0, 0, -t6[0], -t6[1], -t6[2], t6[0], t6[1], t6[2]};
    double e1a = pow(a1 * (r1 - r2), mu - 1);
    double e1b = (a1 * (r1 - r2)) * e1a;
    double xi = exp(e1b);
    double den = dd + (h1 - cosine) * (h1 - cosine);
    double gee = 1 + cc / dd - cc / den;
    double nxi[9], ngee[9];
    for (int q = 0; q < 9; ++q)
        nxi[q] = tm2[q] * xi + decay * (exp(e1b) * a1 * mu * e1a) * (nr1[q] - nr2[q]);
    xi = xi * decay;
    for (int q = 0; q < 9; ++q)
        ngee[q] = -cc / (den * den) * 2.0 * (h1 - cosine) * ncos[q];
    for (int q = 0; q < 9; ++q) grad[q] = nxi[q] * gee + xi * ngee[q];
}

typedef struct {
    int n;
    const int *sp;
    const double *pos, *q, *cell, *inv;
    const int *nbstart, *nbidx, *nboff;
    int npot, nbond;
    pot_t *pot;
    bond_t *bond;
    /* per-species lists, built from the descriptors (Core.f90:1034-1125) */
    double *maxpot, *maxbond; /* indexed by atom */
} sys_t;

static int orc_unpack(sys_t *S, int npot, const double *pd, int nbond, const double *bd)
{
    S->npot = npot;
    S->nbond = nbond;
    S->pot = calloc(npot > 0 ? npot : 1, sizeof(pot_t));
    S->bond = calloc(nbond > 0 ? nbond : 1, sizeof(bond_t));
    if (!S->pot || !S->bond) return ORC_ENOMEM;
    int m = 0;
    fac_t *last = NULL;   /* the factor a continuation row completes */
    pot_t *lastbend = NULL;
    for (int k = 0; k < npot; ++k) {
        const double *d = pd + ORC_POT_STRIDE * k;
        int type = (int)d[0];
        if (type < 1 || type > 15) return ORC_EINVAL;
        if (d[11] == 2.0 && type == 15) { /* the third target and the original targets */
            if (!lastbend) return ORC_EINVAL;
            lastbend->e[2] = (int)d[1];
            for (int q = 0; q < 3; ++q) lastbend->o[q] = (int)d[2 + q];
            lastbend = NULL;
            continue;
        }
        if (lastbend) return ORC_EINVAL;
        if (d[11] == 2.0) { /* parameters 5.. and original_elements(1) of the row above */
            if (!last || last->type != type || !orc_is_charge(type)) return ORC_EINVAL;
            for (int q = 0; q < 9; ++q) last->p[4 + q] = d[2 + q];
            last->orig = (int)d[1];
            if (type == 12) { /* calculate_derived_parameters_charge_exp, 3747-3766 */
                double *pp = last->p;
                last->eta[0] = log(pp[1] / (pp[1] - pp[2])) / log(pp[3] / (pp[3] - pp[4]));
                last->eta[1] = log(pp[5] / (pp[5] - pp[6])) / log(pp[7] / (pp[7] - pp[8]));
                last->beta[0] = pow(pp[2] - pp[1], 1.0 / last->eta[0]) / (pp[3] - pp[4]);
                last->beta[1] = pow(pp[6] - pp[5], 1.0 / last->eta[1]) / (pp[7] - pp[8]);
            }
            last = NULL;
            continue;
        }
        if (last) return ORC_EINVAL;   /* a charge term without its continuation row */
        fac_t fac;
        memset(&fac, 0, sizeof(fac));
        fac.type = type;
        fac.orig = -1;
        fac.p[0] = d[4]; fac.p[1] = d[5]; fac.p[2] = d[9]; fac.p[3] = d[10];
        if (type == 7) {
            int id = (int)d[4];
            for (int t = 0; t < orc_ntables; ++t)
                if (orc_tables[t].id == id) { fac.tab = orc_tables[t].v; fac.tlen = orc_tables[t].n; }
            if (!fac.tab) return ORC_EINVAL;
        }
        int single = type == 2 || type == 10 || type == 11;
        if (d[11] != 0.0) { /* a multiplier of the previous potential */
            if (m == 0) return ORC_EINVAL;
            pot_t *q = S->pot + m - 1;
            if (q->nf == ORC_MAXF || q->nt != 2 || single || type == 15) return ORC_EINVAL;
            q->f[q->nf++] = fac;
            if (orc_is_charge(type)) last = q->f + q->nf - 1;
            continue;
        }
        pot_t *p = S->pot + m++;
        p->nf = 1;
        p->f[0] = fac;
        if (orc_is_charge(type)) last = p->f;
        if (type == 15) {
            lastbend = p;
            /* calculate_derived_parameters_bond_bending (Potentials.f90:3178-3185) */
            p->f[0].c0 = pow(cos(p->f[0].p[1]), p->f[0].p[2]);
        }
        p->nt = (int)d[1];
        p->e[0] = (int)d[2];
        p->e[1] = (int)d[3];
        p->cut = d[6];
        /* create_potential soft-cutoff rule (Potentials.f90:2382-2388) */
        if (d[7] > 0.0 && d[7] < d[6]) { p->soft = d[7]; p->smooth = 1; }
        else { p->soft = d[6]; p->smooth = 0; }
        p->group = (int)d[8];
        if (p->nt != (single ? 1 : type == 15 ? 3 : 2)) return ORC_EINVAL;
        if (type == 15 && p->group > -1) return ORC_EINVAL;
    }
    if (last || lastbend) return ORC_EINVAL;
    S->npot = m;
    for (int k = 0; k < m; ++k) /* bond-order scaling: pair potentials (and products) */
        if (S->pot[k].group > -1 && S->pot[k].nt != 2) return ORC_EINVAL;
    for (int k = 0; k < nbond; ++k) {
        const double *d = bd + ORC_BOND_STRIDE * k;
        bond_t *b = S->bond + k;
        b->e[0] = (int)d[0]; b->e[1] = (int)d[1]; b->e[2] = (int)d[2];
        b->beta = d[3]; b->eta = d[4]; b->a = d[5]; b->c = d[6]; b->d = d[7];
        b->h = d[8]; b->mu = d[9]; b->cut = d[10]; b->soft = d[11];
        b->group = (int)d[12];
    }
    S->maxpot = calloc(S->n > 0 ? S->n : 1, sizeof(double));
    S->maxbond = calloc(S->n > 0 ? S->n : 1, sizeof(double));
    if (!S->maxpot || !S->maxbond) return ORC_ENOMEM;
    for (int i = 0; i < S->n; ++i) {
        double mp = 0.0, mb = 0.0;
        for (int k = 0; k < S->npot; ++k)
            if (S->pot[k].e[0] == S->sp[i] && S->pot[k].cut > mp) mp = S->pot[k].cut;
        for (int k = 0; k < nbond; ++k)
            if (S->bond[k].e[0] == S->sp[i] && S->bond[k].cut > mb) mb = S->bond[k].cut;
        S->maxpot[i] = mp;
        S->maxbond[i] = mb;
    }
    return ORC_OK;
}

static void orc_free(sys_t *S)
{
    free(S->pot); free(S->bond); free(S->maxpot); free(S->maxbond);
}

/* The first level-2, post-processing bond factor of the atom's species in
 * the group (Core.f90:2763-2807, 4993-5031). */
static const bond_t *orc_postproc(const sys_t *S, int species, int group)
{
    for (int k = 0; k < S->nbond; ++k)
        if (S->bond[k].e[0] == species && S->bond[k].group == group) return S->bond + k;
    return NULL;
}

/*
 * core_calculate_pair_bond_order_factor (Core.f90:2439-2637): raw zeta sums
 * for the pair (i1 -> i2) with separation sep (i1 to i2), both directions.
 * When grad != NULL this instead performs core_add_pair_bond_order_forces
 * (Core.f90:4952-5353): scatters -process_factor*prefactor*grad to forces
 * and accumulates the post-processed virial into stress.
 */
static void orc_pair_bond(const sys_t *S, int i1, int i2, const double *sep, double dist,
                          int group, double *sums, double prefactor, double *forces,
                          double *stress)
{
    const double *cell = S->cell;
    double s1[3], s2[3];
    double virial[2][6] = {{0}};
    double pf[2] = {0, 0};
    const bond_t *pp1 = NULL, *pp2 = NULL;
    if (forces) {
        pp1 = orc_postproc(S, S->sp[i1], group);
        pp2 = orc_postproc(S, S->sp[i2], group);
        if (pp1) pf[0] = orc_tersoff_dbt(sums[0], pp1);
        if (pp2) pf[1] = orc_tersoff_dbt(sums[1], pp2);
    } else {
        sums[0] = sums[1] = 0.0;
    }
    for (int side = 0; side < 2; ++side) {
        int c = side == 0 ? i1 : i2;     /* center atom */
        int p = side == 0 ? i2 : i1;     /* pair partner */
        for (int a = 0; a < 3; ++a) s1[a] = side == 0 ? -sep[a] : sep[a];
        for (int t = S->nbstart[c]; t < S->nbstart[c + 1]; ++t) {
            int k = S->nbidx[t];
            orc_separation(S->pos + 3 * c, S->pos + 3 * k, S->nboff + 3 * t, cell, s2);
            double d2 = sqrt(orc_dot(s2, s2));
            if (!(s1[0] + s2[0] != 0.0 || s1[1] + s2[1] != 0.0 || s1[2] + s2[2] != 0.0))
                continue;
            if (!(d2 < S->maxbond[c])) continue;
            for (int bi = 0; bi < S->nbond; ++bi) {
                const bond_t *b = S->bond + bi;
                if (b->e[0] != S->sp[c] || b->group != group) continue;
                if (b->e[1] != S->sp[p] || b->e[2] != S->sp[k]) continue;
                if (!forces) {
                    sums[side] += orc_tersoff_term(s1, dist, s2, d2, b);
                } else {
                    double g[9];
                    orc_tersoff_grad(s1, dist, s2, d2, b, g);
                    double scale = (side == 0 ? pp1 : pp2) ? pf[side] : 1.0;
                    for (int a = 0; a < 3; ++a) {
                        forces[3 * p + a] += -scale * prefactor * g[a];
                        forces[3 * c + a] += -scale * prefactor * g[3 + a];
                        forces[3 * k + a] += -scale * prefactor * g[6 + a];
                    }
                    double *v = virial[side];
                    v[0] += -s1[0] * g[0] + s2[0] * g[6];
                    v[1] += -s1[1] * g[1] + s2[1] * g[7];
                    v[2] += -s1[2] * g[2] + s2[2] * g[8];
                    v[3] += -s1[1] * g[2] + s2[1] * g[8];
                    v[4] += -s1[0] * g[2] + s2[0] * g[8];
                    v[5] += -s1[0] * g[1] + s2[0] * g[7];
                }
            }
        }
    }
    if (forces) {
        for (int q = 0; q < 6; ++q) {
            double t0 = pp1 ? pf[0] * virial[0][q] : virial[0][q];
            double t1 = pp2 ? pf[1] * virial[1][q] : virial[1][q];
            stress[q] -= prefactor * (t0 + t1);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Ewald reciprocal part (Potentials.f90:1428-2249)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int on;
    double rc, kc, sigma, eps;
    int K[3];
    const double *scaler;
} ewald_t;

static void orc_kvec(const double *rcell, int k1, int k2, int k3, double *kv)
{
    /* k1*b1 + k2*b2 + k3*b3, b_i = 2 pi (row i of inv)  (Potentials.f90:1529-1531) */
    for (int c = 0; c < 3; ++c)
        kv[c] = k1 * rcell[0 + c] + k2 * rcell[3 + c] + k3 * rcell[6 + c];
}

/* calc: 1 energy, 2 forces (+stress), 3 electronegativities.  Real-space
 * part is excluded (it runs inline in the pair loop, Core.f90:3416-3460). */
static int orc_ewald_kspace(const sys_t *S, const ewald_t *E, double volume, int calc,
                            double *energy, double *forces, double *stress, double *enegs)
{
    const int n = S->n;
    const int K1 = E->K[0], K2 = E->K[1], K3 = E->K[2];
    const long nk = (long)(2 * K1 + 1) * (2 * K2 + 1) * (K3 + 1);
    double *sf = calloc(2 * nk, sizeof(double));
    if (!sf) return ORC_ENOMEM;
    double pi = orc_pi();
    double rcell[9];
    for (int q = 0; q < 9; ++q) rcell[q] = 2 * pi * S->inv[q];
    double kcut2 = E->kc * E->kc;
    double inv_eps_4pi = 1.0 / (4.0 * pi * E->eps);
    double inv_eps_2v = 1.0 / (2.0 * volume * E->eps);
    double inv_sigma_sqrt_2pi = 1.0 / (sqrt(2.0 * pi) * E->sigma);
    double s2 = E->sigma * E->sigma;
    double qsum = 0.0, selfsum = 0.0;
#define SFI(a, b, c) ((((long)(a) + K1) * (2 * K2 + 1) + ((b) + K2)) * (K3 + 1) + (c))
    /* structure factors: S(k) = sum q e^{ik.r} (energy) or its conjugate
     * (forces, electronegativities) (Potentials.f90:1526-1548, 1812-1829) */
    double conj = calc == 1 ? 1.0 : -1.0;
    for (int i = 0; i < n; ++i) {
        double q = S->q[i] * E->scaler[i];
        if (calc != 3 && q == 0.0) continue;
        for (int k1 = -K1; k1 <= K1; ++k1)
            for (int k2 = -K2; k2 <= K2; ++k2)
                for (int k3 = 0; k3 <= K3; ++k3) {
                    if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                    double kv[3];
                    orc_kvec(rcell, k1, k2, k3, kv);
                    if (!(orc_dot(kv, kv) < kcut2)) continue;
                    double dot = orc_dot(kv, S->pos + 3 * i);
                    long t = SFI(k1, k2, k3);
                    sf[2 * t] += q * cos(dot);
                    sf[2 * t + 1] += conj * q * sin(dot);
                }
        selfsum += q * q;
        qsum += q;
    }
    if (calc == 1) {
        double erec = 0.0;
        for (int k1 = -K1; k1 <= K1; ++k1)
            for (int k2 = -K2; k2 <= K2; ++k2)
                for (int k3 = 0; k3 <= K3; ++k3) {
                    if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                    double kv[3];
                    orc_kvec(rcell, k1, k2, k3, kv);
                    double k2n = orc_dot(kv, kv);
                    if (!(k2n < kcut2)) continue;
                    double ex = exp(-s2 * k2n * 0.5) / k2n;
                    long t = SFI(k1, k2, k3);
                    double m = sf[2 * t] * sf[2 * t] + sf[2 * t + 1] * sf[2 * t + 1];
                    erec += (k3 == 0 ? 1.0 : 2.0) * ex * m;
                }
        /* Potentials.f90:1643-1651 (no dipole correction on the main path) */
        double e2 = -selfsum * inv_eps_4pi * inv_sigma_sqrt_2pi;
        double e3 = erec * inv_eps_2v;
        double e4 = -qsum * qsum * 0.5 * inv_eps_2v * s2;
        *energy = e2 + e3 + e4;
    } else if (calc == 2) {
        double st[6] = {0};
        for (int i = 0; i < n; ++i) {
            double q = S->q[i] * E->scaler[i];
            double f[3] = {0, 0, 0};
            if (q != 0.0) {
                for (int k1 = -K1; k1 <= K1; ++k1)
                    for (int k2 = -K2; k2 <= K2; ++k2)
                        for (int k3 = 0; k3 <= K3; ++k3) {
                            if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                            double kv[3];
                            orc_kvec(rcell, k1, k2, k3, kv);
                            double k2n = orc_dot(kv, kv);
                            if (!(k2n < kcut2)) continue;
                            double dot = orc_dot(kv, S->pos + 3 * i);
                            double sd = sin(dot), cd = cos(dot);
                            long t = SFI(k1, k2, k3);
                            double w = (k3 == 0 ? 2.0 : 4.0);
                            double pre = inv_eps_2v * exp(-s2 * k2n * 0.5) / k2n * w;
                            for (int c = 0; c < 3; ++c) {
                                double kq = q * kv[c];
                                double n1 = -sd * kq, n2 = cd * kq;
                                f[c] = f[c] - pre * (sf[2 * t] * n1 - sf[2 * t + 1] * n2);
                            }
                        }
            }
            for (int c = 0; c < 3; ++c) forces[3 * i + c] += f[c];
            /* reciprocal "stress" with absolute positions (Potentials.f90:1940-1945) */
            const double *r = S->pos + 3 * i;
            st[0] += r[0] * f[0];
            st[1] += r[1] * f[1];
            st[2] += r[2] * f[2];
            st[3] += r[1] * f[2];
            st[4] += r[0] * f[2];
            st[5] += r[0] * f[1];
        }
        for (int q = 0; q < 6; ++q) stress[q] += st[q];
    } else {
        for (int i = 0; i < n; ++i) {
            double q = S->q[i] * E->scaler[i];
            double x = 2.0 * q * inv_eps_4pi * inv_sigma_sqrt_2pi;  /* self, 2126 */
            x = x - E->scaler[i] * inv_eps_2v * (-s2 * qsum);      /* background, 2166 */
            for (int k1 = -K1; k1 <= K1; ++k1)
                for (int k2 = -K2; k2 <= K2; ++k2)
                    for (int k3 = 0; k3 <= K3; ++k3) {
                        if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                        double kv[3];
                        orc_kvec(rcell, k1, k2, k3, kv);
                        double k2n = orc_dot(kv, kv);
                        if (!(k2n < kcut2)) continue;
                        double dot = orc_dot(kv, S->pos + 3 * i);
                        long t = SFI(k1, k2, k3);
                        double w = (k3 == 0 ? 2.0 : 4.0);
                        x = x - inv_eps_2v * exp(-s2 * k2n * 0.5) / k2n * w *
                                    (sf[2 * t] * cos(dot) - sf[2 * t + 1] * sin(dot));
                    }
            enegs[i] += x;
        }
    }
#undef SFI
    free(sf);
    return ORC_OK;
}

/*
 * core_loop_over_local_interactions (Core.f90:3249-4082) for
 * calc = 1 (energy), 2 (forces + stress), 3 (electronegativities).
 *
 * Neighbor lists are CSR (nbstart[N+1], nbidx, nboff) as produced by
 * orc_neighbor_list.  ewald_desc = {on, rc, kc, K1, K2, K3, sigma, eps};
 * scaler may be NULL (all ones).
 */
/* bond_bend on the triplet (a, v, b) with v the vertex, s1 = r_v - r_a,
 * s2 = r_b - r_v (evaluate_energy_bond_bending and
 * evaluate_force_bond_bending, Potentials.f90:3196-3328): zero unless the
 * vertex is the original second target.  Returns E; with f, the forces on
 * a, v, b. */
static double orc_bend(const pot_t *P, int spv, const double *s1, double r1, const double *s2,
                       double r2, double f[3][3])
{
    const fac_t *fc = P->f;
    if (f) memset(f, 0, 9 * sizeof(double));
    if (P->o[1] != spv) return 0.0;
    int n = (int)fc->p[2], m = (int)fc->p[3];
    double t1[3] = {-s1[0], -s1[1], -s1[2]};
    double r3 = orc_dot(t1, s2);
    double e = fc->p[0] * orc_ipow(orc_ipow(r3 / (r1 * r2), n) - fc->c0, m);
    if (f) {
        double t3[3], t4[3];
        for (int a = 0; a < 3; ++a) t3[a] = s2[a] - r3 / (r1 * r1) * t1[a];
        for (int a = 0; a < 3; ++a) t4[a] = t1[a] - r3 / (r2 * r2) * s2[a];
        double ratio = 1.0 / (r1 * r2);
        double cosplain = r3 * ratio;
        double cospower = orc_ipow(cosplain, n - 1);
        double r6 = fc->p[0] * (double)(m * n) * orc_ipow(cospower * cosplain - fc->c0, m - 1) *
                    cospower;
        for (int a = 0; a < 3; ++a) {
            f[0][a] = -t3[a] * ratio * r6;
            f[1][a] = (t3[a] + t4[a]) * ratio * r6;
            f[2][a] = -t4[a] * ratio * r6;
        }
    }
    return e;
}

/* One triplet of core_evaluate_local_triplet_B (Core.f90:6656-7028) for
 * every many-body potential of the pair that matches it: atoms (a, v, b),
 * s1 = r_v - r_a with direction d1, s2 = r_b - r_v with direction d2.
 * Bond-order scaling of triplets is not supported (no coordinator). */
static void orc_triplet(const sys_t *S, int calc, const int *mb, int nmb, int ia, int iv, int ib,
                        const double *s1, const double *d1, double r1, const double *s2,
                        const double *d2, double r2, double *en, double *F, double *st)
{
    for (int k = 0; k < nmb; ++k) {
        const pot_t *P = S->pot + mb[k];
        if (P->e[2] != S->sp[ib] || !(P->cut > r2)) continue;
        double fc1 = 1.0, fc2 = 1.0;
        if (P->smooth) {
            fc1 = orc_fc(r1, P->cut, P->soft);
            fc2 = orc_fc(r2, P->cut, P->soft);
        }
        if (calc == 1) {
            double e = orc_bend(P, S->sp[iv], s1, r1, s2, r2, NULL);
            *en = *en + e * fc1 * fc2 * (1.0 + 1.0 + 1.0) / 3.0;
        } else if (calc == 2) {
            double e = P->smooth ? orc_bend(P, S->sp[iv], s1, r1, s2, r2, NULL) : 0.0;
            double cg1[3] = {0, 0, 0}, cg2[3] = {0, 0, 0};
            if (P->smooth) {
                double g1 = orc_dfc(r1, P->cut, P->soft), g2 = orc_dfc(r2, P->cut, P->soft);
                for (int a = 0; a < 3; ++a) { cg1[a] = d1[a] * g1; cg2[a] = d2[a] * g2; }
            }
            double f[3][3], tf[3][3];
            orc_bend(P, S->sp[iv], s1, r1, s2, r2, f);
            for (int a = 0; a < 3; ++a) {
                tf[0][a] = f[0][a] * fc1 * fc2 + cg1[a] * fc2 * e;
                tf[1][a] = f[1][a] * fc1 * fc2 + (-cg1[a] * fc2 + cg2[a] * fc1) * e;
                tf[2][a] = f[2][a] * fc1 * fc2 - cg2[a] * fc1 * e;
                F[3 * ia + a] += tf[0][a];
                F[3 * iv + a] += tf[1][a];
                F[3 * ib + a] += tf[2][a];
            }
            st[0] = st[0] + s1[0] * tf[1][0] + (s1[0] + s2[0]) * tf[2][0];
            st[1] = st[1] + s1[1] * tf[1][1] + (s1[1] + s2[1]) * tf[2][1];
            st[2] = st[2] + s1[2] * tf[1][2] + (s1[2] + s2[2]) * tf[2][2];
            st[3] = st[3] + s1[1] * tf[1][2] + (s1[1] + s2[1]) * tf[2][2];
            st[4] = st[4] + s1[0] * tf[1][2] + (s1[0] + s2[0]) * tf[2][2];
            st[5] = st[5] + s1[0] * tf[1][1] + (s1[0] + s2[0]) * tf[2][1];
        }
        /* bond_bend carries no charge: nothing for the electronegativities */
    }
}

int orc_evaluate(int calc, int n, const int *species, const double *pos, const double *charge,
                 const double *cell, const double *inv, const int *nbstart, const int *nbidx,
                 const int *nboff, int npot, const double *potdesc, int nbond,
                 const double *bonddesc, const double *ewald_desc, const double *scaler,
                 double *energy, double *forces, double *stress, double *enegs)
{
    sys_t S;
    memset(&S, 0, sizeof(S));
    S.n = n; S.sp = species; S.pos = pos; S.q = charge; S.cell = cell; S.inv = inv;
    S.nbstart = nbstart; S.nbidx = nbidx; S.nboff = nboff;
    int rc = orc_unpack(&S, npot, potdesc, nbond, bonddesc);
    if (rc) { orc_free(&S); return rc; }
    ewald_t E;
    memset(&E, 0, sizeof(E));
    double *ones = NULL;
    if (ewald_desc && ewald_desc[0] != 0.0) {
        E.on = 1; E.rc = ewald_desc[1]; E.kc = ewald_desc[2];
        E.K[0] = (int)ewald_desc[3]; E.K[1] = (int)ewald_desc[4]; E.K[2] = (int)ewald_desc[5];
        E.sigma = ewald_desc[6]; E.eps = ewald_desc[7];
        if (!scaler) {
            ones = malloc((n > 0 ? n : 1) * sizeof(double));
            for (int i = 0; i < n; ++i) ones[i] = 1.0;
            E.scaler = ones;
        } else E.scaler = scaler;
    }
    double pi = orc_pi();
    double volume;
    {
        const double *v1 = cell, *v2 = cell + 3, *v3 = cell + 6;
        double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                        v1[0] * v2[1] - v1[1] * v2[0]};
        volume = fabs(orc_dot(cr, v3));
    }
    double inv_eps_4pi = E.on ? 1.0 / (4.0 * pi * E.eps) : 0.0;
    double inv_sigma_sqrt_2 = E.on ? 1.0 / (sqrt(2.0) * E.sigma) : 0.0;
    double inv_sigma_sqrt_2pi = E.on ? 1.0 / (sqrt(2.0 * pi) * E.sigma) : 0.0;
    double inv_sigma_sq_2 = inv_sigma_sqrt_2 * inv_sigma_sqrt_2;
    double inv_sigma_sqrt_2perpi = 2.0 * inv_sigma_sqrt_2pi;

    double en = 0.0, st[6] = {0, 0, 0, 0, 0, 0};
    double *F = NULL, *X = NULL;
    if (calc == 2) { F = forces; memset(F, 0, 3 * (size_t)n * sizeof(double)); }
    if (calc == 3) { X = enegs; memset(X, 0, (size_t)n * sizeof(double)); }

    for (int i1 = 0; i1 < n; ++i1) {
        /* core_evaluate_local_singlet (Core.f90:4095-4239) */
        for (int k = 0; k < S.npot; ++k) {
            const pot_t *p = S.pot + k;
            if (p->nt != 1 || p->e[0] != species[i1]) continue;
            const double *pp = p->f[0].p;
            if (p->f[0].type == 2) {
                int nn = (int)pp[1];
                double q = charge[i1];
                /* Potentials.f90:4341-4372: eps*q^n, -n*eps*q^(n-1) (integer powers) */
                if (calc == 1) en += pp[0] * pow(q, nn);
                if (calc == 3) X[i1] += -pp[0] * nn * pow(q, nn - 1);
            } else if (p->f[0].type == 10) {
                /* constant potential (Potentials.f90:3942-3966): E = V, no force */
                if (calc == 1) en += pp[0];
            } else if (p->f[0].type == 11) {
                /* constant force (Potentials.f90:3122-3147): F, E = -F . r with
                 * the dot product of Quaternions.f90:475-483 */
                const double *r = pos + 3 * i1;
                if (calc == 1) en += (-pp[0] * r[0] + -pp[1] * r[1]) + -pp[2] * r[2];
                if (calc == 2)
                    for (int a = 0; a < 3; ++a) F[3 * i1 + a] += pp[a];
            }
        }
        for (int t = nbstart[i1]; t < nbstart[i1 + 1]; ++t) {
            int i2 = nbidx[t];
            const int *off = nboff + 3 * t;
            if (!orc_pick(i1, i2, off)) continue;
            double sep[3], dir[3];
            orc_separation(pos + 3 * i1, pos + 3 * i2, off, cell, sep);
            double dist = sqrt(orc_dot(sep, sep));
            if (dist == 0.0) dir[0] = dir[1] = dir[2] = 0.0;
            else for (int a = 0; a < 3; ++a) dir[a] = sep[a] / dist;

            /* inline real-space Ewald with unscaled charges (Core.f90:3416-3460) */
            if (E.on && dist < E.rc) {
                double q1 = charge[i1], q2 = charge[i2];
                if (calc == 1) {
                    en = en + q1 * q2 * (1.0 - erf(dist * inv_sigma_sqrt_2)) / dist * inv_eps_4pi;
                } else if (calc == 2) {
                    double id = 1 / dist;
                    double tf[3];
                    double pre = -inv_eps_4pi * q1 * q2 *
                                 ((1 - erf(dist * inv_sigma_sqrt_2)) * id * id +
                                  inv_sigma_sqrt_2perpi * exp(-dist * dist * inv_sigma_sq_2) * id);
                    for (int a = 0; a < 3; ++a) tf[a] = pre * sep[a] * id;
                    for (int a = 0; a < 3; ++a) { F[3 * i1 + a] += tf[a]; F[3 * i2 + a] -= tf[a]; }
                    st[0] -= sep[0] * tf[0];
                    st[1] -= sep[1] * tf[1];
                    st[2] -= sep[2] * tf[2];
                    st[3] -= sep[1] * tf[2];
                    st[4] -= sep[0] * tf[2];
                    st[5] -= sep[0] * tf[1];
                } else {
                    double te = -inv_eps_4pi / dist * (1.0 - erf(dist * inv_sigma_sqrt_2));
                    X[i1] += q2 * te;
                    X[i2] += q1 * te;
                }
            }
            if (!(dist < S.maxpot[i1])) continue;
            /* the many-body potentials the doublet routine finds for this
             * pair: first target atom 1, second atom 2, cutoff > r
             * (Core.f90:4293-4378) */
            int mb[64], nmb = 0;
            for (int k = 0; k < S.npot && nmb < 64; ++k) {
                const pot_t *p = S.pot + k;
                if (p->nt == 3 && p->e[0] == species[i1] && p->cut > dist &&
                    p->e[1] == species[i2])
                    mb[nmb++] = k;
            }
            /* doublets (Core.f90:4257-4385, 5371-5606) */
            for (int k = 0; k < S.npot; ++k) {
                const pot_t *p = S.pot + k;
                if (p->nt != 2 || p->e[0] != species[i1]) continue;
                if (!(p->cut > dist)) continue;
                if (p->e[1] != species[i2]) continue;
                double bo[2] = {0, 0}, sums[2] = {0, 0};
                if (p->group > -1) {
                    orc_pair_bond(&S, i1, i2, sep, dist, p->group, sums, 0, NULL, NULL);
                    const bond_t *b1 = orc_postproc(&S, species[i1], p->group);
                    const bond_t *b2 = orc_postproc(&S, species[i2], p->group);
                    bo[0] = b1 ? orc_tersoff_bt(sums[0], b1) : sums[0];
                    bo[1] = b2 ? orc_tersoff_bt(sums[1], b2) : sums[1];
                }
                double cutf = p->smooth ? orc_fc(dist, p->cut, p->soft) : 1.0;
                const qpair_t qp = {species[i1], charge[i1], charge[i2]};
                if (calc == 1) {
                    double v = orc_product_energy(p, dist, &qp);
                    if (p->group > -1) en = en + v * cutf * (bo[0] + bo[1]) * 0.5;
                    else en = en + v * cutf;
                } else if (calc == 2) {
                    double v = (p->group > -1 || p->smooth) ? orc_product_energy(p, dist, &qp) : 0.0;
                    double cg[3] = {0, 0, 0};
                    if (p->smooth) {
                        double d = orc_dfc(dist, p->cut, p->soft);
                        for (int a = 0; a < 3; ++a) cg[a] = dir[a] * d;
                    }
                    if (p->group > -1)
                        orc_pair_bond(&S, i1, i2, sep, dist, p->group, sums, 0.5 * v * cutf, F, st);
                    double f1[3];
                    orc_product_force(p, sep, dist, &qp, f1);
                    double pf1[3], pf2[3];
                    double bsum = (bo[0] + bo[1]) * 0.5;
                    for (int a = 0; a < 3; ++a) {
                        pf1[a] = f1[a] * cutf + v * cg[a];
                        pf2[a] = -f1[a] * cutf - v * cg[a];
                        if (p->group > -1) { pf1[a] *= bsum; pf2[a] *= bsum; }
                        F[3 * i1 + a] += pf1[a];
                        F[3 * i2 + a] += pf2[a];
                    }
                    st[0] += sep[0] * pf2[0];
                    st[1] += sep[1] * pf2[1];
                    st[2] += sep[2] * pf2[2];
                    st[3] += sep[1] * pf2[2];
                    st[4] += sep[0] * pf2[2];
                    st[5] += sep[0] * pf2[1];
                } else if (calc == 3) {
                    /* core_evaluate_local_doublet_electronegativities_B
                     * (Core.f90:5622-5758): charge_exp / charge_abs terms,
                     * times the smoothing factor and, for a coordinated
                     * potential, the pair's bond-order factor
                     * (b_i + b_j + B_ij + B_ji) / 2 with b_i = 0 for the
                     * level-2 'tersoff' factor (Core.f90:5717-5729) */
                    double x[2];
                    orc_product_eneg(p, dist, &qp, x);
                    double bsum = p->group > -1 ? (bo[0] + bo[1]) * 0.5 : 1.0;
                    X[i1] += x[0] * cutf * bsum;
                    X[i2] += x[1] * cutf * bsum;
                }
            }
            if (nmb == 0 || calc == 3) continue;
            /* triplets (Core.f90:3567-3610, 3760-3800): atom 1 as the vertex,
             * (atom2, atom1, atom3) with pick(atom2, atom3); then atom 2 as
             * the vertex, (atom1, atom2, atom3) with pick(atom1, atom3) */
            double ms[3] = {-sep[0], -sep[1], -sep[2]}, md[3] = {-dir[0], -dir[1], -dir[2]};
            for (int t3 = nbstart[i1]; t3 < nbstart[i1 + 1]; ++t3) {
                int i3 = nbidx[t3];
                const int *o3 = nboff + 3 * t3;
                int toff[3] = {o3[0] - off[0], o3[1] - off[1], o3[2] - off[2]};
                if (!orc_pick(i2, i3, toff)) continue;
                double s2[3], d2[3];
                orc_separation(pos + 3 * i1, pos + 3 * i3, o3, cell, s2);
                double r2 = sqrt(orc_dot(s2, s2));
                if (r2 == 0.0) d2[0] = d2[1] = d2[2] = 0.0;
                else for (int a = 0; a < 3; ++a) d2[a] = s2[a] / r2;
                if (r2 < S.maxpot[i1])
                    orc_triplet(&S, calc, mb, nmb, i2, i1, i3, ms, md, dist, s2, d2, r2, &en, F,
                                st);
            }
            for (int t3 = nbstart[i2]; t3 < nbstart[i2 + 1]; ++t3) {
                int i3 = nbidx[t3];
                const int *o3 = nboff + 3 * t3;
                int toff[3] = {o3[0] + off[0], o3[1] + off[1], o3[2] + off[2]};
                if (!orc_pick(i1, i3, toff)) continue;
                double s2[3], d2[3];
                orc_separation(pos + 3 * i2, pos + 3 * i3, o3, cell, s2);
                double r2 = sqrt(orc_dot(s2, s2));
                if (r2 == 0.0) d2[0] = d2[1] = d2[2] = 0.0;
                else for (int a = 0; a < 3; ++a) d2[a] = s2[a] / r2;
                if (r2 < S.maxpot[i2])
                    orc_triplet(&S, calc, mb, nmb, i1, i2, i3, sep, dir, dist, s2, d2, r2, &en, F,
                                st);
            }
        }
    }
    if (E.on) {
        double ek = 0.0;
        rc = orc_ewald_kspace(&S, &E, volume, calc, &ek, F, st, X);
        en += ek;
    }
    if (calc == 1) *energy = en;
    if (calc == 2) for (int q = 0; q < 6; ++q) stress[q] = st[q];
    free(ones);
    orc_free(&S);
    return rc;
}

/*
 * ChargeRelaxation 'dynamic' (charges/relaxation.py:417-450): damped
 * dynamics of the charges driven by the electronegativity differences
 * chi - mean(chi).  rates (N) carries the charge velocities in and out
 * (relaxation.py:431).  Stops when max(dchi) <= tolerance or after n_steps.
 * Returns the number of steps taken; *err receives the last max(dchi).
 */
int orc_relax_dynamic(int n, const int *species, const double *pos, double *charge,
                      const double *cell, const double *inv, const int *nbstart,
                      const int *nbidx, const int *nboff, int npot, const double *potdesc,
                      const double *ewald_desc, const double *scaler, int n_steps, double dt,
                      double inertia, double friction, double tolerance, double *rates,
                      double *err)
{
    double dt2 = dt * dt, inv_mq = 1.0 / inertia;
    double future = 1.0 / (1.0 + 0.5 * friction * dt);
    double *chi = malloc((n > 0 ? n : 1) * sizeof(double));
    if (!chi) return ORC_ENOMEM;
    int rc;
#define DCHI()                                                                              \
    do {                                                                                    \
        rc = orc_evaluate(3, n, species, pos, charge, cell, inv, nbstart, nbidx, nboff,     \
                          npot, potdesc, 0, NULL, ewald_desc, scaler, NULL, NULL, NULL,     \
                          chi);                                                             \
        if (rc) { free(chi); return rc; }                                                   \
        double s = 0.0;                                                                     \
        for (int i = 0; i < n; ++i) s += chi[i];                                            \
        double avg = s / n;                                                                 \
        for (int i = 0; i < n; ++i) chi[i] = chi[i] - avg;                                  \
    } while (0)
    DCHI();
    double error = 2.0 * tolerance;
    int step = 0;
    while (error > tolerance && step < n_steps) {
        for (int i = 0; i < n; ++i) {
            charge[i] += rates[i] * dt + (chi[i] - friction * rates[i]) * inv_mq * dt2;
            rates[i] = (1.0 - 0.5 * friction * dt) * rates[i] + chi[i] * 0.5 * inv_mq * dt;
        }
        DCHI();
        error = chi[0];
        for (int i = 0; i < n; ++i) {
            rates[i] = future * (rates[i] + chi[i] * 0.5 * inv_mq * dt);
            if (chi[i] > error) error = chi[i];
        }
        step++;
    }
#undef DCHI
    if (err) *err = error;
    free(chi);
    return step;
}
