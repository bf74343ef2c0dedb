// ewald.h -- Ewald summation and electronegativity equalization on device.
#pragma once

#include "system.h"

#include <vector>

struct EwaldParams {
    int on = 0;
    double rc = 0, kc = 0, sigma = 0, eps = 0;
    int K[3] = {0, 0, 0};
    int has_scaler = 0;
    std::vector<double> scaler;  // original order
};

struct EwaldState {
    bool dirty = true;
    int nk = 0;
    DBuf<double4> kvec;     // kx, ky, kz, coefficient c_k
    DBuf<double2> sk;       // sum q s cos(k.r), sum q s sin(k.r)
    DBuf<double> scaler;    // current (sorted) order
    DBuf<double> chi;       // electronegativities (sorted order)
    DBuf<double> part;      // block partials
    DBuf<double> red;       // reduced scalars
    DBuf<double2> skpart;   // structure-factor partials over atom chunks
    DBuf<double4> rpart;    // per-atom reciprocal partials over k chunks
    DBuf<double4> th;       // per-atom fractional coordinates and q s
    DBuf<double2> pw;       // per-atom q s exp(2 pi i l u3), 16 per atom
    DBuf<double2> wl;       // per-atom exp(2 pi i 16 u3)
    DBuf<double4> zw;       // coefficients of the per-atom sums (plan order)
    DBuf<int> kpk;          // packed (k1, k2, l0) per plan slot
    DBuf<unsigned long long> hpart;   // state fingerprint partials
    // sk and th hold the structure factors of the state whose fingerprint
    // is sk_hash
    bool sk_valid = false;
    unsigned long long sk_hash = 0;
    // k-vector segments: runs of at most 16 consecutive k3 at fixed (k1, k2),
    // {k1, k2, k3_start | len << 16, index of the first k}, natural order;
    // the structure-factor kernel's work plan over them (16 x 4 segments
    // per wave) and its continuation flags
    DBuf<int4> seg;
    DBuf<int> plan, pflags;
    int nseg = 0, nwaves = 0;
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // reciprocal vectors 2 pi M^-1 (rows)
    double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // M^-1 (rows): fractional coordinates
    double volume = 0;
    int skchunks = 0;
};

void ewald_release(EwaldState &E);
// Energy of the charge_self singlets + Ewald (real, reciprocal, self, background).
double ewald_eem_energy(System &S, const EwaldParams &P, EwaldState &E);
// get_ewald_energy (Core.f90:8077-8089): the Ewald sum alone, real space
// included, with the dipole term (the caller passes unit scalers and a state
// of its own, so the cached structure factors of the model's Ewald stay).
double ewald_standalone_energy(System &S, const EwaldParams &P, EwaldState &E);
// Adds Ewald forces into S.frc; stress[6] receives the Ewald part.
void ewald_add_forces(System &S, const EwaldParams &P, EwaldState &E, double *stress);
// Electronegativities (charge_self + Ewald) into E.chi, sorted order.
DBuf<double> &ewald_eem_chi(System &S, const EwaldParams &P, EwaldState &E);
// Damped charge dynamics on device; S.q and rates (sorted order) updated.
void ewald_relax(System &S, const EwaldParams &P, EwaldState &E, double *rates, int n_steps,
                 double dt, double inertia, double friction, double tolerance, int *steps,
                 double *error);
