// learned.h -- the charge warm-start network: radial descriptors and a small
// BF16 MLP with FP32 accumulation, trained and evaluated on the device.
//
// Descriptors (per atom i, per neighbor species s, per Gaussian m):
//   G[i][s*16 + m] = sum_{j != i, r_ij < rc, species_j = s}
//                    exp(-eta_m (r_ij - mu_m)^2) * 0.5 (cos(pi r_ij / rc) + 1)
// with 16 Gaussians per neighbor species and at most two species, so there
// are 32 features (zero-padded for one species).
//
// Network (one per system; the center species enters through per-species
// biases of the first and the output layer):
//   x  = bf16((G - mean) * inv_scale)
//   h1 = bf16(tanh(W1^T x + b1[s]))        W1: 32 x 64, BF16
//   h2 = bf16(tanh(W2^T h1 + b2))          W2: 64 x 64, BF16
//   y  = w3 . h2 + b3[s]                   w3: 64, BF16
// Every product takes BF16 operands (v_dot2 on packed pairs) and accumulates
// in FP32; biases are FP32.  Training keeps FP32 master weights (Adam) and
// re-rounds the BF16 copies after every step.  Loss = mean (y - t)^2.
#pragma once

#include "system.h"

#include <stdint.h>

constexpr int MLP_G = 16;   // Gaussians per neighbor species
constexpr int MLP_F = 32;   // features
constexpr int MLP_H = 64;   // hidden width (both layers)
constexpr int MLP_S = 4;    // species slots of the species biases

// offsets into the flat FP32 parameter vector
constexpr int MLP_OFF_W1 = 0;                                // [F][H]  (k * H + j)
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_F * MLP_H;       // [S][H]
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_S * MLP_H;       // [H][H]  (k * H + j)
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;       // [H]
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;               // [H]
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;               // [S]
constexpr int MLP_NP = MLP_OFF_B3 + MLP_S;                   // 6532

// packed BF16 weights: pairs along the input index k (element k in the low half)
constexpr int MLP_PK_W1 = 0;                                 // [F/2][H]
constexpr int MLP_PK_W2 = MLP_PK_W1 + (MLP_F / 2) * MLP_H;   // [H/2][H]
constexpr int MLP_PK_W3 = MLP_PK_W2 + (MLP_H / 2) * MLP_H;   // [H/2]
constexpr int MLP_NPK = MLP_PK_W3 + MLP_H / 2;               // 3104 words

constexpr int MLP_TB = 64;  // samples per training block (one wavefront)

// Residual corrector: stages that move the predicted charges towards the
// relaxed ones, each fed by the electronegativities chi of the current
// charges and by the previous stage's change d.  With the descriptor
// Gaussians phi_m (same rc, mu, eta) and the pair class p = s_i + s_j of the
// species (0, 1 or 2), stage k computes
//   dq_i = sum_{j != i, r_ij < rc} sum_m phi_m(r_ij) [ a_k[p][m] (chi_j - chi_i)
//                                                   + b_k[p][m] (d_j - d_i) ]
//          + c_k d_i
//   q_i += dq_i,  d_i = dq_i            (d = 0 before the first stage)
// The weights are symmetric in i and j, so sum_i dq_i = c_k sum_i d_i = 0:
// the total charge is kept, and a uniform shift of chi changes nothing.
constexpr int CORR_G = MLP_G;
constexpr int CORR_NC = 2 * 3 * CORR_G + 1;  // a[3][16] | b[3][16] | c = 97 per stage
constexpr int CORR_MAX_STAGES = 16;

struct DescParams {
    float rc;
    float mu[MLP_G];
    float eta[MLP_G];
};

struct MlpState {
    int device = 0;
    hipStream_t stream = nullptr;
    int nspecies = 0;
    DescParams dp{};
    bool configured = false, has_params = false;
    DBuf<float> norm;        // mean[F], inv_scale[F]
    DBuf<float> P;           // FP32 master parameters
    DBuf<uint32_t> W;        // packed BF16 weights
    DBuf<float> m1, m2;      // Adam moments
    DBuf<float> grad;        // reduced gradient
    DBuf<float> gpart;       // per-block partial gradients [blocks][NP]
    DBuf<float> lpart;       // per-block loss partials
    DBuf<float> losses;      // per-step losses of one train call
    int adam_t = 0;
    // training set, uploaded once
    int ntrain = 0;
    DBuf<float> feat;        // [ntrain][F]
    DBuf<int> spec;          // [ntrain]
    DBuf<float> targ;        // [ntrain]
    DBuf<int> order;         // sample indices of one call
    DBuf<float> y;           // predictions
    // inference on a system
    DBuf<float> sfeat;       // [n][F], current (cell-sorted) order of the system
    DBuf<int> sspec;
    DBuf<double> q;          // predicted charges, original order
    // residual corrector
    int corr_stages = 0;
    DBuf<double> ccoef;          // [corr_stages][CORR_NC]
    DBuf<double> cprev, cnext;   // d of the previous stage, current order (ping-pong)
    DBuf<double> cfeat;          // [n][CORR_NC] corrector features, current order
    void release_all();
};

// mlp.hip launchers (all on m.stream unless a stream is given)
void mlp_descriptors(System &S, const MlpState &m, hipStream_t st);      // -> m.sfeat, m.sspec
void mlp_pack(MlpState &m);                                             // P -> W (BF16)
void mlp_forward(MlpState &m, int count, const int *order, const float *feat, const int *spec,
                 float *y, hipStream_t st);
// gradient of mean (y - t)^2 over the listed samples into m.grad; loss into loss_dev[0]
void mlp_gradient(MlpState &m, int count, const int *order, float *loss_dev);
void mlp_adam(MlpState &m, float lr, float beta1, float beta2, float eps);
// y (current order, FP32) -> q[id[s]] (original order, FP64)
void mlp_scatter_charges(System &S, const float *y, double *q, hipStream_t st);
// y (current order) -> S.q = y - mean(y) (current order, FP64); zeroes prev
void mlp_neutral_charges(System &S, MlpState &m, const float *y, double *prev, hipStream_t st);
// corrector features of chi and prev on S (current order) into m.cfeat
void mlp_corrector_features(System &S, MlpState &m, const double *chi, const double *prev,
                            hipStream_t st);
// one corrector stage: S.q += dq, next = dq
void mlp_corrector_stage(System &S, const MlpState &m, const double *coef, const double *chi,
                         const double *prev, double *next, hipStream_t st);
