/*
 * pysic_learned.h -- C-ABI of the learned charge warm start (libpysic_hip.so).
 *
 * No counterpart in the reference: pysic starts its charge relaxation from
 * the charges stored on the atoms (relaxation.py:417-450 reads them through
 * calculator.get_atoms().get_charges()).  These calls predict those starting
 * charges with a small network evaluated on the device; the relaxation then
 * runs unchanged, to the same tolerance, through pysic_relax_charges.
 *
 * Descriptors (per atom i, neighbor species s in {0,1}, Gaussian m < 16):
 *   G[i][16 s + m] = sum_{j != i, r_ij < rc, species_j = s}
 *                    exp(-eta_m (r_ij - mu_m)^2) * (cos(pi r_ij / rc) + 1) / 2
 * (32 features; 16..31 are zero for a single species).
 *
 * Network, with FP32 master parameters in one flat vector of
 * pysic_mlp_param_count() = 6532 floats:
 *   W1[32][64] | b1[4][64] | W2[64][64] | b2[64] | w3[64] | b3[4]
 *   x  = bf16((G - mean) * inv_scale)
 *   h1 = bf16(tanh(W1^T x + b1[species]))
 *   h2 = bf16(tanh(W2^T h1 + b2))
 *   y  = w3 . h2 + b3[species]
 * W1, W2, w3 are used as BF16 (rounded to nearest even from the master
 * copy); every product has BF16 operands and an FP32 accumulator.
 * Training minimises mean (y - target)^2 with Adam on the master copy.
 *
 * All calls return 0 or a negative PYSIC_E* code (pysic_hip.h), with a message
 * from pysic_mlp_last_error().
 */
#ifndef PYSIC_LEARNED_H
#define PYSIC_LEARNED_H

#include "pysic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pysic_mlp pysic_mlp;

int pysic_mlp_create(pysic_mlp **out, int device);
void pysic_mlp_destroy(pysic_mlp *m);
const char *pysic_mlp_last_error(pysic_mlp *m);
int pysic_mlp_param_count(void);

/* Descriptor and input-normalisation set-up: nspecies 1 or 2, cutoff rc (A),
 * mu[16], eta[16] of the Gaussians, mean[32] and inv_scale[32]. */
int pysic_mlp_configure(pysic_mlp *m, int nspecies, double rc, const double *mu,
                        const double *eta, const float *mean, const float *inv_scale);
/* FP32 master parameters in and out (resets the Adam state on set). */
int pysic_mlp_set_params(pysic_mlp *m, const float *params);
int pysic_mlp_get_params(pysic_mlp *m, float *params);

/* Descriptors of the system held by h (its neighbor lists must reach rc):
 * feat[n][32] and species[n] in the original atom order. */
int pysic_mlp_descriptors(pysic_mlp *m, pysic_ctx *h, int n, float *feat, int *species);
/* Predicted charges of the system held by h (original order).  With
 * neutralize != 0 the mean is subtracted so that the total charge is zero,
 * as the relaxation conserves it. */
int pysic_mlp_predict_charges(pysic_mlp *m, pysic_ctx *h, int n, int neutralize,
                              double *charges);

/* Residual corrector.  Stages that move the predicted charges towards the
 * relaxed ones, each fed by the electronegativities chi (the quantity the
 * relaxation drives to uniformity) of the current charges, evaluated by the
 * system's own model (EEM terms + Ewald), and by the previous stage's change
 * d (zero before the first stage).  With the descriptor Gaussians
 * phi_m(r) = exp(-eta_m (r - mu_m)^2) (cos(pi r / rc) + 1) / 2 and the pair
 * class p = species_i + species_j (0, 1 or 2), stage k computes
 *   dq_i = sum_{j != i, r_ij < rc} sum_m phi_m(r_ij) [ a[p][m] (chi_j - chi_i)
 *                                                   + b[p][m] (d_j - d_i) ]
 *          + c d_i,
 *   q_i += dq_i,   d_i = dq_i,
 * with the stage's PYSIC_CORRECTOR_NCOEF coefficients stored as
 * a[3][16] | b[3][16] | c.  The weights are symmetric in i and j: the total
 * charge is kept, and chi enters only through differences.  Each stage costs
 * one evaluation of chi. */
#define PYSIC_CORRECTOR_NCOEF 97
#define PYSIC_CORRECTOR_MAX_STAGES 16

/* The corrector's stages: coef[n_stages][PYSIC_CORRECTOR_NCOEF]; 0 stages
 * removes it. */
int pysic_mlp_set_corrector(pysic_mlp *m, int n_stages, const double *coef);
/* Number of stages held (or a negative code for a null handle). */
int pysic_mlp_corrector_stages(pysic_mlp *m);
/* Training features of one stage on the system held by h: sets its charges
 * to `charges`, evaluates chi, and writes, per atom (original order),
 * features[n][PYSIC_CORRECTOR_NCOEF] such that dq_i = features_i . coef of
 * the stage formula above with d = `previous`.  enegs[n] (may be NULL)
 * receives chi. */
int pysic_mlp_corrector_features(pysic_mlp *m, pysic_ctx *h, int n, const double *charges,
                                 const double *previous, double *features, double *enegs);
/* The warm start: predicted charges, made neutral (mean subtracted), then the
 * first n_stages corrector stages, all on the device.  The system's charges
 * are left at the result, which is also written to charges[n] (original
 * order).  n_stages = 0 gives pysic_mlp_predict_charges with neutralize. */
int pysic_mlp_warm_start(pysic_mlp *m, pysic_ctx *h, int n, int n_stages, double *charges);

/* Training set: n samples of 32 features, species and target charge, kept on
 * the device until replaced. */
int pysic_mlp_set_training_data(pysic_mlp *m, int n, const float *feat, const int *species,
                                const float *targets);
/* Network outputs for the listed training samples. */
int pysic_mlp_evaluate(pysic_mlp *m, int count, const int *samples, float *y);
/* Gradient of mean (y - t)^2 over the listed training samples, no update. */
int pysic_mlp_gradients(pysic_mlp *m, int count, const int *samples, float *grad,
                        float *loss);
/* nsteps Adam steps; step k uses samples order[k*batch .. (k+1)*batch-1].
 * losses[k] receives the batch loss before step k's update. */
int pysic_mlp_train(pysic_mlp *m, int nsteps, int batch, const int *order, float lr,
                    float beta1, float beta2, float eps, float *losses);

#ifdef __cplusplus
}
#endif

#endif
