// mlp.hip -- descriptors, BF16 MLP forward, backward and Adam on the MI355X.
//
// This part has no counterpart in the reference: it predicts starting
// charges for the damped charge relaxation (relaxation.py:417-450), which
// then runs unchanged to its tolerance.  The network and descriptors are
// defined in learned.h.
//
//  * k_descriptors: one thread per atom over its full neighbor list
//    (separations in FP64 as the force kernels form them, Gaussians in FP32);
//  * k_forward: one thread per sample, all BF16 weights staged in LDS
//    (12.4 KB, read as wave-wide broadcasts), products by v_dot2_f32_bf16 on
//    packed BF16 pairs with FP32 accumulation;
//  * k_train: one wavefront per 64 samples; the forward pass and the
//    per-sample backward pass run one sample per lane, then each lane owns
//    one hidden unit and reduces the outer products dW = sum_s a_s g_s^T over
//    the block's samples from LDS; block partials are summed in a fixed
//    order by k_grad_reduce (deterministic, no atomics);
//  * k_adam + k_pack: FP32 master update and BF16 re-rounding.
#include "learned.h"

#include <cmath>

namespace {

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ uint32_t pack2(float lo, float hi)
{
    return bf16_rne(lo) | (bf16_rne(hi) << 16);
}
__device__ __forceinline__ float lo16(uint32_t p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float hi16(uint32_t p) { return __uint_as_float(p & 0xffff0000u); }
// c + a.lo * b.lo + a.hi * b.hi, BF16 operands, FP32 accumulate
__device__ __forceinline__ float dot2(uint32_t a, uint32_t b, float c)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, a),
                                           __builtin_bit_cast(bf16x2_t, b), c, false);
}

template <int NS>
__global__ __launch_bounds__(256) void k_descriptors(int n, CellGeom g,
                                                     const double4 *__restrict__ pos,
                                                     const int *__restrict__ nbr,
                                                     const int *__restrict__ nimg,
                                                     const int *__restrict__ ncnt, DescParams dp,
                                                     float *__restrict__ feat,
                                                     int *__restrict__ spec)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc[NS][MLP_G];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int m = 0; m < MLP_G; ++m) acc[s][m] = 0.0f;
    const double4 pi = pos[i];
    const double rc2 = (double)dp.rc * (double)dp.rc;
    const float pirc = 3.14159265358979f / dp.rc;
    const int nf = ncnt[i];
    for (int t = 0; t < nf; ++t) {
        const int j = nbr[(size_t)t * n + i];
        const int im = nimg[(size_t)t * n + i];
        const double4 pj = pos[j];
        const double3 d3 = separation64(pi, pj, im, g.v);
        const double r2 = dot64(d3);
        if (!(r2 < rc2)) continue;
        const float r = (float)sqrt(r2);
        const float fc = 0.5f * (cosf(r * pirc) + 1.0f);
        const int sj = (int)pj.w;
#pragma unroll
        for (int m = 0; m < MLP_G; ++m) {
            const float dr = r - dp.mu[m];
            const float v = expf(-dp.eta[m] * dr * dr) * fc;
            if (NS == 1) {
                acc[0][m] += v;
            } else {
                acc[0][m] += (sj == 0) ? v : 0.0f;
                acc[NS - 1][m] += (sj == 1) ? v : 0.0f;
            }
        }
    }
    float4 *o = reinterpret_cast<float4 *>(feat + (size_t)i * MLP_F);
#pragma unroll
    for (int q = 0; q < MLP_F / 4; ++q) {
        float f[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int k = 4 * q + c;
            f[c] = (k < NS * MLP_G) ? acc[k / MLP_G][k % MLP_G] : 0.0f;
        }
        o[q] = make_float4(f[0], f[1], f[2], f[3]);
    }
    spec[i] = (int)pi.w;
}

// normalized, BF16-rounded input pairs of one sample
__device__ __forceinline__ void load_input(const float *__restrict__ feat, int smp,
                                           const float *sN, uint32_t *xp)
{
    const float4 *f4 = reinterpret_cast<const float4 *>(feat + (size_t)smp * MLP_F);
#pragma unroll
    for (int q = 0; q < MLP_F / 4; ++q) {
        const float4 v = f4[q];
        const int k = 4 * q;
        xp[2 * q] = pack2((v.x - sN[k]) * sN[MLP_F + k], (v.y - sN[k + 1]) * sN[MLP_F + k + 1]);
        xp[2 * q + 1] =
            pack2((v.z - sN[k + 2]) * sN[MLP_F + k + 2], (v.w - sN[k + 3]) * sN[MLP_F + k + 3]);
    }
}

// The three layers of one sample.  Each lane keeps its packed hidden
// activations in its own LDS rows (h1row, and h2row when not null), so that
// the loops over output chunks can run without dynamically indexed register
// arrays; the inner loops over inputs are unrolled with static indices.
__device__ __forceinline__ float forward_rows(const uint32_t *xp, int s, const uint32_t *sW,
                                              const float *sB1, const float *sB2,
                                              const float *sB3, uint32_t *h1row,
                                              uint32_t *h2row)
{
#pragma unroll 1
    for (int jc = 0; jc < MLP_H / 8; ++jc) {
        float z[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) z[c] = sB1[s * MLP_H + 8 * jc + c];
#pragma unroll
        for (int kp = 0; kp < MLP_F / 2; ++kp) {
            const uint32_t *w = sW + MLP_PK_W1 + kp * MLP_H + 8 * jc;
#pragma unroll
            for (int c = 0; c < 8; ++c) z[c] = dot2(xp[kp], w[c], z[c]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) h1row[4 * jc + q] = pack2(tanhf(z[2 * q]), tanhf(z[2 * q + 1]));
    }
    uint32_t h1[MLP_H / 2];
#pragma unroll
    for (int kp = 0; kp < MLP_H / 2; ++kp) h1[kp] = h1row[kp];
    float out = sB3[s];
#pragma unroll 1
    for (int jc = 0; jc < MLP_H / 8; ++jc) {
        float z[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) z[c] = sB2[8 * jc + c];
#pragma unroll
        for (int kp = 0; kp < MLP_H / 2; ++kp) {
            const uint32_t *w = sW + MLP_PK_W2 + kp * MLP_H + 8 * jc;
#pragma unroll
            for (int c = 0; c < 8; ++c) z[c] = dot2(h1[kp], w[c], z[c]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t hp = pack2(tanhf(z[2 * q]), tanhf(z[2 * q + 1]));
            if (h2row) h2row[4 * jc + q] = hp;
            out = dot2(hp, sW[MLP_PK_W3 + 4 * jc + q], out);
        }
    }
    return out;
}

__device__ __forceinline__ void stage_params(const uint32_t *__restrict__ W,
                                             const float *__restrict__ P,
                                             const float *__restrict__ norm, uint32_t *sW,
                                             float *sB1, float *sB2, float *sB3, float *sN)
{
    for (int k = threadIdx.x; k < MLP_NPK; k += blockDim.x) sW[k] = W[k];
    for (int k = threadIdx.x; k < MLP_S * MLP_H; k += blockDim.x) sB1[k] = P[MLP_OFF_B1 + k];
    for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) sB2[k] = P[MLP_OFF_B2 + k];
    for (int k = threadIdx.x; k < MLP_S; k += blockDim.x) sB3[k] = P[MLP_OFF_B3 + k];
    for (int k = threadIdx.x; k < 2 * MLP_F; k += blockDim.x) sN[k] = norm[k];
}

constexpr int FWD_B = 128;  // threads per forward block

__global__ __launch_bounds__(FWD_B) void k_forward(int count, const int *__restrict__ order,
                                                 const float *__restrict__ feat,
                                                 const int *__restrict__ spec,
                                                 const float *__restrict__ norm,
                                                 const uint32_t *__restrict__ W,
                                                 const float *__restrict__ P,
                                                 float *__restrict__ y)
{
    __shared__ uint32_t sW[MLP_NPK];
    __shared__ float sB1[MLP_S * MLP_H], sB2[MLP_H], sB3[MLP_S], sN[2 * MLP_F];
    __shared__ uint32_t sh1[FWD_B][MLP_H / 2 + 1];
    stage_params(W, P, norm, sW, sB1, sB2, sB3, sN);
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int smp = order ? order[t] : t;
    const int s = spec[smp];
    uint32_t xp[MLP_F / 2];
    load_input(feat, smp, sN, xp);
    y[t] = forward_rows(xp, s, sW, sB1, sB2, sB3, sh1[threadIdx.x], nullptr);
}

// One wavefront, MLP_TB samples: partial gradient of sum_s (y_s - t_s)^2 * inv_count.
__global__ __launch_bounds__(MLP_TB) void k_train(int count, const int *__restrict__ order,
                                                  const float *__restrict__ feat,
                                                  const int *__restrict__ spec,
                                                  const float *__restrict__ targ, float inv_count,
                                                  const float *__restrict__ norm,
                                                  const uint32_t *__restrict__ W,
                                                  const float *__restrict__ P,
                                                  float *__restrict__ gpart,
                                                  float *__restrict__ lpart)
{
    __shared__ uint32_t sW[MLP_NPK];
    __shared__ float sB1[MLP_S * MLP_H], sB2[MLP_H], sB3[MLP_S], sN[2 * MLP_F];
    __shared__ uint32_t sx[MLP_TB][MLP_F / 2 + 1];
    __shared__ uint32_t sh1[MLP_TB][MLP_H / 2 + 1];
    __shared__ uint32_t sh2[MLP_TB][MLP_H / 2 + 1];
    __shared__ float sg1[MLP_TB][MLP_H + 1];
    __shared__ float sg2[MLP_TB][MLP_H + 1];
    __shared__ float sgy[MLP_TB], sl[MLP_TB];
    __shared__ int ssp[MLP_TB];
    stage_params(W, P, norm, sW, sB1, sB2, sB3, sN);
    __syncthreads();

    // ---- per-sample forward and backward (lane = sample) ----
    {
        const int t = threadIdx.x;
        const int idx = blockIdx.x * MLP_TB + t;
        const bool valid = idx < count;
        const int smp = valid ? order[idx] : 0;
        const int s = valid ? spec[smp] : 0;
        uint32_t xp[MLP_F / 2];
        load_input(feat, smp, sN, xp);
#pragma unroll
        for (int kp = 0; kp < MLP_F / 2; ++kp) sx[t][kp] = xp[kp];
        const float yv = forward_rows(xp, s, sW, sB1, sB2, sB3, sh1[t], sh2[t]);
        const float res = valid ? (yv - targ[smp]) : 0.0f;
        const float gy = 2.0f * res * inv_count;
        sgy[t] = gy;
        sl[t] = res * res;
        ssp[t] = s;
        // dL/dz2_j = gy * w3_j * (1 - h2_j^2)
        float g2[MLP_H];
#pragma unroll
        for (int jp = 0; jp < MLP_H / 2; ++jp) {
            const uint32_t w = sW[MLP_PK_W3 + jp];
            const uint32_t hw = sh2[t][jp];
            const float a = lo16(hw), b = hi16(hw);
            g2[2 * jp] = gy * lo16(w) * (1.0f - a * a);
            g2[2 * jp + 1] = gy * hi16(w) * (1.0f - b * b);
        }
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) sg2[t][j] = g2[j];
        // dL/dh1_k = sum_j W2[k][j] dL/dz2_j ; dL/dz1_k = that * (1 - h1_k^2)
#pragma unroll 1
        for (int kc = 0; kc < MLP_H / 8; ++kc) {
            float gh[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) gh[c] = 0.0f;
#pragma unroll
            for (int j = 0; j < MLP_H; ++j) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t w = sW[MLP_PK_W2 + (4 * kc + q) * MLP_H + j];
                    gh[2 * q] += lo16(w) * g2[j];
                    gh[2 * q + 1] += hi16(w) * g2[j];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t hw = sh1[t][4 * kc + q];
                const float a = lo16(hw), b = hi16(hw);
                sg1[t][8 * kc + 2 * q] = gh[2 * q] * (1.0f - a * a);
                sg1[t][8 * kc + 2 * q + 1] = gh[2 * q + 1] * (1.0f - b * b);
            }
        }
    }
    __syncthreads();

    // ---- outer-product reductions over the block's samples (lane = unit j) ----
    const int j = threadIdx.x;
    float *gp = gpart + (size_t)blockIdx.x * MLP_NP;
    {
        float a[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) a[k] = 0.0f;
        float db2 = 0.0f, dw3 = 0.0f;
        for (int s = 0; s < MLP_TB; ++s) {
            const float g = sg2[s][j];
            db2 += g;
            const uint32_t hw = sh2[s][j >> 1];
            dw3 += sgy[s] * ((j & 1) ? hi16(hw) : lo16(hw));
#pragma unroll
            for (int kp = 0; kp < MLP_H / 2; ++kp) {
                const uint32_t w = sh1[s][kp];
                a[2 * kp] += lo16(w) * g;
                a[2 * kp + 1] += hi16(w) * g;
            }
        }
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) gp[MLP_OFF_W2 + k * MLP_H + j] = a[k];
        gp[MLP_OFF_B2 + j] = db2;
        gp[MLP_OFF_W3 + j] = dw3;
    }
    {
        float c[MLP_F];
#pragma unroll
        for (int k = 0; k < MLP_F; ++k) c[k] = 0.0f;
        float d1[MLP_S] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < MLP_TB; ++s) {
            const float g = sg1[s][j];
            const int sp = ssp[s];
#pragma unroll
            for (int q = 0; q < MLP_S; ++q) d1[q] += (sp == q) ? g : 0.0f;
#pragma unroll
            for (int kp = 0; kp < MLP_F / 2; ++kp) {
                const uint32_t w = sx[s][kp];
                c[2 * kp] += lo16(w) * g;
                c[2 * kp + 1] += hi16(w) * g;
            }
        }
#pragma unroll
        for (int k = 0; k < MLP_F; ++k) gp[MLP_OFF_W1 + k * MLP_H + j] = c[k];
#pragma unroll
        for (int q = 0; q < MLP_S; ++q) gp[MLP_OFF_B1 + q * MLP_H + j] = d1[q];
    }
    if (j < MLP_S) {
        float b3 = 0.0f;
        for (int s = 0; s < MLP_TB; ++s) b3 += (ssp[s] == j) ? sgy[s] : 0.0f;
        gp[MLP_OFF_B3 + j] = b3;
    }
    if (j == 0) {
        float l = 0.0f;
        for (int s = 0; s < MLP_TB; ++s) l += sl[s];
        lpart[blockIdx.x] = l;
    }
}

// grad[p] = sum over blocks in block order; thread MLP_NP writes the loss
__global__ void k_grad_reduce(int nblk, const float *__restrict__ gpart,
                              const float *__restrict__ lpart, float inv_count,
                              float *__restrict__ grad, float *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NP) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += gpart[(size_t)b * MLP_NP + p];
        grad[p] = s;
    } else if (p == MLP_NP) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += lpart[b];
        loss[0] = s * inv_count;
    }
}

__global__ void k_adam(float *__restrict__ P, float *__restrict__ m1, float *__restrict__ m2,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    const float gm = g[p];
    const float m = b1 * m1[p] + (1.0f - b1) * gm;
    const float v = b2 * m2[p] + (1.0f - b2) * gm * gm;
    m1[p] = m;
    m2[p] = v;
    P[p] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

__global__ void k_pack(const float *__restrict__ P, uint32_t *__restrict__ W)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= MLP_NPK) return;
    float lo, hi;
    if (w < MLP_PK_W2) {
        const int kp = w / MLP_H, j = w % MLP_H;
        lo = P[MLP_OFF_W1 + (2 * kp) * MLP_H + j];
        hi = P[MLP_OFF_W1 + (2 * kp + 1) * MLP_H + j];
    } else if (w < MLP_PK_W3) {
        const int v = w - MLP_PK_W2;
        const int kp = v / MLP_H, j = v % MLP_H;
        lo = P[MLP_OFF_W2 + (2 * kp) * MLP_H + j];
        hi = P[MLP_OFF_W2 + (2 * kp + 1) * MLP_H + j];
    } else {
        const int jp = w - MLP_PK_W3;
        lo = P[MLP_OFF_W3 + 2 * jp];
        hi = P[MLP_OFF_W3 + 2 * jp + 1];
    }
    W[w] = pack2(lo, hi);
}

__global__ void k_scatter_charges(int n, const int *__restrict__ id, const float *__restrict__ y,
                                  double *__restrict__ q)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) q[id[s]] = (double)y[s];
}

// ---- residual corrector (learned.h) -----------------------------------------

// The corrector Gaussians of one pair, as k_descriptors forms them: FP64
// separation, FP32 distance, cutoff function and exponentials.
__device__ __forceinline__ void corr_basis(float r, const DescParams &dp, float pirc,
                                           float *phi)
{
    const float fc = 0.5f * (cosf(r * pirc) + 1.0f);
#pragma unroll
    for (int m = 0; m < CORR_G; ++m) {
        const float dr = r - dp.mu[m];
        phi[m] = expf(-dp.eta[m] * dr * dr) * fc;
    }
}

// One thread per atom: the 97 features of the stage formula, so that
// dq_i = features_i . coefficients.  Training-only (FP64 accumulators).
template <int NS>
__global__ __launch_bounds__(128) void k_corr_features(int n, CellGeom g,
                                                       const double4 *__restrict__ pos,
                                                       const int *__restrict__ nbr,
                                                       const int *__restrict__ nimg,
                                                       const int *__restrict__ ncnt, DescParams dp,
                                                       const double *__restrict__ chi,
                                                       const double *__restrict__ prev,
                                                       double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ar[NS][CORR_G], ad[NS][CORR_G];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int m = 0; m < CORR_G; ++m) ar[s][m] = ad[s][m] = 0.0;
    const double4 pi = pos[i];
    const double rc2 = (double)dp.rc * (double)dp.rc;
    const float pirc = 3.14159265358979f / dp.rc;
    const double ci = chi[i], di = prev[i];
    const int nf = ncnt[i];
    for (int t = 0; t < nf; ++t) {
        const int j = nbr[(size_t)t * n + i];
        const int im = nimg[(size_t)t * n + i];
        const double4 pj = pos[j];
        const double r2 = dot64(separation64(pi, pj, im, g.v));
        if (!(r2 < rc2)) continue;
        float phi[CORR_G];
        corr_basis((float)sqrt(r2), dp, pirc, phi);
        const double dc = chi[j] - ci, dd = prev[j] - di;
        const int sj = (int)pj.w;
#pragma unroll
        for (int m = 0; m < CORR_G; ++m) {
            const double v = (double)phi[m];
            if (NS == 1) {
                ar[0][m] += v * dc;
                ad[0][m] += v * dd;
            } else {
                ar[0][m] += (sj == 0) ? v * dc : 0.0;
                ad[0][m] += (sj == 0) ? v * dd : 0.0;
                ar[NS - 1][m] += (sj == 1) ? v * dc : 0.0;
                ad[NS - 1][m] += (sj == 1) ? v * dd : 0.0;
            }
        }
    }
    // pair class p = s_i + s_j: the row holds ar/ad at p in {s_i, s_i + 1}
    const int si = (int)pi.w;
    double *o = out + (size_t)i * CORR_NC;
    for (int k = 0; k < CORR_NC - 1; ++k) o[k] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int m = 0; m < CORR_G; ++m) {
            o[(si + s) * CORR_G + m] = ar[s][m];
            o[3 * CORR_G + (si + s) * CORR_G + m] = ad[s][m];
        }
    o[CORR_NC - 1] = di;
}

// One corrector stage, CORR_LANES lanes per atom (the atom's neighbor list
// split over the lanes, partial sums reduced by cross-lane shuffles); the
// stage's coefficients are staged in LDS and read per pair class.
constexpr int CORR_LANES = 16;

__global__ __launch_bounds__(256) void k_corr_stage(int n, CellGeom g,
                                                    const double4 *__restrict__ pos,
                                                    const int *__restrict__ nbr,
                                                    const int *__restrict__ nimg,
                                                    const int *__restrict__ ncnt, DescParams dp,
                                                    const double *__restrict__ coef,
                                                    const double *__restrict__ chi,
                                                    const double *__restrict__ prev,
                                                    double *__restrict__ q,
                                                    double *__restrict__ next)
{
    __shared__ double sc[CORR_NC];
    for (int k = threadIdx.x; k < CORR_NC; k += blockDim.x) sc[k] = coef[k];
    __syncthreads();
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / CORR_LANES;
    const int lane = threadIdx.x % CORR_LANES;
    double acc = 0.0;
    if (i < n) {
        const double4 pi = pos[i];
        const int si = (int)pi.w;
        const double rc2 = (double)dp.rc * (double)dp.rc;
        const float pirc = 3.14159265358979f / dp.rc;
        const double ci = chi[i], di = prev[i];
        const int nf = ncnt[i];
        for (int t = lane; t < nf; t += CORR_LANES) {
            const int j = nbr[(size_t)t * n + i];
            const int im = nimg[(size_t)t * n + i];
            const double4 pj = pos[j];
            const double r2 = dot64(separation64(pi, pj, im, g.v));
            if (!(r2 < rc2)) continue;
            float phi[CORR_G];
            corr_basis((float)sqrt(r2), dp, pirc, phi);
            const double *a = sc + (si + (int)pj.w) * CORR_G;
            const double *b = a + 3 * CORR_G;
            double wa = 0.0, wb = 0.0;
#pragma unroll
            for (int m = 0; m < CORR_G; ++m) {
                wa += a[m] * (double)phi[m];
                wb += b[m] * (double)phi[m];
            }
            acc += wa * (chi[j] - ci) + wb * (prev[j] - di);
        }
    }
    // every lane of the group takes part in the shuffles (no early return)
#pragma unroll
    for (int w = CORR_LANES / 2; w > 0; w >>= 1) acc += __shfl_xor(acc, w, CORR_LANES);
    if (i < n && lane == 0) {
        const double dq = acc + sc[CORR_NC - 1] * prev[i];
        q[i] += dq;
        next[i] = dq;
    }
}

// S.q = y - mean(y) and prev = 0, in one block (the mean in FP64, fixed order)
__global__ __launch_bounds__(1024) void k_neutral_charges(int n, const float *__restrict__ y,
                                                          double *__restrict__ q,
                                                          double *__restrict__ prev)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 1024) s += (double)y[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double mean = red[0] / n;
    for (int k = threadIdx.x; k < n; k += 1024) {
        q[k] = (double)y[k] - mean;
        prev[k] = 0.0;
    }
}

}  // namespace

void MlpState::release_all()
{
    norm.release();
    P.release();
    W.release();
    m1.release();
    m2.release();
    grad.release();
    gpart.release();
    lpart.release();
    losses.release();
    feat.release();
    spec.release();
    targ.release();
    order.release();
    y.release();
    sfeat.release();
    sspec.release();
    q.release();
    ccoef.release();
    cprev.release();
    cnext.release();
    cfeat.release();
    corr_stages = 0;
}

void mlp_descriptors(System &S, const MlpState &m, hipStream_t st)
{
    const int n = S.n;
    const int blocks = div_up(n, 256);
    if (m.nspecies == 1)
        k_descriptors<1><<<blocks, 256, 0, st>>>(n, S.geom, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                                 m.dp, m.sfeat.p, m.sspec.p);
    else
        k_descriptors<2><<<blocks, 256, 0, st>>>(n, S.geom, S.pos.p, S.nbr.p, S.nimg.p, S.ncnt.p,
                                                 m.dp, m.sfeat.p, m.sspec.p);
    HIP_TRY(hipGetLastError());
}

void mlp_pack(MlpState &m)
{
    k_pack<<<div_up(MLP_NPK, 256), 256, 0, m.stream>>>(m.P.p, m.W.p);
    HIP_TRY(hipGetLastError());
}

void mlp_forward(MlpState &m, int count, const int *order, const float *feat, const int *spec,
                 float *y, hipStream_t st)
{
    if (count <= 0) return;
    k_forward<<<div_up(count, FWD_B), FWD_B, 0, st>>>(count, order, feat, spec, m.norm.p, m.W.p, m.P.p,
                                                  y);
    HIP_TRY(hipGetLastError());
}

void mlp_gradient(MlpState &m, int count, const int *order, float *loss_dev)
{
    const int nblk = div_up(count, MLP_TB);
    m.gpart.ensure((size_t)nblk * MLP_NP);
    m.lpart.ensure(nblk);
    const float inv = 1.0f / (float)count;
    k_train<<<nblk, MLP_TB, 0, m.stream>>>(count, order, m.feat.p, m.spec.p, m.targ.p, inv,
                                           m.norm.p, m.W.p, m.P.p, m.gpart.p, m.lpart.p);
    HIP_TRY(hipGetLastError());
    k_grad_reduce<<<div_up(MLP_NP + 1, 256), 256, 0, m.stream>>>(nblk, m.gpart.p, m.lpart.p, inv,
                                                                 m.grad.p, loss_dev);
    HIP_TRY(hipGetLastError());
}

void mlp_adam(MlpState &m, float lr, float beta1, float beta2, float eps)
{
    m.adam_t += 1;
    const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)m.adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)m.adam_t));
    k_adam<<<div_up(MLP_NP, 256), 256, 0, m.stream>>>(m.P.p, m.m1.p, m.m2.p, m.grad.p, lr, beta1,
                                                      beta2, eps, c1, c2);
    HIP_TRY(hipGetLastError());
    mlp_pack(m);
}

void mlp_scatter_charges(System &S, const float *y, double *q, hipStream_t st)
{
    k_scatter_charges<<<div_up(S.n, 256), 256, 0, st>>>(S.n, S.id.p, y, q);
    HIP_TRY(hipGetLastError());
}

void mlp_neutral_charges(System &S, MlpState &m, const float *y, double *prev, hipStream_t st)
{
    (void)m;
    k_neutral_charges<<<1, 1024, 0, st>>>(S.n, y, S.q.p, prev);
    HIP_TRY(hipGetLastError());
}

void mlp_corrector_features(System &S, MlpState &m, const double *chi, const double *prev,
                            hipStream_t st)
{
    const int n = S.n;
    const int blocks = div_up(n, 128);
    if (m.nspecies == 1)
        k_corr_features<1><<<blocks, 128, 0, st>>>(n, S.geom, S.pos.p, S.nbr.p, S.nimg.p,
                                                    S.ncnt.p, m.dp, chi, prev, m.cfeat.p);
    else
        k_corr_features<2><<<blocks, 128, 0, st>>>(n, S.geom, S.pos.p, S.nbr.p, S.nimg.p,
                                                    S.ncnt.p, m.dp, chi, prev, m.cfeat.p);
    HIP_TRY(hipGetLastError());
}

void mlp_corrector_stage(System &S, const MlpState &m, const double *coef, const double *chi,
                         const double *prev, double *next, hipStream_t st)
{
    const long threads = (long)S.n * CORR_LANES;
    k_corr_stage<<<(int)div_up(threads, 256L), 256, 0, st>>>(S.n, S.geom, S.pos.p, S.nbr.p, S.nimg.p,
                                                   S.ncnt.p, m.dp, coef, chi, prev, S.q.p, next);
    HIP_TRY(hipGetLastError());
}
