// gp_encoder.hip -- amortization encoder: an MLP from a cell's expression
// profile to a Gaussian over its latent position, used to warm-start X.
// No counterpart in the reference (which starts X from an external PCA file,
// ggplvm.cu:1088-1095 reading argv[4]).
//
//   a0 = log1p(y_i. / s_i)                     (F features, one per gene)
//   z1 = W1 a0 + b1, h1 = gelu(z1)             (H1)
//   z2 = W2 h1 + b2, h2 = gelu(z2)             (H2)
//   z3 = W3 h2 + b3 = [mu (Q) | log-variance (Q)]
//   loss = (1/n) sum_i sum_q 1/2 [ (x_iq - mu_iq)^2 exp(-s_iq) + s_iq ],
//   s = log-variance clamped to [-10, 10]
//
// gelu is the exact form z Phi(z).  Weights are FP32 masters with BF16 copies
// for the products; activations and backward signals enter the products as
// BF16, every product accumulates in FP32 (v_mfma_f32_16x16x32_bf16); Adam
// runs on the FP32 masters.
//
// Every product is one kernel shape, C[i][j] = sum_k A[i][k] B[j][k] with
// both operands k-contiguous, so an MFMA fragment (8 consecutive k of one
// row) is one 16-byte load for A and for B alike.  The forward epilogues
// write the activation row-major (the next layer's A operand) and transposed
// (the weight gradient's B operand); the backward epilogues do the same for
// the GELU-masked signal.  All dimensions are padded to multiples of 64
// (rows, columns) and the batch to a multiple of 64, so no tile is partial.
#include "gp_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <string>
#include <vector>

typedef __bf16 bf16;

struct gp_enc {
    int F = 0, Fp = 0, H1 = 0, H2 = 0, Q = 0, O = 64, B = 0;
    long long cap = 0, n = 0;
    size_t np = 0;                                   // parameter count (padded layout)
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    long long t = 0;                                 // Adam steps taken
    std::mt19937_64 rng;
    hipStream_t stream = nullptr;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *W1b = nullptr, *W2b = nullptr, *W2t = nullptr, *W3b = nullptr, *W3t = nullptr;
    bf16 *feat = nullptr;                            // cap x Fp
    float *tgt = nullptr;                            // cap x Q
    int *perm = nullptr;                             // identity 0 .. max(cap, B) - 1
    bf16 *A0 = nullptr, *A0t = nullptr;              // B x Fp, Fp x B
    float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr;
    bf16 *H1b = nullptr, *H1t = nullptr, *H2b = nullptr, *H2t = nullptr;
    float *dZ1 = nullptr, *dZ2 = nullptr, *dZ3 = nullptr;
    bf16 *dZ1t = nullptr, *dZ2b = nullptr, *dZ2t = nullptr, *dZ3b = nullptr, *dZ3t = nullptr;
    double *lpart = nullptr, *lsum = nullptr;
    std::vector<int> perm_h;
};

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int ET = 256;

enum { E_STORE = 0, E_BIAS = 1, E_BIAS_GELU = 2, E_GELU_BWD = 3 };

struct Epi {
    float *out;
    int ldo;
    const float *bias;     // E_BIAS, E_BIAS_GELU
    const float *zin;      // E_GELU_BWD: saved pre-activation (ld = ldo)
    bf16 *hb;              // BF16 row-major copy (ld = ldo), may be null
    bf16 *ht;              // BF16 transposed copy, may be null
    int ldt;
};

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C (R x C) = A (R x K) . B (C x K)^T; one workgroup per 64 x 64 tile, wave w
// owns the 32 x 32 quarter (w / 2, w % 2) as 2 x 2 MFMA tiles.  Fragment
// layout of v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + e],
// B[8 (l / 16) + e][l % 16], and result rows 4 (l / 16) + r of column l % 16.
template <int E>
__global__ void __launch_bounds__(ET) k_gemm_nt(int K, const bf16 *__restrict__ A, int lda,
                                                const bf16 *__restrict__ Bm, int ldb, Epi e) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r0 = blockIdx.x * 64 + 32 * (w >> 1), c0 = blockIdx.y * 64 + 32 * (w & 1);
    const bf16 *pa = A + (size_t)(r0 + (l & 15)) * lda + 8 * (l >> 4);
    const bf16 *pb = Bm + (size_t)(c0 + (l & 15)) * ldb + 8 * (l >> 4);
    const size_t sa = (size_t)16 * lda, sb = (size_t)16 * ldb;
    f4 acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++) acc[t][u] = f4{0.f, 0.f, 0.f, 0.f};
    bf8 a0 = *(const bf8 *)pa, a1 = *(const bf8 *)(pa + sa);
    bf8 b0 = *(const bf8 *)pb, b1 = *(const bf8 *)(pb + sb);
    for (int k = 32; k < K; k += 32) {
        const bf8 na0 = *(const bf8 *)(pa + k), na1 = *(const bf8 *)(pa + sa + k);
        const bf8 nb0 = *(const bf8 *)(pb + k), nb1 = *(const bf8 *)(pb + sb + k);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
        a0 = na0; a1 = na1; b0 = nb0; b1 = nb1;
    }
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 2; t++) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int i0 = r0 + 16 * t + 4 * (l >> 4), j = c0 + 16 * u + (l & 15);
            bf4 tv;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const size_t ij = (size_t)(i0 + r) * e.ldo + j;
                float val = acc[t][u][r];
                if (E == E_BIAS || E == E_BIAS_GELU) val += e.bias[j];
                if (E == E_GELU_BWD) val *= gelu_d(e.zin[ij]);
                e.out[ij] = val;
                if (E == E_BIAS_GELU) {
                    const float h = gelu(val);
                    e.hb[ij] = (bf16)h;
                    tv[r] = (bf16)h;
                } else if (E == E_GELU_BWD) {
                    if (e.hb) e.hb[ij] = (bf16)val;
                    tv[r] = (bf16)val;
                }
            }
            if (E == E_BIAS_GELU || E == E_GELU_BWD) *(bf4 *)(e.ht + (size_t)j * e.ldt + i0) = tv;
        }
    }
}

// Batch rows r < n of A0 (B x Fp) and A0t (Fp x B) from the feature rows
// idx[r]; rows n..B-1 are zero.  One workgroup per 64 x 64 tile, transposed
// through LDS.  The values are moved as raw 16-bit patterns (integer
// vectors): element access on BF16 vectors is not used for data movement.
__global__ void __launch_bounds__(ET) k_enc_gather(int Fp, int B, int n, const int *__restrict__ idx,
                                                   const bf16 *__restrict__ feat, bf16 *__restrict__ A0,
                                                   bf16 *__restrict__ A0t) {
    __shared__ unsigned short tile[64][66];
    const int f0 = blockIdx.x * 64, rb = blockIdx.y * 64, tid = threadIdx.x;
    {
        const int r = tid >> 2, c = (tid & 3) * 16;
        us8 x0 = 0, x1 = 0;
        if (rb + r < n) {
            const us8 *src = (const us8 *)(feat + (size_t)idx[rb + r] * Fp + f0 + c);
            x0 = src[0];
            x1 = src[1];
        }
        us8 *dst = (us8 *)(A0 + (size_t)(rb + r) * Fp + f0 + c);
        dst[0] = x0;
        dst[1] = x1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            tile[r][c + q] = x0[q];
            tile[r][c + 8 + q] = x1[q];
        }
    }
    __syncthreads();
    {
        const int f = tid >> 2, c = (tid & 3) * 16;
        us8 y0, y1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            y0[q] = tile[c + q][f];
            y1[q] = tile[c + 8 + q][f];
        }
        us8 *dst = (us8 *)(A0t + (size_t)(f0 + f) * B + rb + c);
        dst[0] = y0;
        dst[1] = y1;
    }
}

// Gaussian negative log-likelihood of the targets (without the 1/2 log 2 pi
// constant) and its gradient with respect to z3, over batch rows r < n
// (rows >= n and columns >= 2Q get zero).  The log-variance is clamped to
// [-LV_CLAMP, LV_CLAMP] (zero gradient outside) so that exp(-s) stays finite
// in FP32.  One thread per row.
constexpr float LV_CLAMP = 10.0f;

__global__ void __launch_bounds__(ET) k_enc_loss(int B, int n, int Q, int O, const int *__restrict__ idx,
                                                 const float *__restrict__ tgt, const float *__restrict__ Z3,
                                                 float *__restrict__ dZ3, bf16 *__restrict__ dZ3b,
                                                 bf16 *__restrict__ dZ3t, double *__restrict__ lpart) {
    __shared__ double sh[ET / 64];
    const int r = blockIdx.x * ET + threadIdx.x;
    double ls = 0.0;
    if (r < B) {
        const float inv_n = 1.0f / (float)n;
        for (int q = 0; q < O; q++) {
            float d = 0.0f;
            if (r < n && q < 2 * Q) {
                const int qq = q < Q ? q : q - Q;
                const float x = tgt[(size_t)idx[r] * Q + qq];
                const float mu = Z3[(size_t)r * O + qq], sr = Z3[(size_t)r * O + Q + qq];
                const float s = fminf(fmaxf(sr, -LV_CLAMP), LV_CLAMP);
                const float res = x - mu, es = expf(-s);
                if (q < Q) {
                    d = -res * es * inv_n;
                    ls += 0.5 * ((double)res * res * es + s);
                } else if (sr > -LV_CLAMP && sr < LV_CLAMP) {
                    d = 0.5f * (1.0f - res * res * es) * inv_n;
                }
            }
            dZ3[(size_t)r * O + q] = d;
            dZ3b[(size_t)r * O + q] = (bf16)d;
            dZ3t[(size_t)q * B + r] = (bf16)d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) ls += __shfl_down(ls, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ls;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < ET / 64; k++) s += sh[k];
        lpart[blockIdx.x] = s;
    }
}

// *lsum += sum of the nb partials (one thread: fixed order, reproducible).
__global__ void k_enc_lsum(int nb, const double *__restrict__ lpart, double *__restrict__ lsum) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < nb; k++) s += lpart[k];
        *lsum += s;
    }
}

// g[j] = sum over the B rows of X[r][j] (bias gradients).
__global__ void __launch_bounds__(ET) k_enc_colsum(int B, int C, const float *__restrict__ X,
                                                   float *__restrict__ g) {
    const int j = blockIdx.x * ET + threadIdx.x;
    if (j >= C) return;
    float s = 0.0f;
    for (int r = 0; r < B; r++) s += X[(size_t)r * C + j];
    g[j] = s;
}

// Adam on the FP32 masters (all parameters in one array).
__global__ void __launch_bounds__(ET) k_enc_adam(size_t np, float lr, float b1, float b2, float eps,
                                                 float c1, float c2, float *__restrict__ w,
                                                 const float *__restrict__ g, float *__restrict__ m,
                                                 float *__restrict__ v) {
    const size_t k = (size_t)blockIdx.x * ET + threadIdx.x;
    if (k >= np) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// BF16 copies of a row-major R x C master: wb (row-major) and, if wt is
// given, its transpose (C x R).
__global__ void __launch_bounds__(ET) k_enc_cast(int R, int C, const float *__restrict__ w,
                                                 bf16 *__restrict__ wb, bf16 *__restrict__ wt) {
    const size_t k = (size_t)blockIdx.x * ET + threadIdx.x;
    if (k >= (size_t)R * C) return;
    const int r = (int)(k / C), c = (int)(k % C);
    const bf16 x = (bf16)w[k];
    wb[k] = x;
    if (wt) wt[(size_t)c * R + r] = x;
}

// Features of cells base .. base + N - 1: feat[(base + i) Fp + j] =
// log1p(y_ij / s_i) for every stored count (the rest was zeroed).  One
// workgroup per gene of the gene-major CSR counts.
__global__ void __launch_bounds__(ET) k_enc_feat(int Fp, long long base, const int *__restrict__ Yp,
                                                 const int *__restrict__ Yi, const double *__restrict__ Yx,
                                                 const double *__restrict__ ls, bf16 *__restrict__ feat) {
    const int j = blockIdx.x;
    for (int k = Yp[j] + threadIdx.x; k < Yp[j + 1]; k += ET) {
        const int i = Yi[k];
        feat[(size_t)(base + i) * Fp + j] = (bf16)(float)log1p(Yx[k] / ls[i]);
    }
}

inline unsigned nblk(size_t n) { return (unsigned)((n + ET - 1) / ET); }

template <int E>
int gemm(gp_enc *en, int R, int C, int K, const bf16 *A, int lda, const bf16 *Bm, int ldb, Epi e) {
    if (R % 64 || C % 64 || K % 32 || K <= 0) {
        gp_set_error("encoder product shape not tiled: " + std::to_string(R) + " x " +
                     std::to_string(C) + " x " + std::to_string(K));
        return 1;
    }
    k_gemm_nt<E><<<dim3(R / 64, C / 64), ET, 0, en->stream>>>(K, A, lda, Bm, ldb, e);
    GP_HIP(hipGetLastError());
    return 0;
}

int refresh(gp_enc *en) {
    k_enc_cast<<<nblk((size_t)en->H1 * en->Fp), ET, 0, en->stream>>>(en->H1, en->Fp, en->w + en->oW1,
                                                                     en->W1b, nullptr);
    k_enc_cast<<<nblk((size_t)en->H2 * en->H1), ET, 0, en->stream>>>(en->H2, en->H1, en->w + en->oW2,
                                                                     en->W2b, en->W2t);
    k_enc_cast<<<nblk((size_t)en->O * en->H2), ET, 0, en->stream>>>(en->O, en->H2, en->w + en->oW3,
                                                                    en->W3b, en->W3t);
    GP_HIP(hipGetLastError());
    return 0;
}

// Forward pass of the batch already gathered into A0 / A0t.
int forward(gp_enc *en) {
    const int B = en->B;
    Epi e{};
    e = Epi{en->Z1, en->H1, en->w + en->ob1, nullptr, en->H1b, en->H1t, B};
    GP_TRY(gemm<E_BIAS_GELU>(en, B, en->H1, en->Fp, en->A0, en->Fp, en->W1b, en->Fp, e));
    e = Epi{en->Z2, en->H2, en->w + en->ob2, nullptr, en->H2b, en->H2t, B};
    GP_TRY(gemm<E_BIAS_GELU>(en, B, en->H2, en->H1, en->H1b, en->H1, en->W2b, en->H1, e));
    e = Epi{en->Z3, en->O, en->w + en->ob3, nullptr, nullptr, nullptr, 0};
    GP_TRY(gemm<E_BIAS>(en, B, en->O, en->H2, en->H2b, en->H2, en->W3b, en->H2, e));
    return 0;
}

// Loss (added to *lsum) and gradients of the batch rows r < n (feature rows idx[r]).
int loss_grad(gp_enc *en, int n, const int *idx) {
    const int B = en->B, nb = (B + ET - 1) / ET;
    k_enc_gather<<<dim3(en->Fp / 64, B / 64), ET, 0, en->stream>>>(en->Fp, B, n, idx, en->feat, en->A0,
                                                                   en->A0t);
    GP_HIP(hipGetLastError());
    GP_TRY(forward(en));
    k_enc_loss<<<nb, ET, 0, en->stream>>>(B, n, en->Q, en->O, idx, en->tgt, en->Z3, en->dZ3, en->dZ3b,
                                          en->dZ3t, en->lpart);
    k_enc_lsum<<<1, 64, 0, en->stream>>>(nb, en->lpart, en->lsum);
    GP_HIP(hipGetLastError());
    Epi e{};
    e = Epi{en->g + en->oW3, en->H2, nullptr, nullptr, nullptr, nullptr, 0};
    GP_TRY(gemm<E_STORE>(en, en->O, en->H2, B, en->dZ3t, B, en->H2t, B, e));
    e = Epi{en->dZ2, en->H2, nullptr, en->Z2, en->dZ2b, en->dZ2t, B};
    GP_TRY(gemm<E_GELU_BWD>(en, B, en->H2, en->O, en->dZ3b, en->O, en->W3t, en->O, e));
    e = Epi{en->g + en->oW2, en->H1, nullptr, nullptr, nullptr, nullptr, 0};
    GP_TRY(gemm<E_STORE>(en, en->H2, en->H1, B, en->dZ2t, B, en->H1t, B, e));
    e = Epi{en->dZ1, en->H1, nullptr, en->Z1, nullptr, en->dZ1t, B};
    GP_TRY(gemm<E_GELU_BWD>(en, B, en->H1, en->H2, en->dZ2b, en->H2, en->W2t, en->H2, e));
    e = Epi{en->g + en->oW1, en->Fp, nullptr, nullptr, nullptr, nullptr, 0};
    GP_TRY(gemm<E_STORE>(en, en->H1, en->Fp, B, en->dZ1t, B, en->A0t, B, e));
    k_enc_colsum<<<nblk(en->O), ET, 0, en->stream>>>(B, en->O, en->dZ3, en->g + en->ob3);
    k_enc_colsum<<<nblk(en->H2), ET, 0, en->stream>>>(B, en->H2, en->dZ2, en->g + en->ob2);
    k_enc_colsum<<<nblk(en->H1), ET, 0, en->stream>>>(B, en->H1, en->dZ1, en->g + en->ob1);
    GP_HIP(hipGetLastError());
    return 0;
}

int adam(gp_enc *en, double lr) {
    en->t++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)en->t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)en->t)));
    k_enc_adam<<<nblk(en->np), ET, 0, en->stream>>>(en->np, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2,
                                                    en->w, en->g, en->m, en->v);
    GP_HIP(hipGetLastError());
    return refresh(en);
}

// Features (and, if X is given, FP32 targets) of N cells into rows base.. of feat / tgt.
int load_cells(gp_enc *en, bf16 *feat, long long base, int N, int J, const int *Yp, const int *Yi,
               const double *Yx, const double *ls, const double *X) {
    const int nnz = Yp[J];
    int *dYp = nullptr, *dYi = nullptr;
    double *dYx = nullptr, *dls = nullptr;
    hipError_t err = hipSuccess;
    auto done = [&]() {
        (void)hipStreamSynchronize(en->stream);
        for (void *p : {(void *)dYp, (void *)dYi, (void *)dYx, (void *)dls})
            if (p) (void)hipFree(p);
    };
#define EN_CHK(call)                                                                  \
    do {                                                                              \
        if ((err = (call)) != hipSuccess) {                                           \
            gp_set_error(std::string(#call) + ": " + hipGetErrorString(err));         \
            done();                                                                   \
            return (int)err;                                                          \
        }                                                                             \
    } while (0)
    EN_CHK(hipMalloc((void **)&dYp, (size_t)(J + 1) * sizeof(int)));
    EN_CHK(hipMalloc((void **)&dYi, (size_t)std::max(nnz, 1) * sizeof(int)));
    EN_CHK(hipMalloc((void **)&dYx, (size_t)std::max(nnz, 1) * sizeof(double)));
    EN_CHK(hipMalloc((void **)&dls, (size_t)N * sizeof(double)));
    EN_CHK(hipMemcpyAsync(dYp, Yp, (size_t)(J + 1) * sizeof(int), hipMemcpyHostToDevice, en->stream));
    if (nnz > 0) {
        EN_CHK(hipMemcpyAsync(dYi, Yi, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, en->stream));
        EN_CHK(hipMemcpyAsync(dYx, Yx, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, en->stream));
    }
    EN_CHK(hipMemcpyAsync(dls, ls, (size_t)N * sizeof(double), hipMemcpyHostToDevice, en->stream));
    EN_CHK(hipMemsetAsync(feat + (size_t)base * en->Fp, 0, (size_t)N * en->Fp * sizeof(bf16), en->stream));
    k_enc_feat<<<J, ET, 0, en->stream>>>(en->Fp, base, dYp, dYi, dYx, dls, feat);
    EN_CHK(hipGetLastError());
    if (X) {
        std::vector<float> t((size_t)N * en->Q);
        for (int i = 0; i < N; i++)
            for (int q = 0; q < en->Q; q++) t[(size_t)i * en->Q + q] = (float)X[i + (size_t)q * N];
        EN_CHK(hipMemcpyAsync(en->tgt + (size_t)base * en->Q, t.data(), t.size() * sizeof(float),
                              hipMemcpyHostToDevice, en->stream));
        EN_CHK(hipStreamSynchronize(en->stream));   // t goes out of scope
    }
#undef EN_CHK
    done();
    return 0;
}

int check_counts(int N, int J, int F, const int *Yp, const int *Yi, const double *Yx, const double *ls) {
    if (N <= 0 || J <= 0) { gp_set_error("encoder: N and J must be positive"); return 1; }
    if (J != F) {
        gp_set_error("encoder: the counts have " + std::to_string(J) + " genes, the encoder " +
                     std::to_string(F) + " features");
        return 1;
    }
    if (Yp[0] != 0) { gp_set_error("encoder: Yp[0] must be 0"); return 1; }
    for (int j = 0; j < J; j++)
        if (Yp[j + 1] < Yp[j]) { gp_set_error("encoder: Yp must be non-decreasing"); return 1; }
    for (int k = 0; k < Yp[J]; k++) {
        if (Yi[k] < 0 || Yi[k] >= N) { gp_set_error("encoder: cell index out of range"); return 1; }
        if (!(Yx[k] >= 0.0)) { gp_set_error("encoder: negative or NaN count"); return 1; }
    }
    for (int i = 0; i < N; i++)
        if (!(ls[i] > 0.0)) { gp_set_error("encoder: size factors must be positive"); return 1; }
    return 0;
}

}  // namespace

extern "C" int gp_enc_destroy(gp_enc *en) {
    if (!en) return 0;
    if (en->stream) (void)hipStreamSynchronize(en->stream);
    for (void *p : {(void *)en->w, (void *)en->g, (void *)en->m, (void *)en->v, (void *)en->W1b,
                    (void *)en->W2b, (void *)en->W2t, (void *)en->W3b, (void *)en->W3t, (void *)en->feat,
                    (void *)en->tgt, (void *)en->perm, (void *)en->A0, (void *)en->A0t, (void *)en->Z1,
                    (void *)en->Z2, (void *)en->Z3, (void *)en->H1b, (void *)en->H1t, (void *)en->H2b,
                    (void *)en->H2t, (void *)en->dZ1, (void *)en->dZ2, (void *)en->dZ3, (void *)en->dZ1t,
                    (void *)en->dZ2b, (void *)en->dZ2t, (void *)en->dZ3b, (void *)en->dZ3t,
                    (void *)en->lpart, (void *)en->lsum})
        if (p) (void)hipFree(p);
    if (en->stream) (void)hipStreamDestroy(en->stream);
    delete en;
    return 0;
}

extern "C" int gp_enc_create(gp_enc **out, int F, int H1, int H2, int Q, int batch, long long cap,
                             unsigned long long seed) {
    if (!out) { gp_set_error("NULL argument"); return 1; }
    *out = nullptr;
    if (F <= 0 || Q <= 0 || 2 * Q > 64) { gp_set_error("encoder: need F > 0 and 0 < Q <= 32"); return 1; }
    if (H1 <= 0 || H2 <= 0 || H1 % 64 || H2 % 64) {
        gp_set_error("encoder: hidden widths must be positive multiples of 64");
        return 1;
    }
    if (batch <= 0 || batch % 64) { gp_set_error("encoder: batch must be a positive multiple of 64"); return 1; }
    if (cap < 0) { gp_set_error("encoder: negative capacity"); return 1; }
    gp_enc *en = new gp_enc();
    en->F = F;
    en->Fp = (F + 63) / 64 * 64;
    en->H1 = H1;
    en->H2 = H2;
    en->Q = Q;
    en->B = batch;
    en->cap = cap;
    en->rng.seed(seed);
    en->oW1 = 0;
    en->ob1 = en->oW1 + (size_t)H1 * en->Fp;
    en->oW2 = en->ob1 + H1;
    en->ob2 = en->oW2 + (size_t)H2 * H1;
    en->oW3 = en->ob2 + H2;
    en->ob3 = en->oW3 + (size_t)en->O * H2;
    en->np = en->ob3 + en->O;
    const size_t B = batch, Fp = en->Fp, O = en->O;
    hipError_t e = hipStreamCreateWithFlags(&en->stream, hipStreamNonBlocking);
#define EN_ALLOC(p, n) if (e == hipSuccess) e = hipMalloc((void **)&(p), std::max<size_t>((size_t)(n), 1) * sizeof(*(p)))
    EN_ALLOC(en->w, en->np);
    EN_ALLOC(en->g, en->np);
    EN_ALLOC(en->m, en->np);
    EN_ALLOC(en->v, en->np);
    EN_ALLOC(en->W1b, (size_t)H1 * Fp);
    EN_ALLOC(en->W2b, (size_t)H2 * H1);
    EN_ALLOC(en->W2t, (size_t)H2 * H1);
    EN_ALLOC(en->W3b, O * H2);
    EN_ALLOC(en->W3t, O * H2);
    EN_ALLOC(en->feat, (size_t)cap * Fp);
    EN_ALLOC(en->tgt, (size_t)cap * Q);
    EN_ALLOC(en->perm, std::max<long long>(cap, B));
    EN_ALLOC(en->A0, B * Fp);
    EN_ALLOC(en->A0t, B * Fp);
    EN_ALLOC(en->Z1, B * H1);
    EN_ALLOC(en->Z2, B * H2);
    EN_ALLOC(en->Z3, B * O);
    EN_ALLOC(en->H1b, B * H1);
    EN_ALLOC(en->H1t, B * H1);
    EN_ALLOC(en->H2b, B * H2);
    EN_ALLOC(en->H2t, B * H2);
    EN_ALLOC(en->dZ1, B * H1);
    EN_ALLOC(en->dZ2, B * H2);
    EN_ALLOC(en->dZ3, B * O);
    EN_ALLOC(en->dZ1t, B * H1);
    EN_ALLOC(en->dZ2b, B * H2);
    EN_ALLOC(en->dZ2t, B * H2);
    EN_ALLOC(en->dZ3b, B * O);
    EN_ALLOC(en->dZ3t, B * O);
    EN_ALLOC(en->lpart, (B + ET - 1) / ET);
    EN_ALLOC(en->lsum, 1);
#undef EN_ALLOC
    if (e != hipSuccess) {
        gp_set_error(std::string("encoder allocation: ") + hipGetErrorString(e));
        gp_enc_destroy(en);
        return (int)e;
    }
    // Linear-layer default: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and
    // biases; padded rows and columns stay zero (their gradients are zero).
    std::vector<float> w(en->np, 0.0f);
    auto fill = [&](size_t ow, size_t ob, int rows, int cols, int ld, int fan_in) {
        std::uniform_real_distribution<double> u(-1.0 / std::sqrt((double)fan_in), 1.0 / std::sqrt((double)fan_in));
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) w[ow + (size_t)r * ld + c] = (float)u(en->rng);
        for (int r = 0; r < rows; r++) w[ob + r] = (float)u(en->rng);
    };
    fill(en->oW1, en->ob1, H1, F, en->Fp, F);
    fill(en->oW2, en->ob2, H2, H1, H1, H1);
    fill(en->oW3, en->ob3, 2 * Q, H2, H2, H2);
    std::vector<int> iden(std::max<long long>(cap, B));
    std::iota(iden.begin(), iden.end(), 0);
    e = hipMemcpy(en->w, w.data(), en->np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(en->g, 0, en->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(en->m, 0, en->np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(en->v, 0, en->np * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(en->perm, iden.data(), iden.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        gp_set_error(std::string("encoder init: ") + hipGetErrorString(e));
        gp_enc_destroy(en);
        return (int)e;
    }
    int rc = refresh(en);
    if (rc == 0 && (e = hipStreamSynchronize(en->stream)) != hipSuccess) {
        gp_set_error(std::string("encoder init: ") + hipGetErrorString(e));
        rc = (int)e;
    }
    if (rc) { gp_enc_destroy(en); return rc; }
    *out = en;
    return 0;
}

extern "C" int gp_enc_dims(gp_enc *en, long long *dims) {
    if (!en || !dims) { gp_set_error("NULL argument"); return 1; }
    const long long d[12] = {en->F, en->Fp, en->H1, en->H2, en->Q, en->O, en->B, en->n,
                             (long long)en->np, en->cap, en->t, 0};
    std::copy(d, d + 12, dims);
    return 0;
}

extern "C" int gp_enc_add(gp_enc *en, int N, int J, const int *Yp, const int *Yi, const double *Yx,
                          const double *ls, const double *X) {
    if (!en || !Yp || !Yi || !Yx || !ls || !X) { gp_set_error("NULL argument"); return 1; }
    GP_TRY(check_counts(N, J, en->F, Yp, Yi, Yx, ls));
    if (en->n + N > en->cap) {
        gp_set_error("encoder: " + std::to_string(en->n + N) + " cells exceed the capacity " +
                     std::to_string(en->cap));
        return 1;
    }
    GP_TRY(load_cells(en, en->feat, en->n, N, J, Yp, Yi, Yx, ls, X));
    en->n += N;
    return 0;
}

extern "C" int gp_enc_loss_grad(gp_enc *en, int n, double *loss) {
    if (!en || !loss) { gp_set_error("NULL argument"); return 1; }
    if (n <= 0 || n > en->B || n > en->n) {
        gp_set_error("encoder: n must be in [1, min(batch, cells loaded)]");
        return 1;
    }
    GP_HIP(hipMemsetAsync(en->lsum, 0, sizeof(double), en->stream));
    GP_TRY(loss_grad(en, n, en->perm));   // perm holds the identity: cells 0..n-1 in load order
    double s = 0.0;
    GP_HIP(hipMemcpyAsync(&s, en->lsum, sizeof(double), hipMemcpyDeviceToHost, en->stream));
    GP_HIP(hipStreamSynchronize(en->stream));
    *loss = s / n;
    return 0;
}

extern "C" int gp_enc_adam_step(gp_enc *en, double lr) {
    if (!en) { gp_set_error("NULL argument"); return 1; }
    if (!(lr > 0.0)) { gp_set_error("encoder: learning rate must be positive"); return 1; }
    GP_TRY(adam(en, lr));
    GP_HIP(hipStreamSynchronize(en->stream));
    return 0;
}

extern "C" int gp_enc_train(gp_enc *en, int epochs, double lr, double *epoch_loss) {
    if (!en) { gp_set_error("NULL argument"); return 1; }
    if (epochs < 0 || !(lr > 0.0)) { gp_set_error("encoder: epochs >= 0 and lr > 0 required"); return 1; }
    if (en->n <= 0) { gp_set_error("encoder: no training cells loaded"); return 1; }
    const long long n = en->n;
    en->perm_h.resize(n);
    int *perm_d = nullptr;
    GP_HIP(hipMalloc((void **)&perm_d, (size_t)n * sizeof(int)));
    int rc = 0;
    for (int ep = 0; ep < epochs && rc == 0; ep++) {
        std::iota(en->perm_h.begin(), en->perm_h.end(), 0);
        std::shuffle(en->perm_h.begin(), en->perm_h.end(), en->rng);
        hipError_t e = hipMemcpyAsync(perm_d, en->perm_h.data(), (size_t)n * sizeof(int),
                                      hipMemcpyHostToDevice, en->stream);
        if (e == hipSuccess) e = hipMemsetAsync(en->lsum, 0, sizeof(double), en->stream);
        if (e != hipSuccess) { gp_set_error(std::string("encoder train: ") + hipGetErrorString(e)); rc = (int)e; break; }
        for (long long s0 = 0; s0 < n && rc == 0; s0 += en->B) {
            const int nb = (int)std::min<long long>(en->B, n - s0);
            rc = loss_grad(en, nb, perm_d + s0);
            if (rc == 0) rc = adam(en, lr);
        }
        double s = 0.0;
        if (rc == 0) {
            e = hipMemcpyAsync(&s, en->lsum, sizeof(double), hipMemcpyDeviceToHost, en->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(en->stream);
            if (e != hipSuccess) { gp_set_error(std::string("encoder train: ") + hipGetErrorString(e)); rc = (int)e; }
        }
        // per-batch losses are batch means: the epoch value is their cell-weighted mean
        if (rc == 0 && epoch_loss) epoch_loss[ep] = s / (double)n;
        if (rc == 0 && !std::isfinite(s)) {
            gp_set_error("encoder training diverged: non-finite loss in epoch " + std::to_string(ep));
            rc = 3;
        }
    }
    (void)hipStreamSynchronize(en->stream);
    (void)hipFree(perm_d);
    return rc;
}

extern "C" int gp_enc_predict(gp_enc *en, int N, int J, const int *Yp, const int *Yi, const double *Yx,
                              const double *ls, double *mu, double *logvar) {
    if (!en || !Yp || !Yi || !Yx || !ls || !mu) { gp_set_error("NULL argument"); return 1; }
    GP_TRY(check_counts(N, J, en->F, Yp, Yi, Yx, ls));
    bf16 *feat = nullptr;
    GP_HIP(hipMalloc((void **)&feat, (size_t)N * en->Fp * sizeof(bf16)));
    int rc = load_cells(en, feat, 0, N, J, Yp, Yi, Yx, ls, nullptr);
    std::vector<float> z((size_t)en->B * en->O);
    for (int s0 = 0; s0 < N && rc == 0; s0 += en->B) {
        const int nb = std::min(en->B, N - s0);
        k_enc_gather<<<dim3(en->Fp / 64, en->B / 64), ET, 0, en->stream>>>(
            en->Fp, en->B, nb, en->perm, feat + (size_t)s0 * en->Fp, en->A0, en->A0t);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { gp_set_error(std::string("encoder gather: ") + hipGetErrorString(e)); rc = (int)e; break; }
        rc = forward(en);
        if (rc) break;
        e = hipMemcpyAsync(z.data(), en->Z3, z.size() * sizeof(float), hipMemcpyDeviceToHost, en->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(en->stream);
        if (e != hipSuccess) { gp_set_error(std::string("encoder predict: ") + hipGetErrorString(e)); rc = (int)e; break; }
        for (int r = 0; r < nb; r++)
            for (int q = 0; q < en->Q; q++) {
                mu[(size_t)(s0 + r) + (size_t)q * N] = z[(size_t)r * en->O + q];
                if (logvar) logvar[(size_t)(s0 + r) + (size_t)q * N] = z[(size_t)r * en->O + en->Q + q];
            }
    }
    (void)hipStreamSynchronize(en->stream);
    (void)hipFree(feat);
    return rc;
}

// "params", "grad", "adam_m", "adam_v": FP32 arrays of np values in the
// padded layout [W1 | b1 | W2 | b2 | W3 | b3].
extern "C" int gp_enc_get(gp_enc *en, const char *name, float *host, size_t n) {
    if (!en || !name || !host) { gp_set_error("NULL argument"); return 1; }
    const std::string s(name);
    const float *src = s == "params" ? en->w : s == "grad" ? en->g : s == "adam_m" ? en->m
                     : s == "adam_v" ? en->v : nullptr;
    if (!src) { gp_set_error("encoder: unknown array " + s); return 1; }
    if (n != en->np) { gp_set_error("encoder: array has " + std::to_string(en->np) + " values"); return 1; }
    GP_HIP(hipStreamSynchronize(en->stream));
    GP_HIP(hipMemcpy(host, src, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

// "params" only; the BF16 copies are refreshed and the Adam state is kept.
extern "C" int gp_enc_set(gp_enc *en, const char *name, const float *host, size_t n) {
    if (!en || !name || !host) { gp_set_error("NULL argument"); return 1; }
    if (std::string(name) != "params") { gp_set_error("encoder: only \"params\" can be set"); return 1; }
    if (n != en->np) { gp_set_error("encoder: array has " + std::to_string(en->np) + " values"); return 1; }
    GP_HIP(hipStreamSynchronize(en->stream));
    GP_HIP(hipMemcpy(en->w, host, n * sizeof(float), hipMemcpyHostToDevice));
    GP_TRY(refresh(en));
    GP_HIP(hipStreamSynchronize(en->stream));
    return 0;
}
