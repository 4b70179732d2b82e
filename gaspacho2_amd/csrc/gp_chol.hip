// gp_chol.hip -- batched FP64 Cholesky, triangular solve and inverse.
//
// Cholesky: left-looking, blocked by 16-column panels held in LDS; one
// workgroup per matrix.  Input is the lower triangle (the Gram writes both
// triangles), L overwrites lower+diag and the strict upper triangle keeps the
// original matrix -- the layout choldc (chol.cu:19-43) leaves behind.  A
// matrix that meets a non-positive pivot is zero-filled (chol.cu:26-29) and
// counted.
#include "gp_internal.h"

namespace {

constexpr int CT = 256;   // threads per workgroup
constexpr int NB = 16;    // panel width
constexpr int KC = 64;    // k-chunk of previous columns staged in LDS
constexpr int MAXN = 640; // largest n handled (panel rows in LDS)

__global__ void __launch_bounds__(CT) k_chol(double *__restrict__ Aall, int n, size_t stride,
                                             int *__restrict__ info) {
    extern __shared__ double lds[];
    double *pan = lds;                 // [n][NB] panel, row-major by panel row
    double *lc = lds + (size_t)n * NB; // [NB][KC] rows kb..kb+NB of L, k-chunk
    __shared__ int bad;
    double *a = Aall + (size_t)blockIdx.x * stride;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    __syncthreads();

    for (int kb = 0; kb < n; kb += NB) {
        const int nbk = min(NB, n - kb);
        const int rows = n - kb;
        // 1. load the panel's lower part
        for (int t = tid; t < rows * nbk; t += CT) {
            const int r = t % rows, cc = t / rows;
            pan[r * NB + cc] = (r >= cc) ? a[(kb + r) + (size_t)(kb + cc) * n] : 0.0;
        }
        __syncthreads();
        // 2. subtract the contribution of the columns already factored
        for (int r0 = 0; r0 < rows; r0 += CT) {
            const int r = r0 + tid;
            double acc[NB];
#pragma unroll
            for (int cc = 0; cc < NB; cc++) acc[cc] = 0.0;
            for (int k0 = 0; k0 < kb; k0 += KC) {
                const int kn = min(KC, kb - k0);
                __syncthreads();
                for (int t = tid; t < nbk * kn; t += CT) {
                    const int cc = t / kn, kk = t % kn;
                    lc[cc * KC + kk] = a[(kb + cc) + (size_t)(k0 + kk) * n];
                }
                __syncthreads();
                if (r < rows) {
                    for (int kk = 0; kk < kn; kk++) {
                        const double lr = a[(kb + r) + (size_t)(k0 + kk) * n];
#pragma unroll
                        for (int cc = 0; cc < NB; cc++) acc[cc] += lr * lc[cc * KC + kk];
                    }
                }
            }
            if (r < rows) {
#pragma unroll
                for (int cc = 0; cc < NB; cc++)
                    if (cc < nbk && cc <= r) pan[r * NB + cc] -= acc[cc];
            }
        }
        __syncthreads();
        // 3. factor the panel in LDS
        for (int cc = 0; cc < nbk; cc++) {
            const double piv = pan[cc * NB + cc];
            if (piv <= 0.0) {           // same test as chol.cu:26
                if (tid == 0) bad = 1;
                break;
            }
            const double d = sqrt(piv);
            __syncthreads();
            for (int r = cc + 1 + tid; r < rows; r += CT) pan[r * NB + cc] /= d;
            if (tid == 0) pan[cc * NB + cc] = d;
            __syncthreads();
            for (int r = cc + 1 + tid; r < rows; r += CT) {
                const double lrc = pan[r * NB + cc];
                const int cmax = min(nbk - 1, r);
                for (int c2 = cc + 1; c2 <= cmax; c2++) pan[r * NB + c2] -= lrc * pan[c2 * NB + cc];
            }
            __syncthreads();
        }
        __syncthreads();
        if (bad) break;
        // 4. write the factored panel back
        for (int t = tid; t < rows * nbk; t += CT) {
            const int r = t % rows, cc = t / rows;
            if (r >= cc) a[(kb + r) + (size_t)(kb + cc) * n] = pan[r * NB + cc];
        }
        __syncthreads();
    }
    if (bad) {
        for (size_t t = tid; t < (size_t)n * n; t += CT) a[t] = 0.0;
        if (tid == 0) atomicAdd(info, 1);
    }
}

// x_j = L_j^{-T} L_j^{-1} b_j, one workgroup per right-hand side.
__global__ void __launch_bounds__(CT) k_cholsl(const double *__restrict__ Lall, int n,
                                               size_t lstride, const double *__restrict__ b,
                                               double *__restrict__ x) {
    extern __shared__ double xs[];
    const int j = blockIdx.x, tid = threadIdx.x;
    const double *L = Lall + (size_t)j * lstride;
    for (int k = tid; k < n; k += CT) xs[k] = b[(size_t)j * n + k];
    for (int i = 0; i < n; i++) {       // L y = b
        __syncthreads();
        const double xi = xs[i] / L[i + (size_t)i * n];
        __syncthreads();
        if (tid == 0) xs[i] = xi;
        for (int r = i + 1 + tid; r < n; r += CT) xs[r] -= L[r + (size_t)i * n] * xi;
    }
    for (int i = n - 1; i >= 0; i--) {  // L^T x = y
        __syncthreads();
        const double xi = xs[i] / L[i + (size_t)i * n];
        __syncthreads();
        if (tid == 0) xs[i] = xi;
        for (int r = tid; r < i; r += CT) xs[r] -= L[i + (size_t)r * n] * xi;
    }
    __syncthreads();
    for (int k = tid; k < n; k += CT) x[(size_t)j * n + k] = xs[k];
}

__global__ void k_identity(double *S, int n, int batch, size_t stride) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t nn = (size_t)n * n;
    if (idx >= nn * batch) return;
    const size_t j = idx / nn, e = idx % nn;
    S[j * stride + e] = ((e % n) == (e / n)) ? 1.0 : 0.0;
}

}  // namespace

int gpk_chol_batched(gp_ctx *c, double *A, int n, int batch, size_t stride, int *info) {
    if (n > MAXN) { gp_set_error("Cholesky size exceeds 640"); return 1; }
    const size_t shm = ((size_t)n * NB + (size_t)NB * KC) * sizeof(double);
    if (shm > 65536)
        GP_HIP(hipFuncSetAttribute((const void *)k_chol, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    k_chol<<<batch, CT, shm, c->stream>>>(A, n, stride, info);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_cholsl_batched(gp_ctx *c, const double *L, int n, size_t lstride, const double *b,
                       double *x, int batch) {
    k_cholsl<<<batch, CT, (size_t)n * sizeof(double), c->stream>>>(L, n, lstride, b, x);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_chol_inverse_batched(gp_ctx *c, const double *L, int n, int batch, size_t stride,
                             double *S) {
    const size_t tot = (size_t)n * n * batch;
    k_identity<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(S, n, batch, stride);
    GP_HIP(hipGetLastError());
    const double one = 1.0;
    GP_BLAS(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_lower,
                                          rocblas_operation_none, rocblas_diagonal_non_unit, n, n,
                                          &one, L, n, (rocblas_stride)stride, S, n,
                                          (rocblas_stride)stride, batch));
    GP_BLAS(rocblas_dtrsm_strided_batched(c->blas, rocblas_side_left, rocblas_fill_lower,
                                          rocblas_operation_transpose, rocblas_diagonal_non_unit, n,
                                          n, &one, L, n, (rocblas_stride)stride, S, n,
                                          (rocblas_stride)stride, batch));
    return 0;
}

int gpk_tri_inverse(gp_ctx *c, const double *L, int n, double *V) {
    const size_t tot = (size_t)n * n;
    k_identity<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(V, n, 1, 0);
    GP_HIP(hipGetLastError());
    const double one = 1.0;
    GP_BLAS(rocblas_dtrsm(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                          rocblas_diagonal_non_unit, n, n, &one, L, n, V, n));
    return 0;
}
