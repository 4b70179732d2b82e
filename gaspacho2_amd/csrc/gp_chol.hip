// gp_chol.hip -- batched FP64 Cholesky, triangular solve and inverse.
//
// Inverse (solvePsi, ggplvm.cu:46-53): V = L^{-1} column by column, then the
// symmetric product S = V^T V -- both hand-written, one workgroup per gene.
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



// One matrix: column j of V = L^{-1} by column-oriented forward substitution
// (one workgroup per column; the strict upper triangle of V is written as 0).
__global__ void __launch_bounds__(CT) k_trinv_cols(const double *__restrict__ L, int n,
                                                   double *__restrict__ V) {
    extern __shared__ double x[];      // running right-hand side of column j
    const int j = blockIdx.x, tid = threadIdx.x;
    double *v = V + (size_t)j * n;
    for (int r = tid; r < n; r += CT) {
        x[r] = (r == j) ? 1.0 : 0.0;
        if (r < j) v[r] = 0.0;
    }
    __syncthreads();
    for (int k = j; k < n; k++) {
        const double vk = x[k] / L[k + (size_t)k * n];
        if (tid == 0) v[k] = vk;
        for (int r = k + 1 + tid; r < n; r += CT) x[r] -= L[r + (size_t)k * n] * vk;
        __syncthreads();
    }
}

// V_j = L_j^{-1} (lower triangle of V; the strict upper triangle is not
// written).  Columns are produced right to left: with L = [[l, 0], [c, L2]],
// V[j+1:, j] = -V2 c / l where V2 = V[j+1:, j+1:] is already known.
__global__ void __launch_bounds__(CT) k_trinv(const double *__restrict__ Lall, int n,
                                              size_t stride, double *__restrict__ Vall) {
    extern __shared__ double lcol[];
    const double *L = Lall + (size_t)blockIdx.x * stride;
    double *V = Vall + (size_t)blockIdx.x * stride;
    const int tid = threadIdx.x;
    for (int j = n - 1; j >= 0; j--) {
        for (int k = j + 1 + tid; k < n; k += CT) lcol[k] = L[k + (size_t)j * n];
        __syncthreads();
        const double vjj = 1.0 / L[j + (size_t)j * n];
        for (int i = j + 1 + tid; i < n; i += CT) {
            double s = 0.0;
            for (int k = j + 1; k <= i; k++) s += V[i + (size_t)k * n] * lcol[k];
            V[i + (size_t)j * n] = -s * vjj;
        }
        if (tid == 0) V[j + (size_t)j * n] = vjj;
        __syncthreads();
    }
}

// Blocked V_j = L_j^{-1} (lower triangle; the strict upper triangle is not
// written), 32 x 32 blocks.  From L V = I, block row I > J of block column J:
//   L_II V_IJ = -sum_{K=J}^{I-1} L_IK V_KJ   =>   V_IJ = -V_II T_IJ,
// so with the diagonal inverses V_II known, every block column is
// independent: one workgroup per (block column, gene) walks I downwards
// keeping the column's finished blocks in LDS.  L is read once per block
// product (from L2), V is written once.
constexpr int TB = 32;

// V_II = L_II^{-1} for every diagonal block: lane c solves column c by
// forward substitution in LDS.
__global__ void __launch_bounds__(64) k_trinv_diag(const double *__restrict__ Lall, int n, size_t stride,
                                                   double *__restrict__ Vall) {
    __shared__ double Ld[TB][TB + 1];
    __shared__ double Vd[TB][TB + 1];
    const double *L = Lall + (size_t)blockIdx.y * stride;
    double *V = Vall + (size_t)blockIdx.y * stride;
    const int r0 = TB * blockIdx.x, bs = min(TB, n - r0), tid = threadIdx.x;
    for (int idx = tid; idx < TB * TB; idx += 64) {
        const int r = idx % TB, c = idx / TB;
        Ld[r][c] = (r < bs && c < bs && r >= c) ? L[(r0 + r) + (size_t)(r0 + c) * n] : 0.0;
    }
    __syncthreads();
    if (tid < bs) {
        const int c = tid;
        for (int r = c; r < bs; r++) {
            double sum = (r == c) ? 1.0 : 0.0;
            for (int k = c; k < r; k++) sum -= Ld[r][k] * Vd[k][c];
            Vd[r][c] = sum / Ld[r][r];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < TB * TB; idx += 64) {
        const int r = idx % TB, c = idx / TB;
        if (r < bs && c < bs && r >= c) V[(r0 + r) + (size_t)(r0 + c) * n] = Vd[r][c];
    }
}

// Off-diagonal blocks of block column J = blockIdx.x (all but the last block
// column).  Thread t owns the 4 entries (t / 8, 4 (t % 8) + q) of a block.
__global__ void __launch_bounds__(CT) k_trinv_panel(const double *__restrict__ Lall, int n, size_t stride,
                                                    double *__restrict__ Vall) {
    extern __shared__ double sm[];
    double (*Lb)[TB + 1] = (double (*)[TB + 1])sm;                  // L_IK
    double (*Vjj)[TB + 1] = (double (*)[TB + 1])(sm + TB * (TB + 1));   // V_JJ
    double (*Tt)[TB + 1] = (double (*)[TB + 1])(sm + 2 * TB * (TB + 1)); // T, then V_II
    double (*Vp)[TB + 1] = (double (*)[TB + 1])(sm + 3 * TB * (TB + 1)); // finished V_KJ, K > J
    const double *L = Lall + (size_t)blockIdx.y * stride;
    double *V = Vall + (size_t)blockIdx.y * stride;
    const int NB = (n + TB - 1) / TB, J = blockIdx.x, c0 = TB * J, tid = threadIdx.x;
    const int tr = tid / 8, tc = 4 * (tid % 8);
    for (int idx = tid; idx < TB * TB; idx += CT) {
        const int r = idx % TB, c = idx / TB;
        Vjj[r][c] = (r >= c) ? V[(c0 + r) + (size_t)(c0 + c) * n] : 0.0;   // block J is full (J < NB - 1)
    }
    for (int I = J + 1; I < NB; I++) {
        const int rI = TB * I, bsI = min(TB, n - rI);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int K = J; K < I; K++) {
            __syncthreads();                      // Lb free (and Vjj / Vp writes visible)
            for (int idx = tid; idx < TB * TB; idx += CT) {
                const int r = idx % TB, k = idx / TB;
                Lb[r][k] = (r < bsI) ? L[(rI + r) + (size_t)(TB * K + k) * n] : 0.0;   // K < I: full width
            }
            __syncthreads();
            const double (*Bk)[TB + 1] = (K == J) ? (const double (*)[TB + 1])Vjj
                                                  : (const double (*)[TB + 1])&Vp[TB * (K - J - 1)];
#pragma unroll 8
            for (int k = 0; k < TB; k++) {
                const double l = Lb[tr][k];
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] += l * Bk[k][tc + q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) Tt[tr][tc + q] = acc[q];
        // V_II into Lb (lower triangle, zero elsewhere)
        for (int idx = tid; idx < TB * TB; idx += CT) {
            const int r = idx % TB, k = idx / TB;
            Lb[r][k] = (r < bsI && k < bsI && r >= k) ? V[(rI + r) + (size_t)(rI + k) * n] : 0.0;
        }
        __syncthreads();
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int k = 0; k < TB; k++) {
            const double l = Lb[tr][k];
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] -= l * Tt[k][tc + q];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) Vp[TB * (I - J - 1) + tr][tc + q] = v[q];
        __syncthreads();
        for (int idx = tid; idx < TB * TB; idx += CT) {
            const int r = idx % TB, c = idx / TB;
            if (r < bsI) V[(rI + r) + (size_t)(c0 + c) * n] = Vp[TB * (I - J - 1) + r][c];
        }
    }
}

// S_j = V_j^T V_j (full symmetric) from the lower-triangular V_j: 64x64 tiles
// of the upper triangle, 32-row chunks of V staged in LDS, 4x4 per thread.
// Entry (a, b), a <= b, sums rows k >= b only (V[k, b] = 0 above the diagonal).
__global__ void __launch_bounds__(CT) k_vtv(const double *__restrict__ Vall, int n, size_t stride,
                                            int ntile, double *__restrict__ Sall) {
    __shared__ double As[32][65];
    __shared__ double Bs[32][65];
    int ta = 0, t = blockIdx.x;
    while (t >= ntile - ta) { t -= ntile - ta; ta++; }
    const int tb = ta + t;
    const int a0 = ta * 64, b0 = tb * 64;
    const double *V = Vall + (size_t)blockIdx.y * stride;
    double *S = Sall + (size_t)blockIdx.y * stride;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r][q] = 0.0;
    for (int k0 = b0; k0 < n; k0 += 32) {
        __syncthreads();
        const int row = tid % 32, cg = tid / 32;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int col = cg * 8 + u, k = k0 + row;
            const int ga = a0 + col, gb = b0 + col;
            As[row][col] = (k < n && ga < n && k >= ga) ? V[k + (size_t)ga * n] : 0.0;
            Bs[row][col] = (k < n && gb < n && k >= gb) ? V[k + (size_t)gb * n] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < 32; kk++) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) av[r] = As[kk][ty * 4 + r];
#pragma unroll
            for (int q = 0; q < 4; q++) bv[q] = Bs[kk][tx * 4 + q];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[r][q] += av[r] * bv[q];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ga = a0 + ty * 4 + r, gb = b0 + tx * 4 + q;
            if (ga < n && gb < n && ga <= gb) {
                S[ga + (size_t)gb * n] = acc[r][q];
                S[gb + (size_t)ga * n] = acc[r][q];
            }
        }
}

}  // namespace

int gpk_chol_batched(gp_ctx *c, double *A, int n, int batch, size_t stride, int *info) {
    if (n > MAXN) { gp_set_error("Cholesky size exceeds 640"); return 1; }
    const size_t shm = ((size_t)n * NB + (size_t)NB * KC) * sizeof(double);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
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
    // V = L^{-1} into the scratch buffer, then S = V^T V
    if ((size_t)batch * (stride ? stride : (size_t)n * n) > c->vtmp_len) {
        if (c->Vtmp) (void)hipFree(c->Vtmp);
        c->Vtmp = nullptr;
        c->vtmp_len = (size_t)batch * (stride ? stride : (size_t)n * n);
        GP_HIP(hipMalloc((void **)&c->Vtmp, c->vtmp_len * sizeof(double)));
    }
    const size_t st = stride ? stride : (size_t)n * n;
    const int nb = (n + TB - 1) / TB;
    const size_t shm = (size_t)(3 + (nb - 1)) * TB * (TB + 1) * sizeof(double);
    if (shm > 160 * 1024) {   // the block column does not fit in LDS: column-by-column form
        k_trinv<<<batch, CT, (size_t)n * sizeof(double), c->stream>>>(L, n, st, c->Vtmp);
        GP_HIP(hipGetLastError());
    } else {
        k_trinv_diag<<<dim3(nb, batch), 64, 0, c->stream>>>(L, n, st, c->Vtmp);
        GP_HIP(hipGetLastError());
        if (nb > 1) {
            if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
                GP_HIP(hipFuncSetAttribute((const void *)k_trinv_panel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            k_trinv_panel<<<dim3(nb - 1, batch), CT, shm, c->stream>>>(L, n, st, c->Vtmp);
            GP_HIP(hipGetLastError());
        }
    }
    const int nt = (n + 63) / 64;
    dim3 g(nt * (nt + 1) / 2, batch);
    k_vtv<<<g, CT, 0, c->stream>>>(c->Vtmp, n, st, nt, S);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_vtv(gp_ctx *c, const double *V, int n, int batch, size_t stride, double *S) {
    const size_t st = stride ? stride : (size_t)n * n;
    const int nt = (n + 63) / 64;
    dim3 g(nt * (nt + 1) / 2, batch);
    k_vtv<<<g, CT, 0, c->stream>>>(V, n, st, nt, S);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_tri_inverse(gp_ctx *c, const double *L, int n, double *V) {
    k_trinv_cols<<<n, CT, (size_t)n * sizeof(double), c->stream>>>(L, n, V);
    GP_HIP(hipGetLastError());
    return 0;
}
