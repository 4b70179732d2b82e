// gp_chol.hip -- batched FP64 Cholesky, triangular solve and inverse.
//
// Inverse (solvePsi, ggplvm.cu:46-53): V = L^{-1} in 32 x 32 blocks (diagonal
// inverses, then the block columns on FP64 MFMA), then the symmetric product
// S = V^T V -- all hand-written.
//
// Cholesky: left-looking, blocked by 8-column panels held in LDS (16-column
// panels as an option); one workgroup per matrix.  Input is the lower triangle (the Gram writes both
// triangles), L overwrites lower+diag and the strict upper triangle keeps the
// original matrix -- the layout choldc (chol.cu:19-43) leaves behind.  A
// matrix that meets a non-positive pivot is zero-filled (chol.cu:26-29) and
// counted.
#include "gp_internal.h"

namespace {

constexpr int CT = 256;   // threads per workgroup
constexpr int MAXN = 640; // largest n handled (panel rows in LDS)
constexpr int PLD = 8;    // panel-load batch per thread

// Panel width NB and k-chunk KC of the batched Cholesky.  NB = 16, KC = 48
// keeps a P = 257 workgroup at 39 KB of LDS, four per CU (KC = 64 made it
// 41 KB, three per CU); NB = 8, KC = 32 takes 18.5 KB, eight per CU, so 2,048
// matrices are in flight at once (twice the panels, half the work per panel).
template <int NB, int KC>
__global__ void __launch_bounds__(CT, NB == 8 ? 8 : 1) k_chol(double *__restrict__ Aall, int n, size_t stride,
                                             int *__restrict__ info) {
    static_assert(KC % NB == 0, "k-chunks hold whole panels");
    static_assert(NB * KC % CT == 0, "the k-chunk stage splits evenly over the threads");
    constexpr int PL = NB == 8 ? 4 : PLD;   // panel loads in flight (the 8-column form fits 64 VGPRs)
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
        // 1. load the panel's lower part, PL loads in flight per thread
        for (int t0 = tid; t0 < rows * nbk; t0 += CT * PL) {
            double v[PL];
#pragma unroll
            for (int u = 0; u < PL; u++) {
                const int t = t0 + u * CT;
                const int r = t % rows, cc = t / rows;
                v[u] = (t < rows * nbk && r >= cc) ? a[(kb + r) + (size_t)(kb + cc) * n] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < PL; u++) {
                const int t = t0 + u * CT;
                if (t < rows * nbk) pan[(t % rows) * NB + t / rows] = v[u];
            }
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
                {
                    // nbk * kn <= NB * KC = 4 * CT: every load in flight at once
                    double v[NB * KC / CT];
#pragma unroll
                    for (int u = 0; u < NB * KC / CT; u++) {
                        const int t = tid + u * CT;
                        const int cc = t / kn, kk = t % kn;
                        v[u] = (t < nbk * kn) ? a[(kb + cc) + (size_t)(k0 + kk) * n] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < NB * KC / CT; u++) {
                        const int t = tid + u * CT;
                        if (t < nbk * kn) lc[(t / kn) * KC + t % kn] = v[u];
                    }
                }
                __syncthreads();
                if (r < rows) {
                    // kn is a multiple of NB (kb and KC are): NB independent
                    // loads of row r in flight per step instead of one
                    for (int kk = 0; kk < kn; kk += NB) {
                        double lr[NB];
#pragma unroll
                        for (int u = 0; u < NB; u++) lr[u] = a[(kb + r) + (size_t)(k0 + kk + u) * n];
#pragma unroll
                        for (int u = 0; u < NB; u++)
#pragma unroll
                            for (int cc = 0; cc < NB; cc++) acc[cc] += lr[u] * lc[cc * KC + kk + u];
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
        // 3. factor the panel: the nbk x nbk diagonal block on wave 0 (lane r
        //    holds row r, pivots and columns broadcast by shuffles), then each
        //    row below by forward substitution against it -- the same
        //    operations in the same order as column-by-column elimination,
        //    with two barriers instead of two per column
        if (tid < 64) {
            const int r = tid;
            double x[NB];
#pragma unroll
            for (int cc = 0; cc < NB; cc++) x[cc] = (r < nbk) ? pan[min(r, nbk - 1) * NB + cc] : 0.0;
            int fail = 0;
#pragma unroll
            for (int cc = 0; cc < NB; cc++) {
                if (cc >= nbk || fail) break;
                const double piv = __shfl(x[cc], cc);
                if (piv <= 0.0) { fail = 1; break; }   // same test as chol.cu:26
                const double d = sqrt(piv);
                if (r == cc) x[cc] = d;
                else if (r > cc) x[cc] /= d;
#pragma unroll
                for (int c2 = cc + 1; c2 < NB; c2++) {
                    const double l2 = __shfl(x[cc], c2);
                    if (r > cc && r >= c2 && c2 < nbk) x[c2] -= x[cc] * l2;
                }
            }
            if (r < nbk) {
#pragma unroll
                for (int cc = 0; cc < NB; cc++)
                    if (cc <= r) pan[r * NB + cc] = x[cc];
            }
            if (tid == 0 && fail) bad = 1;
        }
        __syncthreads();
        if (bad) break;
        for (int r = nbk + tid; r < rows; r += CT) {
            double x[NB];
#pragma unroll
            for (int cc = 0; cc < NB; cc++) x[cc] = pan[r * NB + cc];
#pragma unroll
            for (int cc = 0; cc < NB; cc++) {
                if (cc < nbk) {
                    double v = x[cc];
#pragma unroll
                    for (int c = 0; c < cc; c++) v -= x[c] * pan[cc * NB + c];
                    x[cc] = v / pan[cc * NB + cc];
                }
            }
#pragma unroll
            for (int cc = 0; cc < NB; cc++) pan[r * NB + cc] = x[cc];
        }
        __syncthreads();
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

// One matrix (the Kmm factor): right-looking blocked Cholesky on a
// single 1024-thread workgroup.  Per 32-column panel: the panel's rows are
// staged in LDS, the 32 x 32 diagonal block is factored on wave 0 (rows in
// lanes, pivots and columns by shuffles), the rows below are solved against
// it, the panel is written back, and the trailing lower triangle takes the
// panel's rank-32 update in 4 x 4 register blocks.  The batched kernel gives
// a lone matrix one 256-thread workgroup and reloads every earlier column per
// panel; here the trailing update runs over all the workgroup's threads.
// Same layout and zero-fill rule as k_chol (reads the lower triangle, L to
// lower + diag, the strict upper triangle untouched; a non-positive pivot
// zero-fills the matrix and counts it).
constexpr int C1T = 1024;   // threads
constexpr int NB1 = 32;     // panel width
constexpr int PS1 = NB1 + 1;  // LDS row stride of the panel (doubles)
constexpr int MAXN1 = 512;  // largest n (the panel's rows take n * 33 * 8 bytes of LDS)

__global__ void __launch_bounds__(C1T) k_chol1(double *__restrict__ a, int n, int *__restrict__ info) {
    extern __shared__ double pan[];    // [rows][PS1]
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int kb = 0; kb < n; kb += NB1) {
        const int nb = min(NB1, n - kb);
        const int rows = n - kb;
        // 1. the panel's lower part (earlier panels' updates already applied)
        for (int t = tid; t < rows * nb; t += C1T) {
            const int r = t % rows, cc = t / rows;
            pan[r * PS1 + cc] = (r >= cc) ? a[(kb + r) + (size_t)(kb + cc) * n] : 0.0;
        }
        __syncthreads();
        // 2. the diagonal block on wave 0
        if (tid < 64) {
            const int r = tid;
            double x[NB1];
#pragma unroll
            for (int cc = 0; cc < NB1; cc++) x[cc] = (r < nb && cc < nb) ? pan[r * PS1 + cc] : 0.0;
            // the pivot is the same on every lane, so fail is wave-uniform;
            // after a failure the remaining columns are skipped
            int fail = 0;
#pragma unroll
            for (int cc = 0; cc < NB1; cc++) {
                const double piv = __shfl(x[cc], cc);
                if (cc < nb && piv <= 0.0) fail = 1;    // same test as chol.cu:26
                if (cc < nb && !fail) {
                    const double d = sqrt(piv);
                    if (r == cc) x[cc] = d;
                    else if (r > cc) x[cc] /= d;
                    const double lc = x[cc];
#pragma unroll
                    for (int c2 = cc + 1; c2 < NB1; c2++) {
                        const double l2 = __shfl(lc, c2);
                        if (r >= c2 && c2 < nb) x[c2] -= lc * l2;
                    }
                }
            }
            if (r < nb) {
#pragma unroll
                for (int cc = 0; cc < NB1; cc++)
                    if (cc <= r) pan[r * PS1 + cc] = x[cc];
            }
            if (tid == 0 && fail) bad = 1;
        }
        __syncthreads();
        if (bad) break;
        // 3. the rows below the diagonal block: forward substitution
        for (int r = nb + tid; r < rows; r += C1T) {
            double x[NB1];
#pragma unroll
            for (int cc = 0; cc < NB1; cc++) x[cc] = (cc < nb) ? pan[r * PS1 + cc] : 0.0;
#pragma unroll
            for (int cc = 0; cc < NB1; cc++) {
                if (cc < nb) {
                    double v = x[cc];
#pragma unroll
                    for (int c = 0; c < cc; c++) v -= x[c] * pan[cc * PS1 + c];
                    x[cc] = v / pan[cc * PS1 + cc];
                }
            }
#pragma unroll
            for (int cc = 0; cc < NB1; cc++)
                if (cc < nb) pan[r * PS1 + cc] = x[cc];
        }
        __syncthreads();
        // 4. the factored panel back to the matrix
        for (int t = tid; t < rows * nb; t += C1T) {
            const int r = t % rows, cc = t / rows;
            if (r >= cc) a[(kb + r) + (size_t)(kb + cc) * n] = pan[r * PS1 + cc];
        }
        // 5. trailing update A[i, j] -= sum_c L[i, c] L[j, c], kb + nb <= j <= i < n,
        //    in 4 x 4 blocks (block row bi >= block column bj of the trailing part)
        const int m = rows - nb;
        const int nbl = (m + 3) / 4;
        const int nblk = nbl * (nbl + 1) / 2;
        for (int t = tid; t < nblk; t += C1T) {
            int bi = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
            while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
            while (bi * (bi + 1) / 2 > t) bi--;
            const int bj = t - bi * (bi + 1) / 2;
            const int i0 = nb + 4 * bi, j0 = nb + 4 * bj;
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
            for (int c = 0; c < nb; c++) {
                double pi[4], pj[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    pi[u] = (i0 + u < rows) ? pan[(i0 + u) * PS1 + c] : 0.0;
                    pj[u] = (j0 + u < rows) ? pan[(j0 + u) * PS1 + c] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[u][v] += pi[u] * pj[v];
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + v;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int i = i0 + u;
                    if (i < rows && j <= i) a[(kb + i) + (size_t)(kb + j) * n] -= acc[u][v];
                }
            }
        }
        __syncthreads();   // the updates are visible to the next panel's loads
    }
    if (bad) {
        for (size_t t = tid; t < (size_t)n * n; t += C1T) a[t] = 0.0;
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



// x_j = L_j^{-T} L_j^{-1} b_j blocked by 16 rows: each 16 x 16 diagonal block
// is solved on wave 0 (row r in lane r, each solved value broadcast by a
// shuffle), then every thread updates the rows beyond the block with the 16
// new values -- two barriers per 16 rows instead of two per row.
constexpr int SB = 16;
__global__ void __launch_bounds__(CT) k_cholsl_blk(const double *__restrict__ Lall, int n,
                                                   size_t lstride, const double *__restrict__ b,
                                                   double *__restrict__ x) {
    extern __shared__ double xs[];
    const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const double *L = Lall + (size_t)j * lstride;
    for (int k = tid; k < n; k += CT) xs[k] = b[(size_t)j * n + k];
    __syncthreads();
    // L y = b
    for (int kb = 0; kb < n; kb += SB) {
        const int nb = min(SB, n - kb);
        if (tid < 64) {
            const int r = lane;
            double t = (r < nb) ? xs[kb + r] : 0.0;
            double lrow[SB];
#pragma unroll
            for (int c = 0; c < SB; c++)
                lrow[c] = (r < nb && c <= r && c < nb) ? L[(kb + r) + (size_t)(kb + c) * n] : 0.0;
            double y = 0.0;
#pragma unroll
            for (int c = 0; c < SB; c++) {
                if (r == c && c < nb) y = t / lrow[c];
                const double yc = __shfl(y, c);
                if (r > c && c < nb) t -= lrow[c] * yc;
            }
            if (r < nb) xs[kb + r] = y;
        }
        __syncthreads();
        double yb[SB];
#pragma unroll
        for (int c = 0; c < SB; c++) yb[c] = (c < nb) ? xs[kb + c] : 0.0;
        for (int r = kb + nb + tid; r < n; r += CT) {
            double v = xs[r];
#pragma unroll
            for (int c = 0; c < SB; c++)
                if (c < nb) v -= L[r + (size_t)(kb + c) * n] * yb[c];
            xs[r] = v;
        }
        __syncthreads();
    }
    // L^T x = y
    const int nblk = (n + SB - 1) / SB;
    for (int bk = nblk - 1; bk >= 0; bk--) {
        const int kb = bk * SB, nb = min(SB, n - kb);
        if (tid < 64) {
            const int r = lane;
            double t = (r < nb) ? xs[kb + r] : 0.0;
            double lcol[SB];   // L[kb + c, kb + r], c >= r
#pragma unroll
            for (int c = 0; c < SB; c++)
                lcol[c] = (r < nb && c >= r && c < nb) ? L[(kb + c) + (size_t)(kb + r) * n] : 0.0;
            double xv = 0.0;
#pragma unroll
            for (int c = SB - 1; c >= 0; c--) {
                if (r == c && c < nb) xv = t / lcol[c];
                const double xc = __shfl(xv, c);
                if (r < c && c < nb) t -= lcol[c] * xc;
            }
            if (r < nb) xs[kb + r] = xv;
        }
        __syncthreads();
        double xb[SB];
#pragma unroll
        for (int c = 0; c < SB; c++) xb[c] = (c < nb) ? xs[kb + c] : 0.0;
        for (int r = tid; r < kb; r += CT) {
            double v = xs[r];
            const double *Lc = L + (size_t)r * n + kb;    // L[kb .. kb + nb - 1, r]
#pragma unroll
            for (int c = 0; c < SB; c++)
                if (c < nb) v -= Lc[c] * xb[c];
            xs[r] = v;
        }
        __syncthreads();
    }
    for (int k = tid; k < n; k += CT) x[(size_t)j * n + k] = xs[k];
}

// Blocked V_j = L_j^{-1} (lower triangle; the strict upper triangle is not
// written), 32 x 32 blocks.  From L V = I, block row I > J of block column J:
//   L_II V_IJ = -sum_{K=J}^{I-1} L_IK V_KJ   =>   V_IJ = -V_II T_IJ,
// so with the diagonal inverses V_II known, every block column is
// independent: one workgroup per (block column, gene) walks I downwards.
constexpr int TB = 32;

// Workgroup -> (x, matrix y) for the batched inverse kernels: the hardware
// deals consecutive workgroup ids round robin over the 8 XCDs, so the id
// xcd + 8 (x + nx q) puts every x of matrix y = 8 q + xcd on XCD xcd, where
// they share its L2 (with (x, y) as a 2-D grid a matrix's workgroups span
// all eight L2s and each fetches the matrix from HBM).  False past the last
// matrix; the grid is xcd_grid(nx, batch) workgroups.
__device__ __forceinline__ bool xcd_pick(int nx, int batch, int &x, int &y) {
    const int id = blockIdx.x, xcd = id & 7, q = id >> 3;
    x = q % nx;
    y = 8 * (q / nx) + xcd;
    return y < batch;
}
static inline unsigned xcd_grid(int nx, int batch) {
    return (unsigned)nx * (unsigned)((batch + 7) / 8) * 8u;
}

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

// Off-diagonal blocks of block column J (from the workgroup id; all but the last block
// column) on v_mfma_f64_16x16x4f64: 4 waves, wave w owns the 16 x 16 tile
// (w / 2, w % 2) of each 32 x 32 block.  Per block row I: T = sum_K L_IK V_KJ
// with both fragments read from global memory (L from the factorization; V_KJ
// from k_trinv_diag for K = J, else written by this workgroup in an earlier
// step and visible after its barrier), T through LDS, then V_IJ = -V_II T.
// No LDS beyond T, so several workgroups share a CU and hide the load
// latency of each other's chains.  Fragment layouts (lane l):
//   A[m = l % 16][k = l / 16],  B[k = l / 16][n = l % 16],  D[m = 4 r + l / 16][n = l % 16].
typedef double tri_d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(CT) k_trinv_mfma(const double *__restrict__ Lall, int n, size_t stride,
                                                   double *__restrict__ Vall, int batch) {
    __shared__ double Tt[TB][TB + 1];
    const int NB = (n + TB - 1) / TB;
    int J, y;
    if (!xcd_pick(NB - 1, batch, J, y)) return;
    const double *L = Lall + (size_t)y * stride;
    double *V = Vall + (size_t)y * stride;
    const int c0 = TB * J;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tr = w >> 1, tc = w & 1;
    const int lm = lane & 15, lk = lane >> 4;
    const int bcol = 16 * tc + lm;                        // column of V_.J in the block (J < NB - 1: full)
    double *vcol = V + (size_t)(c0 + bcol) * n;
    for (int I = J + 1; I < NB; I++) {
        const int rI = TB * I, bsI = min(TB, n - rI);
        const int am = 16 * tr + lm;                      // A row within block row I
        const bool arow = am < bsI;
        const double *lrow = L + (rI + min(am, bsI - 1));
        tri_d4 acc = tri_d4{0.0, 0.0, 0.0, 0.0};
        for (int K = J; K < I; K++) {                     // K < I <= NB - 1: block K is full
#pragma unroll
            for (int s = 0; s < TB / 4; s++) {
                const int kk = 4 * s + lk, k = TB * K + kk;
                const double lv = lrow[(size_t)k * n];
                const double vv = vcol[k];
                const double a = arow ? lv : 0.0;
                const double b = (K > J || kk >= bcol) ? vv : 0.0;   // V_JJ: lower triangle only
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Tt[16 * tr + 4 * r + lk][16 * tc + lm] = acc[r];
        __syncthreads();
        tri_d4 v = tri_d4{0.0, 0.0, 0.0, 0.0};
        const double *vii = V + (rI + min(am, bsI - 1)) + (size_t)rI * n;
#pragma unroll
        for (int s = 0; s < TB / 4; s++) {
            const int kk = 4 * s + lk;
            const double x = vii[(size_t)min(kk, bsI - 1) * n];
            const double a = (arow && kk <= am) ? x : 0.0;    // V_II: lower triangle, bsI x bsI
            v = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Tt[kk][16 * tc + lm], v, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * tr + 4 * r + lk;
            if (row < bsI) vcol[rI + row] = -v[r];
        }
        __syncthreads();                                  // V_IJ visible; Tt free
    }
}

// S_j = V_j^T V_j (full symmetric) from the lower-triangular V_j: 64x64 tiles
// of the upper triangle, 32-row chunks of V staged in LDS (the next chunk
// prefetched into registers), 4x4 per thread.
// Entry (a, b), a <= b, sums rows k >= b only (V[k, b] = 0 above the diagonal).
__global__ void __launch_bounds__(CT) k_vtv(const double *__restrict__ Vall, int n, size_t stride,
                                            int ntile, double *__restrict__ Sall, int batch) {
    __shared__ double As[32][65];
    __shared__ double Bs[32][65];
    int t, y;
    if (!xcd_pick(ntile * (ntile + 1) / 2, batch, t, y)) return;
    int ta = 0;
    while (t >= ntile - ta) { t -= ntile - ta; ta++; }
    const int tb = ta + t;
    const int a0 = ta * 64, b0 = tb * 64;
    const double *V = Vall + (size_t)y * stride;
    double *S = Sall + (size_t)y * stride;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r][q] = 0.0;
    // the next chunk's 16 loads per thread are issued before the current
    // chunk's products, so their latency hides behind them
    // thread t stages column t % 64 of both tiles, rows 8 (t / 64) .. +7 of
    // the chunk: eight loads off one address per tile
    const int col = tid % 64, rg = 8 * (tid / 64);
    const int ga = a0 + col, gb = b0 + col;
    const double *va = V + (size_t)min(ga, n - 1) * n;
    const double *vb = V + (size_t)min(gb, n - 1) * n;
    double pa[8], pb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + rg + u;
            pa[u] = (k < n && ga < n && k >= ga) ? va[k] : 0.0;
            pb[u] = (k < n && gb < n && k >= gb) ? vb[k] : 0.0;
        }
    };
    fetch(b0);
    for (int k0 = b0; k0 < n; k0 += 32) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; u++) {
            As[rg + u][col] = pa[u];
            Bs[rg + u][col] = pb[u];
        }
        __syncthreads();
        if (k0 + 32 < n) fetch(k0 + 32);
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

// k_vtv on FP64 MFMA (v_mfma_f64_16x16x4f64): the same 64x64 tiles and
// 32-row chunks of V staged in LDS; wave w owns the 16 rows a0 + 16 w .. of
// the tile against its four 16-column blocks, so per 4-row step one A
// fragment and four B fragments feed four MFMAs (the VALU form reads eight
// values from LDS per 16 FMAs).  Fragment layouts (lane l): A[m = l % 16][k =
// l / 16] = V[k0 + 4 s + l / 16][a0 + 16 w + l % 16], B[k = l / 16][n = l % 16]
// = V[k0 + 4 s + l / 16][b0 + 16 cb + l % 16], D[m = 4 r + l / 16][n = l % 16].
__global__ void __launch_bounds__(CT) k_vtv_mfma(const double *__restrict__ Vall, int n, size_t stride,
                                                 int ntile, double *__restrict__ Sall, int batch) {
    __shared__ double As[32][65];
    __shared__ double Bs[32][65];
    int t, y;
    if (!xcd_pick(ntile * (ntile + 1) / 2, batch, t, y)) return;
    int ta = 0;
    while (t >= ntile - ta) { t -= ntile - ta; ta++; }
    const int tb = ta + t;
    const int a0 = ta * 64, b0 = tb * 64;
    const double *V = Vall + (size_t)y * stride;
    double *S = Sall + (size_t)y * stride;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lm = lane & 15, lk = lane >> 4;
    tri_d4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) acc[cb] = tri_d4{0.0, 0.0, 0.0, 0.0};
    // staging as k_vtv: thread t stages column t % 64 of both tiles, rows
    // 8 (t / 64) .. +7 of the chunk, the next chunk loaded during this one
    const int col = tid % 64, rg = 8 * (tid / 64);
    const int ga = a0 + col, gb = b0 + col;
    const double *va = V + (size_t)min(ga, n - 1) * n;
    const double *vb = V + (size_t)min(gb, n - 1) * n;
    double pa[8], pb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + rg + u;
            pa[u] = (k < n && ga < n && k >= ga) ? va[k] : 0.0;
            pb[u] = (k < n && gb < n && k >= gb) ? vb[k] : 0.0;
        }
    };
    fetch(b0);
    for (int k0 = b0; k0 < n; k0 += 32) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; u++) {
            As[rg + u][col] = pa[u];
            Bs[rg + u][col] = pb[u];
        }
        __syncthreads();
        if (k0 + 32 < n) fetch(k0 + 32);
#pragma unroll
        for (int s4 = 0; s4 < 8; s4++) {
            const double av = As[4 * s4 + lk][16 * w + lm];
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Bs[4 * s4 + lk][16 * cb + lm], acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ra = a0 + 16 * w + 4 * r + lk, cbn = b0 + 16 * cb + lm;
            if (ra < n && cbn < n && ra <= cbn) {
                S[ra + (size_t)cbn * n] = acc[cb][r];
                S[cbn + (size_t)ra * n] = acc[cb][r];
            }
        }
}

}  // namespace

int gpk_chol_batched(gp_ctx *c, double *A, int n, int batch, size_t stride, int *info) {
    if (n > MAXN) { gp_set_error("Cholesky size exceeds 640"); return 1; }
    if (batch == 1 && n <= MAXN1 && c->chol1 && A == c->Kchol) {
        const size_t shm = (size_t)n * PS1 * sizeof(double);
        if (shm > 65536 && !c->capturing)
            GP_HIP(hipFuncSetAttribute((const void *)k_chol1, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
        k_chol1<<<1, C1T, shm, c->stream>>>(A, n, info);
        GP_HIP(hipGetLastError());
        return 0;
    }
    if (c->chol_nb == 8) {
        const size_t shm = ((size_t)n * 8 + (size_t)8 * 32) * sizeof(double);
        if (shm > 65536 && !c->capturing)
            GP_HIP(hipFuncSetAttribute((const void *)k_chol<8, 32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
        k_chol<8, 32><<<batch, CT, shm, c->stream>>>(A, n, stride, info);
        GP_HIP(hipGetLastError());
        return 0;
    }
    const size_t shm = ((size_t)n * 16 + (size_t)16 * 48) * sizeof(double);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_chol<16, 48>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm));
    k_chol<16, 48><<<batch, CT, shm, c->stream>>>(A, n, stride, info);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_cholsl_batched(gp_ctx *c, const double *L, int n, size_t lstride, const double *b,
                       double *x, int batch) {
    if (c->cholsl_blk) {
        k_cholsl_blk<<<batch, CT, (size_t)n * sizeof(double), c->stream>>>(L, n, lstride, b, x);
        GP_HIP(hipGetLastError());
        return 0;
    }
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
    k_trinv_diag<<<dim3(nb, batch), 64, 0, c->stream>>>(L, n, st, c->Vtmp);
    GP_HIP(hipGetLastError());
    if (nb > 1) {
        k_trinv_mfma<<<xcd_grid(nb - 1, batch), CT, 0, c->stream>>>(L, n, st, c->Vtmp, batch);
        GP_HIP(hipGetLastError());
    }
    const int nt = (n + 63) / 64;
    const unsigned g = xcd_grid(nt * (nt + 1) / 2, batch);
    if (c->vtv_mfma)
        k_vtv_mfma<<<g, CT, 0, c->stream>>>(c->Vtmp, n, st, nt, S, batch);
    else
        k_vtv<<<g, CT, 0, c->stream>>>(c->Vtmp, n, st, nt, S, batch);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_vtv(gp_ctx *c, const double *V, int n, int batch, size_t stride, double *S) {
    const size_t st = stride ? stride : (size_t)n * n;
    const int nt = (n + 63) / 64;
    const unsigned g = xcd_grid(nt * (nt + 1) / 2, batch);
    if (c->vtv_mfma)
        k_vtv_mfma<<<g, CT, 0, c->stream>>>(V, n, st, nt, S, batch);
    else
        k_vtv<<<g, CT, 0, c->stream>>>(V, n, st, nt, S, batch);
    GP_HIP(hipGetLastError());
    return 0;
}

// V = L^-1 of one matrix (the Kmm factor) by the batched blocked kernels:
// 32 x 32 diagonal inverses, then each block column walked down on FP64
// MFMA -- a few dozen dependent steps instead of the n serial substitution
// steps of a column-by-column substitution, which stretched from 0.09 to
// 0.8 ms when the ELBO's side stream loads HBM beside it.  They write V's lower triangle only: V
// must hold zeros above the diagonal (the context's buffer is allocated
// zero-filled and nothing writes there).
int gpk_tri_inverse(gp_ctx *c, const double *L, int n, double *V) {
    const int nb = (n + TB - 1) / TB;
    const size_t st = (size_t)n * n;
    k_trinv_diag<<<dim3(nb, 1), 64, 0, c->stream>>>(L, n, st, V);
    GP_HIP(hipGetLastError());
    if (nb > 1) {
        k_trinv_mfma<<<xcd_grid(nb - 1, 1), CT, 0, c->stream>>>(L, n, st, V, 1);
        GP_HIP(hipGetLastError());
    }
    return 0;
}
