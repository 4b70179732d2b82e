// gp_batched.hip -- batched per-gene Poisson GLM / Laplace (IRLS) fits, each
// gene on its own subset of cells with its own covariates, on the latent
// space of a gp_ctx (its Knm, Kmm, titsias, Omega, size factors).
//
// Per gene g with cells S_g (n_g of them) the design row of cell i is
//   z_i = [1 | covariates c_gi (p_g - 1 values, FP32) | Knm(i, :)]   (P_g = p_g + M)
// and the prior Delta_g^-1 = diag(1/delta_int, 1/delta_cov ..) (+) Kmm/delta_GP.
// One IRLS step is the reference's getWForBeta / getPsi / cholPsi / getBetaPhi
// / getLkhdPois (ggplvm.cu:418-514, 621-724) restricted to S_g.
//
// Persistent launch: k_bf_pass runs a fixed grid of workgroups that pull genes
// from a queue (largest subsets first).  For one gene, a workgroup streams its
// cells in double-buffered chunks of 32: producer waves gather the design rows
// (Knm rows from a row-major copy, 1 KB each, plus the gene's covariates),
// compute eta = z.beta, W, nu and the likelihood terms and scale the rows to
// x_i = sqrt(W_i/Omega_i) [z_i | nu_i] in LDS, while four MFMA waves
// accumulate X^T X of the previous chunk with v_mfma_f64_16x16x4f64 --
// Psi_g - Delta, b_g = Z^T D nu and nu^T D nu come out of one symmetric Gram
// (its upper 16x16 tiles dealt round robin to the MFMA waves).  Then a batched
// Cholesky (fixed padded size PB, identity padding) and k_bf_beta (solve,
// damped update, phi) finish the step.
#include "gp_internal.h"

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

struct gp_bfit {
    gp_ctx *c = nullptr;
    int J = 0, M = 0, PB = 0;                // genes, inducing points, padded Psi size
    long long L = 0;                          // total subset positions
    std::vector<int> off_h, ncov_h;
    int *off = nullptr, *cell = nullptr, *ncov = nullptr, *order = nullptr, *queue = nullptr;
    long long *covoff = nullptr;
    float *cov = nullptr, *yd = nullptr;
    double *Krm = nullptr, *Beta = nullptr, *Psi = nullptr, *bvec = nullptr, *sov2 = nullptr;
    double *cellc = nullptr;                  // N x 3 per-cell constants of a pass (k_bf_cellc)
    double *phi = nullptr, *lk = nullptr, *obj = nullptr, *delta = nullptr;
    int *info = nullptr, *jtr = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int grid_mult = 2;                        // persistent workgroups per CU of the pass
};

namespace {

constexpr int BT = 10;                 // 16-column tiles of the augmented design (P_g + 1 <= 160)
constexpr int BPA = 16 * BT;
constexpr int BCH = 32;                // cells per chunk
constexpr int BLD = 176;               // LDS row stride in doubles (rows land 32 banks apart)
constexpr int BNP = 4;                 // producer waves of the pass (8 measured slower: spills at 3 waves per SIMD)
constexpr int BTH = 64 * (4 + BNP);     // threads of the pass (4 MFMA waves + the producers)
constexpr int BCPW = 32 / BNP;          // cells per producer wave and chunk
constexpr int BLPC = 64 / BCPW;         // producer lanes per cell
constexpr int KPL = (BPA + 63) / 64;   // Knm values per lane and cell (M < BPA)
constexpr double WT = 0.1;             // ggplvm.cu:20

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline double bsum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// block sum over BTH threads
__device__ inline double bsum8(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int k = 0; k < BTH / 64; k++) r += sh[k];
    __syncthreads();
    return r;
}

// Prior entry (a, b) of gene design width P_g = pg + M.
__device__ inline double prior(int a, int b, int pg, int M, const double *K, const double *delta) {
    if (a >= pg && b >= pg) {
        const int r = min(a, b) - pg, s = max(a, b) - pg;
        return K[r + (size_t)s * M] / delta[2];
    }
    if (a != b) return 0.0;
    return 1.0 / (a == 0 ? delta[0] : delta[1]);
}

// mode 0: initPsi/initBeta pass (unit weights, nu = log((y+0.05)/ls));
// mode 1: IRLS pass at the current beta (also the Poisson log-likelihood).
// Two roles over a double-buffered 32-cell chunk: while waves 0-3 (one per
// SIMD) run the MFMAs of chunk c, the BNP producer waves build chunk c + 1 in
// the other buffer -- producer wave p owns cells BCPW p .. BCPW p + BCPW - 1:
// it writes their design rows (Knm rows prefetched a chunk ahead in
// registers, covariates, intercept), forms eta with BLPC lanes per cell, W,
// nu, the weight and the likelihood term (the per-cell transcendentals spread
// over the cell's lanes), and scales the rows; no step crosses waves, so one
// barrier per chunk suffices.  The
// upper 16x16 tiles of the (P_g+1)-wide Gram are dealt round robin to the
// four MFMA waves (at most BSLOT4 each).
constexpr int BSLOT4 = (BT * (BT + 1) / 2 + 3) / 4;    // upper tiles per MFMA wave (14)

// Per-cell constants of the W / nu formulas (ggplvm.cu:468-514) for a pass:
// cc[3i] = log(ls_i), cc[3i+1] = dt tits_i WT, cc[3i+2] = exp(dt tits_i WT).
__global__ void k_bf_cellc(int N, const double *__restrict__ ls, const double *__restrict__ tits,
                           const double *__restrict__ delta, double *__restrict__ cc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const double c = delta[2] * tits[i] * WT;
    cc[3 * (size_t)i] = log(ls[i]);
    cc[3 * (size_t)i + 1] = c;
    cc[3 * (size_t)i + 2] = exp(c);
}
constexpr int BCV = (16 + BLPC - 1) / BLPC;             // prefetched covariate values per producer lane
static_assert(BSLOT4 <= 14 && BCH == 32, "pass tiling");

__global__ void __launch_bounds__(BTH) k_bf_pass(int mode, int J, int M, int PB,
        const int *__restrict__ order, int *__restrict__ queue,
        const int *__restrict__ off, const int *__restrict__ cell, const int *__restrict__ ncov,
        const long long *__restrict__ covoff, const float *__restrict__ cov,
        const float *__restrict__ yd, const double *__restrict__ Krm, const double *__restrict__ ls,
        const double *__restrict__ Omega, const double *__restrict__ cellc,
        const double *__restrict__ K, const double *__restrict__ delta,
        const double *__restrict__ Beta, double *__restrict__ Psi, double *__restrict__ bvec,
        double *__restrict__ sov2, double *__restrict__ lk) {
    __shared__ double Xs[2][BCH * BLD];
    __shared__ double bs[BPA];
    __shared__ double red[BTH / 64];
    __shared__ int gsh;
    __shared__ unsigned char tta[BT * (BT + 1) / 2], ttb[BT * (BT + 1) / 2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool producer = w >= 4;                 // waves 4.. build chunks, 0-3 run the MFMAs
    const int pw = w - 4;                         // producer index: cells BCPW pw .. BCPW pw + BCPW - 1
    const int cu = lane / BLPC, part = lane % BLPC;   // producer lane: cell BCPW pw + cu, part of it
    const int cb0 = lane & ~(BLPC - 1);           // first lane of this lane's cell
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(queue, 1);
            gsh = (q < J) ? order[q] : -1;
        }
        __syncthreads();
        const int g = gsh;
        __syncthreads();
        if (g < 0) break;
        const int o0 = off[g], n = off[g + 1] - o0, pg = ncov[g], P = pg + M;
        const int PA = P + 1, TA = (PA + 15) / 16, NTL = TA * (TA + 1) / 2;
        const float *cg = cov + covoff[g];
        for (int k = tid; k < BPA; k += BTH) bs[k] = (mode == 1 && k < P) ? Beta[(size_t)g * PB + k] : 0.0;
        if (tid < NTL) {                          // upper tile t -> (ta, tb), once per gene
            int ta = 0, rem = tid;
            while (rem >= TA - ta) { rem -= TA - ta; ta++; }
            tta[tid] = (unsigned char)ta;
            ttb[tid] = (unsigned char)(ta + rem);
        }
        __syncthreads();                          // bs, tta, ttb visible to both roles
        double lks = 0.0;
        const int nch = (n + BCH - 1) / BCH;
        double *Pg = Psi + (size_t)g * PB * PB;
        // The two roles run separate loops with the same barriers (one per
        // chunk), so the accumulators and the prefetch registers are never
        // live together.
        if (producer) {
        // prefetch registers, one chunk ahead: the Knm rows of its BCPW cells,
        // its covariate values and the lane's cell data (count, size factor,
        // Omega, the per-cell constants); the cell indices come one chunk
        // earlier still, so no load waits on another within a chunk
        double kv[BCPW][KPL];
        float cvv[BCV];
        double yv = 0.0, lsv = 1.0, omv = 1.0, l1v = 0.0, cvn = 0.0, e2v = 1.0;
        int cn[BCPW];
        auto cells = [&](int c0) {                // cell indices of chunk c0 (clamped)
#pragma unroll
            for (int u = 0; u < BCPW; u++) cn[u] = cell[o0 + min(c0 + BCPW * pw + u, n - 1)];
        };
        auto fetch = [&](int c0) {                // uses cn of chunk c0, then loads the next chunk's
#pragma unroll
            for (int u = 0; u < BCPW; u++) {
                const bool ok = c0 + BCPW * pw + u < n;
                const double *row = Krm + (size_t)cn[u] * M;
#pragma unroll
                for (int e = 0; e < KPL; e++) {
                    const int col = lane + 64 * e;
                    kv[u][e] = (ok && col < M) ? row[min(col, M - 1)] : 0.0;
                }
            }
            const bool okc = c0 + BCPW * pw + cu < n;
#pragma unroll
            for (int r = 0; r < BCV; r++) {
                const int col = 1 + part + BLPC * r;
                cvv[r] = (okc && col < pg) ? cg[(size_t)(col - 1) * n + c0 + BCPW * pw + cu] : 0.0f;
            }
            int ic = cn[0];
#pragma unroll
            for (int u = 1; u < BCPW; u++) ic = (cu == u) ? cn[u] : ic;
            yv = (double)yd[o0 + min(c0 + BCPW * pw + cu, n - 1)];
            lsv = ls[ic];
            omv = Omega[ic];
            l1v = cellc[3 * (size_t)ic];
            cvn = cellc[3 * (size_t)ic + 1];
            e2v = cellc[3 * (size_t)ic + 2];
            cells(c0 + BCH);
        };
        // build chunk c0 into buffer nb (producer waves only; wave-local)
        auto build = [&](int c0, int nb) {
            double *X = Xs[nb];
#pragma unroll
            for (int u = 0; u < BCPW; u++) {
                const int ci = BCPW * pw + u;
#pragma unroll
                for (int e = 0; e < KPL; e++) {
                    const int col = lane + 64 * e;
                    if (col < M) X[ci * BLD + pg + col] = kv[u][e];
                }
            }
            const int ci = BCPW * pw + cu;
            const bool ok = c0 + ci < n;
            if (part == 0) X[ci * BLD] = ok ? 1.0 : 0.0;
#pragma unroll
            for (int r = 0; r < BCV; r++) {
                const int col = 1 + part + BLPC * r;
                if (col < pg) X[ci * BLD + col] = (double)cvv[r];
            }
            // covariates beyond the prefetched 8 BCV (wider designs), read here
            for (int col = 1 + part + BLPC * BCV; col < pg; col += BLPC)
                X[ci * BLD + col] = ok ? (double)cg[(size_t)(col - 1) * n + min(c0 + ci, n - 1)] : 0.0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            double eta = 0.0;
            if (mode == 1) {                      // 4 loads of each stream in flight, same order
                int k = part;
                for (; k + 3 * BLPC < P; k += 4 * BLPC) {
                    double xv[4], bv[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) { xv[u] = X[ci * BLD + k + BLPC * u]; bv[u] = bs[k + BLPC * u]; }
#pragma unroll
                    for (int u = 0; u < 4; u++) eta += xv[u] * bv[u];
                }
                for (; k < P; k += BLPC) eta += X[ci * BLD + k] * bs[k];
            }
#pragma unroll
            for (int o = 1; o < BLPC; o <<= 1) eta += __shfl_xor(eta, o, 64);
            // W, nu, the weight and the likelihood term of the cell, the
            // transcendentals spread over the cell's lanes in two rounds
            // (the operations and their order are those of one lane)
            const double y = yv;
            double s = 0.0, nu = 0.0;
            if (mode == 0) {
                s = 1.0;
                if (part == 0) nu = log((y + 0.05) / lsv);
            } else {
                const double l1 = l1v, c = cvn, e2 = e2v;
                // first transcendental step: exp(eta) (part 0), log(y) (part 1)
                double r1 = 0.0;
                if (part == 0) r1 = exp(eta);
                else if (part == 1 && y > 0.0) r1 = log(y);
                const double eeta = __shfl(r1, cb0, 64);
                const double ly = __shfl(r1, cb0 + 1, 64);
                const double logW = eta + l1 + c;
                const double Wi = eeta * lsv * e2;
                // second step: exp(log y - logW) (part 0), sqrt(W / Omega) (part 1), log(W) (part 2)
                double r2 = 0.0;
                if (part == 0 && y > 0.0) r2 = exp(ly - logW);
                else if (part == 1) r2 = sqrt(Wi / omv);
                else if (part == 2 && y > 0.0) r2 = log(Wi);
                const double ex = __shfl(r2, cb0, 64);
                s = __shfl(r2, cb0 + 1, 64);
                const double lw = __shfl(r2, cb0 + 2, 64);
                nu = (y > 0.0) ? eta + ex - 1.0 : eta - 1.0;
                // ggplvm.cu:703-724
                if (part == 0 && ok) lks += (y > 0.0 ? (lw - c) * y : 0.0) - Wi;
            }
            if (!ok) s = 0.0;
            s = __shfl(s, cb0, 64);
            nu = __shfl(nu, cb0, 64);
            {                                     // scale the row: loads first, then stores
                int col = part;
                for (; col + 3 * BLPC < P; col += 4 * BLPC) {
                    double v[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) v[u] = X[ci * BLD + col + BLPC * u];
#pragma unroll
                    for (int u = 0; u < 4; u++) X[ci * BLD + col + BLPC * u] = v[u] * s;
                }
                for (; col < P; col += BLPC) X[ci * BLD + col] *= s;
            }
            if (part == 0) X[ci * BLD + P] = s * nu;
        };
        cells(0);
        fetch(0);
        build(0, 0);
        if (nch > 1) fetch(BCH);
        __syncthreads();                          // chunk 0 built
        for (int ch = 0; ch < nch; ch++) {
            if (ch + 1 < nch) {
                build((ch + 1) * BCH, (ch & 1) ^ 1);
                if (ch + 2 < nch) fetch((ch + 2) * BCH);
            }
            __syncthreads();                      // chunk ch + 1 built; buffer ch & 1 free
        }
        } else {
        d4 acc[BSLOT4];
#pragma unroll
        for (int s = 0; s < BSLOT4; s++) acc[s] = d4{0.0, 0.0, 0.0, 0.0};
        __syncthreads();                          // chunk 0 built
        for (int ch = 0; ch < nch; ch++) {
            // X^T X over chunk ch: tile t = w + 4 s of the upper triangle
            const double *X = Xs[ch & 1];
#pragma unroll
            for (int s = 0; s < BSLOT4; s++) {
                const int t = w + 4 * s;
                if (t < NTL) {
                    const int ta = tta[t], tb = ttb[t];
#pragma unroll
                    for (int k0 = 0; k0 < BCH; k0 += 4) {
                        const double a = X[(k0 + (lane >> 4)) * BLD + 16 * ta + (lane & 15)];
                        const double b = X[(k0 + (lane >> 4)) * BLD + 16 * tb + (lane & 15)];
                        acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[s], 0, 0, 0);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);   // one tile's fragments live at a time
            }
            __syncthreads();                      // chunk ch + 1 built; buffer ch & 1 free
        }
        // write Psi_g (+ prior, mirrored), b_g, nu^T D nu
        {
#pragma unroll
            for (int s = 0; s < BSLOT4; s++) {
                const int t = w + 4 * s;
                if (t < NTL) {
                    const int ta = tta[t], tb = ttb[t];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = 16 * ta + 4 * r + (lane >> 4), col = 16 * tb + (lane & 15);
                        const double v = acc[s][r];
                        if (row > col) continue;
                        if (col < P) {
                            const double pv = v + prior(row, col, pg, M, K, delta);
                            Pg[row + (size_t)col * PB] = pv;
                            Pg[col + (size_t)row * PB] = pv;
                        } else if (col == P) {
                            if (row < P) bvec[(size_t)g * PB + row] = v;
                            else sov2[g] = v;
                        }
                    }
                }
            }
        }
        }
        // identity padding
        for (int e = tid; e < PB * PB; e += BTH) {
            const int r = e % PB, cc = e / PB;
            if (r >= P || cc >= P) Pg[e] = (r == cc) ? 1.0 : 0.0;
        }
        for (int k = P + tid; k < PB; k += BTH) bvec[(size_t)g * PB + k] = 0.0;
        lks = bsum8(lks, red);
        if (tid == 0 && mode == 1) lk[g] = lks;
    }
}

// ggplvm.cu:418-464 per gene on its subset (phi over n_g cells);
// mode 0 (initBeta, init.cu:193-221): beta = Psi^-1 b, phi untouched.
__global__ void __launch_bounds__(256) k_bf_beta(int mode, int PB, const int *__restrict__ off,
        const int *__restrict__ jtrp, const double *__restrict__ Psi, const double *__restrict__ bvec,
        const double *__restrict__ sov2, double *__restrict__ Beta, double *__restrict__ phi) {
    extern __shared__ double x[];     // [PB] solve, [PB] L^T beta
    __shared__ double red[4];
    const int g = blockIdx.x, tid = threadIdx.x;
    const double *L = Psi + (size_t)g * PB * PB;
    for (int k = tid; k < PB; k += 256) x[k] = bvec[(size_t)g * PB + k];
    for (int i = 0; i < PB; i++) {
        __syncthreads();
        const double xi = x[i] / L[i + (size_t)i * PB];
        __syncthreads();
        if (tid == 0) x[i] = xi;
        for (int r = i + 1 + tid; r < PB; r += 256) x[r] -= L[r + (size_t)i * PB] * xi;
    }
    for (int i = PB - 1; i >= 0; i--) {
        __syncthreads();
        const double xi = x[i] / L[i + (size_t)i * PB];
        __syncthreads();
        if (tid == 0) x[i] = xi;
        for (int r = tid; r < i; r += 256) x[r] -= L[i + (size_t)r * PB] * xi;
    }
    __syncthreads();
    double *Bg = Beta + (size_t)g * PB;
    if (mode == 0) {
        for (int k = tid; k < PB; k += 256) Bg[k] = x[k];
        return;
    }
    const int jtr = *jtrp;
    const double rr = (jtr > 100) ? 1.0 : (double)(jtr + 1) / 100.0;
    for (int k = tid; k < PB; k += 256) Bg[k] = Bg[k] * (1.0 - rr) + x[k] * rr;
    __syncthreads();
    // |L^T beta|^2 (getRb)
    double q = 0.0;
    for (int k = tid; k < PB; k += 256) {
        double s = 0.0;
        for (int m = k; m < PB; m++) s += L[m + (size_t)k * PB] * Bg[m];
        q += s * s;
    }
    for (int o = 32; o > 0; o >>= 1) q += __shfl_down(q, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = q;
    __syncthreads();
    if (tid == 0) {
        const double n = (double)(off[g + 1] - off[g]);
        double ph = (sov2[g] - (red[0] + red[1] + red[2] + red[3])) / n;
        if (isnan(ph)) ph = 1.0;
        if (ph < 0.0) ph = 1.0;
        if (ph < 0.01) ph = 0.01;
        if (ph > 1000.0) ph = 1000.0;
        phi[g] = ph;
    }
}

__global__ void k_bf_sum(int J, const double *__restrict__ lk, double *__restrict__ obj,
                         int *__restrict__ jtr, int bump) {
    __shared__ double red[4];
    double s = 0.0;
    for (int g = threadIdx.x; g < J; g += 256) s += lk[g];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        *obj = red[0] + red[1] + red[2] + red[3];
        if (bump) *jtr += 1;
    }
}

__global__ void k_bf_transpose(int N, int Npad, int M, const double *__restrict__ Knm,
                               double *__restrict__ Krm) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)N * M) return;
    const int i = (int)(e / M), m = (int)(e % M);
    Krm[e] = Knm[i + (size_t)m * Npad];
}

__global__ void k_bf_dense_y(int J, const int *__restrict__ off, const int *__restrict__ yoff,
                             const int *__restrict__ ypos, const double *__restrict__ yval,
                             float *__restrict__ yd) {
    const int g = blockIdx.x;
    for (int l = yoff[g] + threadIdx.x; l < yoff[g + 1]; l += 256)
        yd[off[g] + ypos[l]] = (float)yval[l];
}

__global__ void k_bf_fill_int(int n, int *p, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// the persistent pass alone: Psi (unfactored), bvec, sov2 and lk at the current beta
int bf_launch_pass(gp_bfit *b, int mode) {
    gp_ctx *c = b->c;
    k_bf_fill_int<<<1, 256, 0, c->stream>>>(1, b->queue, 0);
    GP_HIP(hipGetLastError());
    k_bf_cellc<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->N, c->ls, c->titsias, b->delta, b->cellc);
    GP_HIP(hipGetLastError());
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = std::min(b->J, b->grid_mult * ncu);
    k_bf_pass<<<grid, BTH, 0, c->stream>>>(mode, b->J, b->M, b->PB, b->order, b->queue, b->off,
                                          b->cell, b->ncov, b->covoff, b->cov, b->yd, b->Krm,
                                          c->ls, c->Omega, b->cellc, c->K, b->delta, b->Beta,
                                          b->Psi, b->bvec, b->sov2, b->lk);
    GP_HIP(hipGetLastError());
    return 0;
}

int bf_pass(gp_bfit *b, int mode) {
    gp_ctx *c = b->c;
    GP_TRY(bf_launch_pass(b, mode));
    GP_TRY(gpk_chol_batched(c, b->Psi, b->PB, b->J, (size_t)b->PB * b->PB, b->info));
    k_bf_beta<<<b->J, 256, (size_t)b->PB * sizeof(double), c->stream>>>(
        mode, b->PB, b->off, b->jtr, b->Psi, b->bvec, b->sov2, b->Beta, b->phi);
    GP_HIP(hipGetLastError());
    return 0;
}

int bf_step(gp_bfit *b) {
    GP_TRY(bf_pass(b, 1));
    k_bf_sum<<<1, 256, 0, b->c->stream>>>(b->J, b->lk, b->obj, b->jtr, 1);
    GP_HIP(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" int gp_bfit_create(gp_bfit **out, gp_ctx *c, int J, const int *off, const int *cell,
                              const int *ncov, const float *cov, const int *yoff, const int *ypos,
                              const double *yval) {
    if (!out || !c || !off || !cell || !ncov || !yoff || !ypos || !yval) {
        gp_set_error("NULL argument");
        return 1;
    }
    *out = nullptr;
    if (J <= 0) { gp_set_error("J must be positive"); return 1; }
    const int M = c->M;
    int pmax = 1;
    long long L = off[J], ncovtot = 0, nnz = yoff[J];
    std::vector<long long> covoff(J + 1, 0);
    for (int g = 0; g < J; g++) {
        const int n = off[g + 1] - off[g];
        if (n <= 0) { gp_set_error("gene " + std::to_string(g) + " has an empty cell subset"); return 1; }
        if (ncov[g] < 1) { gp_set_error("ncov must count the intercept (>= 1)"); return 1; }
        pmax = std::max(pmax, ncov[g]);
        covoff[g + 1] = covoff[g] + (long long)(ncov[g] - 1) * n;
        for (int k = off[g]; k < off[g + 1]; k++)
            if (cell[k] < 0 || cell[k] >= c->N) { gp_set_error("cell index out of range"); return 1; }
        for (int l = yoff[g]; l < yoff[g + 1]; l++)
            if (ypos[l] < 0 || ypos[l] >= n) { gp_set_error("count position out of range"); return 1; }
    }
    ncovtot = covoff[J];
    if (pmax + M + 1 > BPA) {
        gp_set_error("design too wide: p_g + M + 1 must be <= " + std::to_string(BPA));
        return 1;
    }
    gp_bfit *b = new gp_bfit();
    b->c = c;
    b->J = J;
    b->M = M;
    b->L = L;
    b->PB = pmax + M;
    b->off_h.assign(off, off + J + 1);
    b->ncov_h.assign(ncov, ncov + J);
    // genes by subset size, largest first (the queue order of the persistent pass)
    std::vector<int> order(J);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return off[x + 1] - off[x] > off[y + 1] - off[y];
    });
    const double delta0[3] = {1.5, 0.01, 0.3};   // init.cu:31-35
    hipError_t e = hipSuccess;
#define BF_ALLOC(p, n) if (e == hipSuccess) e = hipMalloc((void **)&(p), (size_t)(n) * sizeof(*(p)))
    BF_ALLOC(b->off, J + 1);
    BF_ALLOC(b->cell, std::max<long long>(L, 1));
    BF_ALLOC(b->ncov, J);
    BF_ALLOC(b->order, J);
    BF_ALLOC(b->queue, 1);
    BF_ALLOC(b->covoff, J + 1);
    BF_ALLOC(b->cov, std::max<long long>(ncovtot, 1));
    BF_ALLOC(b->yd, std::max<long long>(L, 1));
    BF_ALLOC(b->Krm, (size_t)c->N * M);
    BF_ALLOC(b->cellc, (size_t)c->N * 3);
    BF_ALLOC(b->Beta, (size_t)J * b->PB);
    BF_ALLOC(b->Psi, (size_t)J * b->PB * b->PB);
    BF_ALLOC(b->bvec, (size_t)J * b->PB);
    BF_ALLOC(b->sov2, J);
    BF_ALLOC(b->phi, J);
    BF_ALLOC(b->lk, J);
    BF_ALLOC(b->obj, 1);
    BF_ALLOC(b->delta, 3);
    BF_ALLOC(b->info, 1);
    BF_ALLOC(b->jtr, 1);
#undef BF_ALLOC
    int *dyoff = nullptr, *dypos = nullptr;
    double *dyval = nullptr;
    if (e == hipSuccess) e = hipMalloc((void **)&dyoff, (size_t)(J + 1) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&dypos, (size_t)std::max<long long>(nnz, 1) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&dyval, (size_t)std::max<long long>(nnz, 1) * sizeof(double));
    hipStream_t st = c->stream;
    auto up = [&](void *d, const void *h, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st);
    };
    up(b->off, off, (size_t)(J + 1) * sizeof(int));
    up(b->cell, cell, (size_t)L * sizeof(int));
    up(b->ncov, ncov, (size_t)J * sizeof(int));
    up(b->order, order.data(), (size_t)J * sizeof(int));
    up(b->covoff, covoff.data(), (size_t)(J + 1) * sizeof(long long));
    if (cov) up(b->cov, cov, (size_t)ncovtot * sizeof(float));
    up(b->delta, delta0, 3 * sizeof(double));
    up(dyoff, yoff, (size_t)(J + 1) * sizeof(int));
    up(dypos, ypos, (size_t)nnz * sizeof(int));
    up(dyval, yval, (size_t)nnz * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(b->yd, 0, (size_t)std::max<long long>(L, 1) * sizeof(float), st);
    if (e == hipSuccess) e = hipMemsetAsync(b->Beta, 0, (size_t)J * b->PB * sizeof(double), st);
    if (e == hipSuccess) e = hipMemsetAsync(b->jtr, 0, sizeof(int), st);
    if (e == hipSuccess) e = hipMemsetAsync(b->info, 0, sizeof(int), st);
    if (e == hipSuccess) {
        k_bf_dense_y<<<J, 256, 0, st>>>(J, b->off, dyoff, dypos, dyval, b->yd);
        const size_t nm = (size_t)c->N * M;
        k_bf_transpose<<<(unsigned)((nm + 255) / 256), 256, 0, st>>>(c->N, c->Npad, M,
                                                                   c->Z + (size_t)c->g0 * c->Npad, b->Krm);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dyoff);
    (void)hipFree(dypos);
    (void)hipFree(dyval);
    if (e != hipSuccess) {
        gp_set_error(std::string("gp_bfit_create: ") + hipGetErrorString(e));
        gp_bfit_destroy(b);
        return (int)e;
    }
    *out = b;
    return 0;
}

extern "C" int gp_bfit_destroy(gp_bfit *b) {
    if (!b) return 0;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *bufs[] = {b->off, b->cell, b->ncov, b->order, b->queue, b->covoff, b->cov, b->yd, b->Krm,
                    b->Beta, b->Psi, b->bvec, b->sov2, b->phi, b->lk, b->obj, b->delta, b->info, b->jtr,
                    b->cellc};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    delete b;
    return 0;
}

// init.cu:125-221 per gene: unweighted Gram of the gene's design, Cholesky,
// beta from log((y + 0.05) / s); resets phi to 1 and the step counter.
extern "C" int gp_bfit_init(gp_bfit *b) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    gp_ctx *c = b->c;
    GP_HIP(hipMemsetAsync(b->jtr, 0, sizeof(int), c->stream));
    GP_HIP(hipMemsetAsync(b->info, 0, sizeof(int), c->stream));
    GP_TRY(gpk_fill(c, b->phi, b->J, 1.0));
    GP_TRY(bf_pass(b, 0));
    int bad = 0;
    GP_HIP(hipMemcpyAsync(&bad, b->info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return bad ? -bad : 0;
}

// One IRLS step for every gene (damping from the internal step counter, which
// gp_bfit_init resets); *obj = sum of the genes' Poisson log-likelihoods at the
// weights the step started from (getLkhdPois + collectLkhd with N = 0).
extern "C" int gp_bfit_step(gp_bfit *b, double *obj) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    gp_ctx *c = b->c;
    GP_HIP(hipMemsetAsync(b->info, 0, sizeof(int), c->stream));
    GP_TRY(bf_step(b));
    int bad = 0;
    if (obj) GP_HIP(hipMemcpyAsync(obj, b->obj, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipMemcpyAsync(&bad, b->info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return bad ? -bad : 0;
}

// Times the persistent pass (the Gram, bvec and likelihood terms at the
// current beta; beta and phi are untouched, the stored factors are
// overwritten by unfactored Grams until the next step) with 1, 2, 3 and 4
// persistent workgroups per CU and keeps the fastest.  Every gene is
// processed by one workgroup in a fixed order whatever the grid, so the
// choice does not change any result.  A recorded step graph is dropped (it
// holds the old grid).
extern "C" int gp_bfit_autotune(gp_bfit *b, int verbose, int *grid_mult) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    gp_ctx *c = b->c;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    hipEvent_t e0, e1;
    GP_HIP(hipEventCreate(&e0));
    GP_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int pick = b->grid_mult;
    for (int m = 1; m <= 4; m++) {
        b->grid_mult = m;
        int r = bf_launch_pass(b, 1);             // warm-up at this grid
        if (r) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return r; }
        (void)hipEventRecord(e0, c->stream);
        for (int k = 0; k < 2 && !r; k++) r = bf_launch_pass(b, 1);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return r; }
        if (verbose) std::printf("bfit autotune: %d workgroups per CU: %.3f ms per pass\n", m, ms / 2);
        if (ms < best) { best = ms; pick = m; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->grid_mult = pick;
    if (grid_mult) *grid_mult = pick;
    return 0;
}

extern "C" int gp_bfit_set_grid(gp_bfit *b, int grid_mult) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    if (grid_mult < 1 || grid_mult > 8) { gp_set_error("grid_mult must be in [1, 8]"); return 1; }
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->grid_mult = grid_mult;
    return 0;
}

// The IRLS loop with the reference's stop rule (ggplvm.cu:1156-1181: stop at
// maxit or when the relative change of the objective drops below tol).  With
// use_graph the step is recorded once as a HIP graph and replayed.
extern "C" int gp_bfit_run(gp_bfit *b, int maxit, double tol, int use_graph, int *iters,
                           double *obj) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    gp_ctx *c = b->c;
    if (use_graph && !b->gexec) {
        GP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int r = bf_step(b);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (r > 0) { if (gr) (void)hipGraphDestroy(gr); return r; }
        if (e != hipSuccess) { gp_set_error(std::string("capture: ") + hipGetErrorString(e)); return (int)e; }
        b->graph = gr;
        GP_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    double prev = 0.0, cur = 0.0;
    int it = 0;
    GP_TRY(gpk_fill_int(c, b->info, 1, 0));
    for (it = 0; it < maxit; it++) {
        if (use_graph) GP_HIP(hipGraphLaunch(b->gexec, c->stream));
        else GP_TRY(bf_step(b));
        GP_HIP(hipMemcpyAsync(&cur, b->obj, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        GP_HIP(hipStreamSynchronize(c->stream));
        if (it > 0 && std::fabs(prev - cur) / std::fabs(cur) < tol) { it++; break; }
        prev = cur;
    }
    if (iters) *iters = it;
    if (obj) *obj = cur;
    int bad = 0;
    GP_HIP(hipMemcpyAsync(&bad, b->info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return bad ? -bad : 0;
}

// "Beta" (J x PB, gene g's first p_g + M entries), "phi" (J), "lkhd" (J),
// "Psi" (J x PB x PB factors).
extern "C" int gp_bfit_get(gp_bfit *b, const char *name, double *host, size_t n) {
    if (!b || !name || !host) { gp_set_error("NULL argument"); return 1; }
    const std::string s(name);
    const double *src = nullptr;
    size_t len = 0;
    if (s == "Beta") { src = b->Beta; len = (size_t)b->J * b->PB; }
    else if (s == "phi") { src = b->phi; len = b->J; }
    else if (s == "lkhd") { src = b->lk; len = b->J; }
    else if (s == "Psi") { src = b->Psi; len = (size_t)b->J * b->PB * b->PB; }
    else { gp_set_error("unknown batched-fit buffer " + s); return 1; }
    if (n < len) { gp_set_error("buffer too small"); return 1; }
    GP_HIP(hipStreamSynchronize(b->c->stream));
    GP_HIP(hipMemcpy(host, src, len * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int gp_bfit_dims(gp_bfit *b, int *J, int *PB) {
    if (!b) { gp_set_error("NULL handle"); return 1; }
    if (J) *J = b->J;
    if (PB) *PB = b->PB;
    return 0;
}
