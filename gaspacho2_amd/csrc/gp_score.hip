// gp_score.hip -- score statistics of gene x genotype tests (getScoreStats,
// ggplvm_de.cu:491-532) from the current fit state:
//   b    = Z^T D_j h,  D_j = diag(W_j / Omega)
//   hVh  = h^T D_j h - b^T Psi_j^{-1} b = h^T D_j h - |L_j^{-1} b|^2   (Psi_j = L_j L_j^T)
//   ss   = h^T Omega^{-1} (y_j - W_j)
//   stat = ss / sqrt(phi_j hVh),   p = erfc(|stat| / sqrt 2)  (two-sided normal)
// with h the genotype column g_l, or, given a per-cell factor d, the
// interaction column h = g_l * d - mean_i(g_l * d) that ggplvm_de.cu:989-994
// builds on the host (G_h[i + l N] *= d_h[i]; scale() centres it).
//
// Tests are processed in chunks of TT: k_score_acc walks the cells in 64-cell
// tiles (tiles dealt round robin to NBK workgroups per chunk), stages h_i D_j
// for the chunk's tests in LDS and accumulates, per model-matrix column k (one
// thread per column), the partial b_k of every test of the chunk, plus the
// partial h^T D h and h^T D 1; k_score_fin sums the NBK partials in a fixed
// order, adds the count term over the gene's non-zeros, and does the forward
// substitution with the gene's factor.  Several chunks run side by side when
// the partial buffer holds them.  The cell weights, dosages and model-matrix
// columns are read once per chunk; no library routine is called, so the same
// launcher serves the entry point that allocates nothing (its partials live in
// a static device array).  Works in the basis of the current factor (Z, or the
// whitened Zw: b^T Psi^{-1} b does not depend on the basis).
#include "gp_internal.h"
#include "../../include/gaspacho2.h"

#include <algorithm>

namespace {

constexpr int ST = 256;       // threads per workgroup
constexpr int TT = 16;        // tests per chunk
constexpr int NBK = 128;      // cell-tile workgroups per chunk
constexpr int CT = 64;        // cells per staged tile
constexpr int PMAX = 1024;    // widest design handled (model-matrix columns)

// partial layout per chunk in flight: NBK x TT x (P + 2) doubles
inline size_t chunk_doubles(int P) { return (size_t)NBK * TT * (P + 2); }

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

__device__ inline double gval(const double *G, size_t i) { return G[i]; }
__device__ inline double gval(const unsigned char *G, size_t i) { return (double)G[i]; }

// Genotype of SNP l at cell i: column l of G (N rows), or with a donor index
// the row of the cell's donor in column l of G (ND rows): cells share their
// donor's genotype
template <typename GT>
__device__ inline double gat(const GT *G, const int *donor, int ld, int l, int i) {
    return gval(G + (size_t)l * ld, donor ? donor[i] : i);
}

// cmean[l] = mean_i G[i, l] d[i]  (the centring of ggplvm_de.cu:994, util.c:100-110)
template <typename GT>
__global__ void __launch_bounds__(ST) k_score_colmean(int N, const GT *__restrict__ G,
        const int *__restrict__ donor, int ld, const double *__restrict__ d,
        double *__restrict__ cmean) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < N; i += ST) s += gat(G, donor, ld, blockIdx.x, i) * d[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) cmean[blockIdx.x] = s / (double)N;
}

__device__ inline bool test_ok(int j, int l, int J, int S) {
    return j >= 0 && j < J && l >= 0 && l < S;
}

// grid (NBK, chunks in flight); chunk c of this launch covers tests
// t0 + c*TT .. t0 + c*TT + TT - 1 (those < T).
template <typename GT>
__global__ void __launch_bounds__(ST) k_score_acc(int N, int Npad, int P, int J, int S, int T,
        int t0, const GT *__restrict__ G, const int *__restrict__ donor, int ld,
        const double *__restrict__ d,
        const double *__restrict__ cmean, const int *__restrict__ gene,
        const int *__restrict__ snp, const double *__restrict__ W,
        const double *__restrict__ Omega, const double *__restrict__ Zsrc,
        double *__restrict__ part, int *__restrict__ err) {
    __shared__ double v[CT][TT];          // h_i * o_ij for the tile's cells
    __shared__ int sj[TT], sl[TT];
    __shared__ double sc[TT];
    const int tid = threadIdx.x;
    const int tb = t0 + blockIdx.y * TT;
    if (tb >= T) return;
    const int nt = min(TT, T - tb);
    if (tid < TT) {
        int j = -1, l = -1;
        double c = 0.0;
        if (tid < nt) {
            j = gene[tb + tid];
            l = snp[tb + tid];
            if (!test_ok(j, l, J, S)) {
                if (blockIdx.x == 0) err[0] = 1;
                j = -1;
                l = -1;
            } else if (d) {
                c = cmean[l];
            }
        }
        sj[tid] = j;
        sl[tid] = l;
        sc[tid] = c;
    }
    __syncthreads();
    double *pb = part + ((size_t)blockIdx.y * NBK + blockIdx.x) * TT * (P + 2);
    // staging map: thread -> cell tid % 64 of the tile, tests w, w+4, w+8, w+12 (w = wave)
    const int ci = tid & (CT - 1), w = tid >> 6;
    const int ntile = Npad / CT;
    for (int k0 = 0; k0 < P; k0 += ST) {
        const int k = k0 + tid;
        double acc[TT];
#pragma unroll
        for (int t = 0; t < TT; t++) acc[t] = 0.0;
        double sa[4] = {0.0, 0.0, 0.0, 0.0}, sw[4] = {0.0, 0.0, 0.0, 0.0};
        for (int tile = blockIdx.x; tile < ntile; tile += NBK) {
            const int i = tile * CT + ci;
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = w + 4 * u;
                double val = 0.0;
                if (i < N && sj[t] >= 0) {
                    double h = gat(G, donor, ld, sl[t], i);
                    if (d) h = h * d[i] - sc[t];
                    const double o = W[i + (size_t)sj[t] * Npad] / Omega[i];
                    val = h * o;
                    sa[u] += h * val;
                    sw[u] += val;
                }
                v[ci][t] = val;
            }
            __syncthreads();
            if (k < P) {
                const double *zc = Zsrc + (size_t)k * Npad + (size_t)tile * CT;
#pragma unroll 4
                for (int ii = 0; ii < CT; ii++) {
                    const double z = zc[ii];
#pragma unroll
                    for (int t = 0; t < TT; t++) acc[t] += z * v[ii][t];
                }
            }
        }
        if (k < P) {
#pragma unroll
            for (int t = 0; t < TT; t++) pb[(size_t)t * (P + 2) + k] = acc[t];
        }
        if (k0 == 0) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double a = wave_sum(sa[u]), s = wave_sum(sw[u]);
                if ((tid & 63) == 0) {
                    const int t = w + 4 * u;
                    pb[(size_t)t * (P + 2) + P] = a;
                    pb[(size_t)t * (P + 2) + P + 1] = s;
                }
            }
        }
    }
}

// The same partials with every model-matrix column in one pass over the
// cells: per 64-cell tile the staged h_i o_ij (as above) is the B operand of
// v_mfma_f64_16x16x4f64 and the tile's Z rows the A operand, read straight
// from Zsrc; wave w owns the 16-column tiles w, w + 4, ... (KTW of them, FP64
// accumulators), so the cell weights are staged once per tile instead of
// once per 256 columns.  ZT = float reads the mixed path's FP32 copy of Zw
// (z32 = zw 2^(14 - zc), half the bytes; each column's sum is unscaled by
// zun[k] = 2^(zc - 14) at the end), ZT = double the FP64 model matrix (zun
// null).  Fragment layouts (lane l): A[m = l % 16][k = l / 16]
// = Z[cell 4 s + l / 16][column 16 kt + l % 16], B[k = l / 16][n = l % 16] =
// v[cell 4 s + l / 16][test l % 16], D[m = 4 r + l / 16][n = l % 16].
typedef double sc_d4 __attribute__((ext_vector_type(4)));
template <typename GT, int KTW, typename ZT>
__global__ void __launch_bounds__(ST) k_score_mfma(int N, int Npad, int P, int J, int S, int T,
        int t0, const GT *__restrict__ G, const int *__restrict__ donor, int ld,
        const double *__restrict__ d,
        const double *__restrict__ cmean, const int *__restrict__ gene,
        const int *__restrict__ snp, const double *__restrict__ W,
        const double *__restrict__ Omega, const ZT *__restrict__ Zsrc, const double *__restrict__ zun,
        double *__restrict__ part, int *__restrict__ err) {
    static_assert(TT == 16 && CT == 64 && ST == 256, "fragment layout: 16 tests, 64-cell tiles, 4 waves");
    __shared__ double v[CT][TT];
    __shared__ int sj[TT], sl[TT];
    __shared__ double sc[TT];
    const int tid = threadIdx.x;
    const int tb = t0 + blockIdx.y * TT;
    if (tb >= T) return;
    const int nt = min(TT, T - tb);
    if (tid < TT) {
        int j = -1, l = -1;
        double c = 0.0;
        if (tid < nt) {
            j = gene[tb + tid];
            l = snp[tb + tid];
            if (!test_ok(j, l, J, S)) {
                if (blockIdx.x == 0) err[0] = 1;
                j = -1;
                l = -1;
            } else if (d) {
                c = cmean[l];
            }
        }
        sj[tid] = j;
        sl[tid] = l;
        sc[tid] = c;
    }
    __syncthreads();
    double *pb = part + ((size_t)blockIdx.y * NBK + blockIdx.x) * TT * (P + 2);
    const int ci = tid & (CT - 1), w = tid >> 6, lane = tid & 63;
    const int lm = lane & 15, lk = lane >> 4;
    const int ntile = Npad / CT;
    const int KT = (P + 15) / 16;
    sc_d4 acc[KTW];
#pragma unroll
    for (int q = 0; q < KTW; q++) acc[q] = sc_d4{0.0, 0.0, 0.0, 0.0};
    double sa[4] = {0.0, 0.0, 0.0, 0.0}, sw[4] = {0.0, 0.0, 0.0, 0.0};
    for (int tile = blockIdx.x; tile < ntile; tile += NBK) {
        const int i = tile * CT + ci;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t = w + 4 * u;
            double val = 0.0;
            if (i < N && sj[t] >= 0) {
                double h = gat(G, donor, ld, sl[t], i);
                if (d) h = h * d[i] - sc[t];
                const double o = W[i + (size_t)sj[t] * Npad] / Omega[i];
                val = h * o;
                sa[u] += h * val;
                sw[u] += val;
            }
            v[ci][t] = val;
        }
        __syncthreads();
        double bf[CT / 4];
#pragma unroll
        for (int st = 0; st < CT / 4; st++) bf[st] = v[4 * st + lk][lm];
        const size_t cell0 = (size_t)tile * CT + lk;
#pragma unroll
        for (int q = 0; q < KTW; q++) {
            const int kt = w + 4 * q;
            if (kt < KT) {                             // uniform over the wave
                const int col = min(16 * kt + lm, P - 1);
                const bool colok = 16 * kt + lm < P;
                const ZT *zc = Zsrc + (size_t)col * Npad + cell0;
                double af[CT / 4];
#pragma unroll
                for (int st = 0; st < CT / 4; st++) af[st] = (double)zc[4 * st];
#pragma unroll
                for (int st = 0; st < CT / 4; st++)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(colok ? af[st] : 0.0, bf[st], acc[q], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < KTW; q++) {
        const int kt = w + 4 * q;
        if (kt < KT) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = 16 * kt + 4 * r + lk;
                if (k < P) pb[(size_t)lm * (P + 2) + k] = zun ? acc[q][r] * zun[k] : acc[q][r];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const double a = wave_sum(sa[u]), s = wave_sum(sw[u]);
        if ((tid & 63) == 0) {
            const int t = w + 4 * u;
            pb[(size_t)t * (P + 2) + P] = a;
            pb[(size_t)t * (P + 2) + P + 1] = s;
        }
    }
}

// One workgroup per test of the launch: sum the NBK partials, add the count
// term over the gene's non-zeros, forward substitution, statistic, p-value.
template <typename GT>
__global__ void __launch_bounds__(ST) k_score_fin(int N, int P, int T, int t0,
        const GT *__restrict__ G, const int *__restrict__ donor, int ld,
        const double *__restrict__ d, const double *__restrict__ cmean,
        const int *__restrict__ gene, const int *__restrict__ snp, int J, int S,
        const int *__restrict__ Yp, const int *__restrict__ Yi, const double *__restrict__ Yx,
        const double *__restrict__ Omega, const double *__restrict__ Psi,
        const double *__restrict__ phi, const double *__restrict__ part,
        double *__restrict__ stat, double *__restrict__ pval) {
    extern __shared__ double xs[];        // P + 2
    __shared__ double sh[4];
    const int tl = blockIdx.x, tid = threadIdx.x;
    const int t = t0 + tl;
    if (t >= T) return;
    const int j = gene[t], l = snp[t];
    if (!test_ok(j, l, J, S)) {
        if (tid == 0) {
            stat[t] = __builtin_nan("");
            if (pval) pval[t] = __builtin_nan("");
        }
        return;
    }
    const int c = tl / TT, tt = tl % TT;
    const double *pc = part + (size_t)c * NBK * TT * (P + 2) + (size_t)tt * (P + 2);
    for (int k = tid; k < P + 2; k += ST) {
        double s = 0.0;
        for (int b = 0; b < NBK; b++) s += pc[(size_t)b * TT * (P + 2) + k];
        xs[k] = s;
    }
    const double cm = d ? cmean[l] : 0.0;
    double sy = 0.0;
    {   // the count term over the gene's non-zeros, SYU entries per thread in
        // flight (the cell index, then its genotype, factor and Omega)
        constexpr int SYU = 8;
        const int q1 = Yp[j + 1];
        int q = Yp[j] + tid;
        for (; q + (SYU - 1) * ST < q1; q += SYU * ST) {
            int ii[SYU];
            double yx[SYU], h[SYU], om[SYU], dv[SYU];
#pragma unroll
            for (int u = 0; u < SYU; u++) {
                ii[u] = Yi[q + u * ST];
                yx[u] = Yx[q + u * ST];
            }
#pragma unroll
            for (int u = 0; u < SYU; u++) {
                h[u] = gat(G, donor, ld, l, ii[u]);
                om[u] = Omega[ii[u]];
                dv[u] = d ? d[ii[u]] : 1.0;
            }
#pragma unroll
            for (int u = 0; u < SYU; u++) {
                const double hh = d ? h[u] * dv[u] - cm : h[u];
                sy += hh * yx[u] / om[u];
            }
        }
        for (; q < q1; q += ST) {
            const int i = Yi[q];
            double h = gat(G, donor, ld, l, i);
            if (d) h = h * d[i] - cm;
            sy += h * Yx[q] / Omega[i];
        }
    }
    sy = block_sum(sy, sh);                // also orders the xs writes before the reads below
    const double a = xs[P], sw = xs[P + 1];
    const double *L = Psi + (size_t)j * P * P;
    // L y = b by blocks of SBF rows: the diagonal block on wave 0 (row r in
    // lane r, each y_c broadcast by a shuffle), then every later row takes the
    // block's SBF terms; two barriers per block instead of two per row
    constexpr int SBF = 16;
    const int lane = tid & 63;
    for (int kb = 0; kb < P; kb += SBF) {
        const int nb = min(SBF, P - kb);
        if (tid < 64) {
            const int r = lane;
            double t = (r < nb) ? xs[kb + r] : 0.0;
            double lrow[SBF];
#pragma unroll
            for (int cc = 0; cc < SBF; cc++)
                lrow[cc] = (r < nb && cc <= r) ? L[(kb + r) + (size_t)(kb + cc) * P] : 0.0;
            double y = 0.0;
#pragma unroll
            for (int cc = 0; cc < SBF; cc++) {
                if (r == cc && cc < nb) y = t / lrow[cc];
                const double yc = __shfl(y, cc);
                if (r > cc && cc < nb) t -= lrow[cc] * yc;
            }
            if (r < nb) xs[kb + r] = y;
        }
        __syncthreads();
        double yb[SBF];
#pragma unroll
        for (int cc = 0; cc < SBF; cc++) yb[cc] = (cc < nb) ? xs[kb + cc] : 0.0;
        for (int r = kb + nb + tid; r < P; r += ST) {
            double v = xs[r];
#pragma unroll
            for (int cc = 0; cc < SBF; cc++)
                if (cc < nb) v -= L[r + (size_t)(kb + cc) * P] * yb[cc];
            xs[r] = v;
        }
        __syncthreads();
    }
    double q = 0.0;
    for (int k = tid; k < P; k += ST) q += xs[k] * xs[k];
    q = block_sum(q, sh);
    if (tid == 0) {
        const double st = (sy - sw) / sqrt(phi[j] * (a - q));
        stat[t] = st;
        if (pval) pval[t] = erfc(fabs(st) * 0.70710678118654752440);
    }
}

// Every chunk of TT tests: accumulation and finish, `fly` chunks per launch
// (part holds fly * chunk_doubles(P) doubles).  All pointers are device
// pointers; d / cmean / pval may be null.
template <typename GT>
int score_run(gp_ctx *c, int S, const GT *G, const int *donor, int ld, const double *d,
              const double *cmean, int T, const int *gene, const int *snp, double *stat,
              double *pval, double *part, int fly, int *err) {
    const int P = c->P;
    const double *Zsrc = c->psi_whitened ? c->Zw : c->Z;
    const int per = fly * TT;
    for (int t0 = 0; t0 < T; t0 += per) {
        const int n = std::min(per, T - t0);
        const int nch = (n + TT - 1) / TT;
        const int ktw = ((P + 15) / 16 + 3) / 4;   // 16-column tiles per wave of the MFMA form
        // the FP32 copy of Zw when the mixed path holds it for the current basis
        const bool z32 = c->precision == GP_PREC_MIXED && c->psi_whitened && c->zsplit_valid && c->z32;
#define SCORE_MFMA(K)                                                                                    \
    case K:                                                                                              \
        if (z32)                                                                                         \
            k_score_mfma<GT, K, float><<<dim3(NBK, nch), ST, 0, c->stream>>>(                            \
                c->N, c->Npad, P, c->J, S, T, t0, G, donor, ld, d, cmean, gene, snp, c->W, c->Omega,    \
                c->z32, c->zun, part, err);                                                              \
        else                                                                                             \
            k_score_mfma<GT, K, double><<<dim3(NBK, nch), ST, 0, c->stream>>>(                           \
                c->N, c->Npad, P, c->J, S, T, t0, G, donor, ld, d, cmean, gene, snp, c->W, c->Omega,    \
                Zsrc, nullptr, part, err);                                                               \
        break;
        switch (c->score_mfma ? ktw : 0) {
            SCORE_MFMA(1) SCORE_MFMA(2) SCORE_MFMA(3) SCORE_MFMA(4) SCORE_MFMA(5)
            SCORE_MFMA(6) SCORE_MFMA(7) SCORE_MFMA(8) SCORE_MFMA(9)
        default:
            k_score_acc<GT><<<dim3(NBK, nch), ST, 0, c->stream>>>(c->N, c->Npad, P, c->J, S, T, t0, G, donor,
                                                                 ld, d, cmean, gene, snp, c->W, c->Omega,
                                                                 Zsrc, part, err);
        }
#undef SCORE_MFMA
        GP_HIP(hipGetLastError());
        k_score_fin<GT><<<n, ST, (size_t)(P + 2) * sizeof(double), c->stream>>>(
            c->N, P, T, t0, G, donor, ld, d, cmean, gene, snp, c->J, S, c->Yp, c->Yi, c->Yx, c->Omega, c->Psi,
            c->phi, part, stat, pval);
        GP_HIP(hipGetLastError());
    }
    return 0;
}

__global__ void k_score_clear(int *err) { err[0] = 0; }

// Host-array entry points: device copies of the inputs and a partial buffer
// for up to 16 chunks in flight, freed before returning.
template <typename GT>
int score_tests_host(gp_ctx *c, int S, const GT *G_host, int ND, const int *donor_host,
                     const double *d_host, int T, const int *gene_host, const int *snp_host,
                     double *stat_host, double *pval_host) {
    if (S <= 0 || T <= 0) { gp_set_error("score tests: S and T must be positive"); return 1; }
    GP_TRY(gp_wnu_sync(c));   // the tests read W
    if (donor_host) {
        if (ND <= 0) { gp_set_error("score tests: the donor count must be positive"); return 1; }
        for (int i = 0; i < c->N; i++)
            if (donor_host[i] < 0 || donor_host[i] >= ND) {
                gp_set_error("score tests: donor index out of range at cell " + std::to_string(i));
                return 1;
            }
    }
    if (c->P > PMAX) { gp_set_error("score tests: more than 1024 model-matrix columns"); return 1; }
    for (int t = 0; t < T; t++) {
        if (gene_host[t] < 0 || gene_host[t] >= c->J || snp_host[t] < 0 || snp_host[t] >= S) {
            gp_set_error("score tests: gene or genotype index out of range at test " +
                         std::to_string(t));
            return 1;
        }
    }
    const int N = c->N;
    const int ld = donor_host ? ND : N;          // rows of G
    const int fly = std::max(1, std::min(16, (T + TT - 1) / TT));
    GT *dG = nullptr;
    int *dgene = nullptr, *dsnp = nullptr, *derr = nullptr, *ddon = nullptr;
    double *dd = nullptr, *dcm = nullptr, *part = nullptr, *dstat = nullptr, *dp = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (void *p : {(void *)dG, (void *)dgene, (void *)dsnp, (void *)derr, (void *)dd,
                        (void *)dcm, (void *)part, (void *)dstat, (void *)dp, (void *)ddon})
            if (p) (void)hipFree(p);
    };
#define SC_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            gp_set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            cleanup();                                                                 \
            return (int)e_;                                                            \
        }                                                                              \
    } while (0)
    SC_HIP(hipMalloc((void **)&dG, (size_t)ld * S * sizeof(GT)));
    if (donor_host) {
        SC_HIP(hipMalloc((void **)&ddon, (size_t)N * sizeof(int)));
        SC_HIP(hipMemcpyAsync(ddon, donor_host, (size_t)N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    SC_HIP(hipMalloc((void **)&dgene, (size_t)T * sizeof(int)));
    SC_HIP(hipMalloc((void **)&dsnp, (size_t)T * sizeof(int)));
    SC_HIP(hipMalloc((void **)&derr, sizeof(int)));
    SC_HIP(hipMalloc((void **)&part, (size_t)fly * chunk_doubles(c->P) * sizeof(double)));
    SC_HIP(hipMalloc((void **)&dstat, (size_t)T * sizeof(double)));
    if (pval_host) SC_HIP(hipMalloc((void **)&dp, (size_t)T * sizeof(double)));
    if (d_host) {
        SC_HIP(hipMalloc((void **)&dd, (size_t)N * sizeof(double)));
        SC_HIP(hipMalloc((void **)&dcm, (size_t)S * sizeof(double)));
        SC_HIP(hipMemcpyAsync(dd, d_host, (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    SC_HIP(hipMemcpyAsync(dG, G_host, (size_t)ld * S * sizeof(GT), hipMemcpyHostToDevice, c->stream));
    SC_HIP(hipMemcpyAsync(dgene, gene_host, (size_t)T * sizeof(int), hipMemcpyHostToDevice, c->stream));
    SC_HIP(hipMemcpyAsync(dsnp, snp_host, (size_t)T * sizeof(int), hipMemcpyHostToDevice, c->stream));
    k_score_clear<<<1, 1, 0, c->stream>>>(derr);
    SC_HIP(hipGetLastError());
    if (d_host) {
        k_score_colmean<GT><<<S, ST, 0, c->stream>>>(N, dG, ddon, ld, dd, dcm);
        SC_HIP(hipGetLastError());
    }
    const int rc = score_run<GT>(c, S, dG, ddon, ld, dd, dcm, T, dgene, dsnp, dstat, dp, part, fly, derr);
    if (rc > 0) { cleanup(); return rc; }
    SC_HIP(hipMemcpyAsync(stat_host, dstat, (size_t)T * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (pval_host)
        SC_HIP(hipMemcpyAsync(pval_host, dp, (size_t)T * sizeof(double), hipMemcpyDeviceToHost, c->stream));
#undef SC_HIP
    cleanup();
    return 0;
}

}  // namespace

size_t gpk_score_part_doubles(int P) { return chunk_doubles(P); }
int gpk_score_max_columns() { return PMAX; }
int gpk_score_clear(gp_ctx *c, int *err) {
    k_score_clear<<<1, 1, 0, c->stream>>>(err);
    GP_HIP(hipGetLastError());
    return 0;
}
int gpk_score_run_dosage(gp_ctx *c, int S, const unsigned char *G, int T, const int *gene,
                         const int *snp, double *stat, double *part, int fly, int *err) {
    return score_run<unsigned char>(c, S, G, nullptr, c->N, nullptr, nullptr, T, gene, snp, stat,
                                    nullptr, part, fly, err);
}

extern "C" int gp_score_tests(gp_ctx *c, int S, const double *G, int T, const int *gene,
                              const int *snp, double *stat) {
    if (!c || !G || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests_host<double>(c, S, G, 0, nullptr, nullptr, T, gene, snp, stat, nullptr);
}

extern "C" int gp_score_tests_dosage(gp_ctx *c, int S, const unsigned char *dosage, int T,
                                     const int *gene, const int *snp, double *stat) {
    if (!c || !dosage || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests_host<unsigned char>(c, S, dosage, 0, nullptr, nullptr, T, gene, snp, stat, nullptr);
}

extern "C" int gp_score_tests_interaction(gp_ctx *c, int S, const unsigned char *dosage,
                                          const double *factor, int T, const int *gene,
                                          const int *snp, double *stat, double *pval) {
    if (!c || !dosage || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests_host<unsigned char>(c, S, dosage, 0, nullptr, factor, T, gene, snp, stat, pval);
}

extern "C" int gp_score_tests_donor(gp_ctx *c, int ndonor, const int *donor, int S,
                                    const unsigned char *dosage, const double *factor, int T,
                                    const int *gene, const int *snp, double *stat, double *pval) {
    if (!c || !donor || !dosage || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests_host<unsigned char>(c, S, dosage, ndonor, donor, factor, T, gene, snp, stat, pval);
}
