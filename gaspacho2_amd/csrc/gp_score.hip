// gp_score.hip -- score statistics of gene x genotype tests (getScoreStats,
// ggplvm_de.cu:491-532) from the current fit state:
//   b    = Z^T D_j g,  D_j = diag(W_j / Omega)
//   gVg  = g^T D_j g - b^T Psi_j^{-1} b = g^T D_j g - |L_j^{-1} b|^2   (Psi_j = L_j L_j^T)
//   ss   = g^T Omega^{-1} (y_j - W_j)
//   stat = ss / sqrt(phi_j gVg)
// The tests are processed in chunks: one kernel forms D_j g for every test of
// the chunk (plus the scalar sums g^T D_j g, g^T D_j 1 and g^T Omega^{-1} y_j
// over the gene's non-zeros), one split-K GEMM forms all b = Z^T (D_j g) at
// once, and one kernel per test does the forward substitution with the
// gene's factor.  Works in the basis of the current factor (Z or the
// whitened Zw: b^T Psi^{-1} b does not depend on the basis).
#include "gp_internal.h"

#include <algorithm>

namespace {

constexpr int ST = 256;

__device__ inline double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

__device__ inline double gval(const double *G, size_t i) { return G[i]; }
__device__ inline double gval(const unsigned char *G, size_t i) { return (double)G[i]; }

// One block per test t of the chunk: Gw[:, t] = D_j g (zero on padded rows),
// and the scalars a[t] = g^T D_j g, s[t] = g^T Omega^{-1} (y_j - W_j).
template <typename GT>
__global__ void __launch_bounds__(ST) k_score_prep(int N, int Npad, int t0, int tc,
        const GT *__restrict__ G, const int *__restrict__ gene, const int *__restrict__ snp,
        const double *__restrict__ W, const double *__restrict__ Omega,
        const int *__restrict__ Yp, const int *__restrict__ Yi, const double *__restrict__ Yx,
        double *__restrict__ Gw, double *__restrict__ a, double *__restrict__ s) {
    __shared__ double sh[4];
    const int t = blockIdx.x;
    if (t >= tc) return;
    const int j = gene[t0 + t];
    const GT *g = G + (size_t)snp[t0 + t] * N;
    const double *Wj = W + (size_t)j * Npad;
    double *gw = Gw + (size_t)t * Npad;
    double sa = 0.0, sw = 0.0;
    for (int i = threadIdx.x; i < Npad; i += ST) {
        double v = 0.0;
        if (i < N) {
            const double gi = gval(g, i);
            const double o = Wj[i] / Omega[i];
            v = o * gi;
            sa += gi * v;
            sw += v;
        }
        gw[i] = v;
    }
    double sy = 0.0;
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += ST) {
        const int i = Yi[l];
        sy += gval(g, i) * Yx[l] / Omega[i];
    }
    sa = block_sum(sa, sh);
    sw = block_sum(sw, sh);
    sy = block_sum(sy, sh);
    if (threadIdx.x == 0) {
        a[t] = sa;
        s[t] = sy - sw;
    }
}

// One block per test: |L_j^{-1} b|^2 by forward substitution, then the statistic.
__global__ void __launch_bounds__(ST) k_score_finish(int P, int t0, int tc,
        const int *__restrict__ gene, const double *__restrict__ Psi,
        const double *__restrict__ phi, const double *__restrict__ B,
        const double *__restrict__ a, const double *__restrict__ s, double *__restrict__ stat) {
    extern __shared__ double xs[];
    __shared__ double sh[4];
    const int t = blockIdx.x, tid = threadIdx.x;
    if (t >= tc) return;
    const int j = gene[t0 + t];
    const double *L = Psi + (size_t)j * P * P;
    for (int k = tid; k < P; k += ST) xs[k] = B[(size_t)t * P + k];
    for (int i = 0; i < P; i++) {
        __syncthreads();
        const double xi = xs[i] / L[i + (size_t)i * P];
        __syncthreads();
        if (tid == 0) xs[i] = xi;
        for (int r = i + 1 + tid; r < P; r += ST) xs[r] -= L[r + (size_t)i * P] * xi;
    }
    __syncthreads();
    double q = 0.0;
    for (int k = tid; k < P; k += ST) q += xs[k] * xs[k];
    q = block_sum(q, sh);
    if (tid == 0) stat[t0 + t] = s[t] / sqrt(phi[j] * (a[t] - q));
}

template <typename GT>
int score_tests(gp_ctx *c, int S, const GT *G_host, int T, const int *gene_host,
                const int *snp_host, double *stat_host) {
    if (S <= 0 || T <= 0) { gp_set_error("score tests: S and T must be positive"); return 1; }
    for (int t = 0; t < T; t++) {
        if (gene_host[t] < 0 || gene_host[t] >= c->J || snp_host[t] < 0 || snp_host[t] >= S) {
            gp_set_error("score tests: gene or genotype index out of range at test " +
                         std::to_string(t));
            return 1;
        }
    }
    const int N = c->N, Npad = c->Npad, P = c->P;
    const double *Zsrc = c->psi_whitened ? c->Zw : c->Z;
    // chunk: Gw (Npad x tc) within ~2 GiB
    int tc = (int)std::max<size_t>(1, ((size_t)1 << 28) / (size_t)Npad);
    tc = std::min(tc, T);
    GT *dG = nullptr;
    int *dgene = nullptr, *dsnp = nullptr;
    double *Gw = nullptr, *a = nullptr, *s = nullptr, *B = nullptr, *dstat = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (void *p : {(void *)dG, (void *)dgene, (void *)dsnp, (void *)Gw, (void *)a, (void *)s,
                        (void *)B, (void *)dstat})
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
    SC_HIP(hipMalloc((void **)&dG, (size_t)N * S * sizeof(GT)));
    SC_HIP(hipMalloc((void **)&dgene, (size_t)T * sizeof(int)));
    SC_HIP(hipMalloc((void **)&dsnp, (size_t)T * sizeof(int)));
    SC_HIP(hipMalloc((void **)&Gw, (size_t)Npad * tc * sizeof(double)));
    SC_HIP(hipMalloc((void **)&a, (size_t)tc * sizeof(double)));
    SC_HIP(hipMalloc((void **)&s, (size_t)tc * sizeof(double)));
    SC_HIP(hipMalloc((void **)&B, (size_t)P * tc * sizeof(double)));
    SC_HIP(hipMalloc((void **)&dstat, (size_t)T * sizeof(double)));
    SC_HIP(hipMemcpyAsync(dG, G_host, (size_t)N * S * sizeof(GT), hipMemcpyHostToDevice, c->stream));
    SC_HIP(hipMemcpyAsync(dgene, gene_host, (size_t)T * sizeof(int), hipMemcpyHostToDevice, c->stream));
    SC_HIP(hipMemcpyAsync(dsnp, snp_host, (size_t)T * sizeof(int), hipMemcpyHostToDevice, c->stream));
    for (int t0 = 0; t0 < T; t0 += tc) {
        const int n = std::min(tc, T - t0);
        k_score_prep<GT><<<n, ST, 0, c->stream>>>(N, Npad, t0, n, dG, dgene, dsnp, c->W, c->Omega,
                                                  c->Yp, c->Yi, c->Yx, Gw, a, s);
        SC_HIP(hipGetLastError());
        rc = gpk_gemm_tn_splitk(c, P, n, Npad, Zsrc, Npad, Gw, Npad, B, P);
        if (rc > 0) { cleanup(); return rc; }
        k_score_finish<<<n, ST, (size_t)P * sizeof(double), c->stream>>>(P, t0, n, dgene, c->Psi,
                                                                          c->phi, B, a, s, dstat);
        SC_HIP(hipGetLastError());
    }
    SC_HIP(hipMemcpyAsync(stat_host, dstat, (size_t)T * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
#undef SC_HIP
    cleanup();
    return 0;
}

}  // namespace

extern "C" int gp_score_tests(gp_ctx *c, int S, const double *G, int T, const int *gene,
                              const int *snp, double *stat) {
    if (!c || !G || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests<double>(c, S, G, T, gene, snp, stat);
}

extern "C" int gp_score_tests_dosage(gp_ctx *c, int S, const unsigned char *dosage, int T,
                                     const int *gene, const int *snp, double *stat) {
    if (!c || !dosage || !gene || !snp || !stat) { gp_set_error("NULL argument"); return 1; }
    return score_tests<unsigned char>(c, S, dosage, T, gene, snp, stat);
}
