// gp_pca.hip -- the PCA initial latent the reference's fit starts from
// (it reads it from a file, ggplvm.cu:1088-1095; the recipe is the unused
// helpers getNu0 / getNuTNu / getX, init.cu:304-357):
//   nu0[i, j] = log((y_ij + 0.5) / s_i) - z_i . beta_j     (z_i: intercept + covariates)
//   G = nu0^T nu0 (J x J), top-Q eigenpairs (lambda_k, v_k), largest first
//   X[:, k] = nu0 v_k / sqrt(lambda_k) * sqrt(N)
// beta_j is the least-squares fit of log((y_j + 0.5) / s) on the covariates
// (the reference leaves the choice to the caller of getNu0); dependent
// covariate columns are dropped, so nu0 is the residual of the projection
// onto their span.  Each v_k is
// signed so that its entry of largest magnitude (first such) is positive; the
// reference inherits the sign from its eigensolver.
//
// GPU: dense nu0 (N x J, FP64) built from the CSR counts, Z^T nu0 and the
// covariate fit by rocBLAS, G by SYRK, eigenpairs by rocSOLVER syevd, the
// projection by GEMM.  Only the p x p normal equations and the Q sign flips
// run on the host.
#include "gp_internal.h"
#include "../../include/gaspacho2.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace {

constexpr int PT = 256;

// nu0 = log(0.5 / s_i) everywhere (the zero counts)
__global__ void __launch_bounds__(PT) k_pca_fill(int N, int J, const double *__restrict__ ls,
                                                 double *__restrict__ Y0) {
    const size_t idx = (size_t)blockIdx.x * PT + threadIdx.x;
    if (idx >= (size_t)N * J) return;
    Y0[idx] = log(0.5 / ls[idx % N]);
}

// the stored counts of gene j (one workgroup per gene)
__global__ void __launch_bounds__(PT) k_pca_counts(int N, const int *__restrict__ Yp, const int *__restrict__ Yi,
                                                   const double *__restrict__ Yx, const double *__restrict__ ls,
                                                   double *__restrict__ Y0) {
    const int j = blockIdx.x;
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += PT) {
        const int i = Yi[l];
        Y0[i + (size_t)j * N] = log((Yx[l] + 0.5) / ls[i]);
    }
}

// X[:, k] *= sqrt(N / lambda_k)
__global__ void __launch_bounds__(PT) k_pca_scale(int N, int Q, const double *__restrict__ f,
                                                  double *__restrict__ X) {
    const size_t idx = (size_t)blockIdx.x * PT + threadIdx.x;
    if (idx >= (size_t)N * Q) return;
    X[idx] *= f[idx / N];
}

// Least-squares coefficients of the covariate fit from the normal equations
// A x = b (A = Z^T Z, p x p column-major), in place for nrhs right-hand
// sides.  Columns that are linear combinations of earlier ones (for example
// a full set of one-hot levels next to the intercept) are dropped: their
// coefficient is 0 and the residual is the projection onto the span of all
// columns.  Returns the number of columns kept.
int small_ls_solve(std::vector<double> A, int p, double *B, int nrhs) {
    std::vector<char> keep(p, 1);
    int kept = 0;
    for (int j = 0; j < p; j++) {
        const double a0 = A[j + j * p];
        double d = a0;
        for (int k = 0; k < j; k++)
            if (keep[k]) d -= A[j + k * p] * A[j + k * p];
        if (!(d > 1e-10 * a0) || !(a0 > 0.0)) {
            keep[j] = 0;
            for (int i = j; i < p; i++) A[i + j * p] = 0.0;
            continue;
        }
        kept++;
        d = std::sqrt(d);
        A[j + j * p] = d;
        for (int i = j + 1; i < p; i++) {
            double s = A[i + j * p];
            for (int k = 0; k < j; k++)
                if (keep[k]) s -= A[i + k * p] * A[j + k * p];
            A[i + j * p] = s / d;
        }
    }
    for (int r = 0; r < nrhs; r++) {
        double *b = B + (size_t)r * p;
        for (int i = 0; i < p; i++) {
            if (!keep[i]) { b[i] = 0.0; continue; }
            double s = b[i];
            for (int k = 0; k < i; k++)
                if (keep[k]) s -= A[i + k * p] * b[k];
            b[i] = s / A[i + i * p];
        }
        for (int i = p - 1; i >= 0; i--) {
            if (!keep[i]) continue;
            double s = b[i];
            for (int k = i + 1; k < p; k++)
                if (keep[k]) s -= A[k + i * p] * b[k];
            b[i] = s / A[i + i * p];
        }
    }
    return kept;
}

}  // namespace

extern "C" int gp_pca_init(int N, int J, const int *Yp, const int *Yi, const double *Yx, const double *ls,
                           int p0, const double *Zcov, int Q, double *X, double *eval) {
    if (!Yp || !Yi || !Yx || !ls || !Zcov || !X) { gp_set_error("NULL argument"); return 1; }
    if (N <= 0 || J <= 0 || p0 < 1 || Q < 1 || Q > J) {
        gp_set_error("pca init: need N, J > 0, p >= 1 and 1 <= Q <= J");
        return 1;
    }
    if (Yp[0] != 0) { gp_set_error("pca init: Yp[0] must be 0"); return 1; }
    for (int j = 0; j < J; j++) {
        if (Yp[j + 1] < Yp[j]) { gp_set_error("pca init: Yp must be non-decreasing"); return 1; }
        for (int l = Yp[j]; l < Yp[j + 1]; l++)
            if (Yi[l] < 0 || Yi[l] >= N || (l > Yp[j] && Yi[l] <= Yi[l - 1])) {
                gp_set_error("pca init: row indices of gene " + std::to_string(j) +
                             " must be increasing and in [0, N)");
                return 1;
            }
    }
    for (int i = 0; i < N; i++)
        if (!(ls[i] > 0.0)) { gp_set_error("pca init: size factors must be positive"); return 1; }
    const int L = Yp[J];
    // normal equations of the covariate fit (host, p0 x p0)
    std::vector<double> A((size_t)p0 * p0, 0.0);
    for (int a = 0; a < p0; a++)
        for (int b = 0; b <= a; b++) {
            double s = 0.0;
            for (int i = 0; i < N; i++) s += Zcov[i + (size_t)a * N] * Zcov[i + (size_t)b * N];
            A[a + b * p0] = A[b + a * p0] = s;
        }

    hipStream_t st = nullptr;
    rocblas_handle h = nullptr;
    int *dYp = nullptr, *dYi = nullptr, *dinfo = nullptr;
    double *dYx = nullptr, *dls = nullptr, *dZ = nullptr, *Y0 = nullptr, *dB = nullptr, *G = nullptr;
    double *dD = nullptr, *dE = nullptr, *dV = nullptr, *dX = nullptr, *dF = nullptr;
    int rc = 0;
    auto done = [&]() {
        if (st) (void)hipStreamSynchronize(st);
        for (void *p : {(void *)dYp, (void *)dYi, (void *)dinfo, (void *)dYx, (void *)dls, (void *)dZ, (void *)Y0,
                        (void *)dB, (void *)G, (void *)dD, (void *)dE, (void *)dV, (void *)dX, (void *)dF})
            if (p) (void)hipFree(p);
        if (h) (void)rocblas_destroy_handle(h);
        if (st) (void)hipStreamDestroy(st);
        return rc;
    };
#define PCA_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            gp_set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            rc = (int)e_;                                                               \
            return done();                                                              \
        }                                                                               \
    } while (0)
#define PCA_BLAS(call)                                                                  \
    do {                                                                                \
        rocblas_status s_ = (call);                                                     \
        if (s_ != rocblas_status_success) {                                             \
            gp_set_error(std::string(#call) + ": rocblas status " + std::to_string((int)s_)); \
            rc = 900 + (int)s_;                                                         \
            return done();                                                              \
        }                                                                               \
    } while (0)
    PCA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    PCA_BLAS(rocblas_create_handle(&h));
    PCA_BLAS(rocblas_set_stream(h, st));
    PCA_HIP(hipMalloc((void **)&dYp, (size_t)(J + 1) * sizeof(int)));
    PCA_HIP(hipMalloc((void **)&dYi, (size_t)std::max(L, 1) * sizeof(int)));
    PCA_HIP(hipMalloc((void **)&dYx, (size_t)std::max(L, 1) * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dls, (size_t)N * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dZ, (size_t)N * p0 * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&Y0, (size_t)N * J * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dB, (size_t)p0 * J * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&G, (size_t)J * J * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dD, (size_t)J * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dE, (size_t)J * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dinfo, sizeof(int)));
    PCA_HIP(hipMalloc((void **)&dV, (size_t)J * Q * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dX, (size_t)N * Q * sizeof(double)));
    PCA_HIP(hipMalloc((void **)&dF, (size_t)Q * sizeof(double)));
    PCA_HIP(hipMemcpyAsync(dYp, Yp, (size_t)(J + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    if (L > 0) {
        PCA_HIP(hipMemcpyAsync(dYi, Yi, (size_t)L * sizeof(int), hipMemcpyHostToDevice, st));
        PCA_HIP(hipMemcpyAsync(dYx, Yx, (size_t)L * sizeof(double), hipMemcpyHostToDevice, st));
    }
    PCA_HIP(hipMemcpyAsync(dls, ls, (size_t)N * sizeof(double), hipMemcpyHostToDevice, st));
    PCA_HIP(hipMemcpyAsync(dZ, Zcov, (size_t)N * p0 * sizeof(double), hipMemcpyHostToDevice, st));
    const size_t NJ = (size_t)N * J;
    k_pca_fill<<<(unsigned)((NJ + PT - 1) / PT), PT, 0, st>>>(N, J, dls, Y0);
    PCA_HIP(hipGetLastError());
    k_pca_counts<<<J, PT, 0, st>>>(N, dYp, dYi, dYx, dls, Y0);
    PCA_HIP(hipGetLastError());
    // beta = (Z^T Z)^-1 Z^T Y0 ; nu0 = Y0 - Z beta
    const double one = 1.0, zero = 0.0, mone = -1.0;
    PCA_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, p0, J, N, &one, dZ, N, Y0, N,
                           &zero, dB, p0));
    std::vector<double> B((size_t)p0 * J);
    PCA_HIP(hipMemcpyAsync(B.data(), dB, B.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    PCA_HIP(hipStreamSynchronize(st));
    if (small_ls_solve(A, p0, B.data(), J) == 0) {
        gp_set_error("pca init: every covariate column is zero");
        rc = 1;
        return done();
    }
    PCA_HIP(hipMemcpyAsync(dB, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice, st));
    PCA_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, J, p0, &mone, dZ, N, dB, p0, &one,
                           Y0, N));
    // G = nu0^T nu0 (lower triangle) and its eigenpairs (ascending)
    PCA_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, J, N, &one, Y0, N, &zero, G, J));
    PCA_BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, J, G, J, dD, dE, dinfo));
    int info = 0;
    std::vector<double> lam(Q), V((size_t)J * Q);
    PCA_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
    PCA_HIP(hipMemcpyAsync(lam.data(), dD + (J - Q), (size_t)Q * sizeof(double), hipMemcpyDeviceToHost, st));
    PCA_HIP(hipMemcpyAsync(V.data(), G + (size_t)(J - Q) * J, V.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    PCA_HIP(hipStreamSynchronize(st));
    if (info != 0) { gp_set_error("pca init: eigensolver did not converge"); rc = 1; return done(); }
    // largest first, signed so that the entry of largest magnitude is positive
    std::vector<double> Vs((size_t)J * Q), f(Q);
    for (int k = 0; k < Q; k++) {
        const int src = Q - 1 - k;
        const double lk = lam[src];
        if (!(lk > 0.0)) { gp_set_error("pca init: non-positive eigenvalue among the top Q"); rc = 1; return done(); }
        const double *v = V.data() + (size_t)src * J;
        int am = 0;
        for (int j = 1; j < J; j++)
            if (std::fabs(v[j]) > std::fabs(v[am])) am = j;
        const double sg = v[am] < 0.0 ? -1.0 : 1.0;
        for (int j = 0; j < J; j++) Vs[j + (size_t)k * J] = sg * v[j];
        f[k] = std::sqrt((double)N / lk);
        if (eval) eval[k] = lk;
    }
    PCA_HIP(hipMemcpyAsync(dV, Vs.data(), Vs.size() * sizeof(double), hipMemcpyHostToDevice, st));
    PCA_HIP(hipMemcpyAsync(dF, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, st));
    PCA_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, Q, J, &one, Y0, N, dV, J, &zero,
                           dX, N));
    k_pca_scale<<<(unsigned)(((size_t)N * Q + PT - 1) / PT), PT, 0, st>>>(N, Q, dF, dX);
    PCA_HIP(hipGetLastError());
    PCA_HIP(hipMemcpyAsync(X, dX, (size_t)N * Q * sizeof(double), hipMemcpyDeviceToHost, st));
    PCA_HIP(hipStreamSynchronize(st));
#undef PCA_HIP
#undef PCA_BLAS
    return done();
}
