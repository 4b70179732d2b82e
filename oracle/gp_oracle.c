/*
 * gp_oracle.c -- FP64 CPU oracle for the GASPACHO2 sparse GP-LVM fit.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline; never the
 * product path).  Every function restates one reference routine, cited by
 * file:line relative to the reference's src/ directory, and keeps its loop
 * order so that sums round the same way.  Loops that the reference runs as one
 * GPU thread per gene or per cell run here as OpenMP iterations; each
 * iteration is serial, so results do not depend on the thread count.
 */
#include "gp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_HIST 10

static double *dalloc(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n); abort(); }
    return p;
}

/* ============================ dense primitives ============================ */

int orc_choldc(double *a, int n) {            /* chol.cu:19-43 */
    for (int i = 0; i < n; i++) {
        for (int j = i; j < n; j++) {
            double sum = a[i + (size_t)j * n];
            for (int k = i - 1; k >= 0; k--) sum -= a[i + (size_t)k * n] * a[j + (size_t)k * n];
            if (i == j) {
                if (sum <= 0.0) { memset(a, 0, sizeof(double) * (size_t)n * n); return 1; }
                a[j + (size_t)i * n] = sqrt(sum);
            } else {
                a[j + (size_t)i * n] = sum / a[i + (size_t)i * n];
            }
        }
    }
    return 0;
}

void orc_cholsl(const double *a, int n, const double *b, double *x) {   /* chol.cu:45-56 */
    for (int i = 0; i < n; i++) {
        double sum = b[i];
        for (int k = i - 1; k >= 0; k--) sum -= a[i + (size_t)k * n] * x[k];
        x[i] = sum / a[i + (size_t)i * n];
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = x[i];
        for (int k = i + 1; k < n; k++) sum -= a[k + (size_t)i * n] * x[k];
        x[i] = sum / a[i + (size_t)i * n];
    }
}

void orc_cholslg(const double *a, int n, const double *b, int ldb, double *x, int ldx) { /* chol.cu:86-97 */
    for (int i = 0; i < n; i++) {
        double sum = b[(size_t)i * ldb];
        for (int k = i - 1; k >= 0; k--) sum -= a[i + (size_t)k * n] * x[(size_t)k * ldx];
        x[(size_t)i * ldx] = sum / a[i + (size_t)i * n];
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = x[(size_t)i * ldx];
        for (int k = i + 1; k < n; k++) sum -= a[k + (size_t)i * n] * x[(size_t)k * ldx];
        x[(size_t)i * ldx] = sum / a[i + (size_t)i * n];
    }
}

void orc_getRb(const double *L, double *b, int n) {     /* chol.cu:7-17 */
    for (int j = 0; j < n; j++) {
        double t = 0.0;
        for (int i = j; i < n; i++) t += L[i + (size_t)j * n] * b[i];
        b[j] = t;
    }
}

double orc_quadform(const double *A, const double *b, int n) {   /* util_cuda.cu:442-451 */
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
        sum += A[i + (size_t)i * n] * b[i] * b[i];
        for (int j = i + 1; j < n; j++) sum += 2.0 * A[i + (size_t)j * n] * b[i] * b[j];
    }
    return sum;
}

void orc_mscale(int n, double *x, double m) {            /* util_cuda.cu:546-555 */
    double mh = 0.0;
    for (int i = 0; i < n; i++) mh += x[i];
    mh /= (double)n;
    for (int i = 0; i < n; i++) x[i] *= m / mh;
}

static double fdot(const double *x, int ldx, const double *y, int ldy, int n) { /* util_cuda.cu:166-173 */
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[(size_t)i * ldx] * y[(size_t)i * ldy];
    return s;
}

void orc_lbfgs_two_loop(const double *S, const double *Y, const double *g, int n,
                        int hist, int itr, double *work, double *z) {   /* util_cuda.cu:187-232 */
    int pt = (itr - 1) % hist;
    double *q = work, *a = work + n;
    memcpy(q, g, sizeof(double) * (size_t)n);
    int lo = (itr > hist ? pt - hist + 1 : 0);
    for (int i = pt; i >= lo; i--) {
        int ii = (i + hist) % hist;
        double num = 0.0, den = 0.0;
        for (int k = 0; k < n; k++) {
            num += S[k + (size_t)ii * n] * q[k];
            den += S[k + (size_t)ii * n] * Y[k + (size_t)ii * n];
        }
        a[ii] = num / den;
        for (int k = 0; k < n; k++) q[k] -= a[ii] * Y[k + (size_t)ii * n];
    }
    double num = 0.0, den = 0.0;
    for (int k = 0; k < n; k++) {
        num += S[k + (size_t)pt * n] * Y[k + (size_t)pt * n];
        den += Y[k + (size_t)pt * n] * Y[k + (size_t)pt * n];
    }
    for (int k = 0; k < n; k++) z[k] = (num / den) * q[k];
    for (int i = lo; i <= pt; i++) {
        int ii = (i + hist) % hist;
        double nb = 0.0, db = 0.0;
        for (int k = 0; k < n; k++) {
            nb += Y[k + (size_t)ii * n] * z[k];
            db += S[k + (size_t)ii * n] * Y[k + (size_t)ii * n];
        }
        for (int k = 0; k < n; k++) z[k] += S[k + (size_t)ii * n] * (a[ii] - nb / db);
    }
    for (int k = 0; k < n; k++) {
        if (z[k] > 1.0) z[k] = 1.0;
        if (z[k] < -1.0) z[k] = -1.0;
        z[k] = -z[k];
    }
}

/* ============================ kernel matrices ============================ */

void orc_kernel_mats(int N, int M, int Q, const double *X, const double *T,
                     const double *rho, int ard_only, double *Knm, double *K,
                     double *Kinv) {
    /* kernel_funcs.cu:61-97 (dim 0 periodic) / 221-255 (all dims SE-ARD) */
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < M; j++) {
        for (int i = 0; i < N; i++) {
            double sum = 0.0, t;
            int k0 = 0;
            if (!ard_only) {
                t = sin((X[i] - T[j]) / 2.0);
                sum += t * t / rho[0];
                k0 = 1;
            }
            for (int k = k0; k < Q; k++) {
                t = X[i + (size_t)k * N] - T[j + (size_t)k * M];
                sum += t * t / 2.0 / rho[k];
            }
            Knm[i + (size_t)j * N] = exp(-sum);
        }
        for (int ii = 0; ii < M; ii++) {
            double v;
            if (ii == j) v = 1.0;
            else if (ii > j) v = 0.0;
            else {
                double sum = 0.0, t;
                int k0 = 0;
                if (!ard_only) {
                    t = sin((T[ii] - T[j]) / 2.0);
                    sum += t * t / rho[0];
                    k0 = 1;
                }
                for (int k = k0; k < Q; k++) {
                    t = T[ii + (size_t)k * M] - T[j + (size_t)k * M];
                    sum += t * t / 2.0 / rho[k];
                }
                v = exp(-sum);
            }
            K[ii + (size_t)j * M] = v;
            Kinv[ii + (size_t)j * M] = v;
        }
    }
}

void orc_kinv_kmn(int N, int M, const double *Kchol, const double *Knm,
                  double *KinvKmn, double *titsias) {      /* kernel_funcs.cu:154-162 */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        orc_cholslg(Kchol, M, Knm + i, N, KinvKmn + (size_t)i * M, 1);
        titsias[i] = (1.0 - fdot(Knm + i, N, KinvKmn + (size_t)i * M, 1, M)) / 2.0;
    }
}

void orc_expand_delta(int p, const int *nh, int M, const double *delta,
                      const double *K, double *DeltaInv) {   /* util_cuda.cu:249-269 */
    int P = nh[p];
    memset(DeltaInv, 0, sizeof(double) * (size_t)P * P);
    for (int i = 0; i < p - 1; i++)
        for (int j = nh[i]; j < nh[i + 1]; j++) DeltaInv[j + (size_t)j * P] = 1.0 / delta[i];
    for (int j = 0; j < M; j++) {
        int jj = j + nh[p - 1];
        for (int k = j; k < M; k++) {
            int kk = k + nh[p - 1];
            DeltaInv[jj + (size_t)kk * P] = K[j + (size_t)k * M] / delta[p - 1];
        }
    }
}

/* ============================ DataFrame layout ============================ */

int orc_dataframe_layout(int ncov, const int *c, const int *v, int M, int *nh) {
    /* DataFrame.cu:56-90: variables = intercept, covariates, Knm block */
    int p = ncov + 1 + (M > 0 ? 1 : 0);
    nh[0] = 0;
    nh[1] = 1;
    int NH = 1;
    for (int j = 0; j < ncov; j++) {
        if (c[j] < 1 || v[j] < 1) return -1;
        if (c[j] > 1 && v[j] != 1) return -1;   /* categorical: one int column */
        NH += v[j] * c[j];
        nh[j + 2] = NH;
    }
    if (M > 0) { NH += M; nh[p] = NH; }
    return nh[p];
}

void orc_dataframe_expand(int N, int ncov, const int *c, const int *v,
                          const char *body, const int *nh, double *Zcov) {
    /* getModelMatrix DataFrame.cu:40-50 applied to every covariate column */
    int ncol = nh[ncov + 1];
    for (int i = 0; i < N; i++) Zcov[i] = 1.0;                /* intercept, DataFrame.cu:111 */
    size_t off = 0;
    for (int j = 0; j < ncov; j++) {
        int col0 = nh[j + 1];
        if (c[j] == 1) {
            for (int m = 0; m < v[j]; m++) {
                const char *src = body + off + sizeof(double) * (size_t)N * m;
                for (int i = 0; i < N; i++) {
                    double d; memcpy(&d, src + sizeof(double) * (size_t)i, sizeof(double));
                    Zcov[i + (size_t)(col0 + m) * N] = d;
                }
            }
            off += sizeof(double) * (size_t)N * v[j];
        } else {
            for (int lev = 0; lev < c[j]; lev++)
                for (int i = 0; i < N; i++) {
                    int x; memcpy(&x, body + off + sizeof(int) * (size_t)i, sizeof(int));
                    Zcov[i + (size_t)(col0 + lev) * N] = (x == lev ? 1.0 : 0.0);
                }
            off += sizeof(int) * (size_t)N * v[j];
        }
    }
    (void)ncol;
}

/* ============================ state ============================ */

size_t orc_grad_len(const orc_state *s) {
    return (size_t)s->N + s->p + ((size_t)s->N + s->M + 1) * s->Q;
}

orc_state *orc_create(int N, int J, int M, int Q, int p, const int *nh,
                      const int *Yp, const int *Yi, const double *Yx, int L,
                      const double *Zcov, const double *ls, const double *X0,
                      const double *T0, int ard_only, int nthreads) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->N = N; s->J = J; s->M = M; s->Q = Q; s->p = p; s->L = L;
    s->ard_only = ard_only;
    s->nthreads = nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    s->nh = (int *)calloc((size_t)p + 1, sizeof(int));
    memcpy(s->nh, nh, sizeof(int) * ((size_t)p + 1));
    int P = nh[p];
    s->P = P;
    s->Yp = (int *)calloc((size_t)J + 1, sizeof(int));
    s->Yi = (int *)calloc((size_t)L + 1, sizeof(int));
    s->Yx = dalloc((size_t)L + 1);
    memcpy(s->Yp, Yp, sizeof(int) * ((size_t)J + 1));
    if (L) { memcpy(s->Yi, Yi, sizeof(int) * (size_t)L); memcpy(s->Yx, Yx, sizeof(double) * (size_t)L); }
    /* SparseMat.cu:37-52 allocates L+1 entries; the merge loops read i[L]
     * one past the last entry.  Keep it out of range of any cell index. */
    s->Yi[L] = -1;
    s->Z = dalloc((size_t)N * P);
    memcpy(s->Z, Zcov, sizeof(double) * (size_t)N * nh[p - 1]);
    s->ls = dalloc(N); memcpy(s->ls, ls, sizeof(double) * (size_t)N);
    s->Omega = dalloc(2 * (size_t)N);
    s->delta = dalloc(2 * (size_t)p);
    s->X = dalloc(2 * (size_t)N * Q); memcpy(s->X, X0, sizeof(double) * (size_t)N * Q);
    s->T = dalloc(2 * (size_t)M * Q); memcpy(s->T, T0, sizeof(double) * (size_t)M * Q);
    s->rho = dalloc(2 * (size_t)Q);
    s->phi = dalloc(J);
    s->zeta = dalloc(P);
    s->K = dalloc((size_t)M * M);
    s->Kinv = dalloc((size_t)M * M);
    s->KinvKmn = dalloc((size_t)N * M);
    s->titsias = dalloc(N);
    s->DeltaInv = dalloc((size_t)P * P);
    s->IdenP = dalloc((size_t)P * P);
    s->W = dalloc((size_t)N * (J + 1));
    s->Nu = dalloc((size_t)N * J);
    s->Psi = dalloc(2 * (size_t)J * P * P);
    s->Beta = dalloc((size_t)P * J);
    s->tBeta = dalloc((size_t)P * J);
    s->lkhds = dalloc(J);
    size_t G = orc_grad_len(s);
    s->grad = dalloc(G); s->grad0 = dalloc(G); s->ss = dalloc(G);
    s->SS = dalloc(G * ORC_HIST); s->XX = dalloc(G * ORC_HIST);
    s->gradDeltaTmp = dalloc((size_t)P * J);
    s->gradLogRhoTmp = dalloc((size_t)M * Q);
    s->C1 = dalloc((size_t)N * M);
    s->C2 = dalloc((size_t)M * M);
    s->lkhd = dalloc(50000);
    s->wt = 0.1; s->igap1 = 100.0; s->igb = 100.0;      /* ggplvm.cu:20-22 */
    orc_init_params(s);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    free(s->nh); free(s->Yp); free(s->Yi); free(s->Yx); free(s->Z); free(s->ls);
    free(s->Omega); free(s->delta); free(s->X); free(s->T); free(s->rho);
    free(s->phi); free(s->zeta); free(s->K); free(s->Kinv); free(s->KinvKmn);
    free(s->titsias); free(s->DeltaInv); free(s->IdenP); free(s->W); free(s->Nu);
    free(s->Psi); free(s->Beta); free(s->tBeta); free(s->lkhds); free(s->grad);
    free(s->grad0); free(s->ss); free(s->SS); free(s->XX); free(s->gradDeltaTmp);
    free(s->gradLogRhoTmp); free(s->C1); free(s->C2); free(s->lkhd);
    free(s->tr_obj); free(s->tr_sr); free(s->tr_info);
    free(s);
}

/* ============================ reference kernels ============================ */

void orc_init_params(orc_state *s) {            /* init.cu:27-48 */
    int p = s->p, P = s->P;
    s->delta[0] = 1.5;
    for (int i = 1; i < p - 1; i++) s->delta[i] = 0.01;
    s->delta[p - 1] = 0.3;
    for (int i = 0; i < s->N; i++) s->Omega[i] = 1.0;
    for (int j = 0; j < s->J; j++) s->phi[j] = 1.0;
    s->rho[0] = 5.0;
    for (int k = 1; k < s->Q; k++) s->rho[k] = 10.0;
    memset(s->IdenP, 0, sizeof(double) * (size_t)P * P);
    for (int i = 0; i < P; i++) s->IdenP[i + (size_t)i * P] = 1.0;
}

void orc_get_kernel_mats(orc_state *s) {        /* kernel_funcs.cu:22-24 */
    orc_kernel_mats(s->N, s->M, s->Q, s->X, s->T, s->rho, s->ard_only,
                    s->Z + (size_t)s->nh[s->p - 1] * s->N, s->K, s->Kinv);
}

int orc_chol_kmm(orc_state *s) { return orc_choldc(s->Kinv, s->M); }   /* kernel_funcs.cu:139-143 */

void orc_get_kinv_kmn(orc_state *s) {           /* kernel_funcs.cu:154-162 */
    orc_kinv_kmn(s->N, s->M, s->Kinv, s->Z + (size_t)s->nh[s->p - 1] * s->N, s->KinvKmn, s->titsias);
}

void orc_expand_delta_k(orc_state *s) {         /* util_cuda.cu:249-269 */
    orc_expand_delta(s->p, s->nh, s->M, s->delta, s->K, s->DeltaInv);
}

int orc_chol_delta_inv(orc_state *s) { return orc_choldc(s->DeltaInv, s->P); }  /* ggplvm.cu:695-700 */

void orc_init_psi(orc_state *s) {               /* init.cu:125-191 (unweighted Gram + chol) */
    int N = s->N, P = s->P;
    double *Psi = s->Psi;
    #pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < P; l++) {
        for (int k = 0; k < P; k++) {
            if (k <= l) {
                const double *a = s->Z + (size_t)k * N, *b = s->Z + (size_t)l * N;
                double sum = 0.0;
                for (int i = 0; i < N; i++) sum += a[i] * b[i];
                Psi[k + (size_t)l * P] = sum + s->DeltaInv[k + (size_t)l * P];
            } else {
                Psi[k + (size_t)l * P] = 0.0;
            }
        }
    }
    orc_choldc(Psi, P);
}

void orc_init_beta(orc_state *s) {              /* init.cu:193-221 */
    int N = s->N, P = s->P;
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        double *tmp = dalloc(P);
        int l = s->Yp[j], lmax = s->Yp[j + 1];
        for (int i = 0; i < N; i++) {
            double yi;
            if (l < lmax && i == s->Yi[l]) { yi = log((s->Yx[l] + 0.05) / s->ls[i]); l++; }
            else yi = log(0.05 / s->ls[i]);
            for (int m = 0; m < P; m++) tmp[m] += s->Z[i + (size_t)m * N] * yi;
        }
        orc_cholsl(s->Psi, P, tmp, s->Beta + (size_t)j * P);
        free(tmp);
    }
}

void orc_get_w_for_beta(orc_state *s) {         /* ggplvm.cu:468-514 */
    int N = s->N, P = s->P, p = s->p;
    double dt = s->delta[p - 1];
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        const double *bj = s->Beta + (size_t)j * P;
        double *Wj = s->W + (size_t)j * N, *nj = s->Nu + (size_t)j * N;
        int l = s->Yp[j], lmax = s->Yp[j + 1];
        for (int i = 0; i < N; i++) {
            double sum = 0.0, sum2 = 0.0;
            for (int k = 0; k < P; k++) {
                double z = s->Z[i + (size_t)k * N];
                sum += z * bj[k];
                sum2 += z * (bj[k] - s->zeta[k]);
            }
            double logW = sum + log(s->ls[i]) + dt * s->titsias[i] * s->wt;
            Wj[i] = exp(sum) * s->ls[i] * exp(dt * s->titsias[i] * s->wt);
            if (l < lmax && i == s->Yi[l]) { nj[i] = sum2 + exp(log(s->Yx[l]) - logW) - 1.0; l++; }
            else nj[i] = sum2 - 1.0;
        }
    }
}

/* ggplvm.cu:621-680.  Indicator columns are stored densely (0/1), so the
 * numeric x categorical branches reduce to the numeric formula with exact
 * multiplications by 1.0; products are formed in the reference's order
 * ((z_k * z_l) * W) / Omega. */
static void get_psi_generic(orc_state *s, double *Psi) {
    int N = s->N, P = s->P, J = s->J;
    #pragma omp parallel for schedule(dynamic)
    for (int jl = 0; jl < J * P; jl++) {
        int j = jl / P, l = jl % P;
        const double *Wj = s->W + (size_t)j * N;
        const double *b = s->Z + (size_t)l * N;
        double *Pj = Psi + (size_t)j * P * P;
        for (int k = 0; k < P; k++) {
            if (k <= l) {
                const double *a = s->Z + (size_t)k * N;
                double sum = 0.0;
                for (int i = 0; i < N; i++) sum += a[i] * b[i] * Wj[i] / s->Omega[i];
                Pj[k + (size_t)l * P] = sum + s->DeltaInv[k + (size_t)l * P];
            } else {
                Pj[k + (size_t)l * P] = 0.0;
            }
        }
    }
}

void orc_get_psi(orc_state *s) { get_psi_generic(s, s->Psi); }

int orc_chol_psi(orc_state *s) {                 /* ggplvm.cu:683-692 */
    int P = s->P, bad = 0;
    #pragma omp parallel for schedule(dynamic) reduction(+:bad)
    for (int j = 0; j < s->J; j++) bad += orc_choldc(s->Psi + (size_t)j * P * P, P);
    return bad;
}

/* util_cuda.cu:462-487 getZtOinvjnj (dense model matrix) */
static void zt_oinv_nu(const orc_state *s, const double *Wj, const double *nj, double *b) {
    int N = s->N, P = s->P;
    for (int k = 0; k < P; k++) {
        const double *z = s->Z + (size_t)k * N;
        double sum = 0.0;
        for (int i = 0; i < N; i++) sum += z[i] * Wj[i] / s->Omega[i] * nj[i];
        b[k] = sum;
    }
}

/* ggplvm_de.cu:491-532 getScoreStats, for a list of tests: test t pairs gene
 * gene[t] with genotype column snp[t] of G (N x S, column-major, doubles).
 *   tmp  = Z^T diag(W_j/Omega) g,  tmp2 = Psi_j^{-1} tmp
 *   gVg  = sum_i g_i^2 W_ij/Omega_i - tmp^T tmp2
 *   ss   = sum_i g_i (y_ij - W_ij)/Omega_i   (y_ij = 0 off the non-zeros)
 *   stat = ss / sqrt(phi_j gVg)
 * Uses the current W, Omega, phi and the factors in Psi. */
void orc_score_stats(orc_state *s, int S, const double *G, int T, const int *gene,
                     const int *snp, double *stat) {
    int N = s->N, P = s->P;
    (void)S;
    #pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * P);
        double *tmp2 = tmp + P;
        #pragma omp for schedule(dynamic)
        for (int t = 0; t < T; t++) {
            const int j = gene[t];
            const double *Wj = s->W + (size_t)j * N;
            const double *g = G + (size_t)snp[t] * N;
            zt_oinv_nu(s, Wj, g, tmp);
            orc_cholsl(s->Psi + (size_t)j * P * P, P, tmp, tmp2);
            double gvg = 0.0;
            for (int i = 0; i < N; i++) gvg += g[i] * g[i] * Wj[i] / s->Omega[i];
            for (int k = 0; k < P; k++) gvg -= tmp[k] * tmp2[k];
            int l = s->Yp[j];
            const int lmax = s->Yp[j + 1];
            double ss = 0.0;
            for (int i = 0; i < N; i++) {
                if (l < lmax && s->Yi[l] == i) {
                    ss += g[i] / s->Omega[i] * (s->Yx[l] - Wj[i]);
                    l++;
                } else {
                    ss += g[i] * Wj[i] / s->Omega[i] * (-1.0);
                }
            }
            stat[t] = ss / sqrt(s->phi[j] * gvg);
        }
        free(tmp);
    }
}

void orc_get_beta_phi(orc_state *s, int jtr) {   /* ggplvm.cu:418-464 */
    int N = s->N, P = s->P;
    double r = (double)(jtr + 1) / 100.0;
    if (jtr > 100) r = 1.0;
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        double *b = dalloc(P), *tmp = dalloc(P);
        const double *Wj = s->W + (size_t)j * N, *nj = s->Nu + (size_t)j * N;
        const double *Pj = s->Psi + (size_t)j * P * P;
        double *Bj = s->Beta + (size_t)j * P;
        double phij = 0.0;
        zt_oinv_nu(s, Wj, nj, b);
        for (int i = 0; i < N; i++) phij += nj[i] * nj[i] * Wj[i] / s->Omega[i];
        orc_cholsl(Pj, P, b, tmp);
        for (int k = 0; k < P; k++) {
            Bj[k] = Bj[k] * (1 - r) + tmp[k] * r;
            tmp[k] = Bj[k];
            Bj[k] += s->zeta[k];
        }
        orc_getRb(Pj, tmp, P);
        for (int k = 0; k < P; k++) phij -= tmp[k] * tmp[k];
        double ph = phij / (double)N;
        if (isnan(ph)) ph = 1.0;
        if (ph < 0.0) ph = 1.0;
        if (ph < 0.01) ph = 0.01;
        if (ph > 1000.0) ph = 1000.0;
        s->phi[j] = ph;
        free(b); free(tmp);
    }
}

void orc_get_lkhd_pois(orc_state *s) {          /* ggplvm.cu:703-724 */
    int N = s->N, p = s->p;
    double dt = s->delta[p - 1];
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        const double *Wj = s->W + (size_t)j * N;
        int l = s->Yp[j], lmax = s->Yp[j + 1];
        double sum = 0.0;
        for (int i = 0; i < N; i++) {
            if (l < lmax && s->Yi[l] == i) {
                sum += (log(Wj[i]) - dt * s->titsias[i] * s->wt) * s->Yx[l] - Wj[i];
                l++;
            } else {
                sum -= Wj[i];
            }
        }
        s->lkhds[j] = sum;
    }
}

double orc_collect_lkhd(orc_state *s, int Nterm, int jtr) {   /* ggplvm.cu:788-815 */
    int p = s->p;
    const double *Oinvt = s->W + (size_t)Nterm * s->J;
    double sum = 0.0;
    for (int k = 0; k < s->J; k++) sum += s->lkhds[k];
    for (int i = 0; i < Nterm; i++) sum -= s->wt * s->delta[p - 1] * Oinvt[i] * s->titsias[i];
    for (int i = 0; i < Nterm; i++) sum = sum - s->igb / s->Omega[i] - s->igap1 * log(s->Omega[i]);
    s->lkhd[jtr] = sum;
    return sum;
}

void orc_get_oinvt(orc_state *s) {              /* ggplvm.cu:284-294 */
    int N = s->N, J = s->J;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double o = 0.0;
        for (int j = 0; j < J; j++) o += s->W[i + (size_t)j * N] / s->Omega[i];
        s->W[(size_t)N * J + i] = o;
    }
}

void orc_get_lkhd(orc_state *s) {               /* ggplvm.cu:744-784 (getLkhdj) */
    int N = s->N, P = s->P;
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        const double *Wj = s->W + (size_t)j * N, *nj = s->Nu + (size_t)j * N;
        const double *Pj = s->Psi + (size_t)j * P * P;
        double phij = s->phi[j];
        double *tmp = dalloc(P), *tmp2 = dalloc(P);
        double lk = -((double)N) / 2.0 * log(phij);
        for (int i = 0; i < N; i++) lk -= log(s->Omega[i] / Wj[i]) / 2.0;
        for (int k = 0; k < P; k++) {
            lk -= log(Pj[k + (size_t)k * P]);
            lk += log(s->DeltaInv[k + (size_t)k * P]);
        }
        zt_oinv_nu(s, Wj, nj, tmp);
        for (int i = 0; i < N; i++) {
            double o = Wj[i] / s->Omega[i];
            lk -= nj[i] * nj[i] * o / 2.0 / phij;
        }
        orc_cholsl(Pj, P, tmp, tmp2);
        for (int k = 0; k < P; k++) {
            lk += tmp[k] * tmp2[k] / 2.0 / phij;
            if (isnan(lk)) break;
        }
        s->lkhds[j] = lk;
        free(tmp); free(tmp2);
    }
}

void orc_get_tilde_beta(orc_state *s) {         /* ggplvm.cu:406-414 */
    int N = s->N, P = s->P;
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < s->J; j++) {
        double *b = dalloc(P);
        zt_oinv_nu(s, s->W + (size_t)j * N, s->Nu + (size_t)j * N, b);
        orc_cholsl(s->Psi + (size_t)j * P * P, P, b, s->tBeta + (size_t)j * P);
        free(b);
    }
}

void orc_grad_log_delta(orc_state *s) {         /* ggplvm.cu:210-281 */
    int N = s->N, P = s->P, p = s->p, M = s->M, J = s->J;
    const int *nh = s->nh;
    #pragma omp parallel for schedule(dynamic)
    for (int jk = 0; jk < J * P; jk++) {
        int j = jk / P, k = jk % P;
        double *ek = dalloc(P), *psijk = dalloc(P);
        const double *tbj = s->tBeta + (size_t)j * P;
        double dk = 0.0, sum = 0.0;
        for (int l = 0; l < p; l++)
            if (nh[l] <= k && k < nh[l + 1]) { dk = s->delta[l]; break; }
        if (k < nh[p - 1]) {
            ek[k] = 1.0;
        } else {
            int kk = k - nh[p - 1];
            for (int l = 0; l < kk; l++) ek[l + nh[p - 1]] = s->K[l + (size_t)kk * M];
            for (int l = kk; l < M; l++) ek[l + nh[p - 1]] = s->K[kk + (size_t)l * M];
        }
        orc_cholsl(s->Psi + (size_t)j * P * P, P, ek, psijk);
        sum -= (dk - psijk[k]) / dk / dk / 2.0;
        if (k < nh[p - 1]) {
            sum += tbj[k] * tbj[k] / dk / dk / s->phi[j] / 2.0;
        } else if (k == nh[p - 1]) {
            sum += orc_quadform(s->K, tbj + nh[p - 1], M) / dk / dk / s->phi[j] / 2.0;
        }
        s->gradDeltaTmp[k + (size_t)j * P] = sum * dk;
        free(ek); free(psijk);
    }
    /* collectGradLogDelta ggplvm.cu:256-281 */
    double *gd = s->grad + N;
    const double *Oinvt = s->W + (size_t)N * J;
    for (int l = 0; l < p; l++) {
        double sum = 0.0;
        for (int j = 0; j < J; j++)
            for (int k = nh[l]; k < nh[l + 1]; k++) sum += s->gradDeltaTmp[k + (size_t)j * P];
        if (l == p - 1)
            for (int i = 0; i < N; i++) sum -= s->wt * Oinvt[i] * s->titsias[i] * s->delta[p - 1];
        gd[l] = sum;
    }
}

void orc_grad_log_omega(orc_state *s) {         /* ggplvm.cu:301-376 */
    int N = s->N, P = s->P, p = s->p, M = s->M, J = s->J, Q = s->Q;
    int g0 = s->nh[p - 1];
    double *gO = s->grad, *gX = s->grad + N + p;
    const double *Knm = s->Z + (size_t)g0 * N;
    double dt = s->delta[p - 1];
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
        double *zi = dalloc(P), *u = dalloc(P), *ci = dalloc(M);
        for (int k = 0; k < P; k++) zi[k] = s->Z[i + (size_t)k * N];
        double sum = 0.0, Oii = 0.0, Om = s->Omega[i];
        for (int j = 0; j < J; j++) {
            const double *tbj = s->tBeta + (size_t)j * P;
            double Wij = s->W[i + (size_t)j * N];
            double o = Wij / Om;
            Oii += o;
            orc_cholsl(s->Psi + (size_t)j * P * P, P, zi, u);
            double tn = o * s->Nu[i + (size_t)j * N] - o * fdot(zi, 1, tbj, 1, P);
            sum -= (1.0 / Om - fdot(zi, 1, u, 1, P) / Om * o) / 2.0;
            sum += tn * tn / Wij / s->phi[j] / 2.0;
            for (int k = 0; k < M; k++) ci[k] += (-u[k + g0] * o + tn * tbj[k + g0] / s->phi[j]);
        }
        sum += s->wt * dt * Oii / Om * s->titsias[i];
        gO[i] = sum * Om + s->igb / Om - s->igap1;
        s->W[(size_t)N * J + i] = Oii;
        for (int k = 0; k < M; k++) {
            ci[k] += s->wt * dt * Oii * s->KinvKmn[k + (size_t)i * M];
            ci[k] *= Knm[i + (size_t)k * N];
        }
        double ci1 = 0.0;
        gX[i] = 0.0;
        if (!s->ard_only) {
            for (int k = 0; k < M; k++) {
                ci1 += ci[k];
                double t = (s->X[i] - s->T[k]) / 2.0;
                gX[i] -= ci[k] * sin(t) * cos(t) / s->rho[0];
            }
        } else {
            for (int k = 0; k < M; k++) ci1 += ci[k];
        }
        for (int q = (s->ard_only ? 0 : 1); q < Q; q++)
            gX[i + (size_t)q * N] = (fdot(ci, 1, s->T + (size_t)q * M, 1, M) - ci1 * s->X[i + (size_t)q * N]) / s->rho[q];
        for (int k = 0; k < M; k++) s->C1[i + (size_t)k * N] = ci[k];
        free(zi); free(u); free(ci);
    }
}

void orc_solve_psi(orc_state *s) {              /* ggplvm.cu:46-53 */
    int P = s->P, J = s->J;
    #pragma omp parallel for schedule(dynamic)
    for (int jk = 0; jk < J * P; jk++) {
        int j = jk / P, k = jk % P;
        orc_cholsl(s->Psi + (size_t)j * P * P, P, s->IdenP + (size_t)k * P,
                   s->Psi + (size_t)J * P * P + (size_t)j * P * P + (size_t)k * P);
    }
}

void orc_get_c2(orc_state *s) {                 /* ggplvm.cu:57-100 */
    int N = s->N, P = s->P, p = s->p, M = s->M, J = s->J;
    int g0 = s->nh[p - 1];
    const double *Oinvt = s->W + (size_t)N * J;
    double dt = s->delta[p - 1];
    #pragma omp parallel for schedule(dynamic)
    for (int jk = 0; jk < M * M; jk++) {
        int j = jk % M, k = jk / M;
        if (j > k) continue;
        double *ev = dalloc(M), *ev2 = dalloc(M);
        double sum = 0.0, c;
        for (int i = 0; i < N; i++)
            sum += s->KinvKmn[j + (size_t)i * M] * s->KinvKmn[k + (size_t)i * M] * Oinvt[i];
        c = -s->wt * dt * sum / 2.0;
        sum = 0.0;
        for (int i = 0; i < J; i++)
            sum += s->tBeta[j + g0 + (size_t)i * P] * s->tBeta[k + g0 + (size_t)i * P] / s->phi[i];
        c -= sum / dt / 2.0;
        sum = 0.0;
        for (int i = 0; i < J; i++)
            sum += s->Psi[(size_t)J * P * P + (size_t)i * P * P + (size_t)(k + g0) * P + (j + g0)];
        c -= sum / dt / 2.0;
        ev[k] = 1.0;
        orc_cholsl(s->Kinv, M, ev, ev2);
        c += ((double)J) * ev2[j] / 2.0;
        c *= s->K[j + (size_t)k * M];
        s->C2[j + (size_t)k * M] = c;
        if (j < k) s->C2[k + (size_t)j * M] = c;
        free(ev); free(ev2);
    }
}

void orc_grad_tau(orc_state *s) {               /* ggplvm.cu:103-145 */
    int N = s->N, p = s->p, M = s->M, Q = s->Q;
    double *gT = s->grad + N + p + (size_t)N * Q;
    const double *X = s->X, *T = s->T, *rho = s->rho;
    #pragma omp parallel for schedule(dynamic)
    for (int mq = 0; mq < M * Q; mq++) {
        int m = mq % M, q = mq / M;
        double sum, cm1;
        if (q == 0 && !s->ard_only) {
            sum = 0.0;
            for (int i = 0; i < N; i++) {
                double t = (X[i] - T[m]) / 2.0;
                sum += s->C1[i + (size_t)m * N] * sin(t) * cos(t);
            }
            double g = sum / rho[0];
            sum = 0.0;
            for (int i = 0; i < M; i++) {
                double t = (T[i] - T[m]) / 2.0;
                sum += s->C2[m + (size_t)i * M] * sin(t) * cos(t);
            }
            gT[m] = g + 2.0 * sum / rho[0];
        } else {
            sum = 0.0; cm1 = 0.0;
            for (int i = 0; i < N; i++) {
                sum += s->C1[i + (size_t)m * N] * X[i + (size_t)q * N];
                cm1 += s->C1[i + (size_t)m * N];
            }
            double g = (sum - cm1 * T[m + (size_t)q * M]) / rho[q];
            sum = 0.0; cm1 = 0.0;
            for (int i = 0; i < M; i++) {
                sum += s->C2[m + (size_t)i * M] * T[i + (size_t)q * M];
                cm1 += s->C2[m + (size_t)i * M];
            }
            gT[m + (size_t)q * M] = g + 2.0 * (sum - cm1 * T[m + (size_t)q * M]) / rho[q];
        }
    }
}

void orc_grad_log_rho(orc_state *s) {           /* ggplvm.cu:148-205 */
    int N = s->N, p = s->p, M = s->M, Q = s->Q;
    const double *X = s->X, *T = s->T, *rho = s->rho;
    #pragma omp parallel for schedule(dynamic)
    for (int jq = 0; jq < M * Q; jq++) {
        int j = jq % M, q = jq / M;
        double sum = 0.0, t;
        if (q == 0 && !s->ard_only) {
            for (int i = 0; i < N; i++) {
                t = sin((X[i + (size_t)q * N] - T[j + (size_t)q * M]) / 2.0);
                sum += s->C1[i + (size_t)j * N] * t * t / rho[q] / rho[q];
            }
            for (int i = 0; i < M; i++) {
                t = sin((T[i + (size_t)q * M] - T[j + (size_t)q * M]) / 2.0);
                sum += s->C2[i + (size_t)j * M] * t * t / rho[q] / rho[q];
            }
        } else {
            for (int i = 0; i < N; i++) {
                t = X[i + (size_t)q * N] - T[j + (size_t)q * M];
                sum += s->C1[i + (size_t)j * N] * t * t / rho[q] / rho[q] / 2.0;
            }
            for (int i = 0; i < M; i++) {
                t = T[i + (size_t)q * M] - T[j + (size_t)q * M];
                sum += s->C2[i + (size_t)j * M] * t * t / rho[q] / rho[q] / 2.0;
            }
        }
        s->gradLogRhoTmp[q + (size_t)j * Q] = sum * rho[q];
    }
    double *gr = s->grad + N + p + (size_t)N * Q + (size_t)M * Q;
    for (int q = 0; q < Q; q++) {
        double sum = 0.0;
        for (int k = 0; k < M; k++) sum += s->gradLogRhoTmp[q + (size_t)k * Q];
        gr[q] = sum;
    }
}

void orc_lbfgs(orc_state *s, int itr, int rhoflag) {   /* ggplvm.cu:865-934 */
    int N = s->N, M = s->M, p = s->p, Q = s->Q;
    int pt = (itr - 1) % ORC_HIST, ptp1 = itr % ORC_HIST;
    size_t tot = (size_t)N + p + (size_t)N * Q + (size_t)M * Q + rhoflag;
    double *g = s->grad, *ss = s->ss;
    if (itr == 0) {
        for (size_t k = 0; k < tot; k++) {
            ss[k] = -g[k] / 100000.0;
            if (ss[k] > 1.0) ss[k] = 1.0;
            if (ss[k] < -1.0) ss[k] = -1.0;
            s->grad0[k] = g[k];
        }
    } else {
        for (size_t k = 0; k < tot; k++) {
            s->XX[k + (size_t)pt * tot] = g[k] - s->grad0[k];
            s->grad0[k] = g[k];
        }
        double *work = dalloc(tot + ORC_HIST);
        orc_lbfgs_two_loop(s->SS, s->XX, g, (int)tot, ORC_HIST, itr, work, ss);
        free(work);
    }
    for (int k = 0; k < N; k++) {
        s->Omega[k + N] = s->Omega[k];
        s->Omega[k] = exp(log(s->Omega[k]) - ss[k]);
        if (s->Omega[k] > 1000.) s->Omega[k] = 1000.;
        if (s->Omega[k] < .001) s->Omega[k] = .001;
    }
    for (int k = 0; k < p; k++) {
        s->delta[k + p] = s->delta[k];
        s->delta[k] = exp(log(s->delta[k]) - ss[N + k]);
    }
    for (int i = 0; i < N; i++)
        for (int k = 0; k < Q; k++) {
            size_t a = i + (size_t)k * N;
            s->X[a + (size_t)N * Q] = s->X[a];
            s->X[a] = s->X[a] - ss[N + p + a];
        }
    for (int i = 0; i < M; i++)
        for (int k = 0; k < Q; k++) {
            size_t a = i + (size_t)k * M;
            s->T[a + (size_t)M * Q] = s->T[a];
            s->T[a] = s->T[a] - ss[N + p + (size_t)N * Q + a];
        }
    s->rho[Q] = s->rho[0];
    s->rho[0] = exp(log(s->rho[0]) - ss[N + p + (size_t)N * Q + (size_t)M * Q] * (double)rhoflag);
    orc_mscale(Q - 1, s->rho + 1, 10.0);
    for (size_t k = 0; k < tot; k++) s->SS[k + (size_t)ptp1 * tot] = ss[k];
}

void orc_downdate(orc_state *s, int itr, double r, int rhoflag) {   /* ggplvm.cu:818-859 */
    int N = s->N, M = s->M, p = s->p, Q = s->Q;
    size_t tot = (size_t)N + p + (size_t)N * Q + (size_t)M * Q + rhoflag;
    int ptp1 = itr % ORC_HIST;
    double *ss = s->ss;
    for (int k = 0; k < N; k++) {
        s->Omega[k] = exp(log(s->Omega[k + N]) - r * ss[k]);
        if (s->Omega[k] > 1000.) s->Omega[k] = 1000.;
        if (s->Omega[k] < .001) s->Omega[k] = .001;
    }
    for (int k = 0; k < p; k++) s->delta[k] = exp(log(s->delta[k + p]) - r * ss[N + k]);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < Q; k++) {
            size_t a = i + (size_t)k * N;
            s->X[a] = s->X[a + (size_t)N * Q] - r * ss[N + p + a];
        }
    for (int i = 0; i < M; i++)
        for (int k = 0; k < Q; k++) {
            size_t a = i + (size_t)k * M;
            s->T[a] = s->T[a + (size_t)M * Q] - r * ss[N + p + (size_t)N * Q + a];
        }
    s->rho[0] = exp(log(s->rho[Q]) - r * ss[N + p + (size_t)N * Q + (size_t)M * Q] * (double)rhoflag);
    orc_mscale(Q - 1, s->rho + 1, 10.0);
    for (size_t k = 0; k < tot; k++) s->SS[k + (size_t)ptp1 * tot] = r * ss[k];
}

void orc_beta_from_tilde_beta(orc_state *s) {   /* ggplvm.cu:394-402 */
    size_t n = (size_t)s->P * s->J;
    for (size_t a = 0; a < n; a++) s->Beta[a] = s->tBeta[a] + s->zeta[a % s->P];
}

/* ============================ launch groups ============================ */

void orc_init_fit(orc_state *s) {               /* ggplvm.cu:1111-1137 */
    orc_get_kernel_mats(s);
    orc_chol_kmm(s);
    orc_get_kinv_kmn(s);
    orc_expand_delta_k(s);
    orc_init_psi(s);
    orc_init_beta(s);
}

double orc_irls_step(orc_state *s, int itr, int jtr) {   /* ggplvm.cu:1156-1180 */
    (void)itr;
    orc_get_w_for_beta(s);
    orc_get_psi(s);
    orc_chol_psi(s);
    orc_get_beta_phi(s, jtr);
    orc_get_lkhd_pois(s);
    return orc_collect_lkhd(s, 0, jtr);
}

double orc_ql_start(orc_state *s, int itr) {    /* ggplvm.cu:1205-1211 */
    (void)itr;
    orc_get_oinvt(s);
    orc_chol_delta_inv(s);
    orc_get_lkhd(s);
    return orc_collect_lkhd(s, s->N, 0);
}

double orc_elbo(orc_state *s, int itr, int slot) {   /* ggplvm.cu:1259-1277 */
    (void)itr;
    orc_get_oinvt(s);
    orc_get_kernel_mats(s);
    orc_expand_delta_k(s);
    orc_chol_kmm(s);
    orc_get_kinv_kmn(s);
    orc_get_psi(s);
    orc_chol_psi(s);
    orc_chol_delta_inv(s);
    orc_get_lkhd(s);
    return orc_collect_lkhd(s, s->N, slot);
}

void orc_grad(orc_state *s) {                   /* ggplvm.cu:1215-1240 */
    orc_expand_delta_k(s);
    orc_get_oinvt(s);
    orc_get_tilde_beta(s);
    orc_grad_log_delta(s);
    orc_grad_log_omega(s);
    orc_solve_psi(s);
    orc_get_c2(s);
    orc_grad_tau(s);
    orc_grad_log_rho(s);
}

static void tr_push(orc_state *s, int kind, int itr, int jtr, int ktr, double obj, double sr) {
    if (s->tr_len == s->tr_cap) {
        s->tr_cap = s->tr_cap ? 2 * s->tr_cap : 1024;
        s->tr_obj = (double *)realloc(s->tr_obj, sizeof(double) * (size_t)s->tr_cap);
        s->tr_sr = (double *)realloc(s->tr_sr, sizeof(double) * (size_t)s->tr_cap);
        s->tr_info = (int *)realloc(s->tr_info, sizeof(int) * 4 * (size_t)s->tr_cap);
    }
    int *e = s->tr_info + 4 * (size_t)s->tr_len;
    e[0] = kind; e[1] = itr; e[2] = jtr; e[3] = ktr;
    s->tr_obj[s->tr_len] = obj;
    s->tr_sr[s->tr_len] = sr;
    s->tr_len++;
}

int orc_fit_trace(const orc_state *s, double *obj, double *sr, int *info, int cap) {
    int n = s->tr_len < cap ? s->tr_len : cap;
    for (int k = 0; k < n; k++) {
        if (obj) obj[k] = s->tr_obj[k];
        if (sr) sr[k] = s->tr_sr[k];
        if (info) memcpy(info + 4 * (size_t)k, s->tr_info + 4 * (size_t)k, 4 * sizeof(int));
    }
    return s->tr_len;
}

double orc_fit(orc_state *s, int itrmax, int jtr_cap, int verbose) {   /* ggplvm.cu:1146-1330 */
    int rhoflag = 0, JTRMAX;
    double last = 0.0;
    s->tr_len = 0;
    orc_init_fit(s);
    for (int itr = 0; itr < itrmax; itr++) {
        JTRMAX = (itr == 0 ? 200 : 100);
        if (jtr_cap > 0 && JTRMAX > jtr_cap) JTRMAX = jtr_cap;
        for (int jtr = 0; jtr < JTRMAX; jtr++) {
            double lk = orc_irls_step(s, itr, jtr);
            tr_push(s, 0, itr, jtr, 0, lk, 0.0);
            if (verbose) printf("lkhd[%d,%d]=%lf\n", itr, jtr, lk);
            if (jtr > 0) {
                if (isnan(s->lkhd[jtr])) return NAN;
                if (jtr == JTRMAX - 1 || fabs(s->lkhd[jtr - 1] - s->lkhd[jtr]) / fabs(s->lkhd[jtr]) < 1e-7) break;
            }
        }
        last = orc_ql_start(s, itr);
        tr_push(s, 1, itr, 0, 0, last, 0.0);
        if (verbose) printf("lkhd[%d,0]=%lf\n", itr, last);
        if (itr > 20) rhoflag = 1;
        if (itr == 30) JTRMAX = 10000;
        if (jtr_cap > 0 && JTRMAX > jtr_cap) JTRMAX = jtr_cap;
        for (int jtr = 0; jtr < JTRMAX; jtr++) {
            orc_grad(s);
            orc_lbfgs(s, jtr, rhoflag);
            double sr = 1.0;
            int ktr;
            for (ktr = 1; ktr < 8; ktr++) {
                double lk = orc_elbo(s, itr, jtr + 1);
                if (verbose) printf("lkhd[%d,%d]=%lf\n", itr, jtr + 1, lk);
                if (s->lkhd[jtr + 1] > s->lkhd[jtr] || ktr == 7) break;
                else if (isnan(s->lkhd[jtr + 1]) || isinf(s->lkhd[jtr + 1])) sr *= 0.01;
                else if (s->lkhd[jtr + 1] < s->lkhd[jtr]) sr *= 0.1;
                orc_downdate(s, jtr, sr, rhoflag);
            }
            last = s->lkhd[jtr + 1];
            tr_push(s, 2, itr, jtr, ktr, last, sr);
            if (jtr > 0 && fabs(s->lkhd[jtr + 1] - s->lkhd[jtr]) < fabs(s->lkhd[jtr + 1]) * 1e-8) break;
        }
        orc_expand_delta_k(s);
        orc_beta_from_tilde_beta(s);
    }
    return last;
}

/* ============================ initial inducing points ============================ */

typedef struct { uint32_t mt[624]; int mti; } mt_t;

static void mt_seed(mt_t *g, uint32_t s) {      /* MT19937 init_genrand (2002) */
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->mti = 624;
}

static uint32_t mt_next(mt_t *g) {
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    if (g->mti >= 624) {
        int k;
        uint32_t y;
        for (k = 0; k < 624 - 397; k++) {
            y = (g->mt[k] & 0x80000000u) | (g->mt[k + 1] & 0x7fffffffu);
            g->mt[k] = g->mt[k + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; k < 623; k++) {
            y = (g->mt[k] & 0x80000000u) | (g->mt[k + 1] & 0x7fffffffu);
            g->mt[k] = g->mt[k + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (g->mt[623] & 0x80000000u) | (g->mt[0] & 0x7fffffffu);
        g->mt[623] = g->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

uint32_t orc_mt19937_kat(uint32_t seed, int n) {
    mt_t g; mt_seed(&g, seed);
    uint32_t y = 0;
    for (int i = 0; i < n; i++) y = mt_next(&g);
    return y;
}

static double mt_uniform_pos(mt_t *g) {         /* GSL: x / 2^32, rejecting 0 */
    double x;
    do { x = mt_next(g) / 4294967296.0; } while (x == 0.0);
    return x;
}

static double gsl_gaussian_polar(mt_t *g, double sigma) {   /* GSL polar Box-Muller */
    double x, y, r2;
    do {
        x = -1 + 2 * mt_uniform_pos(g);
        y = -1 + 2 * mt_uniform_pos(g);
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0);
    return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

void orc_init_ta_with_per(int M, int Q, double *T) {   /* util.c:73-87 */
    mt_t g;
    mt_seed(&g, 4357u);                          /* GSL: seed 0 selects 4357 */
    for (int i = 0; i < M; i++) T[i] = ((double)(i + 1) / (double)(M + 1)) * 2.0 * 3.14159265358979;
    for (int k = 1; k < Q; k++)
        for (int i = 0; i < M; i++) T[i + (size_t)k * M] = gsl_gaussian_polar(&g, 1.0);
}

/* ============================ buffer access for tests ============================ */

double *orc_buf(orc_state *s, const char *name, size_t *len) {
    size_t N = s->N, J = s->J, M = s->M, Q = s->Q, P = s->P, p = s->p;
    size_t G = orc_grad_len(s);
#define ORC_B(nm, ptr, n) if (!strcmp(name, nm)) { *len = (n); return (ptr); }
    ORC_B("Z", s->Z, N * P)
    ORC_B("ls", s->ls, N)
    ORC_B("Omega", s->Omega, 2 * N)
    ORC_B("delta", s->delta, 2 * p)
    ORC_B("X", s->X, 2 * N * Q)
    ORC_B("T", s->T, 2 * M * Q)
    ORC_B("rho", s->rho, 2 * Q)
    ORC_B("phi", s->phi, J)
    ORC_B("zeta", s->zeta, P)
    ORC_B("K", s->K, M * M)
    ORC_B("Kinv", s->Kinv, M * M)
    ORC_B("KinvKmn", s->KinvKmn, N * M)
    ORC_B("titsias", s->titsias, N)
    ORC_B("DeltaInv", s->DeltaInv, P * P)
    ORC_B("W", s->W, N * (J + 1))
    ORC_B("Nu", s->Nu, N * J)
    ORC_B("Psi", s->Psi, 2 * J * P * P)
    ORC_B("Beta", s->Beta, P * J)
    ORC_B("tBeta", s->tBeta, P * J)
    ORC_B("lkhds", s->lkhds, J)
    ORC_B("grad", s->grad, G)
    ORC_B("grad0", s->grad0, G)
    ORC_B("ss", s->ss, G)
    ORC_B("SS", s->SS, G * ORC_HIST)
    ORC_B("XX", s->XX, G * ORC_HIST)
    ORC_B("gradDeltaTmp", s->gradDeltaTmp, P * J)
    ORC_B("C1", s->C1, N * M)
    ORC_B("C2", s->C2, M * M)
    ORC_B("lkhd", s->lkhd, 50000)
#undef ORC_B
    *len = 0;
    return NULL;
}
