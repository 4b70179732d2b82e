/*
 * gp_oracle.h -- FP64 CPU oracle for the GASPACHO2 sparse GP-LVM fit.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (natsuhiko/GASPACHO2, src/ggplvm.cu and helpers) used as the
 * correctness checker for the HIP library and as the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libgaspacho2.so) never links or calls it.
 *
 * Layouts are the reference's: column-major, FP64.
 *   Z      N x P      model matrix [intercept | covariates | Knm], Z[i + k*N]
 *   K      M x M      Kmm, upper triangle + unit diagonal, lower = 0
 *   Kinv   M x M      after cholKmm: lower+diag = chol(Kmm), upper = Kmm
 *   KinvKmn M x N     KinvKmn[m + i*M]
 *   W      N x (J+1)  W[i + j*N]; Oinvt at W + N*J
 *   Nu     N x J
 *   Psi    2 x J x P x P  factor (lower+diag) / upper = Psi ; inverse at +J*P*P
 *   Beta, tBeta  P x J
 *   grad   [gOmega N | gdelta p | gX N*Q | gT M*Q | grho Q]
 *
 * Parity status of each function is recorded next to its declaration:
 *   "pinned: ..." names the independent check in tests/ that fixes it;
 *   "parity unpinned" marks functions only checked by following source text.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- dense primitives (chol.cu, util_cuda.cu) ---------------- */

/* chol.cu:19-43 choldc.  Reads the upper triangle, writes L to lower+diag.
 * Non-positive pivot -> whole n*n matrix zero-filled, returns 1.
 * pinned: LAPACK potrf via numpy.linalg.cholesky; zero-fill on indefinite. */
int orc_choldc(double *a, int n);
/* chol.cu:45-56 cholsl.  pinned: numpy.linalg.solve. */
void orc_cholsl(const double *a, int n, const double *b, double *x);
/* chol.cu:86-97 cholslg (strided b/x).  pinned: numpy.linalg.solve. */
void orc_cholslg(const double *a, int n, const double *b, int ldb, double *x, int ldx);
/* chol.cu:7-17 getRb: b <- L^T b.  pinned: dense L^T b. */
void orc_getRb(const double *L, double *b, int n);
/* util_cuda.cu:442-451 quadform using the upper triangle.  pinned: b^T A b. */
double orc_quadform(const double *A, const double *b, int n);
/* util_cuda.cu:546-555 mscale_d.  pinned: mean after rescale == m. */
void orc_mscale(int n, double *x, double m);
/* util_cuda.cu:187-232 lbfgs two-loop recursion with +-1 clip and negation.
 * S, Y: n x hist ring buffers.  Work >= n + hist.
 * pinned: dense inverse-BFGS recursion H_k g (tests/test_oracle_pins.py). */
void orc_lbfgs_two_loop(const double *S, const double *Y, const double *g, int n,
                        int hist, int itr, double *work, double *z);

/* ---------------- kernel matrices (kernel_funcs.cu) ---------------- */

/* kernel_funcs.cu:61-97 getKernelMatsWithPer (ard_only=0) and
 * kernel_funcs.cu:221-255 getKernelMatsOnlyARDSE (ard_only=1).
 * Knm N x M, K and Kinv M x M (upper + unit diag, lower 0).
 * pinned: closed-form kernel values, periodicity, factorisation over dims. */
void orc_kernel_mats(int N, int M, int Q, const double *X, const double *T,
                     const double *rho, int ard_only, double *Knm, double *K,
                     double *Kinv);
/* kernel_funcs.cu:154-162 getKinvKmn (after cholKmm).  pinned: dense solve. */
void orc_kinv_kmn(int N, int M, const double *Kchol, const double *Knm,
                  double *KinvKmn, double *titsias);
/* util_cuda.cu:249-269 expandDelta_k.  pinned: block-diagonal construction. */
void orc_expand_delta(int p, const int *nh, int M, const double *delta,
                      const double *K, double *DeltaInv);

/* ---------------- DataFrame / SparseMat (DataFrame.cu, SparseMat.cu) -------- */

/* DataFrame.cu:56-117 + getModelMatrix DataFrame.cu:40-50: expand the covariate
 * body of a DataFrame file into the dense model-matrix columns 0..nh[p-1]-1
 * (intercept first; categorical levels become indicator columns).
 * ncov = number of covariate variables in the file (excludes intercept/Knm).
 * nh must have room for ncov+3 entries; on return p = ncov+2 and nh[p] = P.
 * body = the covariate bytes exactly as stored in the file.
 * Returns P, or -1 on a malformed header.
 * pinned: tests/test_oracle_pins.py builds one-hot columns independently. */
int orc_dataframe_layout(int ncov, const int *c, const int *v, int M, int *nh);
void orc_dataframe_expand(int N, int ncov, const int *c, const int *v,
                          const char *body, const int *nh, double *Zcov);

/* ---------------- the fit state ---------------- */

typedef struct orc_state {
    int N, J, M, Q, P, p, L;
    int ard_only;
    int *nh;                    /* p+1 */
    int *Yp, *Yi; double *Yx;   /* CSC by gene: p[J+1], i[L], x[L] */
    double *Z;                  /* N*P, GP block starts at column nh[p-1] */
    double *ls;                 /* N */
    double *Omega, *delta, *X, *T, *rho;   /* 2N, 2p, 2NQ, 2MQ, 2Q (value|backup) */
    double *phi, *zeta;         /* J, P (zeta == 0: ggplvm.cu never sets it) */
    double *K, *Kinv, *KinvKmn, *titsias, *DeltaInv, *IdenP;
    double *W, *Nu, *Psi, *Beta, *tBeta, *lkhds;
    double *grad, *grad0, *SS, *XX, *ss, *gradDeltaTmp, *gradLogRhoTmp;
    double *C1, *C2;
    double *lkhd;               /* objective history, 50000 like lkhd_d */
    double wt, igap1, igb;      /* ggplvm.cu:20-22 */
    int nthreads;
    /* trace of the last orc_fit: one entry per IRLS step (kind 0), quasi-
     * likelihood start (1) and accepted L-BFGS step (2) with its objective,
     * outer iteration, step index, line-search trials and final sr */
    double *tr_obj, *tr_sr;
    int *tr_info;               /* 4 ints per entry: kind, itr, jtr, ktr */
    int tr_len, tr_cap;
} orc_state;

/* Allocates the state and runs initParams (init.cu:27-48).
 * Zcov: N x nh[p-1] dense covariate columns (column 0 = intercept ones).
 * T0: initial inducing points M x Q (e.g. from orc_init_ta_with_per).
 * X0: initial latent N x Q. */
orc_state *orc_create(int N, int J, int M, int Q, int p, const int *nh,
                      const int *Yp, const int *Yi, const double *Yx, int L,
                      const double *Zcov, const double *ls, const double *X0,
                      const double *T0, int ard_only, int nthreads);
void orc_destroy(orc_state *s);
size_t orc_grad_len(const orc_state *s);   /* N+p+(N+M+1)Q */

/* ---- reference kernels as state operations (file:line each) ---- */
void orc_init_params(orc_state *s);            /* init.cu:27-48 */
void orc_get_kernel_mats(orc_state *s);        /* kernel_funcs.cu:22-24,61-97 */
int  orc_chol_kmm(orc_state *s);               /* kernel_funcs.cu:139-143 */
void orc_get_kinv_kmn(orc_state *s);           /* kernel_funcs.cu:154-162 */
void orc_expand_delta_k(orc_state *s);         /* util_cuda.cu:249-269 */
int  orc_chol_delta_inv(orc_state *s);         /* ggplvm.cu:695-700 */
void orc_init_psi(orc_state *s);               /* init.cu:125-181 + initPsi2 185-191 */
void orc_init_beta(orc_state *s);              /* init.cu:193-221 */
void orc_get_w_for_beta(orc_state *s);         /* ggplvm.cu:468-514 */
void orc_get_psi(orc_state *s);                /* ggplvm.cu:621-680  pinned: dense Z^T D Z + DeltaInv */
int  orc_chol_psi(orc_state *s);               /* ggplvm.cu:683-692  returns #non-PD */
void orc_get_beta_phi(orc_state *s, int jtr);  /* ggplvm.cu:418-464  pinned: IRLS fixed point */
void orc_get_lkhd_pois(orc_state *s);          /* ggplvm.cu:703-724  pinned: Poisson logpmf + titsias term */
/* ggplvm_de.cu:491-532 getScoreStats over a test list (gene[t], snp[t]) with
 * genotype columns G (N x S column-major).  pinned: dense Woodbury form
 * g^T (D^-1 + Z Delta Z^T)^-1 g of the variance term, dense numerator. */
void orc_score_stats(orc_state *s, int S, const double *G, int T, const int *gene,
                     const int *snp, double *stat);
double orc_collect_lkhd(orc_state *s, int Nterm, int jtr);  /* ggplvm.cu:788-815 */
void orc_get_oinvt(orc_state *s);              /* ggplvm.cu:284-294 */
void orc_get_lkhd(orc_state *s);               /* ggplvm.cu:744-784  pinned: Woodbury dense log N */
void orc_get_tilde_beta(orc_state *s);         /* ggplvm.cu:406-414 */
void orc_grad_log_delta(orc_state *s);         /* ggplvm.cu:210-281  pinned: finite differences */
void orc_grad_log_omega(orc_state *s);         /* ggplvm.cu:301-376  pinned: finite differences */
void orc_solve_psi(orc_state *s);              /* ggplvm.cu:46-53    pinned: numpy inverse */
void orc_get_c2(orc_state *s);                 /* ggplvm.cu:57-100   pinned: FD (via gT, grho) */
void orc_grad_tau(orc_state *s);               /* ggplvm.cu:103-145  pinned: finite differences */
void orc_grad_log_rho(orc_state *s);           /* ggplvm.cu:148-205  pinned: finite differences */
void orc_lbfgs(orc_state *s, int itr, int rhoflag);          /* ggplvm.cu:865-934 */
void orc_downdate(orc_state *s, int itr, double r, int rhoflag); /* ggplvm.cu:818-859 */
void orc_beta_from_tilde_beta(orc_state *s);   /* ggplvm.cu:394-402 */

/* ---- launch groups, in the order main() issues them ---- */
/* ggplvm.cu:1111-1113,1134-1137 */
void orc_init_fit(orc_state *s);
/* ggplvm.cu:1156-1180: one IRLS step; returns the GLM objective */
double orc_irls_step(orc_state *s, int itr, int jtr);
/* ggplvm.cu:1205-1211: objective at the start of the quasi-likelihood phase */
double orc_ql_start(orc_state *s, int itr);
/* ggplvm.cu:1259-1277: one ELBO evaluation at the current parameters */
double orc_elbo(orc_state *s, int itr, int slot);
/* ggplvm.cu:1215-1243: one gradient evaluation (fills s->grad) */
void orc_grad(orc_state *s);
/* ggplvm.cu:1146-1330: the whole outer loop (itrmax outer iterations).
 * jtr_cap > 0 caps both inner loops (0 = reference schedule).  Returns the
 * last objective. */
double orc_fit(orc_state *s, int itrmax, int jtr_cap, int verbose);
/* copies min(len, cap) trace entries of the last orc_fit; returns len */
int orc_fit_trace(const orc_state *s, double *obj, double *sr, int *info, int cap);

/* ---- initial inducing points (util.c:73-87) ----
 * Dim 0 equispaced on (0, 2*pi); other dims N(0,1) from MT19937 seeded as GSL
 * seeds seed 0 (-> 4357) and the GSL polar Box-Muller gaussian.
 * The MT19937 core is pinned by the published KAT (10000th output of the
 * 5489-seeded generator = 4123659995); GSL's seed-0 stream and its uniform
 * (output / 2^32) by the GSL manual's two example outputs (first value
 * 4293858116; Poisson mu = 3 draws 2 5 5 2 1 0 3 4 1 1); the polar Gaussian by
 * a normality test.  Parity unpinned: which coordinate of the accepted polar
 * pair is returned (the second, as GSL's source does; GSL is not in this
 * image, and the two choices have the same distribution). */
void orc_init_ta_with_per(int M, int Q, double *T);
uint32_t orc_mt19937_kat(uint32_t seed, int n);   /* n-th output */

/* Named state buffer (reference layout) and its length in doubles; NULL if unknown. */
double *orc_buf(orc_state *s, const char *name, size_t *len);

#ifdef __cplusplus
}
#endif
#endif
