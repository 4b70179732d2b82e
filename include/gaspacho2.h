/*
 * gaspacho2.h -- C ABI of the MI355X (gfx950) sparse GP-LVM fit library.
 *
 * The reference (natsuhiko/GASPACHO2) has no plugin layer: its boundary is the
 * set of <<<>>> launch groups issued by main() in src/ggplvm.cu.  Each entry
 * point below replaces one such launch group and keeps its argument meaning;
 * the citation on each line names the reference code it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host arrays are in the reference's
 *    layouts (column-major FP64, see gp_get_buffer).
 *  - Every call returns int status: 0 ok; >0 a hipError_t / library error
 *    code (message in gp_last_error()); <0 minus the number of per-gene
 *    Cholesky factorizations that met a non-positive pivot.  Such a factor is
 *    zero-filled, exactly as chol.cu:26-29 does, so the objective becomes
 *    +-inf/NaN the way the reference's does.
 *  - One context = one HIP device, one stream; calls on a context must come
 *    from one host thread at a time.
 */
#ifndef GASPACHO2_H
#define GASPACHO2_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gp_ctx gp_ctx;

/* Precision of the weighted-Gram and gradient contractions. */
enum {
    GP_PREC_FP64 = 0,    /* FP64 throughout (the reference's arithmetic) */
    GP_PREC_MIXED = 1,   /* ELBO/gradient in the whitened basis
                            [covariates | Knm L^-T] (Kmm = L L^T): FP16 hi/lo
                            split MFMA Gram, FP32 accumulate, FP64 segment
                            sums, FP64 Cholesky/log-det */
    GP_PREC_FP64_WHITENED = 2  /* the whitened basis with FP64 arithmetic
                                  (validates the basis change on its own) */
};

/* Library version (major*10000 + minor*100 + patch). */
int gp_version(void);

/* Message for the last non-zero status returned on this host thread. */
const char *gp_last_error(void);

/* Replaces the allocation + upload block of main(): ggplvm.cu:964-968
 * (newSparseMatOnDevice), :981-982 (newDataFrameOnDevice), :1019-1060
 * (cudaMalloc of every buffer, ls / X upload, T upload after initTaWithPer)
 * and initParams<<<1,1>>> ggplvm.cu:1101 / init.cu:27-48.
 *   N cells, J genes, M inducing points, Q latent dims.
 *   p, nh[p+1]: variables and their model-matrix column offsets as built by
 *     loadDataFrame (DataFrame.cu:56-90): variable 0 = intercept, p-1 = Knm.
 *   Yp[J+1], Yi[L], Yx[L]: counts, CSC by gene (SparseMat.cu:37-52); row
 *     indices ascending within a gene.
 *   Zcov: N x nh[p-1] covariate model-matrix columns, column-major, with the
 *     intercept in column 0 and categorical levels as 0/1 indicator columns
 *     (what getModelMatrix, DataFrame.cu:40-50, returns).
 *   ls[N] size factors; X0[N*Q], T0[M*Q] column-major.
 *   ard_only: 0 = periodic dim 0 + SE-ARD (kernel_funcs.cu:61-97, the
 *     reference default); 1 = SE-ARD on all dims (kernel_funcs.cu:221-255).
 *   precision: GP_PREC_FP64 or GP_PREC_MIXED. */
int gp_create(gp_ctx **out, int N, int J, int M, int Q, int p, const int *nh,
              const int *Yp, const int *Yi, const double *Yx, long long L,
              const double *Zcov, const double *ls, const double *X0,
              const double *T0, int ard_only, int precision);
int gp_destroy(gp_ctx *ctx);

/* Switch the contraction precision of an existing context. */
int gp_set_precision(gp_ctx *ctx, int precision);

/* Launch-shape options (no counterpart in the reference, whose launch shapes
 * are fixed in ggplvm.cu; these are the knobs the autotuner sweeps):
 *   "gram_splits" [1, 256]  cell segments of the mixed Gram (FP32 partial
 *                           per segment, FP64 sum across segments)
 *   "gram_waves"  4 or 8    waves per workgroup of the mixed Gram: 8 (two
 *                           blocks of tiles per wave, two waves per SIMD) or
 *                           4 (a quarter of the tiles per wave, fragments
 *                           shared across its blocks, one wave per SIMD)
 *   "gc_groups"   [1, 4096] gene groups of the mixed gradient contraction
 *                           (FP32 partial per group, FP64 sum across groups)
 *   "gc_form"     0/1/2/3   mixed contraction: 0 = 16-cell waves with S_j
 *                           staged in LDS; 1 = 128-cell row blocks with
 *                           each wave's S_j columns in registers (same bits
 *                           as 0); 2 = row blocks with the cell weights
 *                           applied to each gene's FP32 product, so A is
 *                           staged once per gene group; 3 = form 2 on four
 *                           waves of 512 registers (same bits as 2)
 *   "gc_mixed"    0/1       mixed path: contraction on FP16 hi/lo MFMA (1)
 *                           or in FP64 (0)
 *   "irls_mixed"  0/1       mixed path: IRLS steps on the whitened FP16
 *                           hi/lo Gram (1) or in FP64 in the reference
 *                           basis (0)
 *   "irls_fused"  0/1       mixed IRLS steps: the weight update (W/Omega,
 *                           nu, R and the per-gene sums) inside the Gram
 *                           kernel, W and Nu written only when read (1), or
 *                           on separate kernels before the Gram (0)
 *   "chol1"       0/1       the Kmm factor on the one-matrix Cholesky
 *                           kernel (1) or the batched kernel (0)
 *   "cholsl_blk"  0/1       per-gene solves blocked by 16 rows (1) or row
 *                           by row (0)
 *   "tn_f32"      0/1       mixed path: the gradient's z^T beta-tilde and
 *                           C1a products in FP32 (1) or FP64 (0)
 *   "gc_slo"      0/1       intercept-peeled mixed contraction: S_j as
 *                           hi/lo against zw as hi/lo, three MFMAs per tile
 *                           (1, the default), or S_j in FP16 alone, two (0:
 *                           ~35 % faster contraction, gradient ~10x less
 *                           accurate, the inducing-point block most)
 *   "score_mfma"  0/1       score tests: partials on FP64 MFMA over every
 *                           model-matrix column per pass (1, the default;
 *                           P <= 576, the FP32 copy of Zw in a mixed
 *                           context) or one column per thread (0)
 *   "vtv_mfma"    0/1       S = V^T V of the per-gene inverses on FP64
 *                           MFMA (1, the default) or FP64 VALU (0)
 *   "chol_nb"     8/16      panel width of the batched Cholesky of the
 *                           Psi_j (8: eight workgroups per CU, the default)
 *   "gram_wide"   0/1       mixed Gram at M = 512 (intercept-peeled): the
 *                           GP block as two symmetric halves and four
 *                           rectangles on the symmetric kernel's chunk
 *                           loop (1) or on the general tiled kernel (0)
 *   "autotune"    0/1       gp_fit calls gp_autotune before its first
 *                           quasi-likelihood phase (mixed precision only)
 * Unknown names and out-of-range values return a status > 0. */
int gp_set_option(gp_ctx *ctx, const char *name, int value);
int gp_get_option(gp_ctx *ctx, const char *name, int *value);

/* Autotuner (no counterpart in the reference): times the mixed-precision
 * Gram for gram_waves in {8, 4} x gram_splits in {8, 16, 32} and the mixed
 * contraction for gc_form in {2, 1, 0} x gc_groups in {16, 32, 64} on the context's current state and keeps the
 * fastest of each.  Needs GP_PREC_MIXED and a state after gp_ql_start; it
 * recomputes the per-gene factors and inverses at the current parameters. */
int gp_autotune(gp_ctx *ctx, int verbose);

/* Named state buffers in the reference layout (host <-> device copies).
 * Names: Omega delta X T rho phi (value|backup halves where the reference
 * keeps them), Z, K, Kinv, KinvKmn, titsias, DeltaInv, W (incl. Oinvt at
 * +N*J), Nu, Psi (factor, incl. inverse at +J*P*P), Beta, tBeta, lkhds,
 * grad, grad0, ss, SS, XX, C1, C2, lkhd; read-only: U and Uw (the
 * gradient contraction sum_j diag(W_j/Omega) Z Psi_j^-1 in the reference
 * and the whitened basis, N x P).  Lengths in doubles match the
 * reference's cudaMalloc sizes (ggplvm.cu:1024-1060).
 * Replaces readParams (util_cuda.cu:520-544) / writeParams (557-580) copies. */
size_t gp_buffer_len(gp_ctx *ctx, const char *name);
int gp_get_buffer(gp_ctx *ctx, const char *name, double *host, size_t n);
int gp_set_buffer(gp_ctx *ctx, const char *name, const double *host, size_t n);

/* ggplvm.cu:1111-1137: getKernelMats, cholKmm, getKinvKmn, expandDelta_k,
 * initPsi, initPsi2, initBeta. */
int gp_init_fit(gp_ctx *ctx);

/* ggplvm.cu:1156-1181: one IRLS/Laplace step of the GLM phase
 * (getWForBeta, getPsi, cholPsi, getBetaPhi(jtr), getLkhdPois, collectLkhd
 * with N=0).  *obj receives lkhd[jtr].  With "irls_fused" (the default of a
 * mixed context) getWForBeta runs inside the Gram kernel and W / Nu are
 * written on the next call that reads them; every other entry point sees
 * them as the reference leaves them. */
int gp_irls_step(gp_ctx *ctx, int jtr, double *obj);

/* ggplvm.cu:1205-1211: getOinvt, cholDeltaInv, getLkhd, collectLkhd(jtr=0):
 * the objective at the start of the quasi-likelihood phase. */
int gp_ql_start(gp_ctx *ctx, double *obj);

/* ggplvm.cu:1259-1278: one ELBO evaluation at the current parameters
 * (getOinvt, getKernelMats, expandDelta_k, cholKmm, getKinvKmn, getPsi,
 * cholPsi, cholDeltaInv, getLkhd, collectLkhd); stored at lkhd[slot].  The
 * factor of DeltaInv = blockdiag(diag 1/delta, Kmm/delta_GP) is formed from
 * the Kmm factor (cholDeltaInv's result up to rounding). */
int gp_elbo(gp_ctx *ctx, int slot, double *obj);

/* ggplvm.cu:1215-1240: one gradient evaluation (expandDelta_k, getOinvt,
 * getTildeBeta, getGradLogDelta, collectGradLogDelta, getGradLogOmega,
 * solvePsi, getC2, getGradTau, getGradLogRho, collectGradLogRho) into the
 * device grad buffer [gOmega N | gdelta p | gX NQ | gT MQ | grho Q].
 * As in the reference, it uses the factors of the last gp_elbo /
 * gp_ql_start, so it belongs at that call's parameters; when no other
 * state-changing call came in between, getTildeBeta is taken from that
 * call's own solve instead of being repeated (same bits). */
int gp_grad(gp_ctx *ctx);

/* ggplvm.cu:1251 LBFGS<<<1,1>>> (ggplvm.cu:865-934, util_cuda.cu:187-232):
 * L-BFGS direction from the history ring and the parameter update with
 * backups. */
int gp_lbfgs_step(gp_ctx *ctx, int jtr, int rhoflag);

/* ggplvm.cu:1289 Downdate<<<1,1>>> (ggplvm.cu:818-859). */
int gp_downdate(gp_ctx *ctx, int jtr, double sr, int rhoflag);

/* One ELBO + gradient unit -- gp_elbo(slot) followed by gp_grad, i.e.
 * ggplvm.cu:1259-1278 plus 1215-1240 -- recorded once as a HIP graph and
 * replayed (no counterpart in the reference, which launches every kernel
 * from the host).  Replays evaluate at the current parameters; the graph is
 * re-recorded when the precision, the slot or a launch-shape option changes.
 * *obj = lkhd[slot]. */
int gp_eval_graph(gp_ctx *ctx, int slot, double *obj);

/* getScoreStats (ggplvm_de.cu:491-532) for a list of T tests: test t pairs
 * gene gene[t] with genotype column snp[t] of G (N x S, column-major):
 *   stat[t] = g^T Omega^-1 (y_j - W_j) / sqrt(phi_j (g^T D_j g - b^T Psi_j^-1 b)),
 *   D_j = diag(W_j / Omega), b = Z^T D_j g,
 * from the current W, Omega, phi and per-gene factors (call after gp_irls_step
 * or gp_elbo).  The reference launches a (J, L) grid over all pairs with one
 * shared G; a test list covers that (T = J L) and per-gene SNP sets.
 * gp_score_tests_dosage takes 0/1/2 dosages as bytes. */
int gp_score_tests(gp_ctx *ctx, int S, const double *G, int T, const int *gene,
                   const int *snp, double *stat);
int gp_score_tests_dosage(gp_ctx *ctx, int S, const unsigned char *dosage, int T,
                          const int *gene, const int *snp, double *stat);
/* Interaction tests as the differential-expression driver builds them
 * (ggplvm_de.cu:989-994, with util.c:100-110 scale): the test column of SNP l
 * is h = g_l * factor - mean_i(g_l * factor) for a per-cell factor[N] (NULL:
 * h = g_l), then the statistic above with h in place of g
 * (ggplvm_de.cu:491-532, 1206-1210); pval[t] (may be NULL) = the two-sided
 * normal p-value erfc(|stat| / sqrt 2).  dosage: N x S bytes (0/1/2). */
int gp_score_tests_interaction(gp_ctx *ctx, int S, const unsigned char *dosage,
                               const double *factor, int T, const int *gene, const int *snp,
                               double *stat, double *pval);
/* The interaction tests above with genotypes given per donor: cell i carries
 * the genotypes of donor donor[i] (0 <= donor[i] < ndonor), so dosage is
 * ndonor x S bytes and g_l(i) = dosage[donor[i] + l * ndonor]; otherwise as
 * gp_score_tests_interaction (factor may be NULL, pval may be NULL).  The
 * reference expands the genotype matrix to N rows on the host
 * (ggplvm_de.cu:976-996); the per-cell lookup gives the same column. */
int gp_score_tests_donor(gp_ctx *ctx, int ndonor, const int *donor, int S,
                         const unsigned char *dosage, const double *factor, int T,
                         const int *gene, const int *snp, double *stat, double *pval);
/* No counterpart in the reference.  gp_score_tests_dosage for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): dosage_dev (N x S bytes), gene_dev and
 * snp_dev (T ints) and stat_dev (T doubles) are the caller's device buffers,
 * complete before the call; the partial sums live in a static 32 MiB device
 * array that each GPU holds until the process ends, so nothing is allocated.
 * Returns after stat_dev is written.  Designs wider than 1024 model-matrix
 * columns are rejected (status 1).  An out-of-range gene or SNP index sets
 * that test's statistic to NaN and returns status 1.  Calls that overlap on
 * one GPU share the array and race. */
int gp_score_tests_dosage_noalloc(gp_ctx *ctx, int S, const unsigned char *dosage_dev, int T,
                                  const int *gene_dev, const int *snp_dev, double *stat_dev);

/* ---- Batched per-gene GLM / Laplace fits on cell subsets ----
 * No counterpart in the reference, which fits every gene on all N cells
 * (ggplvm.cu:487); each step restates getWForBeta / getPsi / cholPsi /
 * getBetaPhi / getLkhdPois (ggplvm.cu:418-514, 621-724) for gene g on its
 * own cells with design [1 | covariates | Knm], prior deltas
 * (1.5, 0.01, 0.3) as initParams (init.cu:31-35), on the latent space of ctx
 * (its Knm, Kmm, titsias, Omega, size factors: call after gp_init_fit).
 *   off[J+1], cell[off[J]]: gene g's cells are cell[off[g] .. off[g+1]-1]
 *   ncov[J]: p_g = design columns before Knm, intercept included (>= 1)
 *   cov: per gene, n_g x (p_g - 1) FP32 column-major, genes back to back
 *   yoff[J+1], ypos, yval: gene g's non-zero counts at positions (0-based,
 *     within its subset, ascending) ypos[yoff[g] .. yoff[g+1]-1]
 * p_g + M + 1 <= 160. */
typedef struct gp_bfit gp_bfit;
int gp_bfit_create(gp_bfit **out, gp_ctx *ctx, int J, const int *off, const int *cell,
                   const int *ncov, const float *cov, const int *yoff, const int *ypos,
                   const double *yval);
int gp_bfit_destroy(gp_bfit *b);
/* init.cu:125-221 per gene (unweighted Gram, initBeta); phi = 1, step 0. */
int gp_bfit_init(gp_bfit *b);
/* One IRLS step of every gene; *obj = sum of the genes' Poisson
 * log-likelihoods at the step's starting weights (collectLkhd, N = 0). */
int gp_bfit_step(gp_bfit *b, double *obj);
/* Up to maxit steps, stopping when the relative objective change < tol
 * (ggplvm.cu:1156-1181); use_graph records one step as a HIP graph. */
int gp_bfit_run(gp_bfit *b, int maxit, double tol, int use_graph, int *iters, double *obj);
/* No counterpart (config-2 launch shape): times the persistent pass with 1-4
 * workgroups per CU at the current beta and keeps the fastest (*grid_mult,
 * may be NULL; results do not depend on it).  Beta and phi are kept; "Psi"
 * holds unfactored Grams until the next step.  gp_bfit_set_grid sets it. */
int gp_bfit_autotune(gp_bfit *b, int verbose, int *grid_mult);
int gp_bfit_set_grid(gp_bfit *b, int grid_mult);
/* "Beta" (J x PB), "phi" (J), "lkhd" (J), "Psi" (J x PB x PB factors). */
int gp_bfit_get(gp_bfit *b, const char *name, double *host, size_t n);
int gp_bfit_dims(gp_bfit *b, int *J, int *PB);

/* ---- PCA initial latent ----
 * The reference's fit starts from an X read from a file (ggplvm.cu:1088-1095)
 * whose recipe is the helpers getNu0 / getNuTNu / getX (init.cu:304-357):
 *   nu0[i, j] = log((y_ij + 0.5) / s_i) - z_i . beta_j,
 *   top-Q eigenpairs (lambda_k, v_k) of nu0^T nu0, largest first,
 *   X[:, k] = nu0 v_k / sqrt(lambda_k) * sqrt(N).
 * z_i = Zcov row i (N x p0 column-major, intercept first); beta_j is the
 * least-squares fit of log((y_j + 0.5) / s) on Zcov (columns dependent on
 * earlier ones are dropped).  v_k is signed so that
 * its entry of largest magnitude is positive.  Gene-major CSR counts (row
 * indices increasing within a gene).  X: N x Q column-major (the Xinit.bin
 * layout of the driver); eval (may be NULL): the Q eigenvalues.  Standalone:
 * no context; runs on the current device. */
int gp_pca_init(int N, int J, const int *Yp, const int *Yi, const double *Yx, const double *ls,
                int p0, const double *Zcov, int Q, double *X, double *eval);

/* ---- Amortization encoder (warm start of X) ----
 * No counterpart in the reference, which reads the starting X from an
 * external PCA file (ggplvm.cu:1088-1095, argv[4]).  An MLP
 *   a0 = log1p(y_i. / s_i) (F = genes) -> H1 -> H2 -> [mu (Q) | log-variance (Q)]
 * with exact GELU, BF16 weight copies and activations in the products, FP32
 * accumulation, FP32 Adam (0.9, 0.999, 1e-8) on FP32 master weights; loss =
 * mean over cells of sum_q 1/2 [(x - mu)^2 exp(-s) + s], s = log-variance
 * clamped to [-10, 10] (zero gradient outside).  H1, H2 and batch
 * must be multiples of 64, 2Q <= 64.  Parameters: one FP32 array
 * [W1 (H1 x Fp) | b1 | W2 (H2 x H1) | b2 | W3 (64 x H2) | b3 (64)], row-major,
 * Fp = F rounded up to 64 (padding stays zero).
 * dims: F, Fp, H1, H2, Q, 64, batch, cells loaded, parameter count, capacity,
 * Adam steps, 0. */
typedef struct gp_enc gp_enc;
int gp_enc_create(gp_enc **out, int F, int H1, int H2, int Q, int batch, long long capacity,
                  unsigned long long seed);
int gp_enc_destroy(gp_enc *enc);
int gp_enc_dims(gp_enc *enc, long long dims[12]);
/* Append N training cells: gene-major CSR counts (J == F genes), size
 * factors ls[N], targets X (N x Q column-major, as gp_get_buffer "X"). */
int gp_enc_add(gp_enc *enc, int N, int J, const int *Yp, const int *Yi, const double *Yx,
               const double *ls, const double *X);
/* Loss of the first n loaded cells (n <= batch) and its gradient, kept on
 * the device (gp_enc_get "grad"). */
int gp_enc_loss_grad(gp_enc *enc, int n, double *loss);
/* One Adam step from the current gradient. */
int gp_enc_adam_step(gp_enc *enc, double lr);
/* epochs passes over the loaded cells in shuffled batches; epoch_loss[e] is
 * the cell-weighted mean of the batch losses of epoch e (may be NULL).
 * Returns 3 if an epoch's loss is not finite. */
int gp_enc_train(gp_enc *enc, int epochs, double lr, double *epoch_loss);
/* mu and logvar (N x Q column-major; logvar may be NULL) for new cells. */
int gp_enc_predict(gp_enc *enc, int N, int J, const int *Yp, const int *Yi, const double *Yx,
                   const double *ls, double *mu, double *logvar);
/* "params", "grad", "adam_m", "adam_v" (n = parameter count); set: "params". */
int gp_enc_get(gp_enc *enc, const char *name, float *host, size_t n);
int gp_enc_set(gp_enc *enc, const char *name, const float *host, size_t n);

/* ggplvm.cu:1319-1324: expandDelta_k, getBetaFromTildeBeta. */
int gp_end_outer(gp_ctx *ctx);

/* ggplvm.cu:1146-1330: the whole outer loop (phase A, phase B with line
 * search, end-of-iteration), run by the library with the reference's
 * schedule.  jtr_cap > 0 caps every inner loop (0 = reference schedule).
 * on_outer, if non-NULL, is called after each outer iteration (the driver
 * writes X9.gz / phi.gz / Params9.bin there).  *obj = last objective. */
typedef void (*gp_outer_cb)(gp_ctx *ctx, int itr, double obj, void *user);
int gp_fit(gp_ctx *ctx, int itrmax, int jtr_cap, int verbose,
           gp_outer_cb on_outer, void *user, double *obj);

/* Trace of the last gp_fit (no counterpart: the reference prints lkhd[jtr]
 * per step, ggplvm.cu:1170,1276): the objective after every IRLS step
 * (kind 0), quasi-likelihood start (1) and accepted L-BFGS step (2), with
 * host seconds since gp_fit began.  Copies min(*len, cap) entries; any output
 * pointer may be NULL; *len = entries recorded. */
int gp_fit_trace(gp_ctx *ctx, double *obj, double *seconds, int *kind, int cap, int *len);
/* The same entries' positions in the schedule: step[3k..3k+2] = outer
 * iteration itr, inner step jtr and (L-BFGS entries) the line-search trials
 * ktr of ggplvm.cu:1257-1290; sr[k] = the step fraction the line search
 * ended on (1 when the first trial was accepted).  Either pointer may be NULL. */
int gp_fit_trace_steps(gp_ctx *ctx, int *step, double *sr, int cap, int *len);

/* util.c:73-87 initTaWithPer: initial inducing points T (M x Q,
 * column-major).  Dim 0 equispaced on (0, 2*pi); dims 1..Q-1 N(0,1) from
 * MT19937 seeded the way GSL seeds seed 0 (-> 4357), GSL polar Box-Muller.
 * Host-only; needs no device. */
int gp_init_ta_with_per(int M, int Q, double *T);

/* Average device time (ms) per launch of a named kernel group since
 * gp_timing_reset, and the launch count: "gram", "chol", "gradcontract",
 * "kuf", "elbo", "grad", and "gram_kernel" / "gradcontract_kernel" (the main
 * Gram / contraction kernel alone, without the preparation and reduction
 * kernels of its group).  Each timed group synchronizes the stream.
 * Replaces the cudaEvent printf timers of ggplvm.cu:1155-1276. */
int gp_timing_enable(gp_ctx *ctx, int on);
int gp_timing_reset(gp_ctx *ctx);
int gp_timing_get(gp_ctx *ctx, const char *name, double *avg_ms, int *count);

#ifdef __cplusplus
}
#endif
#endif
