// gp_internal.h -- device context and launch wrappers shared by the library's
// translation units.  Host code only sees plain device pointers; every kernel
// lives in exactly one .hip file and is reached through the wrappers below.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#define GP_HIST 10            // L-BFGS history length (ggplvm.cu:872-880)
#define GP_LKHD_SLOTS 50000   // objective history length (ggplvm.cu:1025)
#define GP_GRAPH_KEY 19       // settings a recorded ELBO + gradient graph depends on

struct gp_timer {
    hipEvent_t a = nullptr, b = nullptr;
    double total_ms = 0.0;
    int count = 0;
    bool open = false;
};

constexpr int GP_TN_SPLITS = 4;   // gene ranges of the per-cell tn / E kernel

struct gp_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    // side stream (default priority) and its rocBLAS handle: the gradient's
    // per-cell tn pieces run there while the inverse and the contraction run
    // on the main stream; fork and join through the two events
    hipStream_t stream2 = nullptr;
    rocblas_handle blas2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the ELBO's fork: Oinvt and the FP32 weights beside the kernel matrices
    hipEvent_t ev_efork = nullptr, ev_ejoin = nullptr;
    int efork_pending = 0;   // the ELBO's side chain is forked and not yet joined
    // the end of the likelihood's b = Z^T R product on the ELBO's side stream
    hipEvent_t ev_bdone = nullptr;
    int bfork_pending = 0;
    // the peeled Gram's intercept row on the side stream: fork after the
    // Gram kernel, join before the factorization
    hipEvent_t ev_rfork = nullptr, ev_row0 = nullptr;

    int N = 0, J = 0, M = 0, Q = 0, P = 0, p = 0, g0 = 0, Npad = 0;
    long long L = 0;
    int ard_only = 0, precision = 0;
    std::vector<int> nh;

    // counts, CSC by gene
    int *Yp = nullptr, *Yi = nullptr;
    double *Yx = nullptr;
    int *d_nh = nullptr;

    // model matrix Npad x P (column-major, ld = Npad); GP block at column g0
    double *Z = nullptr;
    double *ls = nullptr;
    // parameters with backup halves (reference layout)
    double *Omega = nullptr, *delta = nullptr, *X = nullptr, *T = nullptr, *rho = nullptr;
    double *phi = nullptr, *zeta = nullptr;

    double *K = nullptr;       // M x M: upper + unit diag, lower 0 (reference K)
    double *Ksym = nullptr;    // M x M full symmetric Kmm
    double *Kchol = nullptr;   // M x M: lower+diag = chol(Kmm), upper = Kmm
    double *Kinv = nullptr;    // M x M explicit inverse of Kmm
    double *A = nullptr;       // Npad x M: KinvKmn transposed
    double *titsias = nullptr; // N
    double *DeltaInv = nullptr;  // P x P reference layout (upper)
    double *DeltaChol = nullptr; // P x P factor of DeltaInv (lower+diag)
    // whitened basis (mixed path): GP columns Knm L^-T with Kmm = L L^T
    double *Linv = nullptr;      // M x M  L^{-1}
    double *LinvT = nullptr;     // M x M  L^{-T}: the whitening product's right operand, untransposed
    double *Zw = nullptr;        // Npad x P  [covariates | Knm L^-T]
    double *DeltaW = nullptr;    // P x P  diag(1/delta_l): the whitened prior
    double *DeltaWChol = nullptr;// P x P  diag(sqrt(1/delta_l))
    double *Uw = nullptr;        // Npad x P  contraction in the whitened basis
    double *tBw = nullptr;       // P x J  beta-tilde in the whitened basis
    int psi_whitened = 0;        // basis of the current Psi factor
    int zw_valid = 0;            // Zw / Linv match the current Kmm factor
    double *skw = nullptr;       // split-K GEMM partials
    size_t skw_len = 0;
    double *Vtmp = nullptr;      // J x P x P scratch: L^{-1} of every factor
    size_t vtmp_len = 0;

    double *W = nullptr;       // Npad x J
    double *Oinvt = nullptr;   // N
    double *Nu = nullptr;      // Npad x J
    double *Psi = nullptr;     // J x P x P: upper = Psi, lower+diag = chol
    double *Sinv = nullptr;    // J x P x P explicit inverse (symmetric)
    double *Beta = nullptr, *tBeta = nullptr, *bvec = nullptr, *tBphi = nullptr;  // P x J
    double *pwork = nullptr;   // P x J scratch (beta-tilde of the likelihood, beta work)
    double *lkhds = nullptr, *gq = nullptr, *gs = nullptr;   // J
    double *grad = nullptr, *grad0 = nullptr, *ss = nullptr, *SS = nullptr, *XX = nullptr;
    double *lbq = nullptr;     // L-BFGS work vector (tot)
    double *gradDeltaTmp = nullptr;  // P x J
    double *gradLogRhoTmp = nullptr; // Q x M
    double *C1 = nullptr;      // Npad x M
    double *C2 = nullptr;      // M x M
    double *U = nullptr;       // Npad x P  sum_j o_ij Psi_j^{-1} z_i
    double *ZB = nullptr;      // Npad x J  Z*beta (IRLS) / Z*tBeta (gradient)
    double *R = nullptr;       // Npad x J  W nu / Omega ; reused for TN/phi
    double *C1a = nullptr;     // Npad x M
    double *AO = nullptr;      // Npad x M  diag(Oinvt) A
    double *G2 = nullptr, *B2 = nullptr, *S2 = nullptr;  // M x M pieces of C2
    double *Ecell = nullptr;   // N
    double *Epart = nullptr;   // GP_TN_SPLITS x Npad partial E sums of k_tn
    double *ci1 = nullptr;     // N: per-cell row sums of C1 (latent gradient)
    double *red = nullptr;     // reduction scratch
    double *scal = nullptr;    // device scalars (L-BFGS alphas etc.)
    double *lkhd = nullptr;    // GP_LKHD_SLOTS
    int *info = nullptr;       // non-PD counters
    // mixed precision staging (fp16 hi/lo of the model matrix, tiled)
    void *Zh = nullptr, *Zl = nullptr;  // FP16 hi/lo of Z (Npad x P16)
    float *z32 = nullptr;      // the same column-scaled model matrix in FP32
    float *w32 = nullptr;      // Npad x J  W/Omega in FP32
    int w32_valid = 0;         // w32 / wsc / wun hold the current W / Omega (cleared by every call that may change them)
    float *wsc = nullptr;      // J  2^-e_j weight scale
    double *wun = nullptr;     // J  2^e_j
    double *zun = nullptr;     // P16  2^-s_a column unscale of the FP16 model matrix
    float *gram_part = nullptr;   // per-segment FP32 partial Gram blocks
    size_t gram_part_bytes = 0;
    int gram_splits = 8;       // cell segments of the mixed Gram
    int gram_waves = 8;        // waves per workgroup of the symmetric mixed Gram (8 or 4)
    int gram_peel = 1;         // symmetric Gram on the GP block only when the intercept is the sole covariate column
    int gram_wide = 1;         // M = 512: the GP block as symmetric halves and rectangles (0: the general tiled kernel)
    float *gram_row0 = nullptr;  // per-segment FP32 intercept-row products (peeled Gram)
    size_t gram_row0_len = 0;
    int zsplit_valid = 0;
    // mixed gradient contraction: per-gene FP16 hi/lo of Sw_j, gene groups
    void *Sh = nullptr, *Sl = nullptr;  // J x P16 x 288 halves, [col][k]
    int *gc_perm = nullptr, *gc_aexp = nullptr, *gc_sexp = nullptr, *gc_bexp = nullptr;
    int *gc_zexp = nullptr;    // per column k of Zw: 2^zc_k >= max |Zw[:,k]|
    int *gc_kappa = nullptr;   // per gene exponent of max o * max |Sw| (group planning)
    int capturing = 0;         // recording a graph: no host reads, syncs or allocations
    double *gc_gun = nullptr;  // per group 2^-C_g
    float *gc_part = nullptr;  // per group FP32 partial U (P16 x Npad)
    size_t gc_part_len = 0;
    int gc_groups = 16;        // gene groups of the mixed contraction
    int gc_mixed = 1;          // mixed path: contraction on FP16 MFMA (0: FP64)
    int gc_form = 2;           // mixed contraction: 0 = 16-cell waves, 1 = 128-cell row blocks, 2 = factored row blocks
    int gc_peel = 1;           // factored form on the GP block only when the intercept is the sole covariate column
    int gc_wave_rows = 8;      // peeled form: row tiles per wave (8 or 4; column tiles 16 / that)
    int gc_waves = 8;          // peeled form: 8 waves (2 column tiles each) or 4 (4 column tiles, 8 row tiles)
    int gc_prefetch = 3;       // peeled form: k-steps of B fragments loaded ahead (0, 1, 3 or 7)
    float *gc_t0 = nullptr;    // J x M  the intercept's k term of each gene (peeled form)
    float *gc_s0 = nullptr;    // J x (M+1)  [Sw_j[0, 1:] | Sw_j[0, 0]] (intercept output column)
    float *gc_y = nullptr;     // Npad x (M+1)  w32 . gc_s0
    int irls_mixed = 1;        // mixed path: IRLS steps on the whitened FP16 hi/lo Gram (0: FP64)
    int irls_fused = 1;        // mixed IRLS steps: weights, nu, R and the sums built inside the Gram kernel
    // fused IRLS pass: per-cell constants [d | Omega | log ls] (3 x Npad), each
    // gene's first non-zero per cell segment, per-(segment, gene) sums and
    // weight unscales, and the segmentation they were made for
    double *irls_cc = nullptr;
    int *irls_segptr = nullptr;
    double *irls_sums = nullptr, *irls_segun = nullptr;
    int irls_nseg = 0, irls_seg = 0;
    // 1 after a fused IRLS step: W and Nu are not written yet; ZB and
    // irls_cc still hold the eta and constants they follow from
    int wnu_pending = 0;
    int autotune = 0;          // gp_fit runs gp_autotune before its first quasi-likelihood phase
    int chol1 = 1;             // the Kmm factor takes the one-matrix Cholesky kernel (0: the batched kernel)
    int cholsl_blk = 1;        // per-gene solves blocked by 16 rows (0: row by row)
    int chol_nb = 8;           // panel width of the batched Cholesky (8 or 16)
    int vtv_mfma = 1;          // S = V^T V of the inverses on FP64 MFMA (1) or VALU (0)
    int score_mfma = 1;        // score tests: partials on FP64 MFMA, all columns per pass (P <= 576)
    int gc_slo = 1;            // peeled mixed contraction: S_j as hi/lo (1, three MFMAs) or in FP16 (0, two)
    // mixed path: the gradient's z^T tBeta and C1a products in FP32 (0: FP64)
    int tn_f32 = 1;
    float *tn_b32 = nullptr;   // (P + M) x J: scaled whitened beta-tilde, GP rows of beta-tilde
    float *tn_c1a32 = nullptr; // Npad x M: C1a in FP32
    // 1 while R, bvec, Oinvt and pwork = Psi^-1 bvec hold the likelihood
    // evaluation (gp_elbo / gp_ql_start) at the current parameters and factors,
    // so gp_grad takes beta-tilde from pwork instead of solving again; every
    // other state-changing entry point clears it
    int lkhd_fresh = 0;

    size_t tot_max = 0;        // N+p+(N+M+1)Q
    int timing = 0;
    gp_timer t_gram, t_chol, t_gc, t_kuf, t_elbo, t_grad;
    gp_timer t_gramk, t_gck;   // the main Gram / contraction kernel alone
    // recorded ELBO + gradient unit
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // precision, slot, gram_splits, gc_groups, gram_waves, gc_form, gc_mixed, gc_peel,
    // gc_wave_rows, gc_prefetch, gram_peel, gc_waves, chol1, cholsl_blk, tn_f32, gram_wide,
    // chol_nb, gc_slo, vtv_mfma
    int graph_key[GP_GRAPH_KEY] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    // objective and elapsed seconds after every accepted step of the last gp_fit
    std::vector<double> fit_obj, fit_sec;
    std::vector<int> fit_kind;           // 0 IRLS step, 1 quasi-likelihood start, 2 L-BFGS step
    std::vector<int> fit_step;           // per entry: outer iteration, inner step, line-search trials
    std::vector<double> fit_sr;          // per entry: the line search's final step fraction
    std::string err;
};

// For its scope, the context's stream and rocBLAS handle are the side ones,
// so the gpk_* launchers enqueue there.
struct SideStream {
    gp_ctx *c;
    explicit SideStream(gp_ctx *cc) : c(cc) { swap_in(); }
    ~SideStream() { swap_in(); }
    void swap_in() {
        hipStream_t s = c->stream;
        c->stream = c->stream2;
        c->stream2 = s;
        rocblas_handle b = c->blas;
        c->blas = c->blas2;
        c->blas2 = b;
    }
};

// ---------------- error handling ----------------
void gp_set_error(const std::string &msg);
#define GP_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            gp_set_error(std::string(#call) + ": " + hipGetErrorString(e_));      \
            return (int)e_;                                                       \
        }                                                                         \
    } while (0)
#define GP_BLAS(call)                                                             \
    do {                                                                          \
        rocblas_status s_ = (call);                                               \
        if (s_ != rocblas_status_success) {                                       \
            gp_set_error(std::string(#call) + ": rocblas status " +               \
                         std::to_string((int)s_));                                \
            return 900 + (int)s_;                                                 \
        }                                                                         \
    } while (0)
#define GP_TRY(call)                                                              \
    do {                                                                          \
        int r_ = (call);                                                          \
        if (r_ > 0) return r_;                                                    \
    } while (0)

// ---------------- launch wrappers (gp_kernels.hip) ----------------
int gpk_kuf(gp_ctx *c);                       // Knm into Z, K, Ksym, Kchol
int gpk_expand_delta(gp_ctx *c);              // DeltaInv and its factor DeltaChol (from Kchol)
int gpk_oinvt(gp_ctx *c);
int gpk_titsias(gp_ctx *c);                   // A = Knm Kinv (rocBLAS) + titsias
int gpk_w_nu(gp_ctx *c);                      // IRLS W / Nu from ZB = Z*Beta
int gpk_r_and_sums(gp_ctx *c, const double *Zsrc);  // R = W nu / Omega, gq, gs, b = Zsrc^T R
// R, gq and gs (the bits of gpk_r_and_sums), and with w32 the mixed Gram's
// FP32 weights and scales (gpk_wprep's values), in one pass over W, nu, Omega
int gpk_r_sums(gp_ctx *c, bool w32);
int gpk_init_y0(gp_ctx *c);                   // R = log((y+0.05)/ls) dense
int gpk_beta_phi(gp_ctx *c, int jtr);
int gpk_beta_phi_in(gp_ctx *c, int jtr, double *Beta);     // damped update of Beta, phi
int gpk_beta_whiten(gp_ctx *c, const double *B, double *Bw);
int gpk_b_whiten(gp_ctx *c);                  // bvec's GP rows <- L^-1 bvec (reference -> whitened basis)
int gpk_beta_unwhiten(gp_ctx *c, const double *Bw, double *B);
int gpk_lkhd_pois(gp_ctx *c);
int gpk_lkhd_genes_with(gp_ctx *c, const double *tb, const double *Dchol);  // tb = Psi_j^{-1} b_j
int gpk_collect(gp_ctx *c, int Nterm, int slot, double *obj);
int gpk_grad_delta(gp_ctx *c, int whitened);
int gpk_whiten(gp_ctx *c);                    // Linv, Zw from the Kmm factor
int gpk_unwhiten_beta(gp_ctx *c);             // tBeta from tBw
int gpk_unwhiten_u(gp_ctx *c);                // U from Uw
int gpk_tn(gp_ctx *c);
int gpk_cell_grad(gp_ctx *c);
int gpk_c2(gp_ctx *c, int whitened);
int gpk_grad_tau_rho(gp_ctx *c);
int gpk_lbfgs(gp_ctx *c, int itr, int rhoflag);
int gpk_downdate(gp_ctx *c, int itr, double r, int rhoflag);
int gpk_beta_from_tilde(gp_ctx *c);
int gpk_fill(gp_ctx *c, double *p, size_t n, double v);
int gpk_fill_int(gp_ctx *c, int *p, size_t n, int v);   // a kernel, not a memset (graph-safe)
// C (M x N, ld ldc) = A^T B with A K x M (lda), B K x N (ldb), K large and
// M x N small: split over K into strided-batched GEMMs, partials summed in FP64.
int gpk_gemm_tn_splitk(gp_ctx *c, int M, int N, int K, const double *A, int lda,
                       const double *B, int ldb, double *C, int ldc);

// ---------------- gp_chol.hip ----------------
// Batched in-place Cholesky of `batch` n x n column-major matrices spaced
// `stride` doubles apart; reads the lower triangle, writes L to lower+diag,
// zero-fills a matrix that meets a non-positive pivot and counts it in *info.
int gpk_chol_batched(gp_ctx *c, double *A, int n, int batch, size_t stride, int *info);
// x_j = A_j^{-1} b_j for factored A_j (stride 0 = one shared factor).
int gpk_cholsl_batched(gp_ctx *c, const double *L, int n, size_t lstride,
                       const double *b, double *x, int batch);
// S_j = A_j^{-1} (full symmetric) from the factors.
int gpk_chol_inverse_batched(gp_ctx *c, const double *L, int n, int batch,
                             size_t stride, double *S);
// V = L^{-1} (lower triangular) of one factored n x n matrix.
int gpk_tri_inverse(gp_ctx *c, const double *L, int n, double *V);
// S_j = V_j^T V_j (full symmetric) for lower-triangular V_j.
int gpk_vtv(gp_ctx *c, const double *V, int n, int batch, size_t stride, double *S);

// ---------------- gp_gram.hip ----------------
// Psi_j = DeltaInv + Z^T diag(W_j / Omega) Z, full symmetric, all genes.
// W == nullptr: unweighted single Gram (initPsi).
int gpk_gram(gp_ctx *c, const double *W, double *Psi, int batch);
// the same from an explicit model matrix and prior (whitened basis)
int gpk_gram_src(gp_ctx *c, const double *W, double *Psi, int batch, const double *Zsrc,
                 const double *Delta);

// ---------------- gp_gram_mixed.hip ----------------
int gpk_split_z(gp_ctx *c, const double *Zsrc);
// weights_ready: w32 / wsc / wun already hold W/Omega for these genes (gpk_wprep)
int gpk_gram_mixed(gp_ctx *c, const double *W, double *Psi, int batch, const double *Zsrc,
                   const double *Delta, bool weights_ready = false);
// w32 = W/Omega in FP32 and the per-gene power-of-two scales (all genes)
int gpk_wprep(gp_ctx *c, const double *W);
// allocates the mixed Gram's buffers (FP16/FP32 model matrix, FP32 weights)
int gpk_mixed_alloc(gp_ctx *c);
// fused IRLS step on the mixed path: eta = Z beta, then one kernel that builds
// W / Omega, nu and R = W nu / Omega from eta and accumulates the whitened
// Gram; writes Psi (whitened, unfactored), R, w32, gq, gs and lkhds
bool gpk_gram_irls_fits(const gp_ctx *c);
int gpk_gram_irls(gp_ctx *c, double *Psi, const double *Delta);
// W and Nu of the last fused IRLS step, in the reference layout
int gpk_irls_wnu(gp_ctx *c);
// writes W and Nu if a fused IRLS step left them pending (every entry point
// that may read them or change what they follow from calls it first)
int gp_wnu_sync(gp_ctx *c);

// ---------------- gp_gradcontract_mixed.hip ----------------
// Uw = sum_j diag(W_j/Omega) Zw Sinv_j on FP16 hi/lo MFMA (whitened basis)
int gpk_gradcontract_mixed(gp_ctx *c);

// ---------------- gp_gradcontract.hip ----------------
// U = sum_j diag(W_j/Omega) Zsrc Sinv_j  (Npad x P)
int gpk_gradcontract(gp_ctx *c, const double *Zsrc, double *Uout);

// ---------------- gp_score.hip ----------------
// doubles of partial sums per chunk of tests in flight, and the widest design
size_t gpk_score_part_doubles(int P);
int gpk_score_max_columns();
int gpk_score_clear(gp_ctx *c, int *err);      // *err = 0 (a kernel, not a memset)
// score statistics of T dosage tests, every pointer on the device; part holds
// fly * gpk_score_part_doubles(P) doubles; *err set to 1 on an index out of range
int gpk_score_run_dosage(gp_ctx *c, int S, const unsigned char *G, int T, const int *gene,
                         const int *snp, double *stat, double *part, int fly, int *err);

// ---------------- gp_batched.hip ----------------
struct gp_bfit;
extern "C" int gp_bfit_destroy(gp_bfit *b);

// ---------------- timing ----------------
void gp_timer_begin(gp_ctx *c, gp_timer &t);
void gp_timer_end(gp_ctx *c, gp_timer &t);
