// gp_ctx.hip -- context, C ABI and the launch groups of the reference driver.
#include "gp_internal.h"
#include "../../include/gaspacho2.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

static thread_local std::string g_err;

void gp_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *gp_last_error(void) { return g_err.c_str(); }
extern "C" int gp_version(void) { return 1 * 10000 + 0 * 100 + 0; }

// ------------------------------------------------------------------ timing
void gp_timer_begin(gp_ctx *c, gp_timer &t) {
    if (!c->timing || c->capturing) return;
    if (!t.a) { (void)hipEventCreate(&t.a); (void)hipEventCreate(&t.b); }
    (void)hipEventRecord(t.a, c->stream);
    t.open = true;
}
void gp_timer_end(gp_ctx *c, gp_timer &t) {
    if (!c->timing || !t.open) return;
    (void)hipEventRecord(t.b, c->stream);
    (void)hipEventSynchronize(t.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    t.total_ms += ms;
    t.count += 1;
    t.open = false;
}

// ------------------------------------------------------------------ memory
template <class T> static int dmalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        gp_set_error(std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                     " bytes: " + hipGetErrorString(e));
        return (int)e;
    }
    e = hipMemset(*p, 0, n * sizeof(T));
    if (e != hipSuccess) { gp_set_error("hipMemset failed"); return (int)e; }
    return 0;
}

// Buffers only the gradient uses (the explicit inverses, the contraction U /
// Uw, the C1 pieces, per-cell sums): allocated, zero-filled, on the first
// call that needs them, so that a context used for IRLS and score tests
// alone (the N = 1M, G = 5K, M = 512 genetic-mapping workload) does not
// hold them.
static int ensure_grad_buffers(gp_ctx *c) {
    if (c->Sinv) return 0;
    const size_t N = c->N, J = c->J, M = c->M, P = c->P, Np = c->Npad;
    int r = 0;
    if ((r = dmalloc(&c->Sinv, J * P * P)) || (r = dmalloc(&c->C1, Np * M)) ||
        (r = dmalloc(&c->U, Np * P)) || (r = dmalloc(&c->Uw, Np * P)) ||
        (r = dmalloc(&c->C1a, Np * M)) || (r = dmalloc(&c->AO, Np * M)) ||
        (r = dmalloc(&c->Ecell, N)) || (r = dmalloc(&c->Epart, (size_t)GP_TN_SPLITS * Np)) ||
        (r = dmalloc(&c->ci1, N)))
        return r;
    return 0;
}

static void drop_graph(gp_ctx *c);
static void free_all(gp_ctx *c) {
    drop_graph(c);
    void *bufs[] = {c->Yp, c->Yi, c->Yx, c->d_nh, c->Z, c->ls, c->Omega, c->delta, c->X, c->T,
                    c->rho, c->phi, c->zeta, c->K, c->Ksym, c->Kchol, c->Kinv, c->A, c->titsias,
                    c->DeltaInv, c->DeltaChol, c->W, c->Oinvt, c->Nu, c->Psi, c->Sinv, c->Beta,
                    c->tBeta, c->bvec, c->tBphi, c->lkhds, c->gq, c->gs, c->grad, c->grad0, c->ss,
                    c->SS, c->XX, c->lbq, c->gradDeltaTmp, c->gradLogRhoTmp, c->C1, c->C2, c->U,
                    c->ZB, c->R, c->C1a, c->AO, c->G2, c->B2, c->S2, c->Ecell, c->Epart, c->ci1, c->red, c->scal,
                    c->lkhd, c->info, c->Zh, c->Zl, c->z32, c->gram_part, c->pwork, c->w32, c->wsc,
                    c->wun, c->zun, c->Linv, c->LinvT, c->Zw, c->DeltaW, c->DeltaWChol, c->Uw, c->tBw,
                    c->Vtmp, c->Sh, c->Sl, c->gc_perm, c->gc_aexp, c->gc_sexp, c->gc_bexp,
                    c->gc_gun, c->gc_part, c->gc_zexp, c->skw, c->gc_kappa, c->gc_t0, c->gc_s0,
                    c->gc_y, c->gram_row0, c->irls_cc, c->irls_segptr, c->irls_sums,
                    c->irls_segun, c->tn_b32, c->tn_c1a32};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    gp_timer *ts[] = {&c->t_gram, &c->t_chol, &c->t_gc, &c->t_kuf, &c->t_elbo, &c->t_grad,
                      &c->t_gramk, &c->t_gck};
    for (gp_timer *t : ts) {
        if (t->a) (void)hipEventDestroy(t->a);
        if (t->b) (void)hipEventDestroy(t->b);
    }
    if (c->blas) rocblas_destroy_handle(c->blas);
    if (c->blas2) rocblas_destroy_handle(c->blas2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_efork) (void)hipEventDestroy(c->ev_efork);
    if (c->ev_ejoin) (void)hipEventDestroy(c->ev_ejoin);
    if (c->ev_bdone) (void)hipEventDestroy(c->ev_bdone);
    if (c->ev_rfork) (void)hipEventDestroy(c->ev_rfork);
    if (c->ev_row0) (void)hipEventDestroy(c->ev_row0);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// ------------------------------------------------------------------ create
extern "C" int gp_create(gp_ctx **out, int N, int J, int M, int Q, int p, const int *nh,
                         const int *Yp, const int *Yi, const double *Yx, long long L,
                         const double *Zcov, const double *ls, const double *X0,
                         const double *T0, int ard_only, int precision) {
    if (!out) { gp_set_error("out is NULL"); return 1; }
    *out = nullptr;
    if (N <= 0 || J <= 0 || M <= 0 || Q <= 0 || p < 2 || !nh) { gp_set_error("bad dimensions"); return 1; }
    if (Q > 64) { gp_set_error("Q > 64 unsupported"); return 1; }
    if (nh[0] != 0 || nh[p] - nh[p - 1] != M) { gp_set_error("nh must end with the M-column Knm block"); return 1; }
    for (int l = 0; l < p; l++)
        if (nh[l + 1] <= nh[l]) { gp_set_error("nh must be increasing"); return 1; }
    if (precision < GP_PREC_FP64 || precision > GP_PREC_FP64_WHITENED) { gp_set_error("bad precision"); return 1; }
    if (L < 0 || (L > 0 && (!Yi || !Yx)) || !Yp || !ls || !X0 || !T0 || !Zcov) {
        gp_set_error("NULL input array"); return 1;
    }
    if (Yp[0] != 0 || Yp[J] != L) { gp_set_error("Yp must start at 0 and end at L"); return 1; }
    for (int j = 0; j < J; j++) {
        if (Yp[j + 1] < Yp[j]) { gp_set_error("Yp must be non-decreasing"); return 1; }
        for (long long l = Yp[j]; l < Yp[j + 1]; l++) {
            if (Yi[l] < 0 || Yi[l] >= N) { gp_set_error("row index out of range"); return 1; }
            if (l > Yp[j] && Yi[l] <= Yi[l - 1]) { gp_set_error("row indices must ascend within a gene"); return 1; }
        }
    }
    gp_ctx *c = new gp_ctx();
    c->N = N; c->J = J; c->M = M; c->Q = Q; c->p = p; c->L = L;
    c->nh.assign(nh, nh + p + 1);
    c->P = nh[p];
    c->g0 = nh[p - 1];
    c->Npad = ((N + 255) / 256) * 256;
    c->ard_only = ard_only ? 1 : 0;
    c->precision = precision;
    const size_t Np = c->Npad, P = c->P;
    c->tot_max = (size_t)N + p + ((size_t)N + M + 1) * Q;
    int r = 0;
#define ALLOC(ptr, n) do { if ((r = dmalloc(&(ptr), (n))) != 0) { free_all(c); delete c; return r; } } while (0)
    hipError_t e = hipGetDevice(&c->dev);
    if (e != hipSuccess) { gp_set_error(std::string("hipGetDevice: ") + hipGetErrorString(e)); delete c; return (int)e; }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { gp_set_error("hipStreamCreate failed"); delete c; return (int)e; }
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) { gp_set_error("rocblas_create_handle failed"); free_all(c); delete c; return 950; }
    rocblas_set_stream(c->blas, c->stream);
    if ((e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_efork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_ejoin, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_bdone, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_rfork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_row0, hipEventDisableTiming)) != hipSuccess) {
        gp_set_error(std::string("side stream: ") + hipGetErrorString(e)); free_all(c); delete c; return (int)e;
    }
    if (rocblas_create_handle(&c->blas2) != rocblas_status_success) { gp_set_error("rocblas_create_handle failed"); free_all(c); delete c; return 950; }
    rocblas_set_stream(c->blas2, c->stream2);
    ALLOC(c->Yp, (size_t)J + 1);
    ALLOC(c->Yi, (size_t)L + 1);
    ALLOC(c->Yx, (size_t)L + 1);
    ALLOC(c->d_nh, (size_t)p + 1);
    ALLOC(c->Z, Np * P);
    ALLOC(c->ls, (size_t)N);
    ALLOC(c->Omega, 2 * (size_t)N);
    ALLOC(c->delta, 2 * (size_t)p);
    ALLOC(c->X, 2 * (size_t)N * Q);
    ALLOC(c->T, 2 * (size_t)M * Q);
    ALLOC(c->rho, 2 * (size_t)Q);
    ALLOC(c->phi, (size_t)J);
    ALLOC(c->zeta, P);
    ALLOC(c->K, (size_t)M * M);
    ALLOC(c->Ksym, (size_t)M * M);
    ALLOC(c->Kchol, (size_t)M * M);
    ALLOC(c->Kinv, (size_t)M * M);
    ALLOC(c->A, Np * M);
    ALLOC(c->titsias, (size_t)N);
    ALLOC(c->DeltaInv, P * P);
    ALLOC(c->DeltaChol, P * P);
    ALLOC(c->Linv, (size_t)M * M);
    ALLOC(c->LinvT, (size_t)M * M);
    ALLOC(c->Zw, Np * P);
    ALLOC(c->DeltaW, P * P);
    ALLOC(c->DeltaWChol, P * P);
    ALLOC(c->tBw, P * J);
    ALLOC(c->W, Np * J);
    ALLOC(c->Oinvt, (size_t)N);
    ALLOC(c->Nu, Np * J);
    ALLOC(c->Psi, (size_t)J * P * P);
    ALLOC(c->Beta, P * J);
    ALLOC(c->tBeta, P * J);
    ALLOC(c->bvec, P * J);
    ALLOC(c->tBphi, P * J);
    ALLOC(c->pwork, P * J);
    ALLOC(c->lkhds, (size_t)J);
    ALLOC(c->gq, (size_t)J);
    ALLOC(c->gs, (size_t)J);
    ALLOC(c->grad, c->tot_max);
    ALLOC(c->grad0, c->tot_max);
    ALLOC(c->ss, c->tot_max);
    ALLOC(c->SS, c->tot_max * GP_HIST);
    ALLOC(c->XX, c->tot_max * GP_HIST);
    ALLOC(c->lbq, c->tot_max);
    ALLOC(c->gradDeltaTmp, P * J);
    ALLOC(c->gradLogRhoTmp, (size_t)M * Q);
    ALLOC(c->C2, (size_t)M * M);
    ALLOC(c->ZB, Np * J);
    ALLOC(c->R, Np * J);
    ALLOC(c->G2, (size_t)M * M);
    ALLOC(c->B2, (size_t)M * M);
    ALLOC(c->S2, (size_t)M * M);
    ALLOC(c->red, 4096);
    ALLOC(c->scal, 64);
    ALLOC(c->lkhd, GP_LKHD_SLOTS);
    ALLOC(c->info, 16);
#undef ALLOC
#define UP(dst, src, n) do { e = hipMemcpy((dst), (src), (n), hipMemcpyHostToDevice); \
        if (e != hipSuccess) { gp_set_error(std::string("upload: ") + hipGetErrorString(e)); free_all(c); delete c; return (int)e; } } while (0)
    UP(c->Yp, Yp, sizeof(int) * ((size_t)J + 1));
    if (L > 0) { UP(c->Yi, Yi, sizeof(int) * (size_t)L); UP(c->Yx, Yx, sizeof(double) * (size_t)L); }
    UP(c->d_nh, nh, sizeof(int) * ((size_t)p + 1));
    e = hipMemcpy2D(c->Z, Np * sizeof(double), Zcov, (size_t)N * sizeof(double),
                    (size_t)N * sizeof(double), c->g0, hipMemcpyHostToDevice);
    if (e != hipSuccess) { gp_set_error("Zcov upload failed"); free_all(c); delete c; return (int)e; }
    UP(c->ls, ls, sizeof(double) * (size_t)N);
    UP(c->X, X0, sizeof(double) * (size_t)N * Q);
    UP(c->T, T0, sizeof(double) * (size_t)M * Q);
    // initParams init.cu:27-48
    {
        std::vector<double> d(2 * (size_t)p, 0.0), om(2 * (size_t)N, 0.0), ph(J, 1.0), rh(2 * (size_t)Q, 0.0);
        d[0] = 1.5;
        for (int i = 1; i < p - 1; i++) d[i] = 0.01;
        d[p - 1] = 0.3;
        for (int i = 0; i < N; i++) om[i] = 1.0;
        rh[0] = 5.0;
        for (int k = 1; k < Q; k++) rh[k] = 10.0;
        UP(c->delta, d.data(), sizeof(double) * d.size());
        UP(c->Omega, om.data(), sizeof(double) * om.size());
        UP(c->phi, ph.data(), sizeof(double) * ph.size());
        UP(c->rho, rh.data(), sizeof(double) * rh.size());
    }
#undef UP
    *out = c;
    return 0;
}

extern "C" int gp_destroy(gp_ctx *c) {
    if (!c) return 0;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return 0;
}

extern "C" int gp_set_precision(gp_ctx *c, int precision) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    c->lkhd_fresh = 0;
    if (precision < GP_PREC_FP64 || precision > GP_PREC_FP64_WHITENED) { gp_set_error("bad precision"); return 1; }
    c->precision = precision;
    return 0;
}

extern "C" int gp_set_option(gp_ctx *c, const char *name, int value) {
    if (!c || !name) { gp_set_error("NULL argument"); return 1; }
    c->lkhd_fresh = 0;
    const std::string n(name);
    if (n == "gram_splits") {
        if (value < 1 || value > 256) { gp_set_error("gram_splits must be in [1, 256]"); return 1; }
        c->gram_splits = value;
    } else if (n == "gc_groups") {
        if (value < 1 || value > 4096) { gp_set_error("gc_groups must be in [1, 4096]"); return 1; }
        c->gc_groups = value;
    } else if (n == "gram_waves") {
        if (value != 4 && value != 8 && value != 12) { gp_set_error("gram_waves must be 4, 8 or 12"); return 1; }
        c->gram_waves = value;
    } else if (n == "gc_form") {
        if (value < 0 || value > 3) { gp_set_error("gc_form must be 0, 1, 2 or 3"); return 1; }
        c->gc_form = value;
    } else if (n == "gc_mixed") {
        c->gc_mixed = value ? 1 : 0;
    } else if (n == "gc_peel") {
        c->gc_peel = value ? 1 : 0;
    } else if (n == "gram_peel") {
        c->gram_peel = value ? 1 : 0;
    } else if (n == "gc_wave_rows") {
        if (value != 8 && value != 4) { gp_set_error("gc_wave_rows must be 8 or 4"); return 1; }
        c->gc_wave_rows = value;
    } else if (n == "gc_waves") {
        if (value != 8 && value != 4) { gp_set_error("gc_waves must be 8 or 4"); return 1; }
        c->gc_waves = value;
    } else if (n == "gc_prefetch") {
        if (value != 0 && value != 1 && value != 3 && value != 7) { gp_set_error("gc_prefetch must be 0, 1, 3 or 7"); return 1; }
        c->gc_prefetch = value;
    } else if (n == "irls_mixed") {
        c->irls_mixed = value ? 1 : 0;
    } else if (n == "irls_fused") {
        c->irls_fused = value ? 1 : 0;
    } else if (n == "chol1") {
        c->chol1 = value ? 1 : 0;
    } else if (n == "cholsl_blk") {
        c->cholsl_blk = value ? 1 : 0;
    } else if (n == "tn_f32") {
        c->tn_f32 = value ? 1 : 0;
    } else if (n == "gram_wide") {
        c->gram_wide = value ? 1 : 0;
    } else if (n == "score_mfma") {
        c->score_mfma = value ? 1 : 0;
    } else if (n == "vtv_mfma") {
        c->vtv_mfma = value ? 1 : 0;
    } else if (n == "gc_slo") {
        c->gc_slo = value ? 1 : 0;
    } else if (n == "chol_nb") {
        if (value != 8 && value != 16) { gp_set_error("chol_nb must be 8 or 16"); return 1; }
        c->chol_nb = value;
    } else if (n == "autotune") {
        c->autotune = value ? 1 : 0;
    } else {
        gp_set_error("unknown option " + n);
        return 1;
    }
    return 0;
}

extern "C" int gp_get_option(gp_ctx *c, const char *name, int *value) {
    if (!c || !name || !value) { gp_set_error("NULL argument"); return 1; }
    const std::string n(name);
    if (n == "gram_splits") *value = c->gram_splits;
    else if (n == "gc_groups") *value = c->gc_groups;
    else if (n == "gc_mixed") *value = c->gc_mixed;
    else if (n == "gc_peel") *value = c->gc_peel;
    else if (n == "gram_peel") *value = c->gram_peel;
    else if (n == "gc_wave_rows") *value = c->gc_wave_rows;
    else if (n == "gc_prefetch") *value = c->gc_prefetch;
    else if (n == "gc_waves") *value = c->gc_waves;
    else if (n == "gram_waves") *value = c->gram_waves;
    else if (n == "gc_form") *value = c->gc_form;
    else if (n == "irls_mixed") *value = c->irls_mixed;
    else if (n == "irls_fused") *value = c->irls_fused;
    else if (n == "chol1") *value = c->chol1;
    else if (n == "cholsl_blk") *value = c->cholsl_blk;
    else if (n == "tn_f32") *value = c->tn_f32;
    else if (n == "gram_wide") *value = c->gram_wide;
    else if (n == "chol_nb") *value = c->chol_nb;
    else if (n == "gc_slo") *value = c->gc_slo;
    else if (n == "vtv_mfma") *value = c->vtv_mfma;
    else if (n == "score_mfma") *value = c->score_mfma;
    else if (n == "autotune") *value = c->autotune;
    else { gp_set_error("unknown option " + n); return 1; }
    return 0;
}

// ------------------------------------------------------------------ buffers
namespace {
enum Kind { DIRECT, ROWS_N, TRANS_A, W_OINVT, PSI2, DELTA_CHOL };
struct BufDesc {
    double *dev = nullptr;
    size_t len = 0;      // doubles in the reference layout
    Kind kind = DIRECT;
    int cols = 0;        // ROWS_N: number of columns
    bool writable = true;
};
BufDesc describe(gp_ctx *c, const char *nm) {
    const size_t N = c->N, J = c->J, M = c->M, Q = c->Q, P = c->P, p = c->p;
    BufDesc d;
    auto is = [&](const char *s) { return std::strcmp(nm, s) == 0; };
    if (is("Omega")) { d.dev = c->Omega; d.len = 2 * N; }
    else if (is("delta")) { d.dev = c->delta; d.len = 2 * p; }
    else if (is("X")) { d.dev = c->X; d.len = 2 * N * Q; }
    else if (is("T")) { d.dev = c->T; d.len = 2 * M * Q; }
    else if (is("rho")) { d.dev = c->rho; d.len = 2 * Q; }
    else if (is("phi")) { d.dev = c->phi; d.len = J; }
    else if (is("zeta")) { d.dev = c->zeta; d.len = P; }
    else if (is("ls")) { d.dev = c->ls; d.len = N; }
    else if (is("K")) { d.dev = c->K; d.len = M * M; }
    else if (is("Kinv")) { d.dev = c->Kchol; d.len = M * M; }
    else if (is("KinvExplicit")) { d.dev = c->Kinv; d.len = M * M; }
    else if (is("KinvKmn")) { d.dev = c->A; d.len = N * M; d.kind = TRANS_A; d.writable = false; }
    else if (is("titsias")) { d.dev = c->titsias; d.len = N; }
    else if (is("DeltaInv")) { d.dev = c->DeltaInv; d.len = P * P; }
    else if (is("DeltaChol")) { d.dev = c->DeltaChol; d.len = P * P; d.kind = DELTA_CHOL; d.writable = false; }
    else if (is("Z")) { d.dev = c->Z; d.len = N * P; d.kind = ROWS_N; d.cols = (int)P; }
    else if (is("W")) { d.dev = c->W; d.len = N * (J + 1); d.kind = W_OINVT; }
    else if (is("Nu")) { d.dev = c->Nu; d.len = N * J; d.kind = ROWS_N; d.cols = (int)J; }
    else if (is("Psi")) { d.dev = c->Psi; d.len = 2 * J * P * P; d.kind = PSI2; }
    else if (is("Beta")) { d.dev = c->Beta; d.len = P * J; }
    else if (is("tBeta")) { d.dev = c->tBeta; d.len = P * J; }
    else if (is("lkhds")) { d.dev = c->lkhds; d.len = J; }
    else if (is("grad")) { d.dev = c->grad; d.len = c->tot_max; }
    else if (is("grad0")) { d.dev = c->grad0; d.len = c->tot_max; }
    else if (is("ss")) { d.dev = c->ss; d.len = c->tot_max; }
    else if (is("SS")) { d.dev = c->SS; d.len = c->tot_max * GP_HIST; }
    else if (is("XX")) { d.dev = c->XX; d.len = c->tot_max * GP_HIST; }
    else if (is("gradDeltaTmp")) { d.dev = c->gradDeltaTmp; d.len = P * J; }
    else if (is("C1")) { d.dev = c->C1; d.len = N * M; d.kind = ROWS_N; d.cols = (int)M; }
    else if (is("C2")) { d.dev = c->C2; d.len = M * M; }
    else if (is("U")) { d.dev = c->U; d.len = N * P; d.kind = ROWS_N; d.cols = (int)P; d.writable = false; }
    else if (is("Uw")) { d.dev = c->Uw; d.len = N * P; d.kind = ROWS_N; d.cols = (int)P; d.writable = false; }
    else if (is("lkhd")) { d.dev = c->lkhd; d.len = GP_LKHD_SLOTS; }
    return d;
}
}  // namespace

extern "C" size_t gp_buffer_len(gp_ctx *c, const char *name) {
    if (!c || !name) return 0;
    return describe(c, name).len;
}

extern "C" int gp_get_buffer(gp_ctx *c, const char *name, double *host, size_t n) {
    if (!c || !name || !host) { gp_set_error("NULL argument"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    if (!std::strcmp(name, "Psi") || !std::strcmp(name, "C1") || !std::strcmp(name, "U") ||
        !std::strcmp(name, "Uw"))
        GP_TRY(ensure_grad_buffers(c));
    BufDesc d = describe(c, name);
    if (!d.dev) { gp_set_error(std::string("unknown buffer ") + name); return 1; }
    if (n != d.len) { gp_set_error(std::string("length mismatch for ") + name); return 1; }
    GP_HIP(hipStreamSynchronize(c->stream));
    const size_t N = c->N, Np = c->Npad, M = c->M, P = c->P, J = c->J;
    switch (d.kind) {
    case DIRECT:
        GP_HIP(hipMemcpy(host, d.dev, n * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case ROWS_N:
        GP_HIP(hipMemcpy2D(host, N * sizeof(double), d.dev, Np * sizeof(double), N * sizeof(double),
                           d.cols, hipMemcpyDeviceToHost));
        break;
    case W_OINVT:
        GP_HIP(hipMemcpy2D(host, N * sizeof(double), c->W, Np * sizeof(double), N * sizeof(double),
                           J, hipMemcpyDeviceToHost));
        GP_HIP(hipMemcpy(host + N * J, c->Oinvt, N * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case TRANS_A: {
        std::vector<double> tmp(N * M);
        GP_HIP(hipMemcpy2D(tmp.data(), N * sizeof(double), c->A, Np * sizeof(double),
                           N * sizeof(double), M, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++)
            for (size_t m = 0; m < M; m++) host[m + i * M] = tmp[i + m * N];
        break;
    }
    case PSI2:
        GP_HIP(hipMemcpy(host, c->Psi, J * P * P * sizeof(double), hipMemcpyDeviceToHost));
        GP_HIP(hipMemcpy(host + J * P * P, c->Sinv, J * P * P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case DELTA_CHOL: {
        std::vector<double> up(P * P);
        GP_HIP(hipMemcpy(host, c->DeltaChol, P * P * sizeof(double), hipMemcpyDeviceToHost));
        GP_HIP(hipMemcpy(up.data(), c->DeltaInv, P * P * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t b = 0; b < P; b++)
            for (size_t a = 0; a < b; a++) host[a + b * P] = up[a + b * P];
        break;
    }
    }
    return 0;
}

extern "C" int gp_set_buffer(gp_ctx *c, const char *name, const double *host, size_t n) {
    if (!c || !name || !host) { gp_set_error("NULL argument"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    if (!std::strcmp(name, "Psi") || !std::strcmp(name, "C1") || !std::strcmp(name, "U") ||
        !std::strcmp(name, "Uw"))
        GP_TRY(ensure_grad_buffers(c));
    c->lkhd_fresh = 0;
    c->w32_valid = 0;
    BufDesc d = describe(c, name);
    if (!d.dev) { gp_set_error(std::string("unknown buffer ") + name); return 1; }
    if (!d.writable) { gp_set_error(std::string("buffer is read-only: ") + name); return 1; }
    if (n != d.len) { gp_set_error(std::string("length mismatch for ") + name); return 1; }
    GP_HIP(hipStreamSynchronize(c->stream));
    const size_t N = c->N, Np = c->Npad, P = c->P, J = c->J;
    switch (d.kind) {
    case DIRECT:
        GP_HIP(hipMemcpy(d.dev, host, n * sizeof(double), hipMemcpyHostToDevice));
        break;
    case ROWS_N:
        GP_HIP(hipMemcpy2D(d.dev, Np * sizeof(double), host, N * sizeof(double), N * sizeof(double),
                           d.cols, hipMemcpyHostToDevice));
        break;
    case W_OINVT:
        GP_HIP(hipMemcpy2D(c->W, Np * sizeof(double), host, N * sizeof(double), N * sizeof(double),
                           J, hipMemcpyHostToDevice));
        GP_HIP(hipMemcpy(c->Oinvt, host + N * J, N * sizeof(double), hipMemcpyHostToDevice));
        break;
    case PSI2:
        GP_HIP(hipMemcpy(c->Psi, host, J * P * P * sizeof(double), hipMemcpyHostToDevice));
        GP_HIP(hipMemcpy(c->Sinv, host + J * P * P, J * P * P * sizeof(double), hipMemcpyHostToDevice));
        break;
    default:
        gp_set_error("buffer not settable");
        return 1;
    }
    return 0;
}

// W and Nu of a fused IRLS step, written from the eta and per-cell constants
// it used (ZB and irls_cc are not changed before this runs: every entry point
// that could change them, or read W / Nu, calls it first)
int gp_wnu_sync(gp_ctx *c) {
    if (!c->wnu_pending) return 0;
    c->wnu_pending = 0;
    return gpk_irls_wnu(c);
}

// ------------------------------------------------------------------ groups
static int read_info(gp_ctx *c, int *bad) {
    int h = 0;
    *bad = 0;
    if (c->capturing) return 0;
    GP_HIP(hipMemcpyAsync(&h, c->info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    *bad = h;
    return 0;
}
static int clear_info(gp_ctx *c) { return gpk_fill_int(c, c->info, 1, 0); }

// Kernel matrices: getKernelMats, cholKmm, explicit Kmm^{-1}, getKinvKmn.
static int kernel_mats(gp_ctx *c, int (*after_kuf)(gp_ctx *) = nullptr) {
    gp_timer_begin(c, c->t_kuf);
    GP_TRY(gpk_kuf(c));
    if (after_kuf) GP_TRY(after_kuf(c));
    c->zsplit_valid = 0;
    c->zw_valid = 0;
    GP_TRY(gpk_chol_batched(c, c->Kchol, c->M, 1, 0, c->info + 1));
    // Kmm^-1 = L^-T L^-1; L^-1 is kept for the whitened basis
    GP_TRY(gpk_tri_inverse(c, c->Kchol, c->M, c->Linv));
    GP_TRY(gpk_vtv(c, c->Linv, c->M, 1, 0, c->Kinv));
    GP_TRY(gpk_titsias(c));
    gp_timer_end(c, c->t_kuf);
    return 0;
}

// Weighted Grams + batched Cholesky.  mixed: FP16 hi/lo MFMA Gram in the
// whitened basis [covariates | Knm L^-T] with the diagonal prior DeltaW, whose
// smallest eigenvalue is min(1/delta) -- well conditioned for the FP32-accumulated
// Gram; otherwise the FP64 Gram in the reference basis.
static int gram_chol(gp_ctx *c, int precision, bool weights_ready = false) {
    gp_timer_begin(c, c->t_gram);
    if (precision != GP_PREC_FP64) {
        if (!c->zw_valid) {
            GP_TRY(gpk_whiten(c));
            c->zw_valid = 1;
            c->zsplit_valid = 0;
        }
        if (precision == GP_PREC_MIXED)
            GP_TRY(gpk_gram_mixed(c, c->W, c->Psi, c->J, c->Zw, c->DeltaW, weights_ready));
        else
            GP_TRY(gpk_gram_src(c, c->W, c->Psi, c->J, c->Zw, c->DeltaW));
        c->psi_whitened = 1;
    } else {
        GP_TRY(gpk_gram(c, c->W, c->Psi, c->J));
        c->psi_whitened = 0;
    }
    gp_timer_end(c, c->t_gram);
    gp_timer_begin(c, c->t_chol);
    GP_TRY(gpk_chol_batched(c, c->Psi, c->P, c->J, (size_t)c->P * c->P, c->info));
    gp_timer_end(c, c->t_chol);
    return 0;
}

extern "C" int gp_init_fit(gp_ctx *c) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;
    c->w32_valid = 0;
    GP_TRY(clear_info(c));
    GP_TRY(kernel_mats(c));
    GP_TRY(gpk_expand_delta(c));
    // initPsi + initPsi2: unweighted Gram in the first slot, factored
    GP_TRY(gpk_gram(c, nullptr, c->Psi, 1));
    GP_TRY(gpk_chol_batched(c, c->Psi, c->P, 1, 0, c->info));
    // initBeta
    GP_TRY(gpk_init_y0(c));
    GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, 0, c->bvec, c->Beta, c->J));
    int bad = 0;
    GP_TRY(read_info(c, &bad));
    return bad ? -bad : 0;
}

extern "C" int gp_irls_step(gp_ctx *c, int jtr, double *obj) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    c->lkhd_fresh = 0;
    c->w32_valid = 0;
    if (jtr < 0 || jtr >= GP_LKHD_SLOTS) { gp_set_error("jtr out of range"); return 1; }
    GP_TRY(clear_info(c));
    // the step writes W and Nu itself (or leaves them pending again)
    c->wnu_pending = 0;
    if (c->precision == GP_PREC_MIXED && c->irls_mixed && c->irls_fused && gpk_gram_irls_fits(c)) {
        // getWForBeta + getPsi in one kernel on the whitened FP16 hi/lo Gram:
        // it builds W / Omega, nu and R = W nu / Omega from eta = Z beta,
        // accumulates the Gram and the sums gq, gs and the GLM objective
        // (getBetaPhi / getLkhdPois), and writes R and the FP32 weights
        // instead of W and nu; those are written only when something reads
        // them (gp_wnu_sync).  The rest is the whitened step below.
        if (!c->zw_valid) {
            GP_TRY(gpk_whiten(c));
            c->zw_valid = 1;
            c->zsplit_valid = 0;
        }
        gp_timer_begin(c, c->t_gram);
        GP_TRY(gpk_gram_irls(c, c->Psi, c->DeltaW));
        gp_timer_end(c, c->t_gram);
        c->psi_whitened = 1;
        gp_timer_begin(c, c->t_chol);
        GP_TRY(gpk_chol_batched(c, c->Psi, c->P, c->J, (size_t)c->P * c->P, c->info));
        gp_timer_end(c, c->t_chol);
        // b_j = Zw^T R_j (getZtOinvjnj, util_cuda.cu:462-487)
        GP_TRY(gpk_gemm_tn_splitk(c, c->P, c->J, c->Npad, c->Zw, c->Npad, c->R, c->Npad, c->bvec, c->P));
        GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, (size_t)c->P * c->P, c->bvec, c->tBphi, c->J));
        GP_TRY(gpk_beta_whiten(c, c->Beta, c->tBw));
        GP_TRY(gpk_beta_phi_in(c, jtr, c->tBw));
        GP_TRY(gpk_beta_unwhiten(c, c->tBw, c->Beta));
        GP_TRY(gpk_collect(c, 0, jtr, obj));
        c->wnu_pending = 1;
        int bad = 0;
        GP_TRY(read_info(c, &bad));
        return bad ? -bad : 0;
    }
    GP_TRY(gpk_w_nu(c));
    if (c->precision == GP_PREC_MIXED && c->irls_mixed) {
        // Same step in the whitened basis T = blockdiag(I, L^-T): Psi_w = T^T Psi T
        // on the FP16 hi/lo Gram, sol_w = Psi_w^-1 Zw^T R = T^-1 sol, and the damped
        // update (linear) and phi (tb^T Psi tb = tb_w^T Psi_w tb_w) are basis-free.
        // beta stays in the reference basis between steps; zeta is zero (it is
        // uninitialised in the reference and read as 0), so adding it in either
        // basis is the same.
        GP_TRY(gram_chol(c, GP_PREC_MIXED));
        GP_TRY(gpk_r_and_sums(c, c->Zw));
        GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, (size_t)c->P * c->P, c->bvec, c->tBphi, c->J));
        GP_TRY(gpk_beta_whiten(c, c->Beta, c->tBw));
        GP_TRY(gpk_beta_phi_in(c, jtr, c->tBw));
        GP_TRY(gpk_beta_unwhiten(c, c->tBw, c->Beta));
    } else {
        GP_TRY(gram_chol(c, GP_PREC_FP64));   // the reference basis, FP64
        GP_TRY(gpk_r_and_sums(c, c->Z));
        GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, (size_t)c->P * c->P, c->bvec, c->tBphi, c->J));
        GP_TRY(gpk_beta_phi(c, jtr));
    }
    GP_TRY(gpk_lkhd_pois(c));
    GP_TRY(gpk_collect(c, 0, jtr, obj));
    int bad = 0;
    GP_TRY(read_info(c, &bad));
    return bad ? -bad : 0;
}

// getLkhd: b_j, beta-tilde_j (scratch), the factor diagonals and the sums.
// r_ready: R, gq and gs already hold the current W, nu and Omega (the ELBO's
// side stream wrote them), so only b_j = Z^T R_j is left; b_ready: b is
// computed too (the ELBO's side stream, joined before the solve).
static int lkhd_genes(gp_ctx *c, bool r_ready = false, bool b_ready = false) {
    const bool w = c->psi_whitened;
    if (b_ready) {
        // nothing left before the solve
    } else if (r_ready) {
        GP_TRY(gpk_gemm_tn_splitk(c, c->P, c->J, c->Npad, w ? c->Zw : c->Z, c->Npad, c->R, c->Npad,
                                  c->bvec, c->P));
    } else {
        GP_TRY(gpk_r_and_sums(c, w ? c->Zw : c->Z));
    }
    GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, (size_t)c->P * c->P, c->bvec, c->pwork, c->J));
    GP_TRY(gpk_lkhd_genes_with(c, c->pwork, w ? c->DeltaWChol : c->DeltaChol));
    c->lkhd_fresh = 1;
    return 0;
}

extern "C" int gp_ql_start(gp_ctx *c, double *obj) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;   // set again by the likelihood evaluation below
    GP_TRY(clear_info(c));
    GP_TRY(gpk_oinvt(c));
    GP_TRY(lkhd_genes(c));
    GP_TRY(gpk_collect(c, c->N, 0, obj));
    return 0;
}

// The ELBO's side chain (see gp_elbo): Oinvt, and the likelihood's R, gq
// and gs together with (mixed) the Gram's FP32 weights, on the side stream
// from the point this is called.
static int fork_weights_impl(gp_ctx *c, bool mixed) {
    if (mixed) GP_TRY(gpk_mixed_alloc(c));   // allocations first (no stream work)
    GP_HIP(hipEventRecord(c->ev_efork, c->stream));
    GP_HIP(hipStreamWaitEvent(c->stream2, c->ev_efork, 0));
    c->efork_pending = 1;
    {
        SideStream side(c);
        GP_TRY(gpk_oinvt(c));
        GP_TRY(gpk_r_sums(c, mixed));
    }
    GP_HIP(hipEventRecord(c->ev_ejoin, c->stream2));    // what the Gram reads
    // then the likelihood's b = Z^T R in the reference basis (Knm is in Z
    // from k_kuf on): beside the Kmm chain and the Gram's first workgroups
    // instead of after the batched Cholesky; joined before the solve
    {
        SideStream side(c);
        GP_TRY(gpk_gemm_tn_splitk(c, c->P, c->J, c->Npad, c->Z, c->Npad, c->R, c->Npad, c->bvec, c->P));
    }
    GP_HIP(hipEventRecord(c->ev_bdone, c->stream2));
    c->bfork_pending = 1;
    return 0;
}
static int fork_weights(gp_ctx *c) { return fork_weights_impl(c, false); }
static int fork_weights_mixed(gp_ctx *c) { return fork_weights_impl(c, true); }

extern "C" int gp_elbo(gp_ctx *c, int slot, double *obj) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    if (!c->capturing) GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;   // set again by the likelihood evaluation below
    if (slot < 0 || slot >= GP_LKHD_SLOTS) { gp_set_error("slot out of range"); return 1; }
    gp_timer_begin(c, c->t_elbo);
    GP_TRY(clear_info(c));
    // the side stream's b product is joined before the solve, or on an early
    // error return, so no side work outlives the call unjoined
    struct BJoin {
        gp_ctx *c;
        ~BJoin() {
            if (c->bfork_pending) (void)hipStreamWaitEvent(c->stream, c->ev_bdone, 0);
            c->bfork_pending = 0;
        }
    } bjoin{c};
    {
        // Oinvt, the likelihood's R = W nu / Omega with its sums and (mixed)
        // the FP32 weights W/Omega depend on neither the kernel matrices nor
        // the basis: forked to the side stream after the
        // HBM-bound Kuf, they run beside the single-matrix Kmm factor and its
        // small follow-ups, which leave the chip idle, and are joined before
        // the Gram (b = Z^T R follows them on the side stream); the join also
        // runs on an early error return
        const bool mixed = c->precision == GP_PREC_MIXED;
        struct Join {
            gp_ctx *c;
            ~Join() {
                if (c->efork_pending) (void)hipStreamWaitEvent(c->stream, c->ev_ejoin, 0);
                c->efork_pending = 0;
            }
        };
        {
            Join join{c};
            GP_TRY(kernel_mats(c, mixed ? fork_weights_mixed : fork_weights));
            GP_TRY(gpk_expand_delta(c));
        }
        GP_TRY(gram_chol(c, c->precision, mixed));
    }
    // b from the side stream, in the reference basis; the whitened basis
    // Zw = Z blockdiag(I, L^-T) takes b_w = blockdiag(I, L^-1) b
    GP_HIP(hipStreamWaitEvent(c->stream, c->ev_bdone, 0));
    c->bfork_pending = 0;
    if (c->psi_whitened) GP_TRY(gpk_b_whiten(c));
    GP_TRY(lkhd_genes(c, true, true));
    GP_TRY(gpk_collect(c, c->N, slot, obj));
    gp_timer_end(c, c->t_elbo);
    int bad = 0;
    GP_TRY(read_info(c, &bad));
    return bad ? -bad : 0;
}


// The gradient has two independent chains once tB is known: tBeta in the
// reference basis plus the per-cell tn pieces (ZB, TN, E, C1a), which need
// only tB, and the inverses, delta gradient and contraction.  The first runs
// on the side stream, forked here, beside the latency-bound inverse and
// contraction set-up kernels that leave most of the chip idle (run beside the
// MFMA-bound contraction kernel instead, it slowed that kernel by more than
// it saved).  It writes tBeta, ZB, R, Epart, Ecell and C1a, none of which the
// main chain reads; the cell gradient joins it.
static int fork_tn_chain(gp_ctx *c) {
    GP_HIP(hipEventRecord(c->ev_fork, c->stream));
    GP_HIP(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    {
        SideStream side(c);
        if (c->psi_whitened) GP_TRY(gpk_unwhiten_beta(c));
        GP_TRY(gpk_tn(c));
    }
    GP_HIP(hipEventRecord(c->ev_join, c->stream2));
    return 0;
}

extern "C" int gp_grad(gp_ctx *c) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    if (!c->capturing) GP_TRY(gp_wnu_sync(c));
    if (!c->capturing) GP_TRY(ensure_grad_buffers(c));
    gp_timer_begin(c, c->t_grad);
    GP_TRY(gpk_expand_delta(c));
    const bool w = c->psi_whitened;
    double *tB = w ? c->tBw : c->tBeta;
    // getTildeBeta (in the basis of the current factor): the likelihood
    // evaluation at these parameters already solved it into pwork (with R,
    // bvec and Oinvt current); otherwise solve it here
    if (c->lkhd_fresh) {
        GP_HIP(hipMemcpyAsync(tB, c->pwork, (size_t)c->P * c->J * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    } else {
        // b as the ELBO forms it (reference basis, then L^-1 on the GP rows),
        // so both paths give the same bits
        GP_TRY(gpk_oinvt(c));
        GP_TRY(gpk_r_and_sums(c, c->Z));
        if (w) GP_TRY(gpk_b_whiten(c));
        GP_TRY(gpk_cholsl_batched(c, c->Psi, c->P, (size_t)c->P * c->P, c->bvec, tB, c->J));
    }
    GP_TRY(fork_tn_chain(c));
    // solvePsi (explicit inverses, needed by the delta gradient, U and C2)
    GP_TRY(gpk_chol_inverse_batched(c, c->Psi, c->P, c->J, (size_t)c->P * c->P, c->Sinv));
    GP_TRY(gpk_grad_delta(c, w));
    gp_timer_begin(c, c->t_gc);
    if (w && c->precision == GP_PREC_MIXED && c->gc_mixed)
        GP_TRY(gpk_gradcontract_mixed(c));
    else
        GP_TRY(gpk_gradcontract(c, w ? c->Zw : c->Z, w ? c->Uw : c->U));
    gp_timer_end(c, c->t_gc);
    if (w) GP_TRY(gpk_unwhiten_u(c));
    GP_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    GP_TRY(gpk_cell_grad(c));
    GP_TRY(gpk_c2(c, w));
    GP_TRY(gpk_grad_tau_rho(c));
    gp_timer_end(c, c->t_grad);
    if (!c->capturing) GP_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ------------------------------------------------------------------ graph
static void drop_graph(gp_ctx *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    for (int &k : c->graph_key) k = -1;
}

// One ELBO + gradient unit (gp_elbo(slot) then gp_grad) recorded once as a
// HIP graph and replayed: every kernel reads the parameters from device
// memory, so a replay evaluates the unit at the current parameters.  The
// first call (and any call after the precision, the slot or a launch-shape
// option changed) runs the unit eagerly once -- which makes every lazy
// allocation -- and then records it.
extern "C" int gp_eval_graph(gp_ctx *c, int slot, double *obj) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    if (slot < 0 || slot >= GP_LKHD_SLOTS) { gp_set_error("slot out of range"); return 1; }
    const int key[GP_GRAPH_KEY] = {c->precision, slot, c->gram_splits, c->gc_groups, c->gram_waves,
                                   c->gc_form, c->gc_mixed, c->gc_peel, c->gc_wave_rows,
                                   c->gc_prefetch, c->gram_peel, c->gc_waves, c->chol1,
                                   c->cholsl_blk, c->tn_f32, c->gram_wide, c->chol_nb, c->gc_slo,
                                   c->vtv_mfma};
    if (!c->graph_exec || !std::equal(key, key + GP_GRAPH_KEY, c->graph_key)) {
        drop_graph(c);
        double tmp = 0.0;
        int r = gp_elbo(c, slot, &tmp);
        if (r > 0) return r;
        if ((r = gp_grad(c)) > 0) return r;
        GP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = 1;
        r = gp_elbo(c, slot, nullptr);
        if (r <= 0) r = gp_grad(c);
        c->capturing = 0;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r > 0) { if (g) (void)hipGraphDestroy(g); return r; }
        if (e != hipSuccess) {
            gp_set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            return (int)e;
        }
        c->graph = g;
        GP_HIP(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        std::copy(key, key + GP_GRAPH_KEY, c->graph_key);
    }
    GP_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    c->lkhd_fresh = 1;   // the replayed unit ran the likelihood at the current parameters
    if (obj) GP_HIP(hipMemcpyAsync(obj, c->lkhd + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    int bad = 0;
    GP_TRY(read_info(c, &bad));
    return bad ? -bad : 0;
}

extern "C" int gp_lbfgs_step(gp_ctx *c, int jtr, int rhoflag) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;
    c->w32_valid = 0;
    if (jtr < 0) { gp_set_error("jtr < 0"); return 1; }
    GP_TRY(gpk_lbfgs(c, jtr, rhoflag ? 1 : 0));
    return 0;
}

extern "C" int gp_downdate(gp_ctx *c, int jtr, double sr, int rhoflag) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;
    c->w32_valid = 0;
    GP_TRY(gpk_downdate(c, jtr, sr, rhoflag ? 1 : 0));
    return 0;
}

extern "C" int gp_end_outer(gp_ctx *c) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;
    GP_TRY(gpk_expand_delta(c));
    GP_TRY(gpk_beta_from_tilde(c));
    GP_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ggplvm.cu:1146-1330
extern "C" int gp_fit(gp_ctx *c, int itrmax, int jtr_cap, int verbose, gp_outer_cb on_outer,
                      void *user, double *obj) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    std::vector<double> lk(GP_LKHD_SLOTS, 0.0);
    const auto t0 = std::chrono::steady_clock::now();
    c->fit_obj.clear();
    c->fit_sec.clear();
    c->fit_kind.clear();
    c->fit_step.clear();
    c->fit_sr.clear();
    auto mark = [&](int kind, double v, int itr, int jtr, int ktr, double sr) {
        c->fit_obj.push_back(v);
        c->fit_sec.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        c->fit_kind.push_back(kind);
        c->fit_step.insert(c->fit_step.end(), {itr, jtr, ktr});
        c->fit_sr.push_back(sr);
    };
    int r = gp_init_fit(c);
    if (r > 0) return r;
    int rhoflag = 0;
    double last = 0.0;
    for (int itr = 0; itr < itrmax; itr++) {
        int JTRMAX = (itr == 0 ? 200 : 100);
        if (jtr_cap > 0 && JTRMAX > jtr_cap) JTRMAX = jtr_cap;
        for (int jtr = 0; jtr < JTRMAX; jtr++) {
            r = gp_irls_step(c, jtr, &lk[jtr]);
            if (r > 0) return r;
            mark(0, lk[jtr], itr, jtr, 0, 0.0);
            if (verbose) { printf("lkhd[%d,%d]=%lf\n", itr, jtr, lk[jtr]); fflush(stdout); }
            if (jtr > 0) {
                if (std::isnan(lk[jtr])) { gp_set_error("GLM objective is NaN"); if (obj) *obj = NAN; return 2; }
                if (jtr == JTRMAX - 1 || std::fabs(lk[jtr - 1] - lk[jtr]) / std::fabs(lk[jtr]) < 1e-7) break;
            }
        }
        r = gp_ql_start(c, &lk[0]);
        if (r > 0) return r;
        mark(1, lk[0], itr, 0, 0, 0.0);
        last = lk[0];
        if (verbose) { printf("lkhd[%d,0]=%lf\n", itr, lk[0]); fflush(stdout); }
        if (itr == 0 && c->autotune && c->precision == GP_PREC_MIXED) {
            r = gp_autotune(c, verbose);
            if (r > 0) return r;
        }
        if (itr > 20) rhoflag = 1;
        if (itr == 30) JTRMAX = 10000;
        if (jtr_cap > 0 && JTRMAX > jtr_cap) JTRMAX = jtr_cap;
        for (int jtr = 0; jtr < JTRMAX && jtr + 1 < GP_LKHD_SLOTS; jtr++) {
            if ((r = gp_grad(c)) > 0) return r;
            if ((r = gp_lbfgs_step(c, jtr, rhoflag)) > 0) return r;
            double sr = 1.0;
            int ktr;
            for (ktr = 1; ktr < 8; ktr++) {
                r = gp_elbo(c, jtr + 1, &lk[jtr + 1]);
                if (r > 0) return r;
                if (verbose) { printf("lkhd[%d,%d]=%lf\n", itr, jtr + 1, lk[jtr + 1]); fflush(stdout); }
                if (lk[jtr + 1] > lk[jtr] || ktr == 7) break;
                else if (std::isnan(lk[jtr + 1]) || std::isinf(lk[jtr + 1])) sr *= 0.01;
                else if (lk[jtr + 1] < lk[jtr]) sr *= 0.1;
                if ((r = gp_downdate(c, jtr, sr, rhoflag)) > 0) return r;
            }
            last = lk[jtr + 1];
            mark(2, last, itr, jtr, ktr, sr);
            if (jtr > 0 && std::fabs(lk[jtr + 1] - lk[jtr]) < std::fabs(lk[jtr + 1]) * 1e-8) break;
        }
        if (on_outer) on_outer(c, itr, last, user);
        if ((r = gp_end_outer(c)) > 0) return r;
    }
    if (obj) *obj = last;
    return 0;
}

extern "C" int gp_fit_trace_steps(gp_ctx *c, int *step, double *sr, int cap, int *len) {
    if (!c || !len) { gp_set_error("NULL argument"); return 1; }
    const int n = (int)c->fit_obj.size();
    *len = n;
    for (int k = 0; k < std::min(n, cap); k++) {
        if (step) std::copy(c->fit_step.begin() + 3 * k, c->fit_step.begin() + 3 * k + 3, step + 3 * k);
        if (sr) sr[k] = c->fit_sr[k];
    }
    return 0;
}

extern "C" int gp_fit_trace(gp_ctx *c, double *obj, double *seconds, int *kind, int cap, int *len) {
    if (!c || !len) { gp_set_error("NULL argument"); return 1; }
    const int n = (int)c->fit_obj.size();
    *len = n;
    for (int k = 0; k < std::min(n, cap); k++) {
        if (obj) obj[k] = c->fit_obj[k];
        if (seconds) seconds[k] = c->fit_sec[k];
        if (kind) kind[k] = c->fit_kind[k];
    }
    return 0;
}

// ------------------------------------------------------------------ autotuner
// Times the mixed Gram (per gram_splits) and the mixed contraction (per
// gc_groups) on the context's current data and keeps the fastest settings.
// Leaves Psi (factor) and Sinv recomputed at the current parameters in the
// whitened basis, as gp_elbo + gp_grad would.
namespace {
int time_ms(gp_ctx *c, int (*fn)(gp_ctx *), int reps, double *ms) {
    hipEvent_t a, b;
    GP_HIP(hipEventCreate(&a));
    GP_HIP(hipEventCreate(&b));
    GP_TRY(fn(c));                       // warm-up
    GP_HIP(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) GP_TRY(fn(c));
    GP_HIP(hipEventRecord(b, c->stream));
    GP_HIP(hipEventSynchronize(b));
    float t = 0.0f;
    GP_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / reps;
    return 0;
}
int tune_gram(gp_ctx *c) { return gpk_gram_mixed(c, c->W, c->Psi, c->J, c->Zw, c->DeltaW); }
int tune_gc(gp_ctx *c) { return gpk_gradcontract_mixed(c); }
}  // namespace

extern "C" int gp_autotune(gp_ctx *c, int verbose) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    GP_TRY(gp_wnu_sync(c));
    c->lkhd_fresh = 0;
    if (c->precision != GP_PREC_MIXED) { gp_set_error("gp_autotune needs GP_PREC_MIXED"); return 1; }
    GP_TRY(ensure_grad_buffers(c));
    static const int splits[] = {8, 16, 32};
    static const int groups[] = {16, 32, 64};
    static const int waves[] = {8, 12, 4};
    static const int forms[] = {2, 1, 0};
    GP_TRY(gpk_expand_delta(c));
    if (!c->zw_valid) {
        GP_TRY(gpk_whiten(c));
        c->zw_valid = 1;
        c->zsplit_valid = 0;
    }
    double best = 1e300;
    int bs = c->gram_splits, bw = c->gram_waves;
    for (int wv : waves)
        for (int sv : splits) {
            c->gram_splits = sv;
            c->gram_waves = wv;
            double ms = 0.0;
            GP_TRY(time_ms(c, tune_gram, 2, &ms));
            if (verbose) std::printf("autotune: gram_waves %d gram_splits %d: %.3f ms\n", wv, sv, ms);
            if (ms < best) { best = ms; bs = sv; bw = wv; }
        }
    c->gram_splits = bs;
    c->gram_waves = bw;
    // the contraction needs the inverses of the current factors
    GP_TRY(clear_info(c));
    GP_TRY(tune_gram(c));
    GP_TRY(gpk_chol_batched(c, c->Psi, c->P, c->J, (size_t)c->P * c->P, c->info));
    c->psi_whitened = 1;
    GP_TRY(gpk_chol_inverse_batched(c, c->Psi, c->P, c->J, (size_t)c->P * c->P, c->Sinv));
    best = 1e300;
    int bg = c->gc_groups, bf = c->gc_form;
    for (int fv : forms)
        for (int gv : groups) {
            c->gc_groups = gv;
            c->gc_form = fv;
            double ms = 0.0;
            GP_TRY(time_ms(c, tune_gc, 2, &ms));
            if (verbose) std::printf("autotune: gc_form %d gc_groups %d: %.3f ms\n", fv, gv, ms);
            if (ms < best) { best = ms; bg = gv; bf = fv; }
        }
    c->gc_groups = bg;
    c->gc_form = bf;
    GP_HIP(hipStreamSynchronize(c->stream));
    int bad = 0;
    GP_TRY(read_info(c, &bad));
    return bad ? -bad : 0;
}

// ------------------------------------------------------------------ timing API
extern "C" int gp_timing_enable(gp_ctx *c, int on) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    c->timing = on ? 1 : 0;
    return 0;
}
extern "C" int gp_timing_reset(gp_ctx *c) {
    if (!c) { gp_set_error("NULL context"); return 1; }
    gp_timer *ts[] = {&c->t_gram, &c->t_chol, &c->t_gc, &c->t_kuf, &c->t_elbo, &c->t_grad,
                      &c->t_gramk, &c->t_gck};
    for (gp_timer *t : ts) { t->total_ms = 0.0; t->count = 0; }
    return 0;
}
extern "C" int gp_timing_get(gp_ctx *c, const char *name, double *avg_ms, int *count) {
    if (!c || !name) { gp_set_error("NULL argument"); return 1; }
    gp_timer *t = nullptr;
    if (!std::strcmp(name, "gram")) t = &c->t_gram;
    else if (!std::strcmp(name, "chol")) t = &c->t_chol;
    else if (!std::strcmp(name, "gradcontract")) t = &c->t_gc;
    else if (!std::strcmp(name, "kuf")) t = &c->t_kuf;
    else if (!std::strcmp(name, "elbo")) t = &c->t_elbo;
    else if (!std::strcmp(name, "grad")) t = &c->t_grad;
    else if (!std::strcmp(name, "gram_kernel")) t = &c->t_gramk;
    else if (!std::strcmp(name, "gradcontract_kernel")) t = &c->t_gck;
    if (!t) { gp_set_error(std::string("unknown timer ") + name); return 1; }
    if (avg_ms) *avg_ms = t->count ? t->total_ms / t->count : 0.0;
    if (count) *count = t->count;
    return 0;
}
