// ggplvm_main.cpp -- host driver with the reference's command line and files
// (ggplvm.cu:951-1354), calling the library through the C ABI.
//
//   ggplvm Y.bin DataFrame.bin ls.bin Xinit.bin [-M m] [-Q q] [--init-params f]
//          [--itrmax n] [--jtr-cap n] [--precision f64|mixed] [--kernel per-ard|ard]
//
// Positional inputs: counts (SparseMat), covariates (DataFrame), size factors
// (N doubles), initial latent (N*Q doubles, column-major).  Defaults M=100,
// Q=51 (ggplvm.cu:975-976).  Outputs in the working directory: Tau.bin once,
// X9.gz / phi.gz / Params9.bin after every outer iteration, params.bin at the
// end (ggplvm.cu:1011-1014, 1302-1316, 1333-1346).  --itrmax / --jtr-cap /
// --precision / --kernel are additions (the reference hard-codes 5000 outer
// iterations, its inner schedule, FP64 and the periodic kernel).
#include "gaspacho2.h"
#include "gp_io.h"

#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Run {
    gp_ctx *ctx = nullptr;
    int N = 0, J = 0, M = 0, P = 0, p = 0, Q = 0;
};

bool get(gp_ctx *c, const char *name, std::vector<double> &out) {
    out.resize(gp_buffer_len(c, name));
    if (gp_get_buffer(c, name, out.data(), out.size()) != 0) {
        std::fprintf(stderr, "ggplvm: %s\n", gp_last_error());
        return false;
    }
    return true;
}

// util_cuda.cu:557-580 writeParams (with header) / ggplvm.cu:1339-1346 (without)
void write_params(const Run &r, const char *fname, bool header) {
    std::vector<double> om, de, x, t, rh, ph;
    if (!get(r.ctx, "Omega", om) || !get(r.ctx, "delta", de) || !get(r.ctx, "X", x) ||
        !get(r.ctx, "T", t) || !get(r.ctx, "rho", rh) || !get(r.ctx, "phi", ph))
        return;
    FILE *f = std::fopen(fname, "wb");
    if (!f) { std::fprintf(stderr, "ggplvm: cannot write %s\n", fname); return; }
    if (header) {
        const int hdr[6] = {r.N, r.J, r.M, r.P, r.p, r.Q};
        std::fwrite(hdr, sizeof(int), 6, f);
    }
    std::fwrite(om.data(), sizeof(double), r.N, f);
    std::fwrite(de.data(), sizeof(double), r.p, f);
    std::fwrite(x.data(), sizeof(double), (size_t)r.N * r.Q, f);
    std::fwrite(t.data(), sizeof(double), (size_t)r.M * r.Q, f);
    std::fwrite(rh.data(), sizeof(double), r.Q, f);
    std::fwrite(ph.data(), sizeof(double), r.J, f);
    std::fclose(f);
    std::printf("Output param=%s\n", fname);
}

// ggplvm.cu:1302-1316: X9.gz (X rows, T rows, rho) and phi.gz (which holds
// Omega[0..J-1], as the reference copies Omega into phi_h there).
void on_outer(gp_ctx *c, int itr, double obj, void *user) {
    Run &r = *static_cast<Run *>(user);
    std::vector<double> x, t, rh, om;
    if (!get(c, "X", x) || !get(c, "T", t) || !get(c, "rho", rh) || !get(c, "Omega", om)) return;
    gzFile f = gzopen("X9.gz", "w6h");
    gzFile fphi = gzopen("phi.gz", "w6h");
    if (f) {
        for (int i = 0; i < r.N; i++) {
            for (int k = 0; k < r.Q; k++) gzprintf(f, "%lf,", x[i + (size_t)k * r.N]);
            gzprintf(f, "\n");
        }
        for (int i = 0; i < r.M; i++) {
            for (int k = 0; k < r.Q; k++) gzprintf(f, "%lf,", t[i + (size_t)k * r.M]);
            gzprintf(f, "\n");
        }
        for (int k = 0; k < r.Q; k++) gzprintf(f, "%lf,", rh[k]);
        gzprintf(f, "\n");
        gzclose(f);
    }
    if (fphi) {
        // the reference reads J values from the 2N-long Omega buffer
        for (int k = 0; k < r.J; k++) gzprintf(fphi, "%lf\n", k < 2 * r.N ? om[k] : 0.0);
        gzclose(fphi);
    }
    write_params(r, "Params9.bin", true);
    std::printf("itr=%d lkhd=%lf\n", itr, obj);
    std::fflush(stdout);
}

int usage() {
    std::fprintf(stderr,
                 "usage: ggplvm Y.bin DataFrame.bin ls.bin Xinit.bin [-M m] [-Q q] "
                 "[--init-params f] [--itrmax n] [--jtr-cap n] [--precision f64|mixed] "
                 "[--kernel per-ard|ard] [--quiet]\n");
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) return usage();
    int M = 100, Q = 51, itrmax = 5000, jtr_cap = 0, precision = GP_PREC_FP64, ard = 0, verbose = 1;
    const char *init_params = nullptr;
    for (int i = 5; i < argc; i++) {
        auto val = [&](void) -> const char * { return (i + 1 < argc) ? argv[++i] : nullptr; };
        if (!std::strcmp(argv[i], "-M")) { const char *v = val(); if (!v) return usage(); M = std::atoi(v); }
        else if (!std::strcmp(argv[i], "-Q")) { const char *v = val(); if (!v) return usage(); Q = std::atoi(v); }
        else if (!std::strcmp(argv[i], "--init-params")) { init_params = val(); if (!init_params) return usage(); }
        else if (!std::strcmp(argv[i], "--itrmax")) { const char *v = val(); if (!v) return usage(); itrmax = std::atoi(v); }
        else if (!std::strcmp(argv[i], "--jtr-cap")) { const char *v = val(); if (!v) return usage(); jtr_cap = std::atoi(v); }
        else if (!std::strcmp(argv[i], "--precision")) {
            const char *v = val(); if (!v) return usage();
            precision = !std::strcmp(v, "mixed") ? GP_PREC_MIXED : GP_PREC_FP64;
        } else if (!std::strcmp(argv[i], "--kernel")) {
            const char *v = val(); if (!v) return usage();
            ard = !std::strcmp(v, "ard");
        } else if (!std::strcmp(argv[i], "--quiet")) verbose = 0;
        else return usage();
    }
    std::string err;
    gpio::SparseMat Y;
    if (!gpio::load_sparsemat(argv[1], Y, err)) { std::fprintf(stderr, "ggplvm: %s\n", err.c_str()); return 1; }
    gpio::DataFrame D;
    if (!gpio::load_dataframe(argv[2], M, D, err)) { std::fprintf(stderr, "ggplvm: %s\n", err.c_str()); return 1; }
    if (D.N != Y.N) { std::fprintf(stderr, "ggplvm: DataFrame has %d cells, counts have %d\n", D.N, Y.N); return 1; }
    const int N = Y.N, J = Y.J;
    std::vector<double> ls, X;
    if (!gpio::read_doubles(argv[3], N, ls, err) || !gpio::read_doubles(argv[4], (size_t)N * Q, X, err)) {
        std::fprintf(stderr, "ggplvm: %s\n", err.c_str());
        return 1;
    }
    std::printf("N=%d, J=%d, P=%d, p=%d, M=%d, Q=%d\n", N, J, D.P, D.p, M, Q);
    std::vector<double> T((size_t)M * Q);
    gp_init_ta_with_per(M, Q, T.data());
    gpio::write_doubles("Tau.bin", T.data(), T.size());

    Run r;
    r.N = N; r.J = J; r.M = M; r.P = D.P; r.p = D.p; r.Q = Q;
    int st = gp_create(&r.ctx, N, J, M, Q, D.p, D.nh.data(), Y.p.data(), Y.i.data(), Y.x.data(),
                       Y.L, D.Zcov.data(), ls.data(), X.data(), T.data(), ard, precision);
    if (st != 0) { std::fprintf(stderr, "ggplvm: gp_create: %s\n", gp_last_error()); return 1; }
    if (init_params) {
        // util_cuda.cu:520-544 readParams (no header)
        std::vector<double> buf;
        const size_t n = (size_t)N + D.p + (size_t)N * Q + (size_t)M * Q + Q + J;
        if (!gpio::read_doubles(init_params, n, buf, err)) { std::fprintf(stderr, "ggplvm: %s\n", err.c_str()); return 1; }
        std::vector<double> om, de, x, t, rh;
        get(r.ctx, "Omega", om); get(r.ctx, "delta", de); get(r.ctx, "X", x);
        get(r.ctx, "T", t); get(r.ctx, "rho", rh);
        size_t o = 0;
        std::copy(buf.begin() + o, buf.begin() + o + N, om.begin()); o += N;
        std::copy(buf.begin() + o, buf.begin() + o + D.p, de.begin()); o += D.p;
        std::copy(buf.begin() + o, buf.begin() + o + (size_t)N * Q, x.begin()); o += (size_t)N * Q;
        std::copy(buf.begin() + o, buf.begin() + o + (size_t)M * Q, t.begin()); o += (size_t)M * Q;
        std::copy(buf.begin() + o, buf.begin() + o + Q, rh.begin()); o += Q;
        std::vector<double> ph(buf.begin() + o, buf.begin() + o + J);
        gp_set_buffer(r.ctx, "Omega", om.data(), om.size());
        gp_set_buffer(r.ctx, "delta", de.data(), de.size());
        gp_set_buffer(r.ctx, "X", x.data(), x.size());
        gp_set_buffer(r.ctx, "T", t.data(), t.size());
        gp_set_buffer(r.ctx, "rho", rh.data(), rh.size());
        gp_set_buffer(r.ctx, "phi", ph.data(), ph.size());
        std::printf("Init param=%s\n", init_params);
    }
    double obj = 0.0;
    st = gp_fit(r.ctx, itrmax, jtr_cap, verbose, on_outer, &r, &obj);
    if (st > 0) { std::fprintf(stderr, "ggplvm: gp_fit: %s\n", gp_last_error()); gp_destroy(r.ctx); return 1; }
    std::printf("Iteration end...\n");
    write_params(r, "params.bin", false);
    std::printf("final lkhd=%.10e\n", obj);
    gp_destroy(r.ctx);
    return 0;
}
