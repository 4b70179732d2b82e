// pcainit_main.cpp -- writes the initial latent the driver reads as
// Xinit.bin (ggplvm.cu:1088-1095), by the recipe of getNu0 / getNuTNu / getX
// (init.cu:304-357) run on the GPU (gp_pca_init).
//
//   pcainit Y.bin DataFrame.bin ls.bin Q Xinit.bin
//
// Inputs in the driver's formats: counts (SparseMat), covariates
// (DataFrame; the intercept and every covariate column enter the fit of the
// per-gene offsets), size factors (N doubles).  Output: N*Q doubles,
// column-major, largest component first.
#include "gaspacho2.h"
#include "gp_io.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: pcainit Y.bin DataFrame.bin ls.bin Q Xinit.bin\n");
        return 1;
    }
    std::string err;
    gpio::SparseMat Y;
    if (!gpio::load_sparsemat(argv[1], Y, err)) { std::fprintf(stderr, "pcainit: %s\n", err.c_str()); return 1; }
    gpio::DataFrame D;
    // with a GP block requested the loader expands every covariate column
    // (the GP block itself is not used here)
    if (!gpio::load_dataframe(argv[2], 1, D, err)) { std::fprintf(stderr, "pcainit: %s\n", err.c_str()); return 1; }
    if (D.N != Y.N) { std::fprintf(stderr, "pcainit: DataFrame has %d cells, counts have %d\n", D.N, Y.N); return 1; }
    const int N = Y.N, J = Y.J, Q = std::atoi(argv[4]);
    std::vector<double> ls;
    if (!gpio::read_doubles(argv[3], N, ls, err)) { std::fprintf(stderr, "pcainit: %s\n", err.c_str()); return 1; }
    const int p0 = D.nh[D.p - 1];
    std::vector<double> X((size_t)N * (Q > 0 ? Q : 1)), ev(Q > 0 ? Q : 1);
    if (gp_pca_init(N, J, Y.p.data(), Y.i.data(), Y.x.data(), ls.data(), p0, D.Zcov.data(), Q, X.data(),
                    ev.data()) != 0) {
        std::fprintf(stderr, "pcainit: %s\n", gp_last_error());
        return 1;
    }
    if (!gpio::write_doubles(argv[5], X.data(), X.size())) {
        std::fprintf(stderr, "pcainit: cannot write %s\n", argv[5]);
        return 1;
    }
    std::printf("N=%d, J=%d, p=%d, Q=%d; eigenvalues", N, J, p0, Q);
    for (int k = 0; k < Q; k++) std::printf(" %.6e", ev[k]);
    std::printf("\nOutput=%s\n", argv[5]);
    return 0;
}
