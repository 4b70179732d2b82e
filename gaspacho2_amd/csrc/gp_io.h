// gp_io.h -- host-side loaders and writers in the reference's file formats.
#pragma once
#include <string>
#include <vector>

namespace gpio {

// SparseMat.cu:37-52: int N, int J, int L, int i[L], int p[J+1], double x[L]
struct SparseMat {
    int N = 0, J = 0, L = 0;
    std::vector<int> i, p;
    std::vector<double> x;
};
bool load_sparsemat(const std::string &path, SparseMat &Y, std::string &err);

// DataFrame.cu:56-117: int N, int ncov, int c[ncov], int v[ncov], body.
// Expanded to the dense covariate model-matrix columns (intercept first,
// categorical levels as 0/1 indicators, what getModelMatrix returns) and the
// variable offsets nh (variable p-1 = the M-column Knm block).
struct DataFrame {
    int N = 0, ncov = 0, p = 0, P = 0;
    std::vector<int> c, v, nh;
    std::vector<double> Zcov;   // N x nh[p-1], column-major
};
bool load_dataframe(const std::string &path, int M, DataFrame &D, std::string &err);

bool read_doubles(const std::string &path, size_t n, std::vector<double> &out, std::string &err);
bool write_doubles(const std::string &path, const double *x, size_t n);

}  // namespace gpio
