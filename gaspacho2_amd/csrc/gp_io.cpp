// gp_io.cpp -- host-side file formats and the initial inducing points.
#include "gp_io.h"
#include "../../include/gaspacho2.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace gpio {

namespace {
struct File {
    FILE *f = nullptr;
    explicit File(const std::string &p, const char *mode) { f = std::fopen(p.c_str(), mode); }
    ~File() { if (f) std::fclose(f); }
};
template <class T> bool rd(FILE *f, T *dst, size_t n) { return std::fread(dst, sizeof(T), n, f) == n; }
}  // namespace

bool load_sparsemat(const std::string &path, SparseMat &Y, std::string &err) {
    File fh(path, "rb");
    if (!fh.f) { err = "cannot open " + path; return false; }
    if (!rd(fh.f, &Y.N, 1) || !rd(fh.f, &Y.J, 1) || !rd(fh.f, &Y.L, 1)) { err = "truncated header in " + path; return false; }
    if (Y.N <= 0 || Y.J <= 0 || Y.L < 0) { err = "bad SparseMat header in " + path; return false; }
    Y.i.resize((size_t)Y.L + 1);
    Y.p.resize((size_t)Y.J + 1);
    Y.x.resize((size_t)Y.L + 1);
    if (!rd(fh.f, Y.i.data(), Y.L) || !rd(fh.f, Y.p.data(), (size_t)Y.J + 1) || !rd(fh.f, Y.x.data(), Y.L)) {
        err = "truncated body in " + path;   // SparseMat.cu:47-49 does not check fread
        return false;
    }
    return true;
}

bool load_dataframe(const std::string &path, int M, DataFrame &D, std::string &err) {
    File fh(path, "rb");
    if (!fh.f) { err = "cannot open " + path; return false; }
    if (!rd(fh.f, &D.N, 1) || !rd(fh.f, &D.ncov, 1)) { err = "truncated header in " + path; return false; }
    if (D.N <= 0 || D.ncov < 0) { err = "bad DataFrame header in " + path; return false; }
    D.c.resize(D.ncov);
    D.v.resize(D.ncov);
    if (D.ncov && (!rd(fh.f, D.c.data(), D.ncov) || !rd(fh.f, D.v.data(), D.ncov))) { err = "truncated header in " + path; return false; }
    D.p = D.ncov + 1 + (M > 0 ? 1 : 0);
    D.nh.assign(D.p + 1, 0);
    D.nh[1] = 1;
    int NH = 1;
    size_t bytes = 0;
    for (int j = 0; j < D.ncov; j++) {
        if (D.c[j] < 1 || D.v[j] < 1 || (D.c[j] > 1 && D.v[j] != 1)) { err = "bad variable spec in " + path; return false; }
        NH += D.v[j] * D.c[j];
        D.nh[j + 2] = NH;
        bytes += (size_t)D.N * D.v[j] * (D.c[j] == 1 ? sizeof(double) : sizeof(int));
    }
    if (M > 0) D.nh[D.p] = NH + M;
    D.P = D.nh[D.p];
    std::vector<char> body(bytes);
    if (bytes && std::fread(body.data(), 1, bytes, fh.f) != bytes) { err = "truncated body in " + path; return false; }
    const int ncol = D.nh[D.p - 1];
    D.Zcov.assign((size_t)D.N * ncol, 0.0);
    for (int i = 0; i < D.N; i++) D.Zcov[i] = 1.0;
    size_t off = 0;
    for (int j = 0; j < D.ncov; j++) {
        const int col0 = D.nh[j + 1];
        if (D.c[j] == 1) {
            for (int m = 0; m < D.v[j]; m++)
                for (int i = 0; i < D.N; i++) {
                    double d;
                    std::memcpy(&d, body.data() + off + sizeof(double) * ((size_t)D.N * m + i), sizeof(double));
                    D.Zcov[i + (size_t)(col0 + m) * D.N] = d;
                }
            off += sizeof(double) * (size_t)D.N * D.v[j];
        } else {
            for (int i = 0; i < D.N; i++) {
                int x;
                std::memcpy(&x, body.data() + off + sizeof(int) * (size_t)i, sizeof(int));
                if (x >= 0 && x < D.c[j]) D.Zcov[i + (size_t)(col0 + x) * D.N] = 1.0;
            }
            off += sizeof(int) * (size_t)D.N * D.v[j];
        }
    }
    return true;
}

bool read_doubles(const std::string &path, size_t n, std::vector<double> &out, std::string &err) {
    File fh(path, "rb");
    if (!fh.f) { err = "cannot open " + path; return false; }
    out.resize(n);
    if (!rd(fh.f, out.data(), n)) { err = "truncated file " + path; return false; }
    return true;
}

bool write_doubles(const std::string &path, const double *x, size_t n) {
    File fh(path, "wb");
    if (!fh.f) return false;
    return std::fwrite(x, sizeof(double), n, fh.f) == n;
}

}  // namespace gpio

// ---------------------------------------------------------------------------
// util.c:73-87 initTaWithPer: dim 0 equispaced on (0, 2pi); dims 1..Q-1 are
// N(0,1) draws from MT19937 seeded as GSL seeds seed 0 (the generator's
// default seed 4357), turned into normals by GSL's polar Box-Muller method
// (uniform = 32-bit output / 2^32, zero rejected).
namespace {
struct MT19937 {
    uint32_t mt[624];
    int mti = 625;
    explicit MT19937(uint32_t s) {
        mt[0] = s;
        for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
        mti = 624;
    }
    uint32_t next() {
        if (mti >= 624) {
            for (int k = 0; k < 624; k++) {
                const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            mti = 0;
        }
        uint32_t y = mt[mti++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    double uniform_pos() {
        double x;
        do { x = next() / 4294967296.0; } while (x == 0.0);
        return x;
    }
    double gaussian() {
        double x, y, r2;
        do {
            x = -1 + 2 * uniform_pos();
            y = -1 + 2 * uniform_pos();
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0);
        return y * std::sqrt(-2.0 * std::log(r2) / r2);
    }
};
}  // namespace

extern "C" int gp_init_ta_with_per(int M, int Q, double *T) {
    if (M <= 0 || Q <= 0 || !T) return 1;
    MT19937 g(4357u);
    for (int i = 0; i < M; i++) T[i] = ((double)(i + 1) / (double)(M + 1)) * 2.0 * 3.14159265358979;
    for (int k = 1; k < Q; k++)
        for (int i = 0; i < M; i++) T[i + (size_t)k * M] = g.gaussian();
    return 0;
}
