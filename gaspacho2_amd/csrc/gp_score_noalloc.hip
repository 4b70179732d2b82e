// gp_score_noalloc.hip -- the score tests of gp_score_tests_dosage for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc): inputs and outputs are the
// caller's device buffers, and the partial sums live in `scratch`, a static
// device array of 32 MiB that every GPU the library runs on holds until the
// process ends.  The kernels are those of gp_score.hip (gpk_score_run_dosage).
//
// Host code never passes `scratch` itself to a kernel: in host code the name
// stands for a host-side shadow, not a device address.  The device address is
// looked up once per device with hipGetSymbolAddress, with that device
// current, and kept in g_scratch_ptr.  Every call on one GPU shares that GPU's
// array and leaves the last call's contents in it; the error flag it holds is
// cleared by a kernel at the start of each call.  Calls that overlap on one
// GPU (two host threads, or two contexts on different streams) race on it.
#include "gp_internal.h"
#include "../../include/gaspacho2.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace {

constexpr size_t SCRATCH_DOUBLES = (size_t)4 << 20;      // 32 MiB
constexpr size_t SCRATCH_PART = SCRATCH_DOUBLES - 32;    // partial sums; the error flag after them
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on device dev (looked up on the first call).
int scratch_for(int dev, double **out) {
    if (dev < 0 || dev >= MAX_DEVICES) { gp_set_error("device index out of range"); return 1; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        GP_HIP(hipGetDevice(&cur));
        if (cur != dev) GP_HIP(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            gp_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return e != hipSuccess ? (int)e : 1;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" int gp_score_tests_dosage_noalloc(gp_ctx *c, int S, const unsigned char *dosage_dev,
                                             int T, const int *gene_dev, const int *snp_dev,
                                             double *stat_dev) {
    if (!c || !dosage_dev || !gene_dev || !snp_dev || !stat_dev) { gp_set_error("NULL argument"); return 1; }
    if (S <= 0 || T <= 0) { gp_set_error("score tests: S and T must be positive"); return 1; }
    GP_TRY(gp_wnu_sync(c));   // the tests read W (a kernel; no allocation)
    const size_t per = gpk_score_part_doubles(c->P);
    if (c->P > gpk_score_max_columns() || per > SCRATCH_PART) {
        gp_set_error("score tests without allocation: " + std::to_string(c->P) +
                     " model-matrix columns; the static buffer holds partial sums for at most " +
                     std::to_string(gpk_score_max_columns()));
        return 1;
    }
    double *sp = nullptr;
    GP_TRY(scratch_for(c->dev, &sp));
    const int fly = (int)std::max<size_t>(1, std::min<size_t>(16, SCRATCH_PART / per));
    int *err = reinterpret_cast<int *>(sp + SCRATCH_PART);
    GP_TRY(gpk_score_clear(c, err));
    GP_TRY(gpk_score_run_dosage(c, S, dosage_dev, T, gene_dev, snp_dev, stat_dev, sp, fly, err));
    int herr = 0;
    GP_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    if (herr) {
        gp_set_error("score tests: a gene or genotype index is out of range (its statistic is NaN)");
        return 1;
    }
    return 0;
}
