// gp_gram_mixed.hip -- mixed-precision weighted Gram
//   Psi_j = DeltaInv + Z^T diag(w_j) Z,   w_j = W_j / Omega
// on FP16 MFMA (v_mfma_f32_16x16x32_f16) with a hi/lo split of both operands,
// FP32 accumulation inside a cell segment and FP64 sums across segments.
//
// Operands.  z~ = z 2^s_a = zh + zl (two FP16 values, pre-split once per Kuf
// build, each column scaled so its maximum sits just under 2^14) and, per gene,
// a = (w 2^-e_j) z~ split the same way in registers, where 2^e_j is a power of
// two >= max_i w_ij (so |a| <= |z~| < 2^14 and nothing overflows FP16).
// Each product uses three MFMAs: ah*zh + ah*zl + al*zh (al*zl ~ 2^-22 dropped).
//
// Work decomposition.  The P16 x P16 output (P16 = 16*ceil(P/16)) is cut into
// 128x128 blocks; workgroup (block pair, gene, segment) computes one block for
// one gene over one segment of cells.  Wave w owns the 64x64 quadrant
// (w/2, w%2) = 4x4 tiles of 16x16; tiles below the diagonal or past P16 are
// skipped.  Per 32-cell chunk:
//   A (weighted) : 128 rows x 32 cells hi/lo, built cooperatively into LDS
//                  (double-buffered), fragment lane l = a[row 16t + l%16][cells 8(l/16)..+7]
//   B (plain z)  : read by each wave straight from global zh/zl (column-major,
//                  cells contiguous -> one 16-byte load per lane per fragment)
//   D            : lane l, r -> block[16t_a + 4(l/16) + r][16t_b + l%16]
// Partials (FP32, per segment) go to a scratch buffer; k_gram_reduce adds the
// segments in FP64, rescales by 2^e_j, adds DeltaInv and writes both triangles.
#include "gp_internal.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int BLK = 128;     // output block edge
constexpr int CH = 32;       // cells per chunk (one MFMA k-step)

// zh/zl (FP16, Npad x P16 column-major) from the FP64 model matrix, each
// column scaled by a power of two 2^s_a that puts its largest magnitude in
// [2^13, 2^14): values down to 2^-16 of the column maximum keep a normal FP16
// lo part, i.e. the full ~22-bit hi/lo precision.  zun[a] = 2^-s_a.
__global__ void __launch_bounds__(256) k_split_z(int Npad, int P, const double *__restrict__ Z,
        _Float16 *__restrict__ zh, _Float16 *__restrict__ zl, float *__restrict__ z32,
        double *__restrict__ zun) {
    __shared__ double sh[4];
    __shared__ float scale;
    const int a = blockIdx.x;
    const double *za = Z + (size_t)a * Npad;
    // each pass keeps SZU loads per thread in flight: with one workgroup per
    // column there is only about one wave per SIMD to hide their latency
    constexpr int SZU = 8;
    double mx = 0.0;
    if (a < P) {
        int i = threadIdx.x;
        for (; i + (SZU - 1) * 256 < Npad; i += SZU * 256) {
            double v[SZU];
#pragma unroll
            for (int u = 0; u < SZU; u++) v[u] = za[i + u * 256];
#pragma unroll
            for (int u = 0; u < SZU; u++) mx = fmax(mx, fabs(v[u]));
        }
        for (; i < Npad; i += 256) mx = fmax(mx, fabs(za[i]));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        int e = 0;
        if (m > 0.0) (void)frexp(m, &e);          // m < 2^e
        const int sa = 14 - e;                     // m * 2^sa < 2^14
        scale = ldexpf(1.0f, sa);
        zun[a] = ldexp(1.0, -sa);
    }
    __syncthreads();
    const float sc = scale;
    _Float16 *h = zh + (size_t)a * Npad, *l = zl + (size_t)a * Npad;
    float *f = z32 + (size_t)a * Npad;
    int i = threadIdx.x;
    if (a < P) {
        for (; i + (SZU - 1) * 256 < Npad; i += SZU * 256) {
            double zv[SZU];
#pragma unroll
            for (int u = 0; u < SZU; u++) zv[u] = za[i + u * 256];
#pragma unroll
            for (int u = 0; u < SZU; u++) {
                const float v = (float)(zv[u] * (double)sc);
                const _Float16 hv = (_Float16)v;
                h[i + u * 256] = hv;
                l[i + u * 256] = (_Float16)(v - (float)hv);
                f[i + u * 256] = v;
            }
        }
    }
    for (; i < Npad; i += 256) {
        const float v = (a < P) ? (float)(za[i] * (double)sc) : 0.0f;
        const _Float16 hv = (_Float16)v;
        h[i] = hv;
        l[i] = (_Float16)(v - (float)hv);
        f[i] = v;
    }
}

// Per gene: w32 = W/Omega (FP32) and the power-of-two scale 2^-e_j.
__global__ void __launch_bounds__(256) k_wprep(int N, int Npad, const double *__restrict__ W,
        const double *__restrict__ Omega, float *__restrict__ w32, float *__restrict__ wsc,
        double *__restrict__ wunscale) {
    __shared__ float sh[4];
    const int j = blockIdx.x;
    const double *Wj = W + (size_t)j * Npad;
    float *wj = w32 + (size_t)j * Npad;
    // WPU cells per thread loaded before any is used (2 WPU loads in flight)
    constexpr int WPU = 4;
    float mx = 0.0f;
    int i = threadIdx.x;
    for (; i + (WPU - 1) * 256 < N; i += WPU * 256) {
        double w[WPU], om[WPU];
#pragma unroll
        for (int u = 0; u < WPU; u++) {
            w[u] = Wj[i + u * 256];
            om[u] = Omega[i + u * 256];
        }
#pragma unroll
        for (int u = 0; u < WPU; u++) {
            const float v = (float)(w[u] / om[u]);
            wj[i + u * 256] = v;
            mx = fmaxf(mx, fabsf(v));
        }
    }
    for (; i < Npad; i += 256) {
        const float v = (i < N) ? (float)(Wj[i] / Omega[i]) : 0.0f;
        wj[i] = v;
        mx = fmaxf(mx, fabsf(v));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
        int e = 0;
        if (m > 0.0f) { (void)frexpf(m, &e); }      // m <= 2^e
        wsc[j] = ldexpf(1.0f, -e);
        wunscale[j] = ldexp(1.0, e);
    }
}

// FP16 hi/lo of FP32 values.  Each value is materialised in FP32 first: left
// to itself the compiler folds the producing multiply into the conversions
// (v_fma_mix), and that form loses accuracy in the lo part (measured: 3e-5
// instead of 2e-7 relative error in the gradient contraction).
// The same split two values at a time: each pair converts with one packed
// round-to-nearest-even conversion (the bits of converting one at a time),
// and hi goes back to FP32 from the packed halves, so no value is converted
// to FP16 twice.  Results as 16-bit patterns, two per word.
__device__ __forceinline__ void split8_pk(float (&v)[8], u4 &hi, u4 &lo) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
        float a = v[2 * e], b = v[2 * e + 1];
        asm volatile("" : "+v"(a), "+v"(b));
        const unsigned hu = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){a, b}, h2));
        const float ha = (float)__builtin_bit_cast(_Float16, (unsigned short)(hu & 0xffffu));
        const float hb = (float)__builtin_bit_cast(_Float16, (unsigned short)(hu >> 16));
        hi[e] = hu;
        lo[e] = __builtin_bit_cast(unsigned, __builtin_convertvector((f2){a - ha, b - hb}, h2));
    }
}

__device__ inline void split8(const float *p, h8 &hi, h8 &lo) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
        float v = p[e];
        asm volatile("" : "+v"(v));
        const _Float16 h = (_Float16)v;
        hi[e] = h;
        lo[e] = (_Float16)(v - (float)h);
    }
}

__global__ void __launch_bounds__(256) k_gram_mixed(int Npad, int P16, int T16, int nbk, int seg,
        const _Float16 *__restrict__ zh, const _Float16 *__restrict__ zl,
        const float *__restrict__ w32, const float *__restrict__ wsc,
        float *__restrict__ part) {
    // LDS: [2 buffers][hi/lo][128 rows][32 cells] halves = 32 KB
    __shared__ __attribute__((aligned(16))) _Float16 As[2][2][BLK][CH];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    // block pair (bi <= bj) from blockIdx.x
    int bi = 0, t = blockIdx.x;
    while (t >= nbk - bi) { t -= nbk - bi; bi++; }
    const int bj = bi + t;
    const int j = blockIdx.y, s = blockIdx.z;   // gene, cell segment
    const int c0 = s * seg, c1 = min(Npad, c0 + seg);
    const float sc = wsc[j];
    const float *wj = w32 + (size_t)j * Npad;
    // quadrant tiles (global tile indices)
    const int qa = w >> 1, qb = w & 1;
    const int ta0 = bi * 8 + qa * 4, tb0 = bj * 8 + qb * 4;
    bool valid[4][4];
    bool any = false;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            valid[x][y] = (ta0 + x < T16) && (tb0 + y < T16) && (ta0 + x <= tb0 + y);
            any |= valid[x][y];
        }
    f4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};

    // staging map: 512 items (128 rows x 4 groups of 8 cells) -> 2 per thread
    const int srow0 = tid >> 2, sgrp = tid & 3;         // items (srow0, sgrp) and (srow0+64, sgrp)
    const int arow0 = bi * BLK + srow0, arow1 = arow0 + 64;
    h8 pzh[2], pzl[2];
    float pw[8];
    auto fetchA = [&](int cc) {
        const size_t off = (size_t)cc + 8 * sgrp;
        pzh[0] = (arow0 < P16) ? *(const h8 *)(zh + (size_t)arow0 * Npad + off) : h8{};
        pzl[0] = (arow0 < P16) ? *(const h8 *)(zl + (size_t)arow0 * Npad + off) : h8{};
        pzh[1] = (arow1 < P16) ? *(const h8 *)(zh + (size_t)arow1 * Npad + off) : h8{};
        pzl[1] = (arow1 < P16) ? *(const h8 *)(zl + (size_t)arow1 * Npad + off) : h8{};
#pragma unroll
        for (int e = 0; e < 8; e++) pw[e] = wj[off + e] * sc;
    };
    auto storeA = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 2; it++) {
            float p[8];
#pragma unroll
            for (int e = 0; e < 8; e++) p[e] = ((float)pzh[it][e] + (float)pzl[it][e]) * pw[e];
            h8 hi, lo;
            split8(p, hi, lo);
            *(h8 *)&As[buf][0][srow0 + 64 * it][8 * sgrp] = hi;
            *(h8 *)&As[buf][1][srow0 + 64 * it][8 * sgrp] = lo;
        }
    };
    // B fragments: lane l loads z[col 16 tb + l%16][cells cc + 8(l/16) ..+7]
    h8 bh[4], bl[4];
    auto fetchB = [&](int cc) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int col = (tb0 + y) * 16 + (lane & 15);
            const size_t off = (size_t)col * Npad + cc + 8 * (lane >> 4);
            const bool ok = (tb0 + y) < T16;
            bh[y] = ok ? *(const h8 *)(zh + off) : h8{};
            bl[y] = ok ? *(const h8 *)(zl + off) : h8{};
        }
    };

    int cur = 0;
    fetchA(c0);
    storeA(0);
    if (any) fetchB(c0);
    __syncthreads();
    for (int cc = c0; cc < c1; cc += CH) {
        const bool more = cc + CH < c1;
        h8 nbh[4], nbl[4];
        if (more) {
            fetchA(cc + CH);
            if (any) {
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    const int col = (tb0 + y) * 16 + (lane & 15);
                    const size_t off = (size_t)col * Npad + cc + CH + 8 * (lane >> 4);
                    const bool ok = (tb0 + y) < T16;
                    nbh[y] = ok ? *(const h8 *)(zh + off) : h8{};
                    nbl[y] = ok ? *(const h8 *)(zl + off) : h8{};
                }
            }
        }
        if (any) {
            h8 ah[4], al[4];
#pragma unroll
            for (int x = 0; x < 4; x++) {
                const int r = qa * 64 + 16 * x + (lane & 15);
                ah[x] = *(const h8 *)&As[cur][0][r][8 * (lane >> 4)];
                al[x] = *(const h8 *)&As[cur][1][r][8 * (lane >> 4)];
            }
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++)
                    if (valid[x][y]) {
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    }
            if (more) {
#pragma unroll
                for (int y = 0; y < 4; y++) { bh[y] = nbh[y]; bl[y] = nbl[y]; }
            }
        }
        if (more) storeA(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // partial block (FP32), row-major 128x128 inside [s][j][block pair]
    const int nbp = nbk * (nbk + 1) / 2;
    float *pb = part + (((size_t)s * gridDim.y + j) * nbp + blockIdx.x) * (size_t)(BLK * BLK);
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = qa * 64 + 16 * x + 4 * (lane >> 4) + r;
                const int col = qb * 64 + 16 * y + (lane & 15);
                pb[row * BLK + col] = valid[x][y] ? acc[x][y][r] : 0.0f;
            }
        }
}

// Psi_j[R,C] = DeltaInv[R,C] + 2^e_j 2^-s_R 2^-s_C sum_s part[s][j][blk][r',c'] for r <= c,
// written to both triangles; the partials cover a Pg x Pg block of the
// design starting at column off (R = off + r, C = off + c): the whole design
// (Pg = P, off = 0) or its GP block (Pg = M, off = 1).
__global__ void k_gram_reduce(int P, int Pg, int off, int nbk, int nseg, int J, const float *__restrict__ part,
                              const double *__restrict__ wun, const double *__restrict__ zun,
                              const double *__restrict__ DeltaInv, double *__restrict__ Psi) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (idx >= (size_t)Pg * Pg) return;
    const int r = (int)(idx % Pg), c = (int)(idx / Pg);
    if (r > c) return;
    const int bi = r / BLK, bj = c / BLK;
    const int bp = bi * nbk - bi * (bi - 1) / 2 + (bj - bi);
    const int nbp = nbk * (nbk + 1) / 2;
    const size_t inblk = (size_t)(r % BLK) * BLK + (c % BLK);
    double s = 0.0;
    for (int g = 0; g < nseg; g++)
        s += (double)part[(((size_t)g * J + j) * nbp + bp) * (size_t)(BLK * BLK) + inblk];
    const int R = off + r, C = off + c;
    const double v = s * (wun[j] * zun[R] * zun[C]) + DeltaInv[R + (size_t)C * P];
    double *Pj = Psi + (size_t)j * P * P;
    Pj[R + (size_t)C * P] = v;
    Pj[C + (size_t)R * P] = v;
}


// ---------------------------------------------------------------------------
// Symmetric form.  With x_i = sqrt(w_ij 2^-e_j) z~_i the weighted Gram is
// X^T X, so one operand serves both sides of every MFMA.  Workgroup =
// (gene, cell segment), 8 waves; the upper triangle of 16x16 tiles is cut
// into blocks over groups of 4 tiles (the last group may be narrower) and
// each wave owns up to two blocks, balanced at compile time (longest block
// first onto the least loaded wave).  Each wave runs its own straight-line
// copy of the chunk loop (the blocks are template arguments), so no
// accumulator is touched by a data-dependent branch.  Per 64-cell chunk (two
// MFMA k-steps) the workgroup builds X hi/lo for all P16 columns in LDS
// (double buffered; the chunk's square-root weights are staged two chunks
// ahead) and every wave runs its blocks out of LDS: per k-step a 4x4 block
// reads 4 + 4 fragments for 16 tiles, a diagonal block 4 for 10.  Two k-steps
// per chunk halve the barriers and the LDS-read ramp at each chunk start per
// MFMA; at P16 = 272 the two buffers take 153 KB of the 160 KB LDS.
constexpr int GW = 8;        // waves per workgroup
constexpr int GSL = 2;       // block slots per wave
constexpr int LDX = 40;      // LDS column stride in halves, four-wave form (32 cells + 8)
constexpr int GKC = 2;       // MFMA k-steps per chunk, eight-wave form
constexpr int GCH = 32 * GKC;    // cells per chunk, eight-wave form
constexpr int GLDX = GCH + 8;    // LDS column stride in halves, eight-wave form
constexpr int MAX_SYM_T16 = 17;  // instantiated widths of the symmetric kernel

struct GramPlan {
    int a[GW * GSL];   // row group of the slot's block (-1: empty)
    int b[GW * GSL];   // column group
};

// Block plans for T16 = 1..17: wave w runs on SIMD w mod 4 and owns the
// blocks in slots 2w, 2w+1 (row group a, column group b; -1 = empty).  Each
// plan was chosen, among random deals of the blocks with at most two per
// wave, for the smallest largest MFMA load of a SIMD (tiles of its two waves)
// and then the fewest LDS fragment reads per chunk (a wave reads each 4-tile
// group it touches once: at P16 = 272, 120 reads instead of 170 when every
// block reads its own).
constexpr GramPlan kGramPlans[MAX_SYM_T16] = {
    {{0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 1: max SIMD load 1 tiles, 2 fragment reads
    {{0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 2: max SIMD load 3 tiles, 4 fragment reads
    {{0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 3: max SIMD load 6 tiles, 6 fragment reads
    {{0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 4: max SIMD load 10 tiles, 8 fragment reads
    {{0, -1, 0, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {0, -1, 1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 5: max SIMD load 10 tiles, 20 fragment reads
    {{0, -1, 0, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {1, -1, 0, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 6: max SIMD load 10 tiles, 24 fragment reads
    {{1, -1, 0, -1, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {1, -1, 0, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 7: max SIMD load 12 tiles, 28 fragment reads
    {{1, -1, 0, -1, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {1, -1, 1, -1, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}},   // T16 = 8: max SIMD load 16 tiles, 32 fragment reads
    {{1, -1, 0, -1, 0, -1, 0, -1, 1, -1, 2, -1, -1, -1, -1, -1}, {2, -1, 2, -1, 0, -1, 1, -1, 1, -1, 2, -1, -1, -1, -1, -1}},   // T16 = 9: max SIMD load 16 tiles, 54 fragment reads
    {{0, -1, 2, -1, 0, -1, 1, -1, 1, -1, 0, -1, -1, -1, -1, -1}, {2, -1, 2, -1, 1, -1, 1, -1, 2, -1, 0, -1, -1, -1, -1, -1}},   // T16 = 10: max SIMD load 16 tiles, 60 fragment reads
    {{2, -1, 0, -1, 0, -1, 0, -1, 1, -1, 1, -1, -1, -1, -1, -1}, {2, -1, 0, -1, 2, -1, 1, -1, 2, -1, 1, -1, -1, -1, -1, -1}},   // T16 = 11: max SIMD load 20 tiles, 66 fragment reads
    {{2, -1, 0, -1, 0, -1, 1, -1, 0, -1, 1, -1, -1, -1, -1, -1}, {2, -1, 1, -1, 0, -1, 2, -1, 2, -1, 1, -1, -1, -1, -1, -1}},   // T16 = 12: max SIMD load 26 tiles, 72 fragment reads
    {{0, 0, 1, -1, 1, -1, 0, -1, 1, -1, 3, -1, 0, -1, 2, 2}, {3, 1, 2, -1, 1, -1, 0, -1, 3, -1, 3, -1, 2, -1, 3, 2}},   // T16 = 13: max SIMD load 26 tiles, 88 fragment reads
    {{0, -1, 2, 0, 0, -1, 2, -1, 0, -1, 3, -1, 1, 1, 1, -1}, {0, -1, 3, 2, 3, -1, 2, -1, 1, -1, 3, -1, 1, 3, 2, -1}},   // T16 = 14: max SIMD load 27 tiles, 96 fragment reads
    {{0, -1, 1, -1, 1, -1, 0, 1, 2, 2, 0, -1, 0, -1, 3, -1}, {0, -1, 3, -1, 2, -1, 1, 1, 3, 2, 2, -1, 3, -1, 3, -1}},   // T16 = 15: max SIMD load 32 tiles, 104 fragment reads
    {{3, -1, 0, -1, 0, -1, 2, -1, 1, 1, 1, -1, 0, 2, 0, -1}, {3, -1, 3, -1, 0, -1, 3, -1, 1, 2, 3, -1, 2, 2, 1, -1}},   // T16 = 16: max SIMD load 36 tiles, 112 fragment reads
    {{4, 1, 2, 0, 1, 2, 3, -1, 2, 1, 0, 3, 0, 0, 1, 0}, {4, 4, 4, 2, 2, 2, 3, -1, 3, 3, 3, 4, 0, 4, 1, 1}},   // T16 = 17: max SIMD load 40 tiles, 120 fragment reads
};

constexpr GramPlan make_gram_plan(int T16) { return kGramPlans[T16 - 1]; }

constexpr bool plan_fits(int T16) {
    const int NG = (T16 + 3) / 4;
    return NG * (NG + 1) / 2 <= GW * GSL;
}

// tiles of block (a, b) of the 4-tile column groups
constexpr int gram_blk_cost(int T16, int a, int b) {
    const int NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    const int sa = a < NG - 1 ? 4 : RL, sb = b < NG - 1 ? 4 : RL;
    return a == b ? sa * (sa + 1) / 2 : sa * sb;
}

// Staging rounds (64 items each, T16 * GKC per chunk) dealt to the waves of
// the eight-wave form: wave W stages rounds r0[W] .. r0[W] + nr[W] - 1.  The
// two waves of a SIMD share its matrix core but each issues its own VALU work,
// so the rounds go to the waves with the least MFMA work (a round is about 200
// VALU cycles against 96 MFMA cycles per tile per chunk): each round to the
// wave whose MFMA plus staging cycles are smallest.  At T16 = 17 the heaviest
// wave (32 tiles) stages nothing and the lightest (5 tiles) a third.
struct GramRounds {
    int r0[GW];
    int nr[GW];
};

// VALU cycles per chunk of wave 0's IRLS work in the fused form (per cell:
// one FP64 exp, one FP64 division, the count scatter and a dozen FP64 ops)
constexpr int IRLS_CYC = 400;

// Wide designs (M = 512, 32 GP tiles): the upper triangle is cut into two
// 16-tile symmetric halves (the T16 = 16 plan on an offset column range) and
// four 8 x 8-tile rectangles of cross terms.  A rectangle stages 128 row and
// 128 column columns as one 256-column operand (4-tile groups 0, 1 = rows,
// 2, 3 = columns) and computes the four off-diagonal 4 x 4 blocks (0|1, 2|3):
// one block on each of waves 0-3, waves 4-7 only staging.
constexpr GramPlan kRectPlan = {{0, -1, 0, -1, 1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
                                {2, -1, 3, -1, 2, -1, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1}};

constexpr GramRounds make_gram_rounds_p(const GramPlan pl, int T16, bool irls) {
    GramRounds g{};
    int t[GW] = {};
    if (irls) t[0] += IRLS_CYC;
    for (int w = 0; w < GW; w++)
        for (int k = 0; k < GSL; k++)
            if (pl.a[w * GSL + k] >= 0) t[w] += 96 * gram_blk_cost(T16, pl.a[w * GSL + k], pl.b[w * GSL + k]);
    for (int r = 0; r < T16 * GKC; r++) {
        int best = 0;
        for (int w = 1; w < GW; w++)
            if (t[w] < t[best]) best = w;
        t[best] += 200;
        g.nr[best]++;
    }
    for (int w = 1; w < GW; w++) g.r0[w] = g.r0[w - 1] + g.nr[w - 1];
    return g;
}

constexpr GramRounds make_gram_rounds(int T16, bool irls = false) {
    return make_gram_rounds_p(make_gram_plan(T16), T16, irls);
}

__device__ inline void mfma3(f4 &acc, const h8 &ah, const h8 &al, const h8 &bh, const h8 &bl) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
}

// one 32-cell k-step of block (A, B) out of the LDS buffer (xh, xl), column
// stride LD halves; lk includes the k-step's offset
template <int T16, int A, int B, int LD>
__device__ __forceinline__ void gram_block(f4 (&acc)[4][4], const _Float16 *xh, const _Float16 *xl,
                                           int lrow, int lk) {
    constexpr int NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    constexpr int SA = A < NG - 1 ? 4 : RL, SB = B < NG - 1 ? 4 : RL;
    h8 ah[SA], al[SA];
#pragma unroll
    for (int x = 0; x < SA; x++) {
        const int col = (4 * A + x) * 16 + lrow;
        ah[x] = *(const h8 *)(xh + col * LD + lk);
        al[x] = *(const h8 *)(xl + col * LD + lk);
    }
    if constexpr (A == B) {
#pragma unroll
        for (int x = 0; x < SA; x++)
#pragma unroll
            for (int y = x; y < SA; y++) mfma3(acc[x][y], ah[x], al[x], ah[y], al[y]);
    } else {
#pragma unroll
        for (int y = 0; y < SB; y++) {
            const int col = (4 * B + y) * 16 + lrow;
            const h8 bh = *(const h8 *)(xh + col * LD + lk);
            const h8 bl = *(const h8 *)(xl + col * LD + lk);
#pragma unroll
            for (int x = 0; x < SA; x++) mfma3(acc[x][y], ah[x], al[x], bh, bl);
        }
    }
}

template <int T16, int A, int B>
__device__ __forceinline__ void gram_block_store(const f4 (&acc)[4][4], float *pg, int lane) {
    constexpr int NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    constexpr int SA = A < NG - 1 ? 4 : RL, SB = B < NG - 1 ? 4 : RL;
#pragma unroll
    for (int x = 0; x < SA; x++)
#pragma unroll
        for (int y = (A == B ? x : 0); y < SB; y++) {
            const int ta = 4 * A + x, tb = 4 * B + y;
            const int t = ta * T16 - ta * (ta - 1) / 2 + (tb - ta);
            *(f4 *)(pg + (size_t)t * 256 + 4 * lane) = acc[x][y];
        }
}

// block (A, B)'s tiles stored at their place in a TGL-tile upper triangle,
// group q of the staged operand starting at tile tmap[q]
template <int T16, int A, int B>
__device__ __forceinline__ void gram_block_store_map(const f4 (&acc)[4][4], float *pg, int lane,
                                                     const int (&tmap)[4], int TGL) {
    constexpr int NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    constexpr int SA = A < NG - 1 ? 4 : RL, SB = B < NG - 1 ? 4 : RL;
#pragma unroll
    for (int x = 0; x < SA; x++)
#pragma unroll
        for (int y = (A == B ? x : 0); y < SB; y++) {
            const int ta = tmap[A] + x, tb = tmap[B] + y;
            const int t = ta * TGL - ta * (ta - 1) / 2 + (tb - ta);
            *(f4 *)(pg + (size_t)t * 256 + 4 * lane) = acc[x][y];
        }
}

struct GramArgs {
    int Npad, c0, nch;
    float sc;
    const float *z32, *wj;
    float *pg;
    // wide designs: the staged 64-column groups' sources, their first tiles
    // in the whole GP block and its tile count (the partials' layout)
    const float *zg[4] = {nullptr, nullptr, nullptr, nullptr};
    int tmap[4] = {0, 4, 8, 12};
    int TGL = 0;
    // fused IRLS pass (k_gram_irls): wave 0 builds each chunk's weights from
    // eta = z^T beta instead of reading them, together with nu, R and the sums
    int N = 0;
    int nz0 = 0, nz1 = 0;                 // the segment's non-zeros: Yi[nz0 .. nz1)
    const double *eta = nullptr;          // eta_j (Npad)
    const double *cd = nullptr, *com = nullptr, *cbl = nullptr;   // per cell: d, Omega, log ls
    const int *Yi = nullptr;
    const double *Yx = nullptr;
    double *Rj = nullptr;                 // R_j = w nu (Npad)
    float *w32o = nullptr;                // w32_j = w in FP32 (Npad)
    double *ywl = nullptr;                // LDS: one chunk's counts by cell
    double *sums = nullptr;               // the (segment, gene)'s sum(w nu^2), sum(d - eta), sum W, sum y(eta + log ls)
};

// Per-cell IRLS quantities (ggplvm.cu:468-514 getWForBeta with zeta = 0, the
// sums of getBetaPhi / getLkhd, ggplvm.cu:418-464, and getLkhdPois,
// ggplvm.cu:703-724) in the form the fused Gram uses.  With the per-cell
// constants d_i = log Omega_i - log ls_i - dt t_i wt and Omega_i:
//   w = W / Omega = exp(eta - d),  W = w Omega,
//   nu = eta + y / W - 1 (y / W = exp(log y - log W); nu = eta - 1 at zeros),
//   R = w nu,  gq += w nu^2,  gs += log(Omega / W) = d - eta,
//   the GLM objective sum_nz y (log W - dt t wt) - sum W = sum_nz y (eta + log ls) - sum W.
struct IrlsLoad {
    double eta, d, om, bl, yx;
    int yi;
};

__device__ __forceinline__ IrlsLoad irls_load(const GramArgs &g, int ch, int nzp, int lane) {
    IrlsLoad v;
    const int i = g.c0 + ch * GCH + lane;
    v.eta = g.eta[i];
    v.d = g.cd[i];
    v.om = g.com[i];
    v.bl = g.cbl[i];
    const int k = nzp + lane;
    v.yi = 0x7fffffff;
    v.yx = 0.0;
    if (k < g.nz1) {
        v.yi = g.Yi[k];
        v.yx = g.Yx[k];
    }
    return v;
}

// wave 0: the chunk's counts by cell (its non-zeros are the next entries of
// the gene's ascending cell list), then the cell's IRLS quantities; returns
// w in FP32 and advances the list pointer
__device__ __forceinline__ float irls_finish(const GramArgs &g, int ch, int &nzp, int lane,
                                             const IrlsLoad &v, double (&acc)[4]) {
    const int cc = g.c0 + ch * GCH;
    const bool in = v.yi < cc + GCH;
    const unsigned long long m = __ballot(in);
    // lane L's slot may be written by another lane: volatile keeps the three
    // accesses in program order and the read from being taken from the
    // lane's own zeroing store; a wave's LDS operations execute in order
    volatile double *yw = g.ywl;
    yw[lane] = 0.0;
    if (in && v.yi >= cc) yw[v.yi - cc] = v.yx;
    const double y = yw[lane];
    nzp += __popcll(m);
    const int i = cc + lane;
    float wf = 0.0f;
    double r = 0.0;
    if (i < g.N) {
        const double w = exp(v.eta - v.d);
        const double Wv = w * v.om;
        const double q = y / Wv;
        const double nu = (v.eta + (y > 0.0 ? q : 0.0)) - 1.0;
        r = w * nu;
        acc[0] += nu * nu * w;
        acc[1] += v.d - v.eta;
        acc[2] += Wv;
        acc[3] += y * (v.eta + v.bl);
        wf = (float)w;
    }
    g.Rj[i] = r;
    g.w32o[i] = wf;
    return wf;
}

// the chunk loop of wave W (its blocks are compile-time constants); IR: the
// fused IRLS form, whose wave 0 builds the weights from eta
// MODE 0: the design's own columns; 1: a 16-tile symmetric half of a wide
// design; 2: a wide design's 8 x 8-tile rectangle (kRectPlan)
template <int T16, int W, bool IR = false, int MODE = 0>
__device__ __forceinline__ void gram_wave(const GramArgs &g, _Float16 *Xs, float (*wsl)[GCH]) {
    constexpr int P16 = 16 * T16;
    static_assert(MODE == 0 || T16 == 16, "wide-design parts stage 256 columns");
    constexpr GramPlan pl = MODE == 2 ? kRectPlan : make_gram_plan(T16);
    constexpr int A0 = pl.a[W * GSL], B0 = pl.b[W * GSL];
    constexpr int A1 = pl.a[W * GSL + 1], B1 = pl.b[W * GSL + 1];
    constexpr int QC = 4 * GKC;                   // 8-cell groups per chunk
    constexpr GramRounds rp = make_gram_rounds_p(pl, T16, IR);
    constexpr int RB = rp.r0[W], IPT = rp.nr[W];  // this wave's staging rounds
    const int tid = threadIdx.x, lane = tid & 63;
    const int lrow = lane & 15, lk = 8 * (lane >> 4);

    f4 acc0[4][4], acc1[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            acc0[x][y] = f4{0.f, 0.f, 0.f, 0.f};
            acc1[x][y] = f4{0.f, 0.f, 0.f, 0.f};
        }

    // The prefetch is straight-line code (chunk clamped, no branches around
    // the loads) and is pinned ahead of the MFMAs by a scheduling barrier, so
    // its wait lands at its first use in store(), after this chunk's MFMAs:
    // the L2/HBM latency hides behind them.  The P16 x QC (column, 8-cell
    // group) items of a chunk come in rounds of 64; this wave's rounds
    // (compile time) hold items 64 (RB + r) + lane, r < IPT.  They are fetched
    // and stored in GKC parts, part s after the MFMAs of k-steps s - 1 and s,
    // so that only one part is held in registers.
    constexpr int IPS = (IPT + GKC - 1) / GKC;    // rounds per part
    f4 pz[IPT > 0 ? IPT : 1][2];
    auto fetch = [&](int cc, int r0, int r1) {
#pragma unroll
        for (int r = r0; r < r1; r++) {
            const int it = 64 * (RB + r) + lane;
            const int col = it / QC;              // (RB + r) / 8 = col / 64 is a compile-time group
            const float *src = (MODE == 0 ? g.z32 + (size_t)col * g.Npad
                                          : g.zg[col >> 6] + (size_t)(col & 63) * g.Npad) +
                               cc + 8 * (it % QC);
            pz[r][0] = *(const f4 *)src;
            pz[r][1] = *(const f4 *)(src + 4);
        }
    };
    auto store = [&](int buf, int wb, int r0, int r1) {
#pragma unroll
        for (int r = r0; r < r1; r++) {
            const int it = 64 * (RB + r) + lane;
            const int col = it / QC, q = it % QC;
            const f4 s0 = *(const f4 *)&wsl[wb][8 * q];
            const f4 s1 = *(const f4 *)&wsl[wb][8 * q + 4];
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (e < 4 ? pz[r][0][e] * s0[e] : pz[r][1][e - 4] * s1[e - 4]);
            u4 hi, lo;
            split8_pk(v, hi, lo);
            *(u4 *)&Xs[((buf * 2 + 0) * P16 + col) * GLDX + 8 * q] = hi;
            *(u4 *)&Xs[((buf * 2 + 1) * P16 + col) * GLDX + 8 * q] = lo;
        }
    };
    auto sqw = [&](int ch) { return sqrtf(g.wj[g.c0 + ch * GCH + tid] * g.sc); };   // tid < GCH
    // fused IRLS form (wave 0): the segment's non-zero pointer and the sums
    int nzp = g.nz0;
    double iacc[4] = {0.0, 0.0, 0.0, 0.0};

    // prologue: weights of chunks 0 and 1, X of chunk 0
    if (tid < GCH) {
        if constexpr (IR) {
            const IrlsLoad v0 = irls_load(g, 0, nzp, lane);
            wsl[0][tid] = sqrtf(irls_finish(g, 0, nzp, lane, v0, iacc) * g.sc);
            if (g.nch > 1) {
                const IrlsLoad v1 = irls_load(g, 1, nzp, lane);
                wsl[1][tid] = sqrtf(irls_finish(g, 1, nzp, lane, v1, iacc) * g.sc);
            }
        } else {
            wsl[0][tid] = sqw(0);
            if (g.nch > 1) wsl[1][tid] = sqw(1);
        }
    }
    fetch(g.c0, 0, IPT);
    __syncthreads();
    store(0, 0, 0, IPT);
    __syncthreads();

    for (int ch = 0; ch < g.nch; ch++) {
        const int buf = ch & 1;
        const int cn = g.c0 + min(ch + 1, g.nch - 1) * GCH;
        float pw = 0.0f;
        [[maybe_unused]] IrlsLoad iv;
        if constexpr (W == 0) {
            if constexpr (IR) iv = irls_load(g, min(ch + 2, g.nch - 1), nzp, lane);
            else pw = g.wj[g.c0 + min(ch + 2, g.nch - 1) * GCH + (lane & (GCH - 1))];
        }
        fetch(cn, 0, min(IPS, IPT));
        __builtin_amdgcn_sched_barrier(0);   // issue the prefetch ahead of this chunk's MFMAs
        const _Float16 *xh = Xs + (size_t)(buf * 2 + 0) * P16 * GLDX;
        const _Float16 *xl = Xs + (size_t)(buf * 2 + 1) * P16 * GLDX;
#pragma unroll
        for (int s = 0; s < GKC; s++) {
            if constexpr (A0 >= 0) gram_block<T16, A0, B0, GLDX>(acc0, xh, xl, lrow, lk + 32 * s);
            if constexpr (A1 >= 0) gram_block<T16, A1, B1, GLDX>(acc1, xh, xl, lrow, lk + 32 * s);
            // unconditional, so the prefetch's only use is in this block and it
            // is not sunk into a branch (the last chunk's extra store is never
            // read); then the next part's loads, ahead of the next k-step
            store(buf ^ 1, (ch + 1) % 3, min(s * IPS, IPT), min((s + 1) * IPS, IPT));
            if (s + 1 < GKC) fetch(cn, min((s + 1) * IPS, IPT), min((s + 2) * IPS, IPT));
            __builtin_amdgcn_sched_barrier(0);
        }
        // wave 0 writes the weights two chunks ahead; past the last chunk the
        // slot holds the last chunk's weights and is never read.  Kept after
        // the MFMAs so that its wait does not stall them.  The fused form
        // builds them (and R, w32, the sums) only for chunks that exist.
        if constexpr (W == 0) {
            if constexpr (IR) {
                if (ch + 2 < g.nch)
                    wsl[(ch + 2) % 3][lane] = sqrtf(irls_finish(g, ch + 2, nzp, lane, iv, iacc) * g.sc);
            } else {
                wsl[(ch + 2) % 3][lane & (GCH - 1)] = sqrtf(pw * g.sc);
            }
        }
        __syncthreads();
    }
    if constexpr (MODE == 0) {
        if constexpr (A0 >= 0) gram_block_store<T16, A0, B0>(acc0, g.pg, lane);
        if constexpr (A1 >= 0) gram_block_store<T16, A1, B1>(acc1, g.pg, lane);
    } else {
        if constexpr (A0 >= 0) gram_block_store_map<T16, A0, B0>(acc0, g.pg, lane, g.tmap, g.TGL);
        if constexpr (A1 >= 0) gram_block_store_map<T16, A1, B1>(acc1, g.pg, lane, g.tmap, g.TGL);
    }
    if constexpr (IR && W == 0) {
        // the wave's sums over the segment, lanes added in a fixed tree order
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double s = iacc[k];
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if (lane == 0) g.sums[k] = s;
        }
    }
}

template <int T16>
__global__ void __launch_bounds__(GW * 64) k_gram_sym(int Npad, int seg, int nseg, int J,
        const float *__restrict__ z32, const float *__restrict__ w32,
        const float *__restrict__ wsc, float *__restrict__ part) {
    constexpr int NT = T16 * (T16 + 1) / 2;       // upper tiles
    extern __shared__ __attribute__((aligned(16))) _Float16 Xs[];   // [2 buf][2 hl][P16][GLDX]
    __shared__ __attribute__((aligned(16))) float wsl[3][GCH];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // workgroup -> (segment, gene); the 8 XCDs take workgroups round robin,
    // so the workgroups resident on one XCD share cell segments in its L2
    int sgi, j;
    {
        const int id = blockIdx.x;
        if (nseg % 8 == 0) {
            const int xcd = id % 8, loc = id / 8, nsx = nseg / 8;
            sgi = xcd + 8 * (loc % nsx);
            j = loc / nsx;
        } else {
            sgi = id % nseg;
            j = id / nseg;
        }
    }
    GramArgs g;
    g.Npad = Npad;
    g.c0 = sgi * seg;
    g.nch = (min(Npad, g.c0 + seg) - g.c0) / GCH;
    g.sc = wsc[j];
    g.z32 = z32;
    g.wj = w32 + (size_t)j * Npad;
    g.pg = part + ((size_t)sgi * J + j) * NT * 256;
    switch (w) {
    case 0: gram_wave<T16, 0>(g, Xs, wsl); break;
    case 1: gram_wave<T16, 1>(g, Xs, wsl); break;
    case 2: gram_wave<T16, 2>(g, Xs, wsl); break;
    case 3: gram_wave<T16, 3>(g, Xs, wsl); break;
    case 4: gram_wave<T16, 4>(g, Xs, wsl); break;
    case 5: gram_wave<T16, 5>(g, Xs, wsl); break;
    case 6: gram_wave<T16, 6>(g, Xs, wsl); break;
    default: gram_wave<T16, 7>(g, Xs, wsl); break;
    }
}

// ---------------------------------------------------------------------------
// Fused IRLS step + weighted Gram (getWForBeta, ggplvm.cu:468-514, and
// getPsi, ggplvm.cu:621-680, in one launch).  Workgroup = (gene, cell
// segment) as in k_gram_sym.  The FP16 hi/lo operands need a power-of-two
// weight scale 2^e >= max w over the cells they accumulate, so the workgroup
// first takes max(eta - d) over its segment (w = exp(eta - d) is monotone in
// it) and scales by 2^-e for that segment alone (k_gram_reduce_sym unscales
// each segment's partial).  Then the chunk loop of the symmetric Gram runs,
// wave 0 building each chunk's weights from eta two chunks ahead -- with the
// chunk's counts, nu, R = w nu, the FP32 weights and the sums -- instead of
// reading them.  W and nu are not stored: R and w32 are what the rest of the
// step reads (b = Zw^T R, the intercept row from w32).
template <int T16>
__global__ void __launch_bounds__(GW * 64) k_gram_irls(int N, int Npad, int seg, int nseg, int J,
        const float *__restrict__ z32, const double *__restrict__ eta,
        const double *__restrict__ cellc, const int *__restrict__ Yi,
        const double *__restrict__ Yx, const int *__restrict__ segptr, float *__restrict__ w32,
        double *__restrict__ R, double *__restrict__ sums, double *__restrict__ segun,
        float *__restrict__ part) {
    extern __shared__ __attribute__((aligned(16))) _Float16 Xs[];
    __shared__ __attribute__((aligned(16))) float wsl[3][GCH];
    __shared__ double ywl[GCH];
    __shared__ double red[GW];
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int sgi, j;
    {
        const int id = blockIdx.x;
        if (nseg % 8 == 0) {
            const int xcd = id % 8, loc = id / 8, nsx = nseg / 8;
            sgi = xcd + 8 * (loc % nsx);
            j = loc / nsx;
        } else {
            sgi = id % nseg;
            j = id / nseg;
        }
    }
    const int c0 = sgi * seg, c1 = min(Npad, c0 + seg);
    const double *ej = eta + (size_t)j * Npad;
    const double *cd = cellc;
    // the segment's largest eta - d over its cells, eight loads in flight
    double m = -INFINITY;
    {
        const int hi = min(c1, N);
        constexpr int MU = 8;
        int i = c0 + tid;
        for (; i + (MU - 1) * GW * 64 < hi; i += MU * GW * 64) {
            double e[MU], d[MU];
#pragma unroll
            for (int u = 0; u < MU; u++) {
                e[u] = ej[i + u * GW * 64];
                d[u] = cd[i + u * GW * 64];
            }
#pragma unroll
            for (int u = 0; u < MU; u++) m = fmax(m, e[u] - d[u]);
        }
        for (; i < hi; i += GW * 64) m = fmax(m, ej[i] - cd[i]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((tid & 63) == 0) red[w] = m;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int k = 1; k < GW; k++) m = fmax(m, red[k]);
    // 2^e >= the largest FP32 weight: the cell at the maximum has w = exp(m)
    const float wmax = (m > -INFINITY) ? (float)exp(m) : 0.0f;
    int e = 0;
    if (wmax > 0.0f) (void)frexpf(wmax, &e);
    if (tid == 0) segun[(size_t)sgi * J + j] = ldexp(1.0, e);

    GramArgs g;
    g.Npad = Npad;
    g.c0 = c0;
    g.nch = (c1 - c0) / GCH;
    g.sc = ldexpf(1.0f, -e);
    g.z32 = z32;
    g.wj = w32 + (size_t)j * Npad;
    g.pg = part + ((size_t)sgi * J + j) * (T16 * (T16 + 1) / 2) * 256;
    g.N = N;
    g.nz0 = segptr[(size_t)j * (nseg + 1) + sgi];
    g.nz1 = segptr[(size_t)j * (nseg + 1) + sgi + 1];
    g.eta = ej;
    g.cd = cellc;
    g.com = cellc + Npad;
    g.cbl = cellc + 2 * (size_t)Npad;
    g.Yi = Yi;
    g.Yx = Yx;
    g.Rj = R + (size_t)j * Npad;
    g.w32o = w32 + (size_t)j * Npad;
    g.ywl = ywl;
    g.sums = sums + ((size_t)sgi * J + j) * 4;
    switch (w) {
    case 0: gram_wave<T16, 0, true>(g, Xs, wsl); break;
    case 1: gram_wave<T16, 1, true>(g, Xs, wsl); break;
    case 2: gram_wave<T16, 2, true>(g, Xs, wsl); break;
    case 3: gram_wave<T16, 3, true>(g, Xs, wsl); break;
    case 4: gram_wave<T16, 4, true>(g, Xs, wsl); break;
    case 5: gram_wave<T16, 5, true>(g, Xs, wsl); break;
    case 6: gram_wave<T16, 6, true>(g, Xs, wsl); break;
    default: gram_wave<T16, 7, true>(g, Xs, wsl); break;
    }
}

// Per-cell constants of the fused IRLS pass: d_i = log Omega_i - log ls_i -
// dt t_i wt (dt = delta[p-1], t = the titsias term), Omega_i and log ls_i;
// zero on the padding cells (which the pass skips).
__global__ void k_irls_cellc(int N, int Npad, const double *__restrict__ Omega,
                             const double *__restrict__ ls, const double *__restrict__ tits,
                             const double *__restrict__ dtp, double wt, double *__restrict__ cellc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Npad) return;
    double d = 0.0, om = 0.0, bl = 0.0;
    if (i < N) {
        const double x = *dtp * tits[i] * wt;
        bl = log(ls[i]);
        om = Omega[i];
        d = (log(om) - bl) - x;
    }
    cellc[i] = d;
    cellc[Npad + i] = om;
    cellc[2 * (size_t)Npad + i] = bl;
}

// segptr[j (nseg + 1) + s] = the first non-zero of gene j at cell >= s seg
// (a lower bound in the gene's ascending cell list)
__global__ void k_irls_segptr(int J, int nseg, int seg, const int *__restrict__ Yp,
                              const int *__restrict__ Yi, int *__restrict__ segptr) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= J * (nseg + 1)) return;
    const int j = t / (nseg + 1), s = t % (nseg + 1);
    const long long c = (long long)s * seg;
    int lo = Yp[j], hi = Yp[j + 1];
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (Yi[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    segptr[t] = lo;
}

// gq_j, gs_j and the GLM objective of gene j from its segments' sums, added
// in segment order
__global__ void k_irls_sums(int J, int nseg, const double *__restrict__ sums, double *__restrict__ gq,
                            double *__restrict__ gs, double *__restrict__ lkhds) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= J) return;
    double q = 0.0, s = 0.0, sw = 0.0, ly = 0.0;
    for (int g = 0; g < nseg; g++) {
        const double *p = sums + ((size_t)g * J + j) * 4;
        q += p[0];
        s += p[1];
        sw += p[2];
        ly += p[3];
    }
    gq[j] = q;
    gs[j] = s;
    lkhds[j] = ly - sw;
}

// W and nu of the fused pass written out in the reference layout (from the
// same eta and per-cell constants, so the values the pass used): dense part
// here, the non-zeros' nu in k_irls_nu_nz
__global__ void k_irls_wnu(int N, int Npad, const double *__restrict__ eta,
                           const double *__restrict__ cellc, double *__restrict__ W,
                           double *__restrict__ Nu) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= Npad) return;
    const size_t a = i + (size_t)j * Npad;
    if (i >= N) { W[a] = 0.0; Nu[a] = 0.0; return; }
    const double e = eta[a];
    W[a] = exp(e - cellc[i]) * cellc[Npad + i];
    Nu[a] = (e + 0.0) - 1.0;
}

__global__ void k_irls_nu_nz(int Npad, const int *__restrict__ Yp, const int *__restrict__ Yi,
                             const double *__restrict__ Yx, const double *__restrict__ eta,
                             const double *__restrict__ W, double *__restrict__ Nu) {
    const int j = blockIdx.x;
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += 256) {
        const size_t a = Yi[l] + (size_t)j * Npad;
        const double y = Yx[l];
        const double q = y / W[a];
        Nu[a] = (eta[a] + (y > 0.0 ? q : 0.0)) - 1.0;
    }
}

// Wide design (the GP block of M = 512 columns, 32 tiles) in six parts per
// (gene, cell segment): MODE 1 = the symmetric halves h = 0, 1 (GP columns
// 256 h .. 256 h + 255), MODE 2 = the rectangles (ra, cb) (GP row columns
// 128 ra .., column columns 256 + 128 cb ..).  zg is the GP block's first
// column (z32 + Npad); partials go to the 528-tile layout k_gram_reduce_sym
// reads.  The segment-major XCD mapping of k_gram_sym, per (gene, part).
template <int MODE>
__global__ void __launch_bounds__(GW * 64) k_gram_wide(int Npad, int seg, int nseg, int J,
        const float *__restrict__ zg, const float *__restrict__ w32,
        const float *__restrict__ wsc, float *__restrict__ part) {
    constexpr int NP = MODE == 1 ? 2 : 4;         // parts of this kind
    constexpr int TGL = 32, NTG = TGL * (TGL + 1) / 2;
    extern __shared__ __attribute__((aligned(16))) _Float16 Xs[];
    __shared__ __attribute__((aligned(16))) float wsl[3][GCH];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int sgi, jp;
    {
        const int id = blockIdx.x;
        if (nseg % 8 == 0) {
            const int xcd = id % 8, loc = id / 8, nsx = nseg / 8;
            sgi = xcd + 8 * (loc % nsx);
            jp = loc / nsx;
        } else {
            sgi = id % nseg;
            jp = id / nseg;
        }
    }
    const int j = jp / NP, pt = jp % NP;
    GramArgs g;
    g.Npad = Npad;
    g.c0 = sgi * seg;
    g.nch = (min(Npad, g.c0 + seg) - g.c0) / GCH;
    g.sc = wsc[j];
    g.z32 = zg;
    g.wj = w32 + (size_t)j * Npad;
    g.pg = part + ((size_t)sgi * J + j) * NTG * 256;
    g.TGL = TGL;
    if constexpr (MODE == 1) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            g.zg[q] = zg + (size_t)(256 * pt + 64 * q) * Npad;
            g.tmap[q] = 16 * pt + 4 * q;
        }
    } else {
        const int ra = pt >> 1, cb = pt & 1;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            g.zg[q] = zg + (size_t)(128 * ra + 64 * q) * Npad;
            g.tmap[q] = 8 * ra + 4 * q;
            g.zg[2 + q] = zg + (size_t)(256 + 128 * cb + 64 * q) * Npad;
            g.tmap[2 + q] = 16 + 8 * cb + 4 * q;
        }
    }
    switch (w) {
    case 0: gram_wave<16, 0, false, MODE>(g, Xs, wsl); break;
    case 1: gram_wave<16, 1, false, MODE>(g, Xs, wsl); break;
    case 2: gram_wave<16, 2, false, MODE>(g, Xs, wsl); break;
    case 3: gram_wave<16, 3, false, MODE>(g, Xs, wsl); break;
    case 4: gram_wave<16, 4, false, MODE>(g, Xs, wsl); break;
    case 5: gram_wave<16, 5, false, MODE>(g, Xs, wsl); break;
    case 6: gram_wave<16, 6, false, MODE>(g, Xs, wsl); break;
    default: gram_wave<16, 7, false, MODE>(g, Xs, wsl); break;
    }
}

// ---- four-wave form: one wave per SIMD.  Each wave owns about a quarter of
// the upper tiles as blocks over 4-tile column groups and loads the fragments
// of every group it touches once per chunk, so blocks of one wave share them
// (at P16 = 272: 102 fragment reads per chunk instead of 170).  Blocks are
// dealt largest first to the least loaded wave, ties going to a wave that
// already touches one of the block's groups.  A single wave per SIMD hides
// latency by interleaving its own LDS reads, staging arithmetic and MFMAs.
constexpr int GW4 = 4;       // waves per workgroup
constexpr int MAXB4 = 8;     // blocks per wave

struct GramPlan4 {
    int nb[GW4];
    int a[GW4][MAXB4];
    int b[GW4][MAXB4];
};

constexpr GramPlan4 make_gram_plan4(int T16) {
    GramPlan4 p{};
    const int NG = (T16 + 3) / 4;
    int ba[16] = {}, bb[16] = {}, bc[16] = {};
    int nb = 0;
    for (int a = 0; a < NG; a++)
        for (int b = a; b < NG; b++) {
            ba[nb] = a; bb[nb] = b; bc[nb] = gram_blk_cost(T16, a, b);
            nb++;
        }
    for (int i = 1; i < nb; i++)
        for (int k = i; k > 0 && bc[k] > bc[k - 1]; k--) {
            int t = ba[k]; ba[k] = ba[k - 1]; ba[k - 1] = t;
            t = bb[k]; bb[k] = bb[k - 1]; bb[k - 1] = t;
            t = bc[k]; bc[k] = bc[k - 1]; bc[k - 1] = t;
        }
    int load[GW4] = {};
    unsigned gm[GW4] = {};
    for (int i = 0; i < nb; i++) {
        const unsigned need = (1u << ba[i]) | (1u << bb[i]);
        int best = -1;
        for (int v = 0; v < GW4; v++) {
            if (p.nb[v] >= MAXB4) continue;
            if (best < 0 || load[v] < load[best] ||
                (load[v] == load[best] && (gm[v] & need) && !(gm[best] & need)))
                best = v;
        }
        p.a[best][p.nb[best]] = ba[i];
        p.b[best][p.nb[best]] = bb[i];
        p.nb[best]++;
        load[best] += bc[i];
        gm[best] |= need;
    }
    return p;
}

constexpr unsigned gram_group_mask4(int T16, int W) {
    const GramPlan4 p = make_gram_plan4(T16);
    unsigned m = 0;
    for (int i = 0; i < p.nb[W]; i++) m |= (1u << p.a[W][i]) | (1u << p.b[W][i]);
    return m;
}

// one 32-cell step of block (A, B) from fragments already in registers
template <int T16, int A, int B>
__device__ __forceinline__ void gram_block4(f4 (&acc)[4][4], const h8 (&ah)[4], const h8 (&al)[4],
                                            const h8 (&bh)[4], const h8 (&bl)[4]) {
    constexpr int NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    constexpr int SA = A < NG - 1 ? 4 : RL, SB = B < NG - 1 ? 4 : RL;
    if constexpr (A == B) {
#pragma unroll
        for (int x = 0; x < SA; x++)
#pragma unroll
            for (int y = x; y < SA; y++) mfma3(acc[x][y], ah[x], al[x], ah[y], al[y]);
    } else {
#pragma unroll
        for (int y = 0; y < SB; y++)
#pragma unroll
            for (int x = 0; x < SA; x++) mfma3(acc[x][y], ah[x], al[x], bh[y], bl[y]);
    }
}

template <int T16, int W, int I>
__device__ __forceinline__ void gram_run4(f4 (&acc)[MAXB4][4][4], const h8 (&fh)[5][4],
                                          const h8 (&fl)[5][4]) {
    constexpr GramPlan4 pl = make_gram_plan4(T16);
    if constexpr (I < pl.nb[W]) {
        constexpr int A = pl.a[W][I], B = pl.b[W][I];
        gram_block4<T16, A, B>(acc[I], fh[A], fl[A], fh[B], fl[B]);
        gram_run4<T16, W, I + 1>(acc, fh, fl);
    }
}

template <int T16, int W, int I>
__device__ __forceinline__ void gram_store4(const f4 (&acc)[MAXB4][4][4], float *pg, int lane) {
    constexpr GramPlan4 pl = make_gram_plan4(T16);
    if constexpr (I < pl.nb[W]) {
        constexpr int A = pl.a[W][I], B = pl.b[W][I];
        gram_block_store<T16, A, B>(acc[I], pg, lane);
        gram_store4<T16, W, I + 1>(acc, pg, lane);
    }
}

template <int T16, int W>
__device__ __forceinline__ void gram_wave4(const GramArgs &g, _Float16 *Xs, float (*wsl)[CH]) {
    constexpr int P16 = 16 * T16, NG = (T16 + 3) / 4, RL = T16 - 4 * (NG - 1);
    constexpr unsigned GM = gram_group_mask4(T16, W);
    constexpr int NITEM = P16 * 4;                // (column, 8-cell group) items per chunk
    constexpr int IPT = (NITEM + GW4 * 64 - 1) / (GW4 * 64);
    const int tid = threadIdx.x, lane = tid & 63;
    const int lrow = lane & 15, lk = 8 * (lane >> 4);

    f4 acc[MAXB4][4][4];
#pragma unroll
    for (int i = 0; i < MAXB4; i++)
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) acc[i][x][y] = f4{0.f, 0.f, 0.f, 0.f};

    // staging as in the eight-wave form: straight-line, whole item rounds
    f4 pz[IPT][2];
    auto fetch = [&](int cc) {
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            if (W + GW4 * r < T16) {
                const int it = tid + GW4 * 64 * r;
                const float *src = g.z32 + (size_t)(it >> 2) * g.Npad + cc + 8 * (it & 3);
                pz[r][0] = *(const f4 *)src;
                pz[r][1] = *(const f4 *)(src + 4);
            }
        }
    };
    auto store = [&](int buf, int wb) {
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            if (!(W + GW4 * r < T16)) continue;
            const int it = tid + GW4 * 64 * r;
            const int col = it >> 2, q = it & 3;
            const f4 s0 = *(const f4 *)&wsl[wb][8 * q];
            const f4 s1 = *(const f4 *)&wsl[wb][8 * q + 4];
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (e < 4 ? pz[r][0][e] * s0[e] : pz[r][1][e - 4] * s1[e - 4]);
            u4 hi, lo;
            split8_pk(v, hi, lo);
            *(u4 *)&Xs[((buf * 2 + 0) * P16 + col) * LDX + 8 * q] = hi;
            *(u4 *)&Xs[((buf * 2 + 1) * P16 + col) * LDX + 8 * q] = lo;
        }
    };
    auto sqw = [&](int ch) { return sqrtf(g.wj[g.c0 + ch * CH + (tid & (CH - 1))] * g.sc); };

    wsl[0][tid & (CH - 1)] = sqw(0);
    wsl[1][tid & (CH - 1)] = sqw(min(1, g.nch - 1));
    fetch(g.c0);
    __syncthreads();
    store(0, 0);
    __syncthreads();

    for (int ch = 0; ch < g.nch; ch++) {
        const int buf = ch & 1;
        const float pw = g.wj[g.c0 + min(ch + 2, g.nch - 1) * CH + (tid & (CH - 1))];
        fetch(g.c0 + min(ch + 1, g.nch - 1) * CH);
        __builtin_amdgcn_sched_barrier(0);
        const _Float16 *xh = Xs + (size_t)(buf * 2 + 0) * P16 * LDX;
        const _Float16 *xl = Xs + (size_t)(buf * 2 + 1) * P16 * LDX;
        h8 fh[5][4], fl[5][4];
#pragma unroll
        for (int gi = 0; gi < NG; gi++) {
            if ((GM >> gi) & 1) {
                const int SG = gi < NG - 1 ? 4 : RL;
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    if (x < SG) {
                        const int col = (4 * gi + x) * 16 + lrow;
                        fh[gi][x] = *(const h8 *)(xh + col * LDX + lk);
                        fl[gi][x] = *(const h8 *)(xl + col * LDX + lk);
                    }
                }
            }
        }
        gram_run4<T16, W, 0>(acc, fh, fl);
        store(buf ^ 1, (ch + 1) % 3);   // past the last chunk: a copy never read
        __builtin_amdgcn_sched_barrier(0);
        wsl[(ch + 2) % 3][tid & (CH - 1)] = sqrtf(pw * g.sc);
        __syncthreads();
    }
    gram_store4<T16, W, 0>(acc, g.pg, lane);
}

template <int T16>
__global__ void __launch_bounds__(GW4 * 64) k_gram_sym4(int Npad, int seg, int nseg, int J,
        const float *__restrict__ z32, const float *__restrict__ w32,
        const float *__restrict__ wsc, float *__restrict__ part) {
    constexpr int NT = T16 * (T16 + 1) / 2;
    extern __shared__ __attribute__((aligned(16))) _Float16 Xs[];   // [2 buf][2 hl][P16][LDX]
    __shared__ __attribute__((aligned(16))) float wsl[3][CH];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int sgi, j;
    {
        const int id = blockIdx.x;
        if (nseg % 8 == 0) {
            const int xcd = id % 8, loc = id / 8, nsx = nseg / 8;
            sgi = xcd + 8 * (loc % nsx);
            j = loc / nsx;
        } else {
            sgi = id % nseg;
            j = id / nseg;
        }
    }
    GramArgs g;
    g.Npad = Npad;
    g.c0 = sgi * seg;
    g.nch = (min(Npad, g.c0 + seg) - g.c0) / CH;
    g.sc = wsc[j];
    g.z32 = z32;
    g.wj = w32 + (size_t)j * Npad;
    g.pg = part + ((size_t)sgi * J + j) * NT * 256;
    switch (w) {
    case 0: gram_wave4<T16, 0>(g, Xs, wsl); break;
    case 1: gram_wave4<T16, 1>(g, Xs, wsl); break;
    case 2: gram_wave4<T16, 2>(g, Xs, wsl); break;
    default: gram_wave4<T16, 3>(g, Xs, wsl); break;
    }
}

// ---- producer / consumer form (option gram_waves 12; 16 GP column tiles,
// M = 256): 12 waves, three per SIMD.  Waves 8-11 (one per SIMD) only stage:
// they hold the next chunk's z rows in registers, scale, split and store
// them to the LDS buffer the consumers are not reading, then load the chunk
// after.  Waves 0-7 (two per SIMD) only run MFMAs out of the other buffer,
// so the staging arithmetic and its global loads overlap the matrix work
// instead of alternating with it; one barrier per 64-cell chunk hands the
// buffers over.  The 136 upper tiles are cut into units (whole 4x4 blocks,
// diagonal blocks, and 4x2 / 2x4 halves of off-diagonal blocks) so that
// every SIMD's two consumers carry 34 tiles; a wave loads each tile's
// fragments it touches once per k-step (60 fragment pairs per k-step).
// Every accumulator sees the MFMA sequence and operands of k_gram_sym, so
// the two forms give the same bits.
constexpr int PC_CW = 8;     // consumer waves
constexpr int PC_PW = 4;     // producer waves
constexpr int PC_T16 = 16;   // the column tiles this form is planned for

struct PcUnit {
    int ra, nr, cb, nc;      // row tiles ra .. ra + nr - 1 x column tiles cb .. cb + nc - 1 (nr = 0: none)
};
struct PcWave {
    PcUnit u0, u1;
};

// wave w pairs with wave w + 4 on SIMD w mod 4: 18 + 16 = 34 tiles each
constexpr PcWave kPcPlan[PC_CW] = {
    {{0, 4, 0, 4}, {0, 4, 4, 2}},      // block (0,0); (0,1) columns 4-5
    {{4, 4, 4, 4}, {4, 4, 8, 2}},      // block (1,1); (1,2) columns 8-9
    {{8, 4, 8, 4}, {8, 4, 12, 2}},     // block (2,2); (2,3) columns 12-13
    {{12, 4, 12, 4}, {0, 2, 12, 4}},   // block (3,3); (0,3) rows 0-1
    {{0, 4, 8, 4}, {0, 0, 0, 0}},      // block (0,2)
    {{4, 4, 12, 4}, {0, 0, 0, 0}},     // block (1,3)
    {{0, 4, 6, 2}, {4, 4, 10, 2}},     // (0,1) columns 6-7; (1,2) columns 10-11
    {{8, 4, 14, 2}, {2, 2, 12, 4}},    // (2,3) columns 14-15; (0,3) rows 2-3
};

constexpr unsigned pc_mask(const PcUnit &u) {
    unsigned m = 0;
    for (int t = 0; t < u.nr; t++) m |= 1u << (u.ra + t);
    for (int t = 0; t < u.nc; t++) m |= 1u << (u.cb + t);
    return u.nr ? m : 0u;
}

// one k-step of unit U from the wave's fragments (fh / fl indexed by tile)
template <int RA, int NR, int CB, int NC>
__device__ __forceinline__ void pc_unit(f4 (&acc)[4][4], const h8 (&fh)[PC_T16], const h8 (&fl)[PC_T16]) {
    if constexpr (NR > 0) {
        if constexpr (RA == CB) {          // diagonal block: upper triangle
#pragma unroll
            for (int x = 0; x < NR; x++)
#pragma unroll
                for (int y = x; y < NC; y++) mfma3(acc[x][y], fh[RA + x], fl[RA + x], fh[CB + y], fl[CB + y]);
        } else {
#pragma unroll
            for (int y = 0; y < NC; y++)
#pragma unroll
                for (int x = 0; x < NR; x++) mfma3(acc[x][y], fh[RA + x], fl[RA + x], fh[CB + y], fl[CB + y]);
        }
    }
}

template <int RA, int NR, int CB, int NC>
__device__ __forceinline__ void pc_unit_store(const f4 (&acc)[4][4], float *pg, int lane) {
    if constexpr (NR > 0) {
#pragma unroll
        for (int x = 0; x < NR; x++)
#pragma unroll
            for (int y = (RA == CB ? x : 0); y < NC; y++) {
                const int ta = RA + x, tb = CB + y;
                const int t = ta * PC_T16 - ta * (ta - 1) / 2 + (tb - ta);
                *(f4 *)(pg + (size_t)t * 256 + 4 * lane) = acc[x][y];
            }
    }
}

template <int W>
__device__ __forceinline__ void pc_consumer(const GramArgs &g, const _Float16 *Xs) {
    constexpr int P16 = 16 * PC_T16;
    constexpr PcUnit U0 = kPcPlan[W].u0, U1 = kPcPlan[W].u1;
    constexpr unsigned MASK = pc_mask(U0) | pc_mask(U1);
    const int lane = threadIdx.x & 63;
    const int lrow = lane & 15, lk = 8 * (lane >> 4);
    f4 acc0[4][4], acc1[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            acc0[x][y] = f4{0.f, 0.f, 0.f, 0.f};
            acc1[x][y] = f4{0.f, 0.f, 0.f, 0.f};
        }
    __syncthreads();                              // chunk 0 staged
    for (int ch = 0; ch < g.nch; ch++) {
        const int buf = ch & 1;
        const _Float16 *xh = Xs + (size_t)(buf * 2 + 0) * P16 * GLDX;
        const _Float16 *xl = Xs + (size_t)(buf * 2 + 1) * P16 * GLDX;
#pragma unroll
        for (int s = 0; s < GKC; s++) {
            h8 fh[PC_T16], fl[PC_T16];
#pragma unroll
            for (int t = 0; t < PC_T16; t++) {
                if ((MASK >> t) & 1u) {              // folded: the loop is unrolled
                    const int off = (t * 16 + lrow) * GLDX + lk + 32 * s;
                    fh[t] = *(const h8 *)(xh + off);
                    fl[t] = *(const h8 *)(xl + off);
                }
            }
            pc_unit<U0.ra, U0.nr, U0.cb, U0.nc>(acc0, fh, fl);
            pc_unit<U1.ra, U1.nr, U1.cb, U1.nc>(acc1, fh, fl);
        }
        __syncthreads();                          // this buffer read; the next one staged
    }
    pc_unit_store<U0.ra, U0.nr, U0.cb, U0.nc>(acc0, g.pg, lane);
    pc_unit_store<U1.ra, U1.nr, U1.cb, U1.nc>(acc1, g.pg, lane);
}

// producer wave PW: columns 64 PW .. 64 PW + 63 of every chunk; lane l
// handles cells 8 (l mod 8) .. + 7 of the columns 8 r + l / 8 (rounds r < 8),
// so its eight square-root weights are the same in every round
template <int PW>
__device__ __forceinline__ void pc_producer(const GramArgs &g, _Float16 *Xs) {
    constexpr int P16 = 16 * PC_T16;
    constexpr int NRD = P16 / PC_PW / 8;         // rounds of 8 columns per wave
    const int lane = threadIdx.x & 63;
    const int q = lane & 7, cl = lane >> 3;
    f4 pz[NRD][2];
    f4 pw[2];
    auto load = [&](int ch) {                     // chunk ch (clamped by the caller)
        const int cell = g.c0 + ch * GCH + 8 * q;
#pragma unroll
        for (int r = 0; r < NRD; r++) {
            const float *src = g.z32 + (size_t)(64 * PW + 8 * r + cl) * g.Npad + cell;
            pz[r][0] = *(const f4 *)src;
            pz[r][1] = *(const f4 *)(src + 4);
        }
        pw[0] = *(const f4 *)(g.wj + cell);
        pw[1] = *(const f4 *)(g.wj + cell + 4);
    };
    auto stage = [&](int buf) {
        float sw[8];
#pragma unroll
        for (int e = 0; e < 8; e++) sw[e] = sqrtf((e < 4 ? pw[0][e] : pw[1][e - 4]) * g.sc);
#pragma unroll
        for (int r = 0; r < NRD; r++) {
            const int col = 64 * PW + 8 * r + cl;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (e < 4 ? pz[r][0][e] : pz[r][1][e - 4]) * sw[e];
            u4 hi, lo;
            split8_pk(v, hi, lo);
            *(u4 *)&Xs[((buf * 2 + 0) * P16 + col) * GLDX + 8 * q] = hi;
            *(u4 *)&Xs[((buf * 2 + 1) * P16 + col) * GLDX + 8 * q] = lo;
        }
    };
    load(0);
    stage(0);
    load(min(1, g.nch - 1));
    __syncthreads();                              // chunk 0 staged
    for (int ch = 0; ch < g.nch; ch++) {
        // chunk ch + 1 into the buffer the consumers are not reading (past
        // the last chunk: a copy of it, never read), then the loads of ch + 2
        stage((ch + 1) & 1);
        load(min(ch + 2, g.nch - 1));
        __syncthreads();
    }
}

__global__ void __launch_bounds__((PC_CW + PC_PW) * 64) k_gram_pc(int Npad, int seg, int nseg, int J,
        const float *__restrict__ z32, const float *__restrict__ w32,
        const float *__restrict__ wsc, float *__restrict__ part) {
    constexpr int NT = PC_T16 * (PC_T16 + 1) / 2;
    extern __shared__ __attribute__((aligned(16))) _Float16 Xs[];   // [2 buf][2 hl][P16][GLDX]
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int sgi, j;
    {
        const int id = blockIdx.x;
        if (nseg % 8 == 0) {
            const int xcd = id % 8, loc = id / 8, nsx = nseg / 8;
            sgi = xcd + 8 * (loc % nsx);
            j = loc / nsx;
        } else {
            sgi = id % nseg;
            j = id / nseg;
        }
    }
    GramArgs g;
    g.Npad = Npad;
    g.c0 = sgi * seg;
    g.nch = (min(Npad, g.c0 + seg) - g.c0) / GCH;
    g.sc = wsc[j];
    g.z32 = z32;
    g.wj = w32 + (size_t)j * Npad;
    g.pg = part + ((size_t)sgi * J + j) * NT * 256;
    switch (w) {
    case 0: pc_consumer<0>(g, Xs); break;
    case 1: pc_consumer<1>(g, Xs); break;
    case 2: pc_consumer<2>(g, Xs); break;
    case 3: pc_consumer<3>(g, Xs); break;
    case 4: pc_consumer<4>(g, Xs); break;
    case 5: pc_consumer<5>(g, Xs); break;
    case 6: pc_consumer<6>(g, Xs); break;
    case 7: pc_consumer<7>(g, Xs); break;
    case 8: pc_producer<0>(g, Xs); break;
    case 9: pc_producer<1>(g, Xs); break;
    case 10: pc_producer<2>(g, Xs); break;
    default: pc_producer<3>(g, Xs); break;
    }
}

constexpr int RSG = 4;       // genes per block of the symmetric reduction
constexpr int GRK_SEG = 64;  // cell segments of the intercept row's FP32 products

// Psi_j from the symmetric-form partials: one 256-thread block per (tile, RSG
// genes), 64 threads per gene, each thread one 16-byte load per segment (the
// four values of one MFMA lane: tile rows 4 (l / 16) + r, column l % 16);
// sums the segments in FP64 in segment order, each rescaled by its weight
// unscale 2^e (wun[s wstride + j]: wstride 0 for one scale per gene, J for the
// fused IRLS pass's per-segment scales; powers of two, so a common scale gives
// the bits of scaling the sum), rescales by 2^-s_r 2^-s_c, adds the prior and
// writes the tile and its mirror (upper-triangle values only, so Psi is
// exactly symmetric).
// off: the tiles cover columns off .. off + 16 T16 - 1 of Psi (1 when the
// intercept row and column come from k_gram_row0)
__global__ void __launch_bounds__(256) k_gram_reduce_sym(int P, int T16, int nseg, int J, int off,
        const float *__restrict__ part, const double *__restrict__ wun, size_t wstride,
        const double *__restrict__ zun, const double *__restrict__ Delta, double *__restrict__ Psi) {
    static_assert(RSG * 64 == 256, "one 64-thread slot per gene of the block");
    __shared__ double tile[RSG][16][17];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int NT = T16 * (T16 + 1) / 2;
    int ta = 0, rem = t;
    while (rem >= T16 - ta) { rem -= T16 - ta; ta++; }
    const int tb = ta + rem;
    const int q = tid >> 6, l = tid & 63;
    const int j = blockIdx.y * RSG + q;
    const int cc = l & 15, r0 = 4 * (l >> 4);
    const int col = off + 16 * tb + cc;
    if (j < J) {
        // the segments' partials in order, GRU 16-byte loads in flight at a time
        constexpr int GRU = 8;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        int g = 0;
        for (; g + GRU <= nseg; g += GRU) {
            float4 pv[GRU];
#pragma unroll
            for (int u = 0; u < GRU; u++)
                pv[u] = *(const float4 *)(part + (((size_t)(g + u) * J + j) * NT + t) * 256 + 4 * l);
#pragma unroll
            for (int u = 0; u < GRU; u++) {
                const double wu = wun[(size_t)(g + u) * wstride + j];
                s[0] += (double)pv[u].x * wu;
                s[1] += (double)pv[u].y * wu;
                s[2] += (double)pv[u].z * wu;
                s[3] += (double)pv[u].w * wu;
            }
        }
        for (; g < nseg; g++) {
            const float4 pv = *(const float4 *)(part + (((size_t)g * J + j) * NT + t) * 256 + 4 * l);
            const double wu = wun[(size_t)g * wstride + j];
            s[0] += (double)pv.x * wu;
            s[1] += (double)pv.y * wu;
            s[2] += (double)pv.z * wu;
            s[3] += (double)pv.w * wu;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = off + 16 * ta + r0 + r;
            double v = 0.0;
            if (row < P && col < P) v = s[r] * (zun[row] * zun[col]) + Delta[row + (size_t)col * P];
            tile[q][r0 + r][cc] = v;
        }
    }
    __syncthreads();
    if (j >= J) return;
    double *Pj = Psi + (size_t)j * P * P;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int e = l + 64 * m, i1 = e & 15, i2 = e >> 4;
        {   // Psi[16ta + i1, 16tb + i2]   (coalesced over i1)
            const int R = off + 16 * ta + i1, C = off + 16 * tb + i2;
            if (R < P && C < P && (ta != tb || i1 <= i2)) Pj[R + (size_t)C * P] = tile[q][i1][i2];
        }
        {   // mirror: Psi[16tb + i1, 16ta + i2] = tile[i2][i1]
            const int R = off + 16 * tb + i1, C = off + 16 * ta + i2;
            if (R < P && C < P && (ta != tb || i2 <= i1)) Pj[R + (size_t)C * P] = tile[q][i2][i1];
        }
    }
}

// The intercept row and column of Psi_j when the symmetric Gram covers only
// the GP block: Psi_j[0, c] = sum_i o_ij zw_ic + Delta[0, c], from the
// per-segment FP32 products part[s][c + j P] = sum_{i in s} z32_ic w32_ij
// (z32 = zw 2^s_c) summed in FP64 and unscaled by zun[c] = 2^-s_c.
__global__ void k_gram_row0(int P, int J, int nsk, const float *__restrict__ part,
                            const double *__restrict__ zun, const double *__restrict__ Delta,
                            double *__restrict__ Psi) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)P * J) return;
    const int c = (int)(idx % P), j = (int)(idx / P);
    double s = 0.0;
    for (int g = 0; g < nsk; g++) s += (double)part[(size_t)g * P * J + idx];
    const double v = s * zun[c] + Delta[(size_t)c * P];
    double *Pj = Psi + (size_t)j * P * P;
    Pj[(size_t)c * P] = v;
    Pj[c] = v;
}

}  // namespace


// Scratch sizes for the mixed Gram (allocated on first use).
static int ensure_mixed(gp_ctx *c, int nseg) {
    const int T16 = (c->P + 15) / 16, P16 = 16 * T16;
    const int nbk = (P16 + BLK - 1) / BLK;
    const size_t nbp = (size_t)nbk * (nbk + 1) / 2;
    const size_t zbytes = (size_t)c->Npad * P16 * sizeof(_Float16);
    const size_t pbytes = (size_t)nseg * c->J * nbp * BLK * BLK * sizeof(float);
    if (!c->Zh) {
        GP_HIP(hipMalloc(&c->Zh, zbytes));
        GP_HIP(hipMalloc(&c->Zl, zbytes));
        GP_HIP(hipMalloc((void **)&c->z32, (size_t)c->Npad * P16 * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->w32, (size_t)c->Npad * c->J * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->wsc, (size_t)c->J * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->wun, (size_t)c->J * sizeof(double)));
        GP_HIP(hipMalloc((void **)&c->zun, (size_t)P16 * sizeof(double)));
    }
    if (c->gram_part_bytes < pbytes) {
        if (c->gram_part) (void)hipFree(c->gram_part);
        c->gram_part = nullptr;
        GP_HIP(hipMalloc((void **)&c->gram_part, pbytes));
        c->gram_part_bytes = pbytes;
    }
    return 0;
}

int gpk_split_z(gp_ctx *c, const double *Zsrc) {
    const int T16 = (c->P + 15) / 16, P16 = 16 * T16;
    GP_TRY(ensure_mixed(c, c->gram_splits));
    k_split_z<<<P16, 256, 0, c->stream>>>(c->Npad, c->P, Zsrc, (_Float16 *)c->Zh,
                                          (_Float16 *)c->Zl, c->z32, c->zun);
    GP_HIP(hipGetLastError());
    c->zsplit_valid = 1;
    return 0;
}

int gpk_mixed_alloc(gp_ctx *c) { return ensure_mixed(c, c->gram_splits); }

int gpk_wprep(gp_ctx *c, const double *W) {
    GP_TRY(ensure_mixed(c, c->gram_splits));
    k_wprep<<<c->J, 256, 0, c->stream>>>(c->N, c->Npad, W, c->Omega, c->w32, c->wsc, c->wun);
    GP_HIP(hipGetLastError());
    c->w32_valid = (W == c->W);
    return 0;
}

template <int T16>
static int launch_gram_sym(gp_ctx *c, int seg, int nseg, int batch, const float *zsrc) {
    constexpr int P16 = 16 * T16;
    static_assert(plan_fits(T16), "block plan does not fit the waves");
    if constexpr (T16 == PC_T16) {
        if (c->gram_waves == 12) {
            const size_t shm = (size_t)2 * 2 * P16 * GLDX * sizeof(_Float16);
            if (!c->capturing)
                GP_HIP(hipFuncSetAttribute((const void *)k_gram_pc,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            k_gram_pc<<<(unsigned)(batch * nseg), (PC_CW + PC_PW) * 64, shm, c->stream>>>(
                c->Npad, seg, nseg, batch, zsrc, c->w32, c->wsc, c->gram_part);
            GP_HIP(hipGetLastError());
            return 0;
        }
    }
    if (c->gram_waves == 4) {
        const size_t shm = (size_t)2 * 2 * P16 * LDX * sizeof(_Float16);
        if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
            GP_HIP(hipFuncSetAttribute((const void *)k_gram_sym4<T16>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        k_gram_sym4<T16><<<(unsigned)(batch * nseg), GW4 * 64, shm, c->stream>>>(
            c->Npad, seg, nseg, batch, zsrc, c->w32, c->wsc, c->gram_part);
    } else {
        const size_t shm = (size_t)2 * 2 * P16 * GLDX * sizeof(_Float16);
        if (shm > 65536 && !c->capturing)
            GP_HIP(hipFuncSetAttribute((const void *)k_gram_sym<T16>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        k_gram_sym<T16><<<(unsigned)(batch * nseg), GW * 64, shm, c->stream>>>(
            c->Npad, seg, nseg, batch, zsrc, c->w32, c->wsc, c->gram_part);
    }
    GP_HIP(hipGetLastError());
    return 0;
}

// The intercept row and column of the peeled forms: GRK_SEG cell segments
// of z32^T w32 (FP32 products over Npad / GRK_SEG cells each), summed in FP64
static int gram_intercept_row(gp_ctx *c, int batch, const double *Delta, double *Psi) {
    const size_t rp = (size_t)GRK_SEG * c->P * batch;
    if (c->gram_row0_len < rp) {
        if (c->gram_row0) (void)hipFree(c->gram_row0);
        c->gram_row0 = nullptr;
        GP_HIP(hipMalloc((void **)&c->gram_row0, rp * sizeof(float)));
        c->gram_row0_len = rp;
    }
    const int kseg = c->Npad / GRK_SEG;
    const float one = 1.0f, zero = 0.0f;
    GP_BLAS(rocblas_sgemm_strided_batched(c->blas, rocblas_operation_transpose, rocblas_operation_none,
                                          c->P, batch, kseg, &one, c->z32, c->Npad, kseg, c->w32,
                                          c->Npad, kseg, &zero, c->gram_row0, c->P,
                                          (rocblas_stride)c->P * batch, GRK_SEG));
    const size_t n0 = (size_t)c->P * batch;
    k_gram_row0<<<(unsigned)((n0 + 255) / 256), 256, 0, c->stream>>>(c->P, batch, GRK_SEG,
                                                                    c->gram_row0, c->zun, Delta, Psi);
    GP_HIP(hipGetLastError());
    return 0;
}

// The intercept row and column (FP32 products of w32 and z32, summed in
// FP64) need only the weights and the model matrix and write only row and
// column 0 of Psi: forked to the side stream after the Gram kernel, they run
// beside the segment reduction (which writes the GP block); the caller's
// factorization waits for them (gram_row_join).
static int gram_row_fork(gp_ctx *c, int batch, const double *Delta, double *Psi) {
    GP_HIP(hipEventRecord(c->ev_rfork, c->stream));
    GP_HIP(hipStreamWaitEvent(c->stream2, c->ev_rfork, 0));
    {
        SideStream side(c);
        GP_TRY(gram_intercept_row(c, batch, Delta, Psi));
    }
    GP_HIP(hipEventRecord(c->ev_row0, c->stream2));
    return 0;
}
static int gram_row_join(gp_ctx *c) {
    GP_HIP(hipStreamWaitEvent(c->stream, c->ev_row0, 0));
    return 0;
}

int gpk_gram_mixed(gp_ctx *c, const double *W, double *Psi, int batch, const double *Zsrc,
                   const double *Delta, bool weights_ready) {
    const int T16 = (c->P + 15) / 16, P16 = 16 * T16;
    const int nbk = (P16 + BLK - 1) / BLK;
    const int nbp = nbk * (nbk + 1) / 2;
    int nseg = c->gram_splits > 0 ? c->gram_splits : 1;
    GP_TRY(ensure_mixed(c, nseg));
    if (!c->zsplit_valid) GP_TRY(gpk_split_z(c, Zsrc));
    // segment length: multiple of the chunk of either symmetric form (GCH is a
    // multiple of CH; Npad is a multiple of 256), covering Npad
    int seg = (c->Npad + nseg - 1) / nseg;
    seg = (seg + GCH - 1) / GCH * GCH;
    nseg = (c->Npad + seg - 1) / seg;
    if (!weights_ready) {
        k_wprep<<<batch, 256, 0, c->stream>>>(c->N, c->Npad, W, c->Omega, c->w32, c->wsc, c->wun);
        GP_HIP(hipGetLastError());
        c->w32_valid = (W == c->W && batch == c->J);
    }
    // intercept-peeled form: the symmetric Gram over the GP block only (M / 16
    // tiles: no 16-column tile for the single intercept column), the
    // intercept row and column from split-K FP32 products Zw^T (W / Omega)
    const int TG = c->M / 16;
    const bool peel = c->gram_peel && c->g0 == 1 && c->M % 16 == 0 && TG <= MAX_SYM_T16 &&
                      TG >= 1 && c->Npad % GRK_SEG == 0;
    if (peel) {
        int r = 0;
        gp_timer_begin(c, c->t_gramk);
        const float *zg = c->z32 + c->Npad;
        switch (TG) {
        case 1: r = launch_gram_sym<1>(c, seg, nseg, batch, zg); break;
        case 2: r = launch_gram_sym<2>(c, seg, nseg, batch, zg); break;
        case 3: r = launch_gram_sym<3>(c, seg, nseg, batch, zg); break;
        case 4: r = launch_gram_sym<4>(c, seg, nseg, batch, zg); break;
        case 5: r = launch_gram_sym<5>(c, seg, nseg, batch, zg); break;
        case 6: r = launch_gram_sym<6>(c, seg, nseg, batch, zg); break;
        case 7: r = launch_gram_sym<7>(c, seg, nseg, batch, zg); break;
        case 8: r = launch_gram_sym<8>(c, seg, nseg, batch, zg); break;
        case 9: r = launch_gram_sym<9>(c, seg, nseg, batch, zg); break;
        case 10: r = launch_gram_sym<10>(c, seg, nseg, batch, zg); break;
        case 11: r = launch_gram_sym<11>(c, seg, nseg, batch, zg); break;
        case 12: r = launch_gram_sym<12>(c, seg, nseg, batch, zg); break;
        case 13: r = launch_gram_sym<13>(c, seg, nseg, batch, zg); break;
        case 14: r = launch_gram_sym<14>(c, seg, nseg, batch, zg); break;
        case 15: r = launch_gram_sym<15>(c, seg, nseg, batch, zg); break;
        case 16: r = launch_gram_sym<16>(c, seg, nseg, batch, zg); break;
        default: r = launch_gram_sym<17>(c, seg, nseg, batch, zg); break;
        }
        if (r) return r;
        gp_timer_end(c, c->t_gramk);
        GP_TRY(gram_row_fork(c, batch, Delta, Psi));
        dim3 gr((unsigned)(TG * (TG + 1) / 2), (unsigned)((batch + RSG - 1) / RSG));
        k_gram_reduce_sym<<<gr, 256, 0, c->stream>>>(c->P, TG, nseg, batch, 1, c->gram_part, c->wun, 0,
                                                      c->zun, Delta, Psi);
        GP_HIP(hipGetLastError());
        return gram_row_join(c);
    }
    if (c->gram_peel && c->gram_wide && c->g0 == 1 && c->M == 512 && c->Npad % GRK_SEG == 0) {
        // M = 512: two symmetric halves and four rectangles of the 32-tile GP
        // block on the symmetric kernel's chunk loop, reduced as one 32-tile
        // triangle at offset 1
        constexpr size_t shm = (size_t)2 * 2 * 256 * GLDX * sizeof(_Float16);
        if (!c->capturing) {
            GP_HIP(hipFuncSetAttribute((const void *)k_gram_wide<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
            GP_HIP(hipFuncSetAttribute((const void *)k_gram_wide<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
        }
        gp_timer_begin(c, c->t_gramk);
        const float *zg = c->z32 + c->Npad;
        k_gram_wide<1><<<(unsigned)(2 * batch * nseg), GW * 64, shm, c->stream>>>(c->Npad, seg, nseg, batch, zg,
                                                                                c->w32, c->wsc, c->gram_part);
        GP_HIP(hipGetLastError());
        k_gram_wide<2><<<(unsigned)(4 * batch * nseg), GW * 64, shm, c->stream>>>(c->Npad, seg, nseg, batch, zg,
                                                                                c->w32, c->wsc, c->gram_part);
        GP_HIP(hipGetLastError());
        gp_timer_end(c, c->t_gramk);
        GP_TRY(gram_row_fork(c, batch, Delta, Psi));
        dim3 gr((unsigned)(32 * 33 / 2), (unsigned)((batch + RSG - 1) / RSG));
        k_gram_reduce_sym<<<gr, 256, 0, c->stream>>>(c->P, 32, nseg, batch, 1, c->gram_part, c->wun, 0,
                                                      c->zun, Delta, Psi);
        GP_HIP(hipGetLastError());
        return gram_row_join(c);
    }
    // the same peeling for other designs wider than the symmetric kernel's
    // tiles: the tiled general kernel over the GP block of the split model
    // matrix (columns 1..M of Zh / Zl), its reduction written at offset 1
    if (c->gram_peel && c->g0 == 1 && c->M % 16 == 0 && TG > MAX_SYM_T16 && c->Npad % GRK_SEG == 0) {
        const int nbk_g = (c->M + BLK - 1) / BLK;
        const int nbp_g = nbk_g * (nbk_g + 1) / 2;
        gp_timer_begin(c, c->t_gramk);
        dim3 g(nbp_g, batch, nseg);
        k_gram_mixed<<<g, 256, 0, c->stream>>>(c->Npad, c->M, TG, nbk_g, seg,
                                                (const _Float16 *)c->Zh + c->Npad,
                                                (const _Float16 *)c->Zl + c->Npad, c->w32, c->wsc,
                                                c->gram_part);
        GP_HIP(hipGetLastError());
        gp_timer_end(c, c->t_gramk);
        dim3 gr((unsigned)(((size_t)c->M * c->M + 255) / 256), batch);
        k_gram_reduce<<<gr, 256, 0, c->stream>>>(c->P, c->M, 1, nbk_g, nseg, batch, c->gram_part, c->wun,
                                                  c->zun, Delta, Psi);
        GP_HIP(hipGetLastError());
        return gram_intercept_row(c, batch, Delta, Psi);
    }
    if (T16 <= MAX_SYM_T16) {
        int r = 0;
        gp_timer_begin(c, c->t_gramk);
        switch (T16) {
        case 1: r = launch_gram_sym<1>(c, seg, nseg, batch, c->z32); break;
        case 2: r = launch_gram_sym<2>(c, seg, nseg, batch, c->z32); break;
        case 3: r = launch_gram_sym<3>(c, seg, nseg, batch, c->z32); break;
        case 4: r = launch_gram_sym<4>(c, seg, nseg, batch, c->z32); break;
        case 5: r = launch_gram_sym<5>(c, seg, nseg, batch, c->z32); break;
        case 6: r = launch_gram_sym<6>(c, seg, nseg, batch, c->z32); break;
        case 7: r = launch_gram_sym<7>(c, seg, nseg, batch, c->z32); break;
        case 8: r = launch_gram_sym<8>(c, seg, nseg, batch, c->z32); break;
        case 9: r = launch_gram_sym<9>(c, seg, nseg, batch, c->z32); break;
        case 10: r = launch_gram_sym<10>(c, seg, nseg, batch, c->z32); break;
        case 11: r = launch_gram_sym<11>(c, seg, nseg, batch, c->z32); break;
        case 12: r = launch_gram_sym<12>(c, seg, nseg, batch, c->z32); break;
        case 13: r = launch_gram_sym<13>(c, seg, nseg, batch, c->z32); break;
        case 14: r = launch_gram_sym<14>(c, seg, nseg, batch, c->z32); break;
        case 15: r = launch_gram_sym<15>(c, seg, nseg, batch, c->z32); break;
        case 16: r = launch_gram_sym<16>(c, seg, nseg, batch, c->z32); break;
        default: r = launch_gram_sym<17>(c, seg, nseg, batch, c->z32); break;
        }
        if (r) return r;
        gp_timer_end(c, c->t_gramk);
        dim3 gr((unsigned)(T16 * (T16 + 1) / 2), (unsigned)((batch + RSG - 1) / RSG));
        k_gram_reduce_sym<<<gr, 256, 0, c->stream>>>(c->P, T16, nseg, batch, 0, c->gram_part, c->wun, 0,
                                                      c->zun, Delta, Psi);
        GP_HIP(hipGetLastError());
        return 0;
    }
    dim3 g(nbp, batch, nseg);
    k_gram_mixed<<<g, 256, 0, c->stream>>>(c->Npad, P16, T16, nbk, seg, (const _Float16 *)c->Zh,
                                            (const _Float16 *)c->Zl, c->w32, c->wsc, c->gram_part);
    GP_HIP(hipGetLastError());
    dim3 gr((unsigned)(((size_t)c->P * c->P + 255) / 256), batch);
    k_gram_reduce<<<gr, 256, 0, c->stream>>>(c->P, c->P, 0, nbk, nseg, batch, c->gram_part, c->wun,
                                              c->zun, Delta, Psi);
    GP_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Fused IRLS step on the mixed path: eta = Z beta (rocBLAS, FP64), the
// per-cell constants, then one launch of k_gram_irls that builds the weights,
// nu, R and the sums and accumulates the symmetric Gram; the per-gene sums,
// the segment reduction (per-segment unscale) and, in the peeled form, the
// intercept row from the w32 it wrote.

// the symmetric eight-wave kernel covers this design: the GP block (peeled)
// or the whole design in at most MAX_SYM_T16 column tiles
static bool irls_peel(const gp_ctx *c) {
    return c->gram_peel && c->g0 == 1 && c->M % 16 == 0 && c->M / 16 <= MAX_SYM_T16 &&
           c->Npad % GRK_SEG == 0;
}

bool gpk_gram_irls_fits(const gp_ctx *c) {
    return irls_peel(c) || (c->P + 15) / 16 <= MAX_SYM_T16;
}

template <int T16>
static int launch_gram_irls(gp_ctx *c, int seg, int nseg, const float *zsrc) {
    constexpr int P16 = 16 * T16;
    const size_t shm = (size_t)2 * 2 * P16 * GLDX * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)
        GP_HIP(hipFuncSetAttribute((const void *)k_gram_irls<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    k_gram_irls<T16><<<(unsigned)(c->J * nseg), GW * 64, shm, c->stream>>>(
        c->N, c->Npad, seg, nseg, c->J, zsrc, c->ZB, c->irls_cc, c->Yi, c->Yx, c->irls_segptr,
        c->w32, c->R, c->irls_sums, c->irls_segun, c->gram_part);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_gram_irls(gp_ctx *c, double *Psi, const double *Delta) {
    if (!gpk_gram_irls_fits(c)) { gp_set_error("fused IRLS Gram: design too wide"); return 1; }
    const int T16 = (c->P + 15) / 16;
    int nseg = c->gram_splits > 0 ? c->gram_splits : 1;
    GP_TRY(ensure_mixed(c, nseg));
    if (!c->zsplit_valid) GP_TRY(gpk_split_z(c, c->Zw));
    int seg = (c->Npad + nseg - 1) / nseg;
    seg = (seg + GCH - 1) / GCH * GCH;
    nseg = (c->Npad + seg - 1) / seg;
    if (!c->irls_cc) {
        GP_HIP(hipMalloc((void **)&c->irls_cc, (size_t)3 * c->Npad * sizeof(double)));
    }
    if (c->irls_nseg != nseg || c->irls_seg != seg) {
        if (c->irls_segptr) (void)hipFree(c->irls_segptr);
        if (c->irls_sums) (void)hipFree(c->irls_sums);
        if (c->irls_segun) (void)hipFree(c->irls_segun);
        c->irls_segptr = nullptr;
        c->irls_sums = c->irls_segun = nullptr;
        c->irls_nseg = c->irls_seg = 0;
        const size_t ns = (size_t)c->J * (nseg + 1);
        GP_HIP(hipMalloc((void **)&c->irls_segptr, ns * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->irls_sums, (size_t)nseg * c->J * 4 * sizeof(double)));
        GP_HIP(hipMalloc((void **)&c->irls_segun, (size_t)nseg * c->J * sizeof(double)));
        k_irls_segptr<<<(unsigned)((ns + 255) / 256), 256, 0, c->stream>>>(c->J, nseg, seg, c->Yp,
                                                                           c->Yi, c->irls_segptr);
        GP_HIP(hipGetLastError());
        c->irls_nseg = nseg;
        c->irls_seg = seg;
    }
    // eta = Z beta for every cell and gene (beta in the reference basis)
    const double one = 1.0, zero = 0.0;
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad, c->J,
                          c->P, &one, c->Z, c->Npad, c->Beta, c->P, &zero, c->ZB, c->Npad));
    k_irls_cellc<<<(unsigned)((c->Npad + 255) / 256), 256, 0, c->stream>>>(
        c->N, c->Npad, c->Omega, c->ls, c->titsias, c->delta + c->p - 1, 0.1, c->irls_cc);
    GP_HIP(hipGetLastError());
    const bool peel = irls_peel(c);
    const int T = peel ? c->M / 16 : T16;
    const float *zg = peel ? c->z32 + c->Npad : c->z32;
    int r = 0;
    gp_timer_begin(c, c->t_gramk);
    switch (T) {
    case 1: r = launch_gram_irls<1>(c, seg, nseg, zg); break;
    case 2: r = launch_gram_irls<2>(c, seg, nseg, zg); break;
    case 3: r = launch_gram_irls<3>(c, seg, nseg, zg); break;
    case 4: r = launch_gram_irls<4>(c, seg, nseg, zg); break;
    case 5: r = launch_gram_irls<5>(c, seg, nseg, zg); break;
    case 6: r = launch_gram_irls<6>(c, seg, nseg, zg); break;
    case 7: r = launch_gram_irls<7>(c, seg, nseg, zg); break;
    case 8: r = launch_gram_irls<8>(c, seg, nseg, zg); break;
    case 9: r = launch_gram_irls<9>(c, seg, nseg, zg); break;
    case 10: r = launch_gram_irls<10>(c, seg, nseg, zg); break;
    case 11: r = launch_gram_irls<11>(c, seg, nseg, zg); break;
    case 12: r = launch_gram_irls<12>(c, seg, nseg, zg); break;
    case 13: r = launch_gram_irls<13>(c, seg, nseg, zg); break;
    case 14: r = launch_gram_irls<14>(c, seg, nseg, zg); break;
    case 15: r = launch_gram_irls<15>(c, seg, nseg, zg); break;
    case 16: r = launch_gram_irls<16>(c, seg, nseg, zg); break;
    default: r = launch_gram_irls<17>(c, seg, nseg, zg); break;
    }
    if (r) return r;
    gp_timer_end(c, c->t_gramk);
    k_irls_sums<<<(unsigned)((c->J + 255) / 256), 256, 0, c->stream>>>(c->J, nseg, c->irls_sums, c->gq,
                                                                      c->gs, c->lkhds);
    GP_HIP(hipGetLastError());
    if (peel) GP_TRY(gram_row_fork(c, c->J, Delta, Psi));
    dim3 gr((unsigned)(T * (T + 1) / 2), (unsigned)((c->J + RSG - 1) / RSG));
    k_gram_reduce_sym<<<gr, 256, 0, c->stream>>>(c->P, T, nseg, c->J, peel ? 1 : 0, c->gram_part,
                                                  c->irls_segun, (size_t)c->J, c->zun, Delta, Psi);
    GP_HIP(hipGetLastError());
    if (peel) GP_TRY(gram_row_join(c));
    return 0;
}

// W and nu of the last fused IRLS pass in the reference layout
int gpk_irls_wnu(gp_ctx *c) {
    dim3 g((unsigned)((c->Npad + 255) / 256), c->J);
    k_irls_wnu<<<g, 256, 0, c->stream>>>(c->N, c->Npad, c->ZB, c->irls_cc, c->W, c->Nu);
    GP_HIP(hipGetLastError());
    k_irls_nu_nz<<<c->J, 256, 0, c->stream>>>(c->Npad, c->Yp, c->Yi, c->Yx, c->ZB, c->W, c->Nu);
    GP_HIP(hipGetLastError());
    return 0;
}
