// gp_gradcontract_mixed.hip -- the gradient contraction of the mixed path,
//   Uw[i,:] = sum_j o_ij Sw_j zw_i          (Sw_j = Psi~_j^{-1}, whitened basis)
// on FP16 MFMA (v_mfma_f32_16x16x32_f16) with hi/lo operands.
//
// Scaling.  Per column k of Zw: 2^zc_k >= max_i |zw_ik|.  Per gene j: 2^e_j >=
// max_i o_ij (from the weight prep) and 2^s_j >= max_{k,c} |Sw_j[k,c]| 2^zc_k.
// kappa_j = e_j + s_j bounds the products; genes are sorted by kappa and cut
// into groups, and each group g gets one exponent C_g = 26 - max_{j in g} kappa_j:
//   A_ik = o_ij zw_ik 2^(12 - e_j - zc_k)        (|A| <= 2^12)
//   B_kc = Sw_j[k,c] 2^(e_j - 12 + C_g + zc_k)    (|B| <= 2^14)
// so sum_k A_ik B_kc = o_ij (zw_i^T Sw_j)_c 2^C_g for every gene of the group,
// the group's FP32 partial is one sum, and k_gc_reduce adds the groups in FP64
// times 2^-C_g.
//
// Work.  Workgroup = 8 waves = 128 cells x all P16 columns for one gene group
// (two waves per SIMD).  Wave w owns cells i0 + 16w .. +15 and keeps their zw
// values (FP32, k = 32 ks + 8(l/16) + e) in registers.  Sw_j^hi/lo are pre-split
// per gene into [col][k] (k contiguous, 288 k per column) and streamed through
// LDS one 32-k step at a time (double buffered); B fragment lane l =
// Sw[k = 8(l/16)..+7][col = l%16].
// D lane l, r -> U[cell 4(l/16) + r][16 ct + l%16].
#include "gp_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int KP = 288;      // padded k extent of the pre-split Sw (multiple of 32)
constexpr int KCH = 32;      // k rows per LDS chunk (one MFMA k step)
constexpr int LDK = 40;      // LDS column stride in halves (32 + 8: conflict-free b128 reads)
constexpr int MAXT16 = 17;   // column tiles handled (P16 <= 272)
constexpr int GCW = 8;       // waves per workgroup (16 cells each)

// per column k of Zw: exponent zc_k with 2^zc_k >= max_i |Zw_ik| (0 beyond P).
// One 1024-thread block per column, eight loads in flight per thread: a column
// is 8 * Npad bytes read once, so the kernel is HBM-bound only with ~64 KB in
// flight per CU.
constexpr int ZX_T = 1024, ZX_U = 8;
__global__ void __launch_bounds__(ZX_T) k_zexp(int Npad, int P, const double *__restrict__ Zw,
                                               int *__restrict__ zexp) {
    __shared__ double sh[ZX_T / 64];
    const int k = blockIdx.x;
    double mx = 0.0;
    if (k < P) {
        const double *col = Zw + (size_t)k * Npad;
        for (int i0 = threadIdx.x; i0 < Npad; i0 += ZX_T * ZX_U) {
            double v[ZX_U];
#pragma unroll
            for (int u = 0; u < ZX_U; ++u) {
                const int i = i0 + u * ZX_T;
                v[u] = i < Npad ? col[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < ZX_U; ++u) mx = fmax(mx, fabs(v[u]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < ZX_T / 64; ++w) m = fmax(m, sh[w]);
        int e = 0;
        if (m > 0.0 && isfinite(m)) (void)frexp(m, &e);
        zexp[k] = e;
    }
}

// per gene: exponent s_j with 2^s_j >= max_{k,c} |Sw_j[k,c]| 2^zc_k.
// The flat index t = k + c P walks the P x P block in steps of 256; its row k
// advances by 256 mod P with one conditional subtraction (no division per
// element), and eight loads are in flight per thread.  A maximum does not
// depend on the order of its terms, so the exponent is the same whatever the
// traversal.
constexpr int SX_U = 8;
__global__ void __launch_bounds__(256) k_smax(int P, const double *__restrict__ S,
                                              const int *__restrict__ zexp, int *__restrict__ sexp) {
    __shared__ double sh[4];
    const int PP = P * P;
    const double *Sj = S + (size_t)blockIdx.x * PP;
    const int step = 256 % P, stepU = (256 * SX_U) % P;
    int k0 = (int)threadIdx.x % P;
    double mx = 0.0;
    for (int t0 = threadIdx.x; t0 < PP; t0 += 256 * SX_U) {
        double v[SX_U];
        int kk[SX_U];
        int k = k0;
#pragma unroll
        for (int u = 0; u < SX_U; ++u) {
            const int t = t0 + u * 256;
            v[u] = t < PP ? Sj[t] : 0.0;
            kk[u] = k;
            k += step;
            if (k >= P) k -= P;
        }
#pragma unroll
        for (int u = 0; u < SX_U; ++u) mx = fmax(mx, ldexp(fabs(v[u]), zexp[kk[u]]));
        k0 += stepU;
        if (k0 >= P) k0 -= P;
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        int e = 0;
        if (m > 0.0 && isfinite(m)) (void)frexp(m, &e);
        sexp[blockIdx.x] = e;
    }
}

// Gene groups of the contraction, on the device (no host round trip, so the
// unit can be captured in a graph): kappa_j = e_j + s_j with 2^e_j = wun[j];
// genes stably sorted by kappa (a counting sort over a small integer range:
// LDS histogram, scan by one thread, block-parallel stable scatter), cut into consecutive groups of gsize; group g gets
// C_g = 26 - max kappa (its last member), gene j the operand exponents
// aexp_j = 12 - e_j and bexp_j = e_j - 12 + C_g, and gun[g] = 2^-C_g.
__global__ void __launch_bounds__(256) k_gc_plan(int J, int gsize, int ngroup,
        const double *__restrict__ wun, const int *__restrict__ sexp, int *__restrict__ perm,
        int *__restrict__ aexp, int *__restrict__ bexp, double *__restrict__ gun,
        int *__restrict__ kappa) {
    constexpr int NB = 1024, KOFF = 512;     // key range [-512, 511]
    __shared__ int cnt[NB];
    __shared__ int chunk[256];
    for (int b = threadIdx.x; b < NB; b += 256) cnt[b] = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < J; j += 256) {
        int e = 0;
        (void)frexp(wun[j], &e);             // wun = 2^e_j exactly -> frexp gives e_j + 1
        const int key = min(max((e - 1) + sexp[j], -KOFF), KOFF - 1);
        kappa[j] = key;
        atomicAdd(&cnt[key + KOFF], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {                  // exclusive scan of the bucket sizes, in LDS
        int run = 0;
        for (int b = 0; b < NB; b++) { const int t = cnt[b]; cnt[b] = run; run += t; }
    }
    __syncthreads();
    // stable scatter, 256 genes at a time: a gene's slot is its bucket's running
    // offset plus the number of earlier genes of the same chunk in that bucket
    for (int base = 0; base < J; base += 256) {
        const int j = base + (int)threadIdx.x;
        const int key = j < J ? kappa[j] : KOFF;   // KOFF lies outside the key range
        chunk[threadIdx.x] = key;
        __syncthreads();
        if (j < J) {
            int r = 0;
            for (int t = 0; t < (int)threadIdx.x; ++t) r += chunk[t] == key;
            perm[cnt[key + KOFF] + r] = j;
        }
        __syncthreads();
        if (j < J) atomicAdd(&cnt[key + KOFF], 1);
        __syncthreads();
    }
    for (int t = threadIdx.x; t < J; t += 256) {
        const int g = t / gsize;
        const int last = min(J, (g + 1) * gsize) - 1;
        const int Cg = 26 - kappa[perm[last]];
        const int j = perm[t];
        int e = 0;
        (void)frexp(wun[j], &e);
        aexp[j] = 12 - (e - 1);
        bexp[j] = (e - 1) - 12 + Cg;
        if (t == g * gsize) gun[g] = ldexp(1.0, -Cg);
    }
    (void)ngroup;
}

// Sw_j[k,c] * 2^(e_j - 12 + C_g + zc_k) split into FP16 hi/lo, layout [col < P16][k < KP]
__global__ void k_ssplit(int P, int P16, const double *__restrict__ S, const int *__restrict__ scale_exp,
                         const int *__restrict__ zexp, _Float16 *__restrict__ sh,
                         _Float16 *__restrict__ sl) {
    const int j = blockIdx.y;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;   // over P16 * KP
    if (idx >= (size_t)P16 * KP) return;
    const int col = (int)(idx / KP), k = (int)(idx % KP);
    const double *Sj = S + (size_t)j * P * P;
    const float v = (col < P && k < P) ? (float)ldexp(Sj[k + (size_t)col * P], scale_exp[j] + zexp[k]) : 0.0f;
    const _Float16 h = (_Float16)v;
    const size_t o = (size_t)j * P16 * KP + idx;
    sh[o] = h;
    sl[o] = (_Float16)(v - (float)h);
}

template <int T16>
__global__ void __launch_bounds__(GCW * 64) k_gc_mixed(int N, int Npad, int P, int gsize, int J,
        const double *__restrict__ Zw, const float *__restrict__ w32,
        const int *__restrict__ perm, const int *__restrict__ aexp, const int *__restrict__ zexp,
        const _Float16 *__restrict__ shi, const _Float16 *__restrict__ slo,
        float *__restrict__ part) {
    constexpr int P16 = 16 * T16;
    constexpr int KS = (P16 + 31) / 32;          // 32-k steps per gene (<= KP/32)
    constexpr int NITEM = 2 * P16 * 4;           // per step: hi/lo x P16 cols x 4 groups of 8 k
    constexpr int IPT = (NITEM + GCW * 64 - 1) / (GCW * 64);
    extern __shared__ __attribute__((aligned(16))) _Float16 Bs[];   // [2 buf][2 hl][P16][LDK]
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.y;
    const int j0 = g * gsize, j1 = min(J, j0 + gsize);
    // this lane's cell (row of the A fragment) and its zw values, k = 32 ks + 8(l/16) + e
    const int cell = blockIdx.x * (16 * GCW) + 16 * w + (lane & 15);
    float z[KS][8];
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = 32 * ks + 8 * (lane >> 4) + e;
            z[ks][e] = (k < P) ? (float)ldexp(Zw[cell + (size_t)k * Npad], -zexp[k]) : 0.0f;
        }
    f4 acc[T16];
#pragma unroll
    for (int ct = 0; ct < T16; ct++) acc[ct] = f4{0.f, 0.f, 0.f, 0.f};

    h8 pf[IPT];
    auto fetch = [&](int j, int ks) {
        const _Float16 *hj = shi + (size_t)j * P16 * KP, *lj = slo + (size_t)j * P16 * KP;
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const int it = tid + GCW * 64 * r;
            if (it < NITEM) {
                const int hl = it / (P16 * 4), rem = it % (P16 * 4);
                pf[r] = *(const h8 *)((hl ? lj : hj) + (size_t)(rem >> 2) * KP + ks * KCH + 8 * (rem & 3));
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const int it = tid + GCW * 64 * r;
            if (it < NITEM) {
                const int hl = it / (P16 * 4), rem = it % (P16 * 4);
                *(h8 *)&Bs[((buf * 2 + hl) * P16 + (rem >> 2)) * LDK + 8 * (rem & 3)] = pf[r];
            }
        }
    };
    if (j0 < j1) {
        fetch(perm[j0], 0);
        store(0);
    }
    __syncthreads();
    int cur = 0;
    const int boff = (lane & 15) * LDK + 8 * (lane >> 4);
    // the cell's weight for the group's first gene; later genes' weights are
    // loaded one gene ahead, so the load's latency is not paid at each gene start
    const int cellc = min(cell, N - 1);
    float wnext = (j0 < j1) ? w32[cellc + (size_t)perm[j0] * Npad] : 0.0f;
    for (int jj = j0; jj < j1; jj++) {
        const int j = perm[jj];
        const int jn = (jj + 1 < j1) ? perm[jj + 1] : j;
        const float o = (cell < N) ? wnext * ldexpf(1.0f, aexp[j]) : 0.0f;
        wnext = w32[cellc + (size_t)jn * Npad];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const bool more = (ks + 1 < KS) || (jj + 1 < j1);
            if (ks + 1 < KS) fetch(j, ks + 1);
            else if (jj + 1 < j1) fetch(jn, 0);
            h8 ah, al;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                // v is materialised in FP32 before the split: a fused
                // multiply-convert (v_fma_mix) loses accuracy in the lo
                // part (measured: 3e-5 instead of 2e-7 relative error)
                float v = z[ks][e] * o;
                asm volatile("" : "+v"(v));
                const _Float16 hv = (_Float16)v;
                ah[e] = hv;
                al[e] = (_Float16)(v - (float)hv);
            }
            const _Float16 *bh = Bs + (size_t)(cur * 2 + 0) * P16 * LDK + boff;
            const _Float16 *bl = Bs + (size_t)(cur * 2 + 1) * P16 * LDK + boff;
#pragma unroll
            for (int ct = 0; ct < T16; ct++) {
                const h8 vbh = *(const h8 *)(bh + 16 * ct * LDK);
                const h8 vbl = *(const h8 *)(bl + 16 * ct * LDK);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, vbh, acc[ct], 0, 0, 0);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, vbl, acc[ct], 0, 0, 0);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, vbh, acc[ct], 0, 0, 0);
            }
            if (more) store(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    }
    // partial of this group: part[g][col][cell] (FP32, ld = Npad)
    float *pg = part + (size_t)g * P16 * Npad;
    const int row0 = blockIdx.x * (16 * GCW) + 16 * w + 4 * (lane >> 4);
#pragma unroll
    for (int ct = 0; ct < T16; ct++)
#pragma unroll
        for (int r = 0; r < 4; r++) pg[(size_t)(16 * ct + (lane & 15)) * Npad + row0 + r] = acc[ct][r];
}

// ---- row-block form: a workgroup owns 128 cells (8 row tiles) of one gene
// group.  Wave w owns the column tiles 2w and 2w+1; with an odd tile count
// the last column tile is shared out by row tile (wave w takes row tiles
// r = w mod NW), so every wave carries the same load.  Each wave keeps its
// slice of S_j (hi/lo, streamed one 32-k step ahead from L2) in registers,
// while A = o z (hi/lo) for the 128 cells is staged once per gene in LDS as
// [cell][k].  Per 32-k step a wave reads 2 A fragments per row tile for 6
// MFMAs (0.33 LDS fragments per MFMA, half of the 16-cell form).  z (scaled
// by 2^-zexp) stays in registers for the whole gene group.  Each accumulator
// sees the same sequence of MFMAs as in the 16-cell form, so both forms give
// the same bits.  (Refilling each 32-k slot for the next gene as soon as it
// is free, to overlap the staging with the MFMAs, measured slower: it needs a
// barrier per step.)
constexpr int GRC = 128;     // cells per workgroup

constexpr int gc_rows_waves(int T16) { return T16 / 2 > 0 ? T16 / 2 : 1; }

template <int T16>
__global__ void __launch_bounds__(gc_rows_waves(T16) * 64) k_gc_rows(int N, int Npad, int P, int gsize, int J,
        const double *__restrict__ Zw, const float *__restrict__ w32,
        const int *__restrict__ perm, const int *__restrict__ aexp, const int *__restrict__ zexp,
        const _Float16 *__restrict__ shi, const _Float16 *__restrict__ slo,
        float *__restrict__ part) {
    constexpr int P16 = 16 * T16;
    constexpr int NPAIR = T16 / 2, ODD = T16 % 2;
    constexpr int NW = gc_rows_waves(T16);
    constexpr int KS = (P16 + 31) / 32;
    constexpr int LDA = 32 * KS + 8;             // halves per cell row (stride 16 B multiple)
    constexpr int NITEM = GRC * KS * 4;          // (cell, 8-k group) items
    constexpr int IPT = (NITEM + NW * 64 - 1) / (NW * 64);
    constexpr int NX = (8 + NW - 1) / NW;        // row tiles of the odd column per wave
    extern __shared__ __attribute__((aligned(16))) _Float16 As[];   // [2 hl][GRC][LDA]
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.y;
    const int j0 = g * gsize, j1 = min(J, j0 + gsize);
    const int cb = blockIdx.x * GRC;
    const int ct0 = 2 * w;

    // z of this thread's items (cell = it % GRC, k = 8 (it / GRC) + e), constant over the group
    float zr[IPT][8];
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int it = min(tid + NW * 64 * r, NITEM - 1);
        const int cl = it % GRC, k0 = 8 * (it / GRC);
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = k0 + e, cell = cb + cl;
            zr[r][e] = (k < P && cell < N) ? (float)ldexp(Zw[cell + (size_t)k * Npad], -zexp[k]) : 0.0f;
        }
    }
    f4 acc[8][2], accx[NX];
#pragma unroll
    for (int r = 0; r < 8; r++) { acc[r][0] = f4{0.f, 0.f, 0.f, 0.f}; acc[r][1] = f4{0.f, 0.f, 0.f, 0.f}; }
#pragma unroll
    for (int x = 0; x < NX; x++) accx[x] = f4{0.f, 0.f, 0.f, 0.f};

    const int lk = (lane & 15), lq = 8 * (lane >> 4);
    const int bcol0 = (16 * min(ct0, T16 - 1) + lk) * KP + lq;
    const int bcol1 = (16 * min(ct0 + 1, T16 - 1) + lk) * KP + lq;
    const int bcolx = (16 * (T16 - 1) + lk) * KP + lq;
    const int arow = lk * LDA + lq;
    _Float16 *Ah = As, *Al = As + (size_t)GRC * LDA;

    for (int jj = j0; jj < j1; jj++) {
        const int j = perm[jj];
        const float sc = ldexpf(1.0f, aexp[j]);
        const float *wj = w32 + (size_t)j * Npad + cb;
        // stage A = o z, split hi/lo (items past the end repeat the last one)
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const int it = min(tid + NW * 64 * r, NITEM - 1);
            const int cl = it % GRC, k0 = 8 * (it / GRC);
            const float o = (cb + cl < N) ? wj[cl] * sc : 0.0f;
            h8 hi, lo;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                float v = zr[r][e] * o;
                asm volatile("" : "+v"(v));
                const _Float16 h = (_Float16)v;
                hi[e] = h;
                lo[e] = (_Float16)(v - (float)h);
            }
            *(h8 *)&Ah[cl * LDA + k0] = hi;
            *(h8 *)&Al[cl * LDA + k0] = lo;
        }
        __syncthreads();
        const _Float16 *hj = shi + (size_t)j * P16 * KP, *lj = slo + (size_t)j * P16 * KP;
        h8 bh0 = *(const h8 *)(hj + bcol0), bl0 = *(const h8 *)(lj + bcol0);
        h8 bh1 = *(const h8 *)(hj + bcol1), bl1 = *(const h8 *)(lj + bcol1);
        h8 bhx = *(const h8 *)(hj + bcolx), blx = *(const h8 *)(lj + bcolx);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int kn = 32 * min(ks + 1, KS - 1);
            const h8 nh0 = *(const h8 *)(hj + bcol0 + kn), nl0 = *(const h8 *)(lj + bcol0 + kn);
            const h8 nh1 = *(const h8 *)(hj + bcol1 + kn), nl1 = *(const h8 *)(lj + bcol1 + kn);
            const h8 nhx = *(const h8 *)(hj + bcolx + kn), nlx = *(const h8 *)(lj + bcolx + kn);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const h8 ah = *(const h8 *)(Ah + 16 * r * LDA + arow + 32 * ks);
                const h8 al = *(const h8 *)(Al + 16 * r * LDA + arow + 32 * ks);
                if constexpr (NPAIR > 0) {
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh0, acc[r][0], 0, 0, 0);
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl0, acc[r][0], 0, 0, 0);
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh0, acc[r][0], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh1, acc[r][1], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl1, acc[r][1], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh1, acc[r][1], 0, 0, 0);
                }
                if constexpr (ODD) {
                    if (r % NW == w) {       // wave-uniform
                        f4 &ax = accx[r / NW];
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bhx, ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, blx, ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bhx, ax, 0, 0, 0);
                    }
                }
            }
            bh0 = nh0; bl0 = nl0; bh1 = nh1; bl1 = nl1; bhx = nhx; blx = nlx;
        }
        __syncthreads();
    }
    // partial of this group: part[g][col][cell]
    float *pg = part + (size_t)g * P16 * Npad;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int row0 = cb + 16 * r + 4 * (lane >> 4);
        if constexpr (NPAIR > 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                pg[(size_t)(16 * ct0 + lk) * Npad + row0 + q] = acc[r][0][q];
                pg[(size_t)(16 * (ct0 + 1) + lk) * Npad + row0 + q] = acc[r][1][q];
            }
        }
        if constexpr (ODD) {
            if (r % NW == w) {
#pragma unroll
                for (int q = 0; q < 4; q++) pg[(size_t)(16 * (T16 - 1) + lk) * Npad + row0 + q] = accx[r / NW][q];
            }
        }
    }
}

// ---- factored row-block form (gc_form 2): the cell weight o is a per-row
// scale of A = o z, so it moves onto the accumulator: per gene the workgroup
// forms C = Z_I S_j with a fresh accumulator and adds o_i C[i, :] into U.
// A = z (hi/lo) is then the same for every gene of the group and is staged
// in LDS once; per gene only the 128 weights are staged (double buffered,
// one barrier per gene).  Same MFMA tiling as the row-block form; the
// rounding differs (o multiplies the FP32 gene sum instead of each operand).
template <int T16>
__global__ void __launch_bounds__(gc_rows_waves(T16) * 64) k_gc_rows2(int N, int Npad, int P, int gsize, int J,
        const double *__restrict__ Zw, const float *__restrict__ w32,
        const int *__restrict__ perm, const int *__restrict__ aexp, const int *__restrict__ zexp,
        const _Float16 *__restrict__ shi, const _Float16 *__restrict__ slo,
        float *__restrict__ part) {
    constexpr int P16 = 16 * T16;
    constexpr int NPAIR = T16 / 2, ODD = T16 % 2;
    constexpr int NW = gc_rows_waves(T16);
    constexpr int KS = (P16 + 31) / 32;
    constexpr int LDA = 32 * KS + 8;
    constexpr int NITEM = GRC * KS * 4;
    constexpr int IPT = (NITEM + NW * 64 - 1) / (NW * 64);
    constexpr int NX = (8 + NW - 1) / NW;
    extern __shared__ __attribute__((aligned(16))) _Float16 As[];   // [2 hl][GRC][LDA]
    __shared__ float os[2][GRC];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.y;
    const int j0 = g * gsize, j1 = min(J, j0 + gsize);
    const int cb = blockIdx.x * GRC;
    const int ct0 = 2 * w;
    _Float16 *Ah = As, *Al = As + (size_t)GRC * LDA;
    // A = z (scaled by 2^-zexp), split hi/lo, once for the group
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int it = min(tid + NW * 64 * r, NITEM - 1);
        const int cl = it % GRC, k0 = 8 * (it / GRC);
        h8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = k0 + e, cell = cb + cl;
            float v = (k < P && cell < N) ? (float)ldexp(Zw[cell + (size_t)k * Npad], -zexp[k]) : 0.0f;
            asm volatile("" : "+v"(v));
            const _Float16 h = (_Float16)v;
            hi[e] = h;
            lo[e] = (_Float16)(v - (float)h);
        }
        *(h8 *)&Ah[cl * LDA + k0] = hi;
        *(h8 *)&Al[cl * LDA + k0] = lo;
    }
    auto weights = [&](int jj, int buf) {
        const int j = perm[min(jj, j1 - 1)];
        const float sc = ldexpf(1.0f, aexp[j]);
        for (int t = tid; t < GRC; t += NW * 64)   // fewer than 128 threads when T16 < 4
            os[buf][t] = (cb + t < N) ? w32[cb + t + (size_t)j * Npad] * sc : 0.0f;
    };
    if (j0 < j1) weights(j0, 0);
    __syncthreads();

    f4 U[8][2], Ux[NX];
#pragma unroll
    for (int r = 0; r < 8; r++) { U[r][0] = f4{0.f, 0.f, 0.f, 0.f}; U[r][1] = f4{0.f, 0.f, 0.f, 0.f}; }
#pragma unroll
    for (int x = 0; x < NX; x++) Ux[x] = f4{0.f, 0.f, 0.f, 0.f};

    const int lk = (lane & 15), lq = 8 * (lane >> 4);
    const int bcol0 = (16 * min(ct0, T16 - 1) + lk) * KP + lq;
    const int bcol1 = (16 * min(ct0 + 1, T16 - 1) + lk) * KP + lq;
    const int bcolx = (16 * (T16 - 1) + lk) * KP + lq;
    const int arow = lk * LDA + lq;
    const int orow = 4 * (lane >> 4);             // this lane's accumulator rows within a tile

    for (int jj = j0; jj < j1; jj++) {
        const int j = perm[jj];
        const int buf = (jj - j0) & 1;
        weights(jj + 1, buf ^ 1);                 // the next gene's weights (read after this gene's barrier)
        const _Float16 *hj = shi + (size_t)j * P16 * KP, *lj = slo + (size_t)j * P16 * KP;
        h8 bh0 = *(const h8 *)(hj + bcol0), bl0 = *(const h8 *)(lj + bcol0);
        h8 bh1 = *(const h8 *)(hj + bcol1), bl1 = *(const h8 *)(lj + bcol1);
        h8 bhx = *(const h8 *)(hj + bcolx), blx = *(const h8 *)(lj + bcolx);
        f4 acc[8][2], accx[NX];
#pragma unroll
        for (int r = 0; r < 8; r++) { acc[r][0] = f4{0.f, 0.f, 0.f, 0.f}; acc[r][1] = f4{0.f, 0.f, 0.f, 0.f}; }
#pragma unroll
        for (int x = 0; x < NX; x++) accx[x] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int kn = 32 * min(ks + 1, KS - 1);
            const h8 nh0 = *(const h8 *)(hj + bcol0 + kn), nl0 = *(const h8 *)(lj + bcol0 + kn);
            const h8 nh1 = *(const h8 *)(hj + bcol1 + kn), nl1 = *(const h8 *)(lj + bcol1 + kn);
            const h8 nhx = *(const h8 *)(hj + bcolx + kn), nlx = *(const h8 *)(lj + bcolx + kn);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const h8 ah = *(const h8 *)(Ah + 16 * r * LDA + arow + 32 * ks);
                const h8 al = *(const h8 *)(Al + 16 * r * LDA + arow + 32 * ks);
                if constexpr (NPAIR > 0) {
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh0, acc[r][0], 0, 0, 0);
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl0, acc[r][0], 0, 0, 0);
                    acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh0, acc[r][0], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh1, acc[r][1], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl1, acc[r][1], 0, 0, 0);
                    acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh1, acc[r][1], 0, 0, 0);
                }
                if constexpr (ODD) {
                    if (r % NW == w) {
                        f4 &ax = accx[r / NW];
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bhx, ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, blx, ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bhx, ax, 0, 0, 0);
                    }
                }
            }
            bh0 = nh0; bl0 = nl0; bh1 = nh1; bl1 = nl1; bhx = nhx; blx = nlx;
        }
        // U += o (rows) * C
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float o = os[buf][16 * r + orow + q];
                if constexpr (NPAIR > 0) {
                    U[r][0][q] += o * acc[r][0][q];
                    U[r][1][q] += o * acc[r][1][q];
                }
                if constexpr (ODD) {
                    if (r % NW == w) Ux[r / NW][q] += o * accx[r / NW][q];
                }
            }
        }
        __syncthreads();                          // os[buf ^ 1] written; os[buf] free
    }
    float *pg = part + (size_t)g * P16 * Npad;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int row0 = cb + 16 * r + orow;
        if constexpr (NPAIR > 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                pg[(size_t)(16 * ct0 + lk) * Npad + row0 + q] = U[r][0][q];
                pg[(size_t)(16 * (ct0 + 1) + lk) * Npad + row0 + q] = U[r][1][q];
            }
        }
        if constexpr (ODD) {
            if (r % NW == w) {
#pragma unroll
                for (int q = 0; q < 4; q++) pg[(size_t)(16 * (T16 - 1) + lk) * Npad + row0 + q] = Ux[r / NW][q];
            }
        }
    }
}

// ---- factored row-block form on one wave per SIMD (gc_form 3): the tiling of
// form 2 with 4 waves of 512 registers instead of 8 of 256.  Wave w keeps
// CT = T16 / 4 column tiles of S_j (columns 16 (w CT + i)) for all 8 row
// tiles, so each A fragment read from LDS feeds 3 CT MFMAs instead of 6
// (half the LDS reads per MFMA); the T16 mod 4 leftover column tiles are
// dealt by row tile (r mod 4 == w).  The B fragments of the next k-step
// are loaded during the current one.  Every accumulator sees form 2's
// MFMA sequence and update, so the two forms give the same bits.
constexpr int GW3 = 4;       // waves per workgroup, one per SIMD

template <int T16>
__global__ void __launch_bounds__(GW3 * 64, 1) k_gc_rows4(int N, int Npad, int P, int gsize, int J,
        const double *__restrict__ Zw, const float *__restrict__ w32,
        const int *__restrict__ perm, const int *__restrict__ aexp, const int *__restrict__ zexp,
        const _Float16 *__restrict__ shi, const _Float16 *__restrict__ slo,
        float *__restrict__ part) {
    constexpr int P16 = 16 * T16;
    constexpr int CT = T16 / GW3, ODD = T16 % GW3;
    constexpr int NX = ODD * (8 / GW3);           // leftover (row tile, column tile) pairs per wave
    constexpr int KS = (P16 + 31) / 32;
    constexpr int LDA = 32 * KS + 8;
    constexpr int NITEM = GRC * KS * 4;
    constexpr int IPT = (NITEM + GW3 * 64 - 1) / (GW3 * 64);
    extern __shared__ __attribute__((aligned(16))) _Float16 As[];   // [2 hl][GRC][LDA]
    __shared__ float os[2][GRC];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.y;
    const int j0 = g * gsize, j1 = min(J, j0 + gsize);
    const int cb = blockIdx.x * GRC;
    _Float16 *Ah = As, *Al = As + (size_t)GRC * LDA;
    // A = z (scaled by 2^-zexp), split hi/lo, once for the group
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int it = min(tid + GW3 * 64 * r, NITEM - 1);
        const int cl = it % GRC, k0 = 8 * (it / GRC);
        h8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = k0 + e, cell = cb + cl;
            float v = (k < P && cell < N) ? (float)ldexp(Zw[cell + (size_t)k * Npad], -zexp[k]) : 0.0f;
            asm volatile("" : "+v"(v));
            const _Float16 h = (_Float16)v;
            hi[e] = h;
            lo[e] = (_Float16)(v - (float)h);
        }
        *(h8 *)&Ah[cl * LDA + k0] = hi;
        *(h8 *)&Al[cl * LDA + k0] = lo;
    }
    auto weights = [&](int jj, int buf) {
        const int j = perm[min(jj, j1 - 1)];
        const float sc = ldexpf(1.0f, aexp[j]);
        for (int t = tid; t < GRC; t += GW3 * 64)
            os[buf][t] = (cb + t < N) ? w32[cb + t + (size_t)j * Npad] * sc : 0.0f;
    };
    if (j0 < j1) weights(j0, 0);
    __syncthreads();

    f4 U[8][CT > 0 ? CT : 1], Ux[NX > 0 ? NX : 1];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
        for (int i = 0; i < CT; i++) U[r][i] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int x = 0; x < NX; x++) Ux[x] = f4{0.f, 0.f, 0.f, 0.f};

    const int lk = (lane & 15), lq = 8 * (lane >> 4);
    int bcol[CT > 0 ? CT : 1], bxc[ODD > 0 ? ODD : 1];
#pragma unroll
    for (int i = 0; i < CT; i++) bcol[i] = (16 * (w * CT + i) + lk) * KP + lq;
#pragma unroll
    for (int e = 0; e < ODD; e++) bxc[e] = (16 * (GW3 * CT + e) + lk) * KP + lq;
    const int arow = lk * LDA + lq;
    const int orow = 4 * (lane >> 4);

    for (int jj = j0; jj < j1; jj++) {
        const int j = perm[jj];
        const int buf = (jj - j0) & 1;
        weights(jj + 1, buf ^ 1);                 // the next gene's weights (read after this gene's barrier)
        const _Float16 *hj = shi + (size_t)j * P16 * KP, *lj = slo + (size_t)j * P16 * KP;
        h8 bh[CT > 0 ? CT : 1], bl[CT > 0 ? CT : 1], xh[ODD > 0 ? ODD : 1], xl[ODD > 0 ? ODD : 1];
#pragma unroll
        for (int i = 0; i < CT; i++) { bh[i] = *(const h8 *)(hj + bcol[i]); bl[i] = *(const h8 *)(lj + bcol[i]); }
#pragma unroll
        for (int e = 0; e < ODD; e++) { xh[e] = *(const h8 *)(hj + bxc[e]); xl[e] = *(const h8 *)(lj + bxc[e]); }
        f4 acc[8][CT > 0 ? CT : 1], accx[NX > 0 ? NX : 1];
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int i = 0; i < CT; i++) acc[r][i] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int x = 0; x < NX; x++) accx[x] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int kn = 32 * min(ks + 1, KS - 1);
            h8 nbh[CT > 0 ? CT : 1], nbl[CT > 0 ? CT : 1], nxh[ODD > 0 ? ODD : 1], nxl[ODD > 0 ? ODD : 1];
#pragma unroll
            for (int i = 0; i < CT; i++) {
                nbh[i] = *(const h8 *)(hj + bcol[i] + kn);
                nbl[i] = *(const h8 *)(lj + bcol[i] + kn);
            }
#pragma unroll
            for (int e = 0; e < ODD; e++) {
                nxh[e] = *(const h8 *)(hj + bxc[e] + kn);
                nxl[e] = *(const h8 *)(lj + bxc[e] + kn);
            }
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const h8 ah = *(const h8 *)(Ah + 16 * r * LDA + arow + 32 * ks);
                const h8 al = *(const h8 *)(Al + 16 * r * LDA + arow + 32 * ks);
#pragma unroll
                for (int i = 0; i < CT; i++) {
                    acc[r][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[i], acc[r][i], 0, 0, 0);
                    acc[r][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[i], acc[r][i], 0, 0, 0);
                    acc[r][i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[i], acc[r][i], 0, 0, 0);
                }
                if (r % GW3 == w) {
#pragma unroll
                    for (int e = 0; e < ODD; e++) {
                        f4 &ax = accx[e * (8 / GW3) + r / GW3];
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, xh[e], ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, xl[e], ax, 0, 0, 0);
                        ax = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, xh[e], ax, 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < CT; i++) { bh[i] = nbh[i]; bl[i] = nbl[i]; }
#pragma unroll
            for (int e = 0; e < ODD; e++) { xh[e] = nxh[e]; xl[e] = nxl[e]; }
        }
        // U += o (rows) * C
#pragma unroll
        for (int r = 0; r < 8; r++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float o = os[buf][16 * r + orow + q];
#pragma unroll
                for (int i = 0; i < CT; i++) U[r][i][q] += o * acc[r][i][q];
                if (r % GW3 == w) {
#pragma unroll
                    for (int e = 0; e < ODD; e++) {
                        const int x = e * (8 / GW3) + r / GW3;
                        Ux[x][q] += o * accx[x][q];
                    }
                }
            }
        }
        __syncthreads();                          // os[buf ^ 1] written; os[buf] free
    }
    float *pg = part + (size_t)g * P16 * Npad;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int row0 = cb + 16 * r + orow;
#pragma unroll
        for (int i = 0; i < CT; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) pg[(size_t)(16 * (w * CT + i) + lk) * Npad + row0 + q] = U[r][i][q];
        if (r % GW3 == w) {
#pragma unroll
            for (int e = 0; e < ODD; e++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    pg[(size_t)(16 * (GW3 * CT + e) + lk) * Npad + row0 + q] = Ux[e * (8 / GW3) + r / GW3][q];
        }
    }
}

__global__ void k_gc_reduce(int Npad, int P, int P16, int ngroup, const float *__restrict__ part,
                            const double *__restrict__ gun, double *__restrict__ Uw) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)Npad * P) return;
    double s = 0.0;
    for (int g = 0; g < ngroup; g++) s += (double)part[(size_t)g * P16 * Npad + idx] * gun[g];
    Uw[idx] = s;
}

}  // namespace

template <int T16>
static int launch_gc_mixed(gp_ctx *c, int gsize, int ngroup);

template <int T16>
static int launch_gc_rows(gp_ctx *c, int gsize, int ngroup) {
    constexpr int P16 = 16 * T16, KS = (P16 + 31) / 32, NW = gc_rows_waves(T16);
    const size_t shm = (size_t)2 * GRC * (32 * KS + 8) * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_rows<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(c->Npad / GRC, ngroup);
    k_gc_rows<T16><<<g, NW * 64, shm, c->stream>>>(c->N, c->Npad, c->P, gsize, c->J, c->Zw, c->w32,
                                                c->gc_perm, c->gc_aexp, c->gc_zexp,
                                                (const _Float16 *)c->Sh, (const _Float16 *)c->Sl,
                                                c->gc_part);
    GP_HIP(hipGetLastError());
    return 0;
}

template <int T16>
static int launch_gc_rows2(gp_ctx *c, int gsize, int ngroup) {
    constexpr int P16 = 16 * T16, KS = (P16 + 31) / 32, NW = gc_rows_waves(T16);
    const size_t shm = (size_t)2 * GRC * (32 * KS + 8) * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_rows2<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(c->Npad / GRC, ngroup);
    k_gc_rows2<T16><<<g, NW * 64, shm, c->stream>>>(c->N, c->Npad, c->P, gsize, c->J, c->Zw, c->w32,
                                                 c->gc_perm, c->gc_aexp, c->gc_zexp,
                                                 (const _Float16 *)c->Sh, (const _Float16 *)c->Sl,
                                                 c->gc_part);
    GP_HIP(hipGetLastError());
    return 0;
}

template <int T16>
static int launch_gc_rows4(gp_ctx *c, int gsize, int ngroup) {
    constexpr int P16 = 16 * T16, KS = (P16 + 31) / 32;
    const size_t shm = (size_t)2 * GRC * (32 * KS + 8) * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_rows4<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(c->Npad / GRC, ngroup);
    k_gc_rows4<T16><<<g, GW3 * 64, shm, c->stream>>>(c->N, c->Npad, c->P, gsize, c->J, c->Zw, c->w32,
                                                  c->gc_perm, c->gc_aexp, c->gc_zexp,
                                                  (const _Float16 *)c->Sh, (const _Float16 *)c->Sl,
                                                  c->gc_part);
    GP_HIP(hipGetLastError());
    return 0;
}

template <int T16>
static int launch_gc(gp_ctx *c, int gsize, int ngroup) {
    if (c->gc_form == 3) return launch_gc_rows4<T16>(c, gsize, ngroup);
    if (c->gc_form == 2) return launch_gc_rows2<T16>(c, gsize, ngroup);
    return c->gc_form == 1 ? launch_gc_rows<T16>(c, gsize, ngroup) : launch_gc_mixed<T16>(c, gsize, ngroup);
}

template <int T16>
static int launch_gc_mixed(gp_ctx *c, int gsize, int ngroup) {
    constexpr int P16 = 16 * T16;
    const size_t shm = (size_t)2 * 2 * P16 * LDK * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_mixed<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(c->Npad / (16 * GCW), ngroup);
    k_gc_mixed<T16><<<g, GCW * 64, shm, c->stream>>>(c->N, c->Npad, c->P, gsize, c->J, c->Zw, c->w32,
                                                 c->gc_perm, c->gc_aexp, c->gc_zexp,
                                                 (const _Float16 *)c->Sh,
                                                 (const _Float16 *)c->Sl, c->gc_part);
    GP_HIP(hipGetLastError());
    return 0;
}

// ---- intercept-peeled factored row-block form (default for a design whose
// only column before the GP block is the intercept).  With zw_i = [1 | g_i]
// (g_i the M GP columns) and Sw_j = [s00 s0G; sG0 SGG]:
//   Uw_i[G] = sum_j o_ij (SGG g_i + sG0),   Uw_i[0] = sum_j o_ij (s00 + s0G . g_i).
// The MFMAs of the factored form run over the GP block only (K = M, M / 16
// column tiles: an even count, so no odd tile, no 16-column tile and no 32-k
// step spent on the single intercept column); the intercept's k term sG0
// (scaled like the product) joins each gene's FP32 product in the update
// U += o (acc + t0).  The intercept output column is one FP32 GEMM
// Y = w32 S0^T over the genes (w32 = W / Omega, S0[j, :] = [s0G | s00]) and
// a per-cell dot Uw_i[0] = g_i . Y_i[0:M] + Y_i[M] in FP64.
//
// Operands of gene j: SGG hi/lo as [col < M][k < M] (k contiguous), scaled as
// the full form's B (2^(bexp_j + zc_{1+k})); t0[j][c] = Sw_j[0, 1+c] 2^bexp_j.
__global__ void k_ssplit_gp(int P, int M, const double *__restrict__ S, const int *__restrict__ bexp,
                            const int *__restrict__ zexp, int J,
                            _Float16 *__restrict__ sh, _Float16 *__restrict__ sl,
                            float *__restrict__ t0, float *__restrict__ s0) {
    const int j = blockIdx.y;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;   // over M * M
    const double *Sj = S + (size_t)j * P * P;
    if (idx < (size_t)M * M) {
        const int col = (int)(idx / M), k = (int)(idx % M);
        const float v = (float)ldexp(Sj[(1 + k) + (size_t)(1 + col) * P], bexp[j] + zexp[1 + k]);
        const _Float16 h = (_Float16)v;
        const size_t o = (size_t)j * M * M + idx;
        sh[o] = h;
        sl[o] = (_Float16)(v - (float)h);
    }
    if (idx < (size_t)M) {
        const double r0 = Sj[(size_t)(1 + idx) * P];              // Sw_j[0, 1+c]
        t0[(size_t)j * M + idx] = (float)ldexp(r0, bexp[j]);
        s0[j + idx * (size_t)J] = (float)r0;
    }
    if (idx == 0) s0[j + (size_t)M * J] = (float)Sj[0];
}

// Wave tiling: the 8 row tiles (128 cells) x TG column tiles of a workgroup
// are cut into blocks of RG row tiles x CG column tiles, one per wave
// ((8 / RG) (TG / CG) waves).  RG = 8, CG = 2 is the factored form's shape
// (8 waves at M = 256, two per SIMD: each wave reads every A row fragment
// and 2 column tiles of B); RG = 4, CG = 4 halves the A fragment reads from
// LDS per MFMA and doubles the B fragment loads from L2 (measured 50 %
// slower at M = 256: the B stream from L2 is the limit); RG = 8, CG = 4
// halves the A reads with the B loads unchanged, on 4 waves of 512
// registers (one per SIMD, the accumulators of the gene and of the running
// sum together take 256).  B fragments are loaded PD k-steps ahead through
// a ring of PD + 1 register sets (PD + 1 divides the k-steps of a gene, so
// the ring lines up at every gene); the last PD k-steps of a gene load the
// first ones of the next.
template <int TG, int RG, int CG>
constexpr int gc_gp_waves() { return (8 / RG) * (TG / CG); }

template <int TG, int RG, int CG, int PD, bool SL = true>
__global__ void __launch_bounds__((gc_gp_waves<TG, RG, CG>() * 64))
k_gc_gp(int N, int Npad, int gsize, int J,
        const float *__restrict__ zg, const float *__restrict__ w32,
        const int *__restrict__ perm, const int *__restrict__ aexp, const int *__restrict__ zexp,
        const _Float16 *__restrict__ shi, const _Float16 *__restrict__ slo,
        const float *__restrict__ t0, float *__restrict__ part) {
    constexpr int M = 16 * TG;
    constexpr int NWC = TG / CG;
    constexpr int NW = gc_gp_waves<TG, RG, CG>();
    static_assert(8 % RG == 0 && TG % CG == 0 && NW >= 1, "wave tiling");
    constexpr int KS = M / 32;
    constexpr int LDA = M + 8;
    constexpr int NITEM = GRC * KS * 4;
    constexpr int IPT = (NITEM + NW * 64 - 1) / (NW * 64);
    extern __shared__ __attribute__((aligned(16))) _Float16 As[];   // [2 hl][GRC][LDA]
    __shared__ float os[2][GRC];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = blockIdx.y;
    const int j0 = g * gsize, j1 = min(J, j0 + gsize);
    const int cb = blockIdx.x * GRC;
    const int rt0 = (w / NWC) * RG, ct0 = (w % NWC) * CG;     // first row / column tile of the wave
    _Float16 *Ah = As, *Al = As + (size_t)GRC * LDA;
    // A = g (the GP block of zw, scaled by 2^-zc), split hi/lo, once for the
    // group; every load of an item is issued before any is used (a row past
    // N reads row N - 1 and is zeroed).  zg is the Gram's FP32 copy of the
    // same columns, zw 2^(14 - zc) rounded once (its column scale comes from
    // the same maximum's exponent), so zg 2^-14 is (float)(zw 2^-zc) exactly
    // and A is read at 4 bytes per element instead of 8.
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int it = min(tid + NW * 64 * r, NITEM - 1);
        const int cl = it % GRC, k0 = 8 * (it / GRC);
        const int cell = cb + cl, cellc = min(cell, N - 1);
        float zv[8];
#pragma unroll
        for (int e = 0; e < 8; e++) zv[e] = zg[cellc + (size_t)(k0 + e) * Npad];
        h8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            float v = (cell < N) ? zv[e] * 0x1p-14f : 0.0f;
            asm volatile("" : "+v"(v));
            const _Float16 h = (_Float16)v;
            hi[e] = h;
            lo[e] = (_Float16)(v - (float)h);
        }
        *(h8 *)&Ah[cl * LDA + k0] = hi;
        *(h8 *)&Al[cl * LDA + k0] = lo;
    }
    // cell weights: gene jj + 1's are loaded into registers at the start of
    // gene jj and stored to LDS at its end, so no store waits on a fresh load
    constexpr int WPT = (GRC + NW * 64 - 1) / (NW * 64);      // weights per thread (1 or 2)
    auto wload = [&](int jj, float (&wv)[WPT]) {
        const int j = perm[min(jj, j1 - 1)];
        const float sc = ldexpf(1.0f, aexp[j]);
#pragma unroll
        for (int u = 0; u < WPT; u++) {
            const int t = tid + u * NW * 64;
            const int cell = min(cb + t, N - 1);
            const float v = w32[cell + (size_t)j * Npad];
            wv[u] = (t < GRC && cb + t < N) ? v * sc : 0.0f;
        }
    };
    auto wstore = [&](const float (&wv)[WPT], int buf) {
#pragma unroll
        for (int u = 0; u < WPT; u++) {
            const int t = tid + u * NW * 64;
            if (t < GRC) os[buf][t] = wv[u];
        }
    };
    float wnext[WPT];
    if (j0 < j1) {
        wload(j0, wnext);
        wstore(wnext, 0);
    }
    __syncthreads();

    f4 U[RG][CG];
#pragma unroll
    for (int r = 0; r < RG; r++)
#pragma unroll
        for (int c = 0; c < CG; c++) U[r][c] = f4{0.f, 0.f, 0.f, 0.f};
    const int lk = (lane & 15), lq = 8 * (lane >> 4);
    int bcol[CG];
#pragma unroll
    for (int c = 0; c < CG; c++) bcol[c] = (16 * (ct0 + c) + lk) * M + lq;
    const int arow = lk * LDA + lq;
    const int orow = 4 * (lane >> 4);

    static_assert(KS % (PD + 1) == 0, "prefetch ring must line up with the gene's k-steps");
    // A fragments are read from LDS AD (k-step, row tile) steps ahead into a
    // ring of AD + 1 register pairs, so each MFMA group finds its operands
    // loaded instead of waiting on the LDS latency; A is the same for every
    // gene of the group, so the ring wraps across genes.  The scheduling
    // barriers keep the compiler from sinking the reads next to their use.
    constexpr int AD = 3, NA = AD + 1;
    constexpr int TS = KS * RG;                  // (k-step, row tile) steps per gene
    static_assert(TS % NA == 0, "A ring must line up with the gene's steps");
    h8 ar_h[NA], ar_l[NA];
    auto aload = [&](int t, int slot) {
        const int ks = (t / RG) % KS, r = t % RG;
        ar_h[slot] = *(const h8 *)(Ah + 16 * (rt0 + r) * LDA + arow + 32 * ks);
        ar_l[slot] = *(const h8 *)(Al + 16 * (rt0 + r) * LDA + arow + 32 * ks);
    };
#pragma unroll
    for (int t = 0; t < AD; t++) aload(t, t);
    h8 bh[PD + 1][CG], bl[PD + 1][CG];
    if (j0 < j1) {
        const int j = perm[j0];
        const _Float16 *hj = shi + (size_t)j * M * M, *lj = slo + (size_t)j * M * M;
#pragma unroll
        for (int d = 0; d < PD; d++)
#pragma unroll
            for (int c = 0; c < CG; c++) {
                bh[d][c] = *(const h8 *)(hj + bcol[c] + 32 * d);
                if constexpr (SL) bl[d][c] = *(const h8 *)(lj + bcol[c] + 32 * d);
            }
    }
    for (int jj = j0; jj < j1; jj++) {
        const int j = perm[jj];
        const int jn = perm[min(jj + 1, j1 - 1)];
        const int buf = (jj - j0) & 1;
        wload(jj + 1, wnext);
        const _Float16 *hj = shi + (size_t)j * M * M, *lj = slo + (size_t)j * M * M;
        const _Float16 *hn = shi + (size_t)jn * M * M, *ln = slo + (size_t)jn * M * M;
        float tv[CG];
#pragma unroll
        for (int c = 0; c < CG; c++) tv[c] = t0[(size_t)j * M + 16 * (ct0 + c) + lk];
        f4 acc[RG][CG];
#pragma unroll
        for (int r = 0; r < RG; r++)
#pragma unroll
            for (int c = 0; c < CG; c++) acc[r][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            // PD k-steps ahead: this gene's, or the next gene's first ones
            constexpr int NB = PD + 1;
            const int sl = (ks + PD) % NB, s0 = ks % NB;
            const int sa = ks + PD;
            const _Float16 *ph = sa < KS ? hj + 32 * sa : hn + 32 * (sa - KS);
            const _Float16 *pl = sa < KS ? lj + 32 * sa : ln + 32 * (sa - KS);
#pragma unroll
            for (int c = 0; c < CG; c++) {
                bh[sl][c] = *(const h8 *)(ph + bcol[c]);
                if constexpr (SL) bl[sl][c] = *(const h8 *)(pl + bcol[c]);
            }
#pragma unroll
            for (int r = 0; r < RG; r++) {
                const int t = ks * RG + r;
                aload(t + AD, (t + AD) % NA);
                __builtin_amdgcn_sched_barrier(0);
                const h8 ah = ar_h[t % NA], al = ar_l[t % NA];
#pragma unroll
                for (int c = 0; c < CG; c++) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[s0][c], acc[r][c], 0, 0, 0);
                    if constexpr (SL)
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[s0][c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[s0][c], acc[r][c], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // U += o (rows) * (C + intercept term)
#pragma unroll
        for (int r = 0; r < RG; r++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float o = os[buf][16 * (rt0 + r) + orow + q];
#pragma unroll
                for (int c = 0; c < CG; c++) U[r][c][q] += o * (acc[r][c][q] + tv[c]);
            }
        }
        wstore(wnext, buf ^ 1);
        __syncthreads();                          // os[buf ^ 1] written; os[buf] free
    }
    float *pg = part + (size_t)g * M * Npad;
#pragma unroll
    for (int r = 0; r < RG; r++) {
        const int row0 = cb + 16 * (rt0 + r) + orow;
#pragma unroll
        for (int c = 0; c < CG; c++) {
            const int col = 16 * (ct0 + c) + lk;
#pragma unroll
            for (int q = 0; q < 4; q++) pg[(size_t)col * Npad + row0 + q] = U[r][c][q];
        }
    }
}

// Uw[:, 1 + c] = sum_g part[g][c][:] 2^-C_g (c < M)
__global__ void k_gc_reduce_gp(int Npad, int M, int ngroup, const float *__restrict__ part,
                               const double *__restrict__ gun, double *__restrict__ Uw) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)Npad * M) return;
    double s = 0.0;
    for (int g = 0; g < ngroup; g++) s += (double)part[(size_t)g * M * Npad + idx] * gun[g];
    Uw[(size_t)Npad + idx] = s;
}

// Uw[i, 0] = sum_{k<M} Zw[i, 1+k] Y[i, k] + Y[i, M]
__global__ void k_gc_u0(int Npad, int M, const double *__restrict__ Zw, const float *__restrict__ Y,
                        double *__restrict__ Uw) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Npad) return;
    double s = (double)Y[i + (size_t)M * Npad];
    for (int k = 0; k < M; k++) s += Zw[i + (size_t)(1 + k) * Npad] * (double)Y[i + (size_t)k * Npad];
    Uw[i] = s;
}

template <int TG, int RG, int CG, int PD, bool SL = true>
static int launch_gc_gp(gp_ctx *c, int gsize, int ngroup) {
    constexpr int M = 16 * TG;
    const size_t shm = (size_t)2 * GRC * (M + 8) * sizeof(_Float16);
    if (shm > 65536 && !c->capturing)
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_gp<TG, RG, CG, PD, SL>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(c->Npad / GRC, ngroup);
    k_gc_gp<TG, RG, CG, PD, SL><<<g, gc_gp_waves<TG, RG, CG>() * 64, shm, c->stream>>>(c->N, c->Npad, gsize, c->J, c->z32 + c->Npad,
                                                           c->w32, c->gc_perm, c->gc_aexp, c->gc_zexp + 1,
                                                           (const _Float16 *)c->Sh, (const _Float16 *)c->Sl,
                                                           c->gc_t0, c->gc_part);
    GP_HIP(hipGetLastError());
    return 0;
}

// prefetch depth: the deepest of 1, 3 and 7 whose ring divides the k-steps
// (M / 32) and does not exceed the option gc_prefetch (at most 1 for the
// 4-wave shape)
template <int TG, int RG, int CG, bool SL>
static int launch_gc_gp_pd(gp_ctx *c, int gsize, int ngroup) {
    constexpr int KS = TG / 2;
    if constexpr (KS % 8 == 0 && RG * CG <= 16) {
        if (c->gc_prefetch >= 7) return launch_gc_gp<TG, RG, CG, 7, SL>(c, gsize, ngroup);
    }
    // 4 B sets of 32 tiles spill with hi/lo B (not with B in FP16 alone)
    if constexpr (KS % 4 == 0 && (RG * CG <= 16 || !SL)) {
        if (c->gc_prefetch >= 3) return launch_gc_gp<TG, RG, CG, 3, SL>(c, gsize, ngroup);
    }
    if constexpr (KS % 2 == 0) {
        if (c->gc_prefetch >= 1) return launch_gc_gp<TG, RG, CG, 1, SL>(c, gsize, ngroup);
    }
    return launch_gc_gp<TG, RG, CG, 0, SL>(c, gsize, ngroup);
}

// wave shape: gc_waves 4 (RG = 8, CG = 4; M a multiple of 64), else
// gc_wave_rows 4 (RG = 4, CG = 4) or the default RG = 8, CG = 2
template <int TG, bool SL>
static int launch_gc_gp_rg(gp_ctx *c, int gsize, int ngroup) {
    if constexpr (TG % 4 == 0) {
        if (c->gc_waves == 4) return launch_gc_gp_pd<TG, 8, 4, SL>(c, gsize, ngroup);
        if (c->gc_wave_rows == 4) return launch_gc_gp_pd<TG, 4, 4, SL>(c, gsize, ngroup);
    }
    return launch_gc_gp_pd<TG, 8, 2, SL>(c, gsize, ngroup);
}

// operand precision of S_j: hi/lo (option gc_slo 1) or FP16 alone
template <int TG>
static int launch_gc_gp_sl(gp_ctx *c, int gsize, int ngroup) {
    if (c->gc_slo) return launch_gc_gp_rg<TG, true>(c, gsize, ngroup);
    return launch_gc_gp_rg<TG, false>(c, gsize, ngroup);
}

// the peeled form applies: the intercept is the only column before the GP
// block, and M is a multiple of 32 with at most 16 column tiles
static bool gc_peel_ok(const gp_ctx *c) {
    return c->gc_peel && c->gc_form == 2 && c->g0 == 1 && c->M % 32 == 0 && c->M / 16 <= 16 &&
           c->Npad % GRC == 0;
}

static int gradcontract_gp(gp_ctx *c, int gsize, int ngroup) {
    const int M = c->M, J = c->J, TG = M / 16;
    if (!c->zsplit_valid) GP_TRY(gpk_split_z(c, c->Zw));     // the FP32 copy of zw the kernel reads
    if (!c->gc_t0) {
        GP_HIP(hipMalloc((void **)&c->gc_t0, (size_t)J * M * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->gc_s0, (size_t)J * (M + 1) * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->gc_y, (size_t)c->Npad * (M + 1) * sizeof(float)));
    }
    dim3 gs((unsigned)(((size_t)M * M + 255) / 256), J);
    k_ssplit_gp<<<gs, 256, 0, c->stream>>>(c->P, M, c->Sinv, c->gc_bexp, c->gc_zexp, J,
                                           (_Float16 *)c->Sh, (_Float16 *)c->Sl, c->gc_t0, c->gc_s0);
    GP_HIP(hipGetLastError());
    // intercept output column: Y = w32 (Npad x J) . S0 (J x (M+1)), FP32
    const float one = 1.0f, zero = 0.0f;
    GP_BLAS(rocblas_sgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad, M + 1, J,
                          &one, c->w32, c->Npad, c->gc_s0, J, &zero, c->gc_y, c->Npad));
    int r = 0;
    gp_timer_begin(c, c->t_gck);
    switch (TG) {
    case 2: r = launch_gc_gp_sl<2>(c, gsize, ngroup); break;
    case 4: r = launch_gc_gp_sl<4>(c, gsize, ngroup); break;
    case 6: r = launch_gc_gp_sl<6>(c, gsize, ngroup); break;
    case 8: r = launch_gc_gp_sl<8>(c, gsize, ngroup); break;
    case 10: r = launch_gc_gp_sl<10>(c, gsize, ngroup); break;
    case 12: r = launch_gc_gp_sl<12>(c, gsize, ngroup); break;
    case 14: r = launch_gc_gp_sl<14>(c, gsize, ngroup); break;
    default: r = launch_gc_gp_sl<16>(c, gsize, ngroup); break;
    }
    if (r) return r;
    gp_timer_end(c, c->t_gck);
    const size_t n = (size_t)c->Npad * M;
    k_gc_reduce_gp<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->Npad, M, ngroup, c->gc_part,
                                                                      c->gc_gun, c->Uw);
    GP_HIP(hipGetLastError());
    k_gc_u0<<<(unsigned)((c->Npad + 255) / 256), 256, 0, c->stream>>>(c->Npad, M, c->Zw, c->gc_y, c->Uw);
    GP_HIP(hipGetLastError());
    return 0;
}

int gpk_gradcontract_mixed(gp_ctx *c) {
    const int T16 = (c->P + 15) / 16, P16 = 16 * T16;
    if (T16 > MAXT16 || c->Npad % 128 != 0) return gpk_gradcontract(c, c->Zw, c->Uw);
    const int J = c->J;
    int ngroup = std::min(c->gc_groups, J);
    const int gsize = (J + ngroup - 1) / ngroup;
    ngroup = (J + gsize - 1) / gsize;
    if (!c->Sh) {
        GP_HIP(hipMalloc(&c->Sh, (size_t)J * P16 * KP * sizeof(_Float16)));
        GP_HIP(hipMalloc(&c->Sl, (size_t)J * P16 * KP * sizeof(_Float16)));
        GP_HIP(hipMalloc((void **)&c->gc_perm, (size_t)J * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->gc_aexp, (size_t)J * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->gc_sexp, (size_t)J * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->gc_bexp, (size_t)J * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->gc_gun, (size_t)J * sizeof(double)));
        GP_HIP(hipMalloc((void **)&c->gc_zexp, (size_t)KP * sizeof(int)));
        GP_HIP(hipMalloc((void **)&c->gc_kappa, (size_t)J * sizeof(int)));
    }
    if (c->gc_part_len < (size_t)ngroup * P16 * c->Npad) {
        if (c->gc_part) (void)hipFree(c->gc_part);
        c->gc_part = nullptr;
        c->gc_part_len = (size_t)ngroup * P16 * c->Npad;
        GP_HIP(hipMalloc((void **)&c->gc_part, c->gc_part_len * sizeof(float)));
    }
    // per-gene exponents: e_j from the weight prep (wun = 2^e_j), s_j from max |Sw_j|;
    // the ELBO's weight prep still holds when W and Omega have not changed since
    if (!c->w32_valid) GP_TRY(gpk_wprep(c, c->W));
    k_zexp<<<KP, ZX_T, 0, c->stream>>>(c->Npad, c->P, c->Zw, c->gc_zexp);
    GP_HIP(hipGetLastError());
    k_smax<<<J, 256, 0, c->stream>>>(c->P, c->Sinv, c->gc_zexp, c->gc_sexp);
    GP_HIP(hipGetLastError());
    k_gc_plan<<<1, 256, 0, c->stream>>>(J, gsize, ngroup, c->wun, c->gc_sexp, c->gc_perm,
                                        c->gc_aexp, c->gc_bexp, c->gc_gun, c->gc_kappa);
    GP_HIP(hipGetLastError());
    if (gc_peel_ok(c)) return gradcontract_gp(c, gsize, ngroup);
    dim3 gs((unsigned)(((size_t)P16 * KP + 255) / 256), J);
    k_ssplit<<<gs, 256, 0, c->stream>>>(c->P, P16, c->Sinv, c->gc_bexp, c->gc_zexp,
                                        (_Float16 *)c->Sh, (_Float16 *)c->Sl);
    GP_HIP(hipGetLastError());
    int r = 0;
    gp_timer_begin(c, c->t_gck);
    switch (T16) {
    case 1: r = launch_gc<1>(c, gsize, ngroup); break;
    case 2: r = launch_gc<2>(c, gsize, ngroup); break;
    case 3: r = launch_gc<3>(c, gsize, ngroup); break;
    case 4: r = launch_gc<4>(c, gsize, ngroup); break;
    case 5: r = launch_gc<5>(c, gsize, ngroup); break;
    case 6: r = launch_gc<6>(c, gsize, ngroup); break;
    case 7: r = launch_gc<7>(c, gsize, ngroup); break;
    case 8: r = launch_gc<8>(c, gsize, ngroup); break;
    case 9: r = launch_gc<9>(c, gsize, ngroup); break;
    case 10: r = launch_gc<10>(c, gsize, ngroup); break;
    case 11: r = launch_gc<11>(c, gsize, ngroup); break;
    case 12: r = launch_gc<12>(c, gsize, ngroup); break;
    case 13: r = launch_gc<13>(c, gsize, ngroup); break;
    case 14: r = launch_gc<14>(c, gsize, ngroup); break;
    case 15: r = launch_gc<15>(c, gsize, ngroup); break;
    case 16: r = launch_gc<16>(c, gsize, ngroup); break;
    default: r = launch_gc<17>(c, gsize, ngroup); break;
    }
    if (r) return r;
    gp_timer_end(c, c->t_gck);
    const size_t n = (size_t)c->Npad * c->P;
    k_gc_reduce<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->Npad, c->P, P16, ngroup,
                                                                   c->gc_part, c->gc_gun, c->Uw);
    GP_HIP(hipGetLastError());
    return 0;
}
