// gp_gram.hip -- per-gene weighted Gram  Psi_j = DeltaInv + Z^T diag(W_j/Omega) Z
// (getPsi, ggplvm.cu:621-680; initPsi, init.cu:125-181 when W is absent).
//
// Generic FP64 path (P16 > 576): 64x64 output tiles of the upper triangle, one gene per
// workgroup, 32-cell chunks of the weighted and plain model-matrix columns
// staged through LDS from coalesced column reads; each thread owns a 4x4
// block.  The result is written to both triangles (the Cholesky reads the
// lower one; the upper one stays as the reference leaves it).
#include "gp_internal.h"

namespace {

constexpr int GT = 256;
constexpr int TILE = 64;
constexpr int KCH = 32;

__device__ inline void tile_pair(int t, int nt, int &ta, int &tb) {
    ta = 0;
    while (t >= nt - ta) { t -= nt - ta; ta++; }
    tb = ta + t;
}

__global__ void __launch_bounds__(GT) k_gram_f64(int N, int Npad, int P, int ntile,
        const double *__restrict__ Z, const double *__restrict__ W,
        const double *__restrict__ Omega, const double *__restrict__ DeltaInv,
        double *__restrict__ Psi) {
    __shared__ double As[KCH][TILE + 1];
    __shared__ double Bs[KCH][TILE + 1];
    __shared__ double ws[KCH];
    int ta, tb;
    tile_pair(blockIdx.x, ntile, ta, tb);
    const int a0 = ta * TILE, b0 = tb * TILE;
    const int j = blockIdx.y;
    const double *Wj = W ? W + (size_t)j * Npad : nullptr;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) acc[r][s] = 0.0;

    for (int i0 = 0; i0 < Npad; i0 += KCH) {
        __syncthreads();
        if (tid < KCH) {
            const int i = i0 + tid;
            ws[tid] = (i < N) ? (Wj ? Wj[i] / Omega[i] : 1.0) : 0.0;
        }
        __syncthreads();
        {
            const int cell = tid % KCH, cg = tid / KCH;   // 8 column groups of 8
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int col = cg * 8 + u;
                const int ga = a0 + col, gb = b0 + col;
                const double za = (ga < P) ? Z[(i0 + cell) + (size_t)ga * Npad] : 0.0;
                const double zb = (gb < P) ? Z[(i0 + cell) + (size_t)gb * Npad] : 0.0;
                As[cell][col] = za * ws[cell];
                Bs[cell][col] = zb;
            }
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < KCH; k++) {
            double av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) av[r] = As[k][ty * 4 + r];
#pragma unroll
            for (int s = 0; s < 4; s++) bv[s] = Bs[k][tx * 4 + s];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc[r][s] += av[r] * bv[s];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int ga = a0 + ty * 4 + r, gb = b0 + tx * 4 + s;
            if (ga < P && gb < P && ga <= gb) {
                const double v = acc[r][s] + DeltaInv[ga + (size_t)gb * P];
                double *Pj = Psi + (size_t)j * P * P;
                Pj[ga + (size_t)gb * P] = v;
                Pj[gb + (size_t)ga * P] = v;
            }
        }
}

// ---------------------------------------------------------------------------
// FP64 MFMA path (v_mfma_f64_16x16x4_f64).  The P x P output is padded to
// P16 = 16*ceil(P/16); its upper-triangle 16x16 tiles are enumerated row by
// row and split into groups of GW*SPW tiles, one workgroup per (group, gene).
// Every 16-cell chunk of the weighted (A) and plain (B) model-matrix columns
// is staged in LDS; each wave keeps SPW tile accumulators in registers and
// reads its A/B fragments straight from LDS:
//   A frag, lane l: (w z)[cell = k0 + l/16][col = 16a + l%16]
//   B frag, lane l:    z [cell = k0 + l/16][col = 16b + l%16]
//   D,      lane l, r: Psi[16a + 4r + l/16][16b + l%16]
// The next chunk is fetched into registers while the current one is
// multiplied, so global latency hides behind the MFMAs.
typedef double gp_d4 __attribute__((ext_vector_type(4)));
constexpr int GW = 4;            // waves per workgroup
constexpr int SPW = 20;          // tiles per wave
constexpr int TPG = GW * SPW;    // tiles per workgroup
constexpr int KCM = 16;          // cells per chunk
constexpr int P16MAX = 576;      // LDS bound: 2*KCM*P16 doubles

// PF >= P16/16: model-matrix columns staged per thread (256 threads, 16 cells)
template <int PF>
__global__ void __launch_bounds__(256) k_gram_mfma_f64(int N, int Npad, int P, int P16, int T16,
        int ntot, const double *__restrict__ Z, const double *__restrict__ W,
        const double *__restrict__ Omega, const double *__restrict__ DeltaInv,
        double *__restrict__ Psi) {
    extern __shared__ double lds[];
    const int LS = P16 + 1;                 // LDS row stride: odd, so the 16 cells of a
                                            // staging store fall in distinct banks
    double *As = lds;                       // [KCM][LS]
    double *Bs = lds + KCM * LS;            // [KCM][LS]
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j = blockIdx.y;
    const double *Wj = W ? W + (size_t)j * Npad : nullptr;
    // this wave's tiles: a contiguous run [t0, t0+SPW) of the row-major
    // upper-triangle enumeration, walked incrementally as (ta, tb)
    const int t0 = blockIdx.x * TPG + w * SPW;
    const int nvalid = min(SPW, max(0, ntot - t0));
    int a0 = 0, rowstart = 0;
    while (a0 < T16 && t0 >= rowstart + (T16 - a0)) { rowstart += T16 - a0; a0++; }
    const int b0 = a0 + (t0 - rowstart);
    gp_d4 acc[SPW];
#pragma unroll
    for (int s = 0; s < SPW; s++) acc[s] = gp_d4{0.0, 0.0, 0.0, 0.0};

    // staging map: thread -> (cell = tid % 16, columns tid/16 + 16 r), so each
    // thread scales its own cell and the global/LDS addresses are one base
    // plus fixed strides
    const int cell = tid % KCM, col0 = tid / KCM;
    const double *zb = Z + cell + (size_t)col0 * Npad;
    const size_t zstride = (size_t)KCM * Npad;
    double *as = As + cell * LS + col0;
    double *bs = Bs + cell * LS + col0;
    double pf[PF];
    double pw = 0.0;
    auto fetch = [&](int i0) {
#pragma unroll
        for (int r = 0; r < PF; r++)
            pf[r] = (col0 + KCM * r < P) ? zb[i0 + r * zstride] : 0.0;
        const int i = i0 + cell;
        pw = (i < N) ? (Wj ? Wj[i] / Omega[i] : 1.0) : 0.0;
    };
    fetch(0);
    for (int i0 = 0; i0 < Npad; i0 += KCM) {
        __syncthreads();                    // previous chunk fully consumed
#pragma unroll
        for (int r = 0; r < PF; r++) {
            if (col0 + KCM * r < P16) {
                as[KCM * r] = pf[r] * pw;
                bs[KCM * r] = pf[r];
            }
        }
        __syncthreads();
        if (i0 + KCM < Npad) fetch(i0 + KCM);
        // per-lane fragment bases of the four 4-cell k-steps
        int base[KCM / 4];
#pragma unroll
        for (int ks = 0; ks < KCM / 4; ks++) {
            base[ks] = (ks * 4 + lane / 16) * LS + (lane % 16);
            asm volatile("" : "+v"(base[ks]));   // formed per chunk, not hoisted
        }
        int ta = a0, tb = b0;
#pragma unroll
        for (int s = 0; s < SPW; s++) {
            // tiles past the end of the run repeat the last tile (never stored)
            const int oa = min(ta, T16 - 1) * 16, ob = min(tb, T16 - 1) * 16;
#pragma unroll
            for (int ks = 0; ks < KCM / 4; ks++) {
                const double av = As[base[ks] + oa];
                const double bv = Bs[base[ks] + ob];
                acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[s], 0, 0, 0);
            }
            if (++tb == T16) { ++ta; tb = ta; }
        }
    }
    double *Pj = Psi + (size_t)j * P * P;
    int ta = a0, tb = b0;
#pragma unroll
    for (int s = 0; s < SPW; s++) {
        if (s < nvalid) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = ta * 16 + 4 * r + lane / 16, col = tb * 16 + lane % 16;
                if (row < P && col < P && row <= col) {
                    const double v = acc[s][r] + DeltaInv[row + (size_t)col * P];
                    Pj[row + (size_t)col * P] = v;
                    Pj[col + (size_t)row * P] = v;
                }
            }
        }
        if (++tb == T16) { ++ta; tb = ta; }
    }
}

}  // namespace

template <int PF>
static int launch_gram_mfma(gp_ctx *c, const double *W, double *Psi, int batch, int P16, int T16,
                            const double *Zs, const double *Dl) {
    const int ntot = T16 * (T16 + 1) / 2;
    const int ng = (ntot + TPG - 1) / TPG;
    const size_t shm = 2 * (size_t)KCM * (P16 + 1) * sizeof(double);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gram_mfma_f64<PF>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    dim3 g(ng, batch);
    gp_timer_begin(c, c->t_gramk);
    k_gram_mfma_f64<PF><<<g, 256, shm, c->stream>>>(c->N, c->Npad, c->P, P16, T16, ntot, Zs, W,
                                                     c->Omega, Dl, Psi);
    GP_HIP(hipGetLastError());
    gp_timer_end(c, c->t_gramk);
    return 0;
}

int gpk_gram(gp_ctx *c, const double *W, double *Psi, int batch) {
    return gpk_gram_src(c, W, Psi, batch, c->Z, c->DeltaInv);
}

int gpk_gram_src(gp_ctx *c, const double *W, double *Psi, int batch, const double *Zs,
                 const double *Dl) {
    const int T16 = (c->P + 15) / 16, P16 = 16 * T16;
    const int pf = T16;   // columns per thread: P16 / 16
    if (P16 <= P16MAX) {
        if (pf <= 4) return launch_gram_mfma<4>(c, W, Psi, batch, P16, T16, Zs, Dl);
        if (pf <= 9) return launch_gram_mfma<9>(c, W, Psi, batch, P16, T16, Zs, Dl);
        if (pf <= 17) return launch_gram_mfma<17>(c, W, Psi, batch, P16, T16, Zs, Dl);
        if (pf <= 25) return launch_gram_mfma<25>(c, W, Psi, batch, P16, T16, Zs, Dl);
        return launch_gram_mfma<36>(c, W, Psi, batch, P16, T16, Zs, Dl);
    }
    const int nt = (c->P + TILE - 1) / TILE;
    dim3 g(nt * (nt + 1) / 2, batch);
    k_gram_f64<<<g, GT, 0, c->stream>>>(c->N, c->Npad, c->P, nt, Zs, W, c->Omega, Dl, Psi);
    GP_HIP(hipGetLastError());
    return 0;
}
