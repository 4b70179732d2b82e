// gp_gradcontract.hip -- the dominant contraction of the ELBO gradient:
//   U[i,:] = sum_j (W_ij / Omega_i) Psi_j^{-1} z_i          (Npad x P)
// getGradLogOmega (ggplvm.cu:301-376) solves Psi_j u = z_i for every
// (cell, gene) pair; this kernel forms the same sums as one contraction of the
// model matrix with the explicit inverses.  z_i^T U[i,:] gives
// sum_j o_ij z_i^T Psi_j^{-1} z_i and the GP columns of U give the first term
// of C1.
//
// FP64 path: 64-cell x 64-column output tiles, 32-row chunks of Sinv_j and of
// the gene-weighted model-matrix columns staged in LDS; 4x4 per thread.
#include "gp_internal.h"

namespace {

constexpr int UT = 256;
constexpr int TI = 64;
constexpr int TN = 64;
constexpr int KL = 32;

__global__ void __launch_bounds__(UT) k_gradcontract_f64(int N, int Npad, int J, int P,
        const double *__restrict__ Z, const double *__restrict__ W,
        const double *__restrict__ Omega, const double *__restrict__ Sinv,
        double *__restrict__ U) {
    __shared__ double As[KL][TI + 1];   // [l][cell]  o_ij z_il
    __shared__ double Bs[KL][TN + 1];   // [l][col]   Sinv_j[l, col]
    __shared__ double os[TI];
    const int i0 = blockIdx.x * TI, n0 = blockIdx.y * TN;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) acc[r][s] = 0.0;

    for (int j = 0; j < J; j++) {
        const double *S = Sinv + (size_t)j * P * P;
        __syncthreads();
        if (tid < TI) {
            const int i = i0 + tid;
            os[tid] = (i < N) ? W[i + (size_t)j * Npad] / Omega[i] : 0.0;
        }
        for (int l0 = 0; l0 < P; l0 += KL) {
            __syncthreads();
            {
                // A chunk: 32 l x 64 cells; thread -> (cell = tid%64, l group = tid/64)
                const int cell = tid % TI, lg = tid / TI;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int l = lg * 8 + u, gl = l0 + l;
                    As[l][cell] = (gl < P) ? Z[(i0 + cell) + (size_t)gl * Npad] * os[cell] : 0.0;
                }
                // B chunk: 32 l x 64 cols; thread -> (l = tid%32, col group = tid/32)
                const int ll = tid % KL, cg = tid / KL;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int col = cg * 8 + u, gc = n0 + col, gl = l0 + ll;
                    Bs[ll][col] = (gl < P && gc < P) ? S[gl + (size_t)gc * P] : 0.0;
                }
            }
            __syncthreads();
#pragma unroll 8
            for (int l = 0; l < KL; l++) {
                double av[4], bv[4];
#pragma unroll
                for (int r = 0; r < 4; r++) av[r] = As[l][ty * 4 + r];
#pragma unroll
                for (int s = 0; s < 4; s++) bv[s] = Bs[l][tx * 4 + s];
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int s = 0; s < 4; s++) acc[r][s] += av[r] * bv[s];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int i = i0 + ty * 4 + r, col = n0 + tx * 4 + s;
            if (col < P) U[i + (size_t)col * Npad] = acc[r][s];
        }
}


// ---------------------------------------------------------------------------
// FP64 MFMA path (v_mfma_f64_16x16x4_f64), P16 = 16*T16 <= 272.
// Workgroup = 4 waves = 64 cells; wave w owns cells i0+16w .. i0+16w+15 and
// all T16 column tiles of U.  Its model-matrix fragments stay in registers for
// the whole kernel:  zf[ks], lane l = Z[cell 16w + l%16][k = 4 ks + l/16].
// For every gene the A fragment is zf[ks] * o_ij (o = W/Omega of the lane's
// cell) and the B fragments are rows of Sinv_j, streamed through two LDS
// buffers of 16 rows x P16 columns (read as columns, Sinv being symmetric):
//   B frag, lane l: Sinv_j[k = 4 ks + l/16][n = 16 ct + l%16]
//   D,      lane l, r: U[cell 16w + 4r + l/16][16 ct + l%16]
typedef double gc_d4 __attribute__((ext_vector_type(4)));

template <int T16>
__global__ void __launch_bounds__(256) k_gc_mfma_f64(int N, int Npad, int J, int P,
        const double *__restrict__ Z, const double *__restrict__ W,
        const double *__restrict__ Omega, const double *__restrict__ Sinv,
        double *__restrict__ U) {
    constexpr int P16 = 16 * T16, KS = 4 * T16;
    extern __shared__ double sl[];               // [2][16][P16]
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i0 = blockIdx.x * 64;
    const int ci = i0 + 16 * w + (lane % 16);    // cell of this lane's A rows
    double zf[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int k = 4 * ks + lane / 16;
        zf[ks] = (k < P) ? Z[ci + (size_t)k * Npad] : 0.0;
    }
    gc_d4 acc[T16];
#pragma unroll
    for (int ct = 0; ct < T16; ct++) acc[ct] = gc_d4{0.0, 0.0, 0.0, 0.0};

    // staging map: thread -> row r = tid/16 of the 16-row chunk, columns tid%16 + 16 t;
    // a running pointer walks the chunks so the loads use immediate offsets
    const int sr = tid / 16, sc = tid % 16;
    const size_t PP = (size_t)P * P;
    const size_t chunk_step = (size_t)16 * P;
    double pf[T16];
    auto fetch = [&](const double *rowp, int k) {
        const bool krow = k < P;
#pragma unroll
        for (int t = 0; t < T16; t++)
            pf[t] = (krow && sc + 16 * t < P) ? rowp[16 * t] : 0.0;
    };
    auto store = [&](int buf) {
        double *d = sl + buf * 16 * P16 + sr * P16 + sc;
#pragma unroll
        for (int t = 0; t < T16; t++) d[16 * t] = pf[t];
    };
    const double *rp = Sinv + (size_t)sr * P + sc;   // row sr of chunk 0, gene 0
    fetch(rp, sr);
    store(0);
    double o = (ci < N) ? W[ci] / Omega[ci] : 0.0;
    double on = 0.0;
    __syncthreads();
    int cur = 0;
    int kn = sr;                                      // row index of the prefetched chunk
    for (int j = 0; j < J; j++) {
#pragma unroll
        for (int kc = 0; kc < T16; kc++) {
            const bool next = (kc + 1 < T16) || (j + 1 < J);
            if (next) {
                if (kc + 1 < T16) { rp += chunk_step; kn += 16; }
                else { rp = Sinv + (size_t)(j + 1) * PP + (size_t)sr * P + sc; kn = sr; }
                asm volatile("" : "+v"(kn));
                fetch(rp, kn);
            }
            if (kc + 1 == T16 && j + 1 < J)
                on = (ci < N) ? W[ci + (size_t)(j + 1) * Npad] / Omega[ci] : 0.0;
            const double *B = sl + cur * 16 * P16 + (lane / 16) * P16 + (lane % 16);
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const double a = zf[4 * kc + kk] * o;
#pragma unroll
                for (int ct = 0; ct < T16; ct++)
                    acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, B[4 * kk * P16 + 16 * ct],
                                                                   acc[ct], 0, 0, 0);
            }
            if (next) store(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
        o = on;
    }
#pragma unroll
    for (int ct = 0; ct < T16; ct++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = i0 + 16 * w + 4 * r + lane / 16, col = 16 * ct + (lane % 16);
            if (col < P) U[row + (size_t)col * Npad] = acc[ct][r];
        }
    }
}

}  // namespace

template <int T16>
static int launch_gc(gp_ctx *c, const double *Zs, double *Uo) {
    const size_t shm = 2 * 16 * (size_t)(16 * T16) * sizeof(double);
    if (shm > 65536 && !c->capturing)   // set before any recording: not a stream operation
        GP_HIP(hipFuncSetAttribute((const void *)k_gc_mfma_f64<T16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    gp_timer_begin(c, c->t_gck);
    k_gc_mfma_f64<T16><<<c->Npad / 64, 256, shm, c->stream>>>(c->N, c->Npad, c->J, c->P, Zs,
                                                               c->W, c->Omega, c->Sinv, Uo);
    GP_HIP(hipGetLastError());
    gp_timer_end(c, c->t_gck);
    return 0;
}

int gpk_gradcontract(gp_ctx *c, const double *Zs, double *Uo) {
    switch ((c->P + 15) / 16) {
    case 1: return launch_gc<1>(c, Zs, Uo);
    case 2: return launch_gc<2>(c, Zs, Uo);
    case 3: return launch_gc<3>(c, Zs, Uo);
    case 4: return launch_gc<4>(c, Zs, Uo);
    case 5: return launch_gc<5>(c, Zs, Uo);
    case 6: return launch_gc<6>(c, Zs, Uo);
    case 7: return launch_gc<7>(c, Zs, Uo);
    case 8: return launch_gc<8>(c, Zs, Uo);
    case 9: return launch_gc<9>(c, Zs, Uo);
    case 10: return launch_gc<10>(c, Zs, Uo);
    case 11: return launch_gc<11>(c, Zs, Uo);
    case 12: return launch_gc<12>(c, Zs, Uo);
    case 13: return launch_gc<13>(c, Zs, Uo);
    case 14: return launch_gc<14>(c, Zs, Uo);
    case 15: return launch_gc<15>(c, Zs, Uo);
    case 16: return launch_gc<16>(c, Zs, Uo);
    case 17: return launch_gc<17>(c, Zs, Uo);
    default: break;
    }
    dim3 g(c->Npad / TI, (c->P + TN - 1) / TN);
    k_gradcontract_f64<<<g, UT, 0, c->stream>>>(c->N, c->Npad, c->J, c->P, Zs, c->W, c->Omega,
                                                c->Sinv, Uo);
    GP_HIP(hipGetLastError());
    return 0;
}
