// gp_gram.hip -- per-gene weighted Gram  Psi_j = DeltaInv + Z^T diag(W_j/Omega) Z
// (getPsi, ggplvm.cu:621-680; initPsi, init.cu:125-181 when W is absent).
//
// FP64 path: 64x64 output tiles of the upper triangle, one gene per
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

}  // namespace

int gpk_gram(gp_ctx *c, const double *W, double *Psi, int batch) {
    const int nt = (c->P + TILE - 1) / TILE;
    dim3 g(nt * (nt + 1) / 2, batch);
    k_gram_f64<<<g, GT, 0, c->stream>>>(c->N, c->Npad, c->P, nt, c->Z, W, c->Omega, c->DeltaInv, Psi);
    GP_HIP(hipGetLastError());
    return 0;
}
