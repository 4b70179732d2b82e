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

}  // namespace

int gpk_gradcontract(gp_ctx *c) {
    dim3 g(c->Npad / TI, (c->P + TN - 1) / TN);
    k_gradcontract_f64<<<g, UT, 0, c->stream>>>(c->N, c->Npad, c->J, c->P, c->Z, c->W, c->Omega,
                                                c->Sinv, c->U);
    GP_HIP(hipGetLastError());
    return 0;
}
