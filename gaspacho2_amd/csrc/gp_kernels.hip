// gp_kernels.hip -- element-wise, reduction and epilogue kernels of the fit:
// the Kuf/Kmm build, IRLS pseudo-data, likelihood assembly, the per-cell and
// per-gene gradient epilogues, C2, the T/rho gradients and the L-BFGS update.
// Every reduction is two-stage with a fixed partition, so results are
// bit-reproducible from run to run.
#include "gp_internal.h"
#include "../../include/gaspacho2.h"

#include <algorithm>

#include <cmath>

namespace {

constexpr int NT = 256;          // threads per block for 1-D kernels
constexpr int RED_BLOCKS = 256;  // stage-1 blocks of the vector reductions

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over a 256-thread block; result valid on thread 0.
__device__ inline double block_sum256(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}

inline unsigned nblk(size_t n, int nt = NT) { return (unsigned)((n + nt - 1) / nt); }

// ---------------------------------------------------------------- Kuf / Kmm
// kernel_funcs.cu:61-97 (periodic dim 0 + SE-ARD) and 221-255 (SE-ARD only).
// 256 cells x 8 inducing points per block.  The block's rows of the latent X
// (Q columns of 256 consecutive cells: one coalesced 2 KB read per column)
// and the 8 inducing points' rows of T are staged in LDS, so each X value
// is read from HBM once per block and from LDS by its cell's thread for
// every inducing point; the arithmetic is the reference's, term by term.
__global__ void __launch_bounds__(NT) k_kuf(int N, int Npad, int M, int Q,
        const double *__restrict__ X, const double *__restrict__ T,
        const double *__restrict__ rho, int ard, double *__restrict__ Knm) {
    extern __shared__ double Xs[];               // [Q][NT]
    __shared__ double Ts[8][64];
    __shared__ double rh[64];
    const int i = blockIdx.x * NT + threadIdx.x;
    const int m0 = blockIdx.y * 8;
    for (int t = threadIdx.x; t < 8 * Q; t += NT) {
        int mm = t % 8, q = t / 8;
        Ts[mm][q] = (m0 + mm < M) ? T[(m0 + mm) + (size_t)q * M] : 0.0;
    }
    for (int q = threadIdx.x; q < Q; q += NT) rh[q] = rho[q];
    for (int q = 0; q < Q; q++) Xs[q * NT + threadIdx.x] = (i < N) ? X[i + (size_t)q * N] : 0.0;
    __syncthreads();
    if (i >= Npad) return;
    const double *xi = Xs + threadIdx.x;
    for (int mm = 0; mm < 8; mm++) {
        const int m = m0 + mm;
        if (m >= M) break;
        double v = 0.0;
        if (i < N) {
            double sum = 0.0, t;
            int k0 = 0;
            if (!ard) {
                t = sin((xi[0] - Ts[mm][0]) / 2.0);
                sum += t * t / rh[0];
                k0 = 1;
            }
            for (int k = k0; k < Q; k++) {
                t = xi[k * NT] - Ts[mm][k];
                sum += t * t / 2.0 / rh[k];
            }
            v = exp(-sum);
        }
        Knm[i + (size_t)m * Npad] = v;
    }
}

__device__ inline double kmm_entry(int a, int b, int M, int Q, const double *T,
                                   const double *rho, int ard) {
    // value for a < b, evaluated exactly as the reference's ii < j branch
    double sum = 0.0, t;
    int k0 = 0;
    if (!ard) {
        t = sin((T[a] - T[b]) / 2.0);
        sum += t * t / rho[0];
        k0 = 1;
    }
    for (int k = k0; k < Q; k++) {
        t = T[a + (size_t)k * M] - T[b + (size_t)k * M];
        sum += t * t / 2.0 / rho[k];
    }
    return exp(-sum);
}

__global__ void k_kmm(int M, int Q, const double *T, const double *rho, int ard,
                      double *K, double *Ksym, double *Kchol) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)M * M) return;
    const int a = (int)(idx % M), b = (int)(idx / M);
    double v, vs;
    if (a == b) { v = 1.0; vs = 1.0; }
    else if (a < b) { v = kmm_entry(a, b, M, Q, T, rho, ard); vs = v; }
    else { v = 0.0; vs = kmm_entry(b, a, M, Q, T, rho, ard); }
    K[idx] = v;
    Ksym[idx] = vs;
    Kchol[idx] = vs;
}

// ---------------------------------------------------------------- DeltaInv
// util_cuda.cu:249-269 expandDelta_k.  DeltaChol is the factor choldc would
// leave for DeltaInv (chol.cu:19-43: L on and below the diagonal, the matrix
// above it): DeltaInv = blockdiag(diag(1/delta_l), K/delta_{p-1}), so its
// factor is blockdiag(diag(sqrt(1/delta_l)), Kchol/sqrt(delta_{p-1})) with
// Kchol the factor of K -- formed here instead of factoring again.
__global__ void k_expand_delta(int P, int p, int M, int g0, const int *nh,
                               const double *delta, const double *K, const double *Kchol,
                               double *DeltaInv, double *DeltaChol, double *DeltaW,
                               double *DeltaWChol) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)P * P) return;
    const int r = (int)(idx % P), c = (int)(idx / P);
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    double sym = 0.0;
    if (lo == hi && lo < g0) {
        int l = 0;
        while (l < p - 1 && !(nh[l] <= lo && lo < nh[l + 1])) l++;
        sym = 1.0 / delta[l];
    } else if (lo >= g0) {
        sym = K[(lo - g0) + (size_t)(hi - g0) * M] / delta[p - 1];
    }
    DeltaInv[idx] = (r <= c) ? sym : 0.0;
    double fac = sym;                                  // above the diagonal: the matrix
    if (r >= c) {                                      // on and below: the factor
        if (c >= g0) fac = Kchol[(r - g0) + (size_t)(c - g0) * M] / sqrt(delta[p - 1]);
        else fac = (r == c) ? sqrt(sym) : 0.0;
    }
    DeltaChol[idx] = fac;
    if (DeltaW) {
        // whitened prior: diag(1/delta_l) (the GP block K/delta becomes I/delta)
        double dw = 0.0;
        if (r == c) {
            int l = 0;
            while (l < p - 1 && !(nh[l] <= r && r < nh[l + 1])) l++;
            dw = 1.0 / delta[l];
        }
        DeltaW[idx] = dw;
        DeltaWChol[idx] = (r == c) ? sqrt(dw) : 0.0;
    }
}

// ---------------------------------------------------------------- Oinvt
// ggplvm.cu:284-294: Oinvt_i = sum_j W_ij / Omega_i (same j order).
__global__ void k_oinvt(int N, int Npad, int J, const double *W, const double *Omega,
                        double *Oinvt) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    const double om = Omega[i];
    double o = 0.0;
    int j = 0;
    for (; j + 8 <= J; j += 8) {          // 8 loads in flight, same summation order
        double w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) w[u] = W[i + (size_t)(j + u) * Npad];
#pragma unroll
        for (int u = 0; u < 8; u++) o += w[u] / om;
    }
    for (; j < J; j++) o += W[i + (size_t)j * Npad] / om;
    Oinvt[i] = o;
}

// ---------------------------------------------------------------- titsias
// kernel_funcs.cu:160: t_i = (1 - k_i^T Kmm^{-1} k_i) / 2, with A = Knm Kmm^{-1}.
__global__ void k_titsias(int N, int Npad, int M, const double *Knm, const double *A,
                          double *tits) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int m = 0; m < M; m++) s += Knm[i + (size_t)m * Npad] * A[i + (size_t)m * Npad];
    tits[i] = (1.0 - s) / 2.0;
}

// ---------------------------------------------------------------- IRLS W / Nu
// ggplvm.cu:468-514 with zeta == 0; eta = z_i^T beta_j precomputed in ZB.
__global__ void k_w_nu(int N, int Npad, const double *ZB, const double *ls,
                       const double *tits, double dtwt_delta, double wt,
                       double *W, double *Nu) {
    const int i = blockIdx.x * NT + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= Npad) return;
    const size_t a = i + (size_t)j * Npad;
    if (i >= N) { W[a] = 0.0; Nu[a] = 0.0; return; }
    const double eta = ZB[a];
    W[a] = exp(eta) * ls[i] * exp(dtwt_delta * tits[i] * wt);
    Nu[a] = eta - 1.0;
}

__global__ void k_nu_sparse(int Npad, const int *Yp, const int *Yi, const double *Yx,
                            const double *ZB, const double *ls, const double *tits,
                            double dtwt_delta, double wt, double *Nu) {
    const int j = blockIdx.x;
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += NT) {
        const int i = Yi[l];
        const size_t a = i + (size_t)j * Npad;
        const double eta = ZB[a];
        const double logW = eta + log(ls[i]) + dtwt_delta * tits[i] * wt;
        Nu[a] = eta + exp(log(Yx[l]) - logW) - 1.0;
    }
}

// R = (W/Omega) nu, gq_j = sum nu^2 W / Omega, gs_j = sum log(Omega / W).
// A thread loads RSU of its cells before using any of them (3 RSU loads in
// flight instead of 3); its sums still run in cell order.
constexpr int RSU = 4;
__global__ void __launch_bounds__(NT) k_r_sums(int N, int Npad, const double *__restrict__ W,
        const double *__restrict__ Nu, const double *__restrict__ Omega, double *__restrict__ R,
        double *__restrict__ gq, double *__restrict__ gs) {
    __shared__ double sh[4];
    const int j = blockIdx.x;
    const double *Wj = W + (size_t)j * Npad, *nj = Nu + (size_t)j * Npad;
    double *Rj = R + (size_t)j * Npad;
    double q = 0.0, s = 0.0;
    int i = threadIdx.x;
    for (; i + (RSU - 1) * NT < N; i += RSU * NT) {
        double w[RSU], nu[RSU], om[RSU];
#pragma unroll
        for (int u = 0; u < RSU; u++) {
            w[u] = Wj[i + u * NT];
            nu[u] = nj[i + u * NT];
            om[u] = Omega[i + u * NT];
        }
#pragma unroll
        for (int u = 0; u < RSU; u++) {
            Rj[i + u * NT] = w[u] / om[u] * nu[u];
            q += nu[u] * nu[u] * w[u] / om[u];
            s += log(om[u] / w[u]);
        }
    }
    for (; i < Npad; i += NT) {
        if (i < N) {
            const double w = Wj[i], nu = nj[i], om = Omega[i];
            Rj[i] = w / om * nu;
            q += nu * nu * w / om;
            s += log(om / w);
        } else {
            Rj[i] = 0.0;
        }
    }
    q = block_sum256(q, sh);
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) { gq[j] = q; gs[j] = s; }
}

// k_r_sums (same traversal and expressions, so the same bits) that also
// writes, when w32 is given, the mixed Gram's FP32 weights W / Omega and
// their per-gene power-of-two scale (the values k_wprep writes): one pass
// over W, nu and Omega serves the likelihood and the Gram.
__global__ void __launch_bounds__(NT) k_r_sums_w32(int N, int Npad, const double *__restrict__ W,
        const double *__restrict__ Nu, const double *__restrict__ Omega, double *__restrict__ R,
        double *__restrict__ gq, double *__restrict__ gs, float *__restrict__ w32,
        float *__restrict__ wsc, double *__restrict__ wun) {
    __shared__ double sh[4];
    __shared__ float shm[4];
    const int j = blockIdx.x;
    const double *Wj = W + (size_t)j * Npad, *nj = Nu + (size_t)j * Npad;
    double *Rj = R + (size_t)j * Npad;
    float *wj = w32 ? w32 + (size_t)j * Npad : nullptr;
    double q = 0.0, s = 0.0;
    float mx = 0.0f;
    int i = threadIdx.x;
    for (; i + (RSU - 1) * NT < N; i += RSU * NT) {
        double w[RSU], nu[RSU], om[RSU];
#pragma unroll
        for (int u = 0; u < RSU; u++) {
            w[u] = Wj[i + u * NT];
            nu[u] = nj[i + u * NT];
            om[u] = Omega[i + u * NT];
        }
#pragma unroll
        for (int u = 0; u < RSU; u++) {
            const double wo = w[u] / om[u];
            Rj[i + u * NT] = wo * nu[u];
            q += nu[u] * nu[u] * w[u] / om[u];
            s += log(om[u] / w[u]);
            if (wj) {
                const float v = (float)wo;
                wj[i + u * NT] = v;
                mx = fmaxf(mx, fabsf(v));
            }
        }
    }
    for (; i < Npad; i += NT) {
        if (i < N) {
            const double w = Wj[i], nu = nj[i], om = Omega[i];
            const double wo = w / om;
            Rj[i] = wo * nu;
            q += nu * nu * w / om;
            s += log(om / w);
            if (wj) {
                const float v = (float)wo;
                wj[i] = v;
                mx = fmaxf(mx, fabsf(v));
            }
        } else {
            Rj[i] = 0.0;
            if (wj) wj[i] = 0.0f;
        }
    }
    q = block_sum256(q, sh);
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) { gq[j] = q; gs[j] = s; }
    if (wj) {
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_down(mx, off, 64));
        if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            const float m = fmaxf(fmaxf(shm[0], shm[1]), fmaxf(shm[2], shm[3]));
            int e = 0;
            if (m > 0.0f) (void)frexpf(m, &e);      // m <= 2^e
            wsc[j] = ldexpf(1.0f, -e);
            wun[j] = ldexp(1.0, e);
        }
    }
}

// init.cu:193-221 right-hand side: log((y + 0.05) / ls) at non-zeros,// init.cu:193-221 right-hand side: log((y + 0.05) / ls) at non-zeros,
// log(0.05 / ls) elsewhere.
__global__ void k_y0_dense(int N, int Npad, const double *ls, double *R) {
    const int i = blockIdx.x * NT + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= Npad) return;
    R[i + (size_t)j * Npad] = (i < N) ? log(0.05 / ls[i]) : 0.0;
}
__global__ void k_y0_sparse(int Npad, const int *Yp, const int *Yi, const double *Yx,
                            const double *ls, double *R) {
    const int j = blockIdx.x;
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += NT) {
        const int i = Yi[l];
        R[i + (size_t)j * Npad] = log((Yx[l] + 0.05) / ls[i]);
    }
}

// ---------------------------------------------------------------- getBetaPhi
// ggplvm.cu:418-464: damped beta update and phi from L^T beta.
__global__ void __launch_bounds__(NT) k_beta_phi(int N, int P, int jtr,
        const double *Psi, const double *tmpsol, const double *zeta,
        const double *gq, double *Beta, double *phi, double *work) {
    __shared__ double sh[4];
    const int j = blockIdx.x;
    double r = (double)(jtr + 1) / 100.0;
    if (jtr > 100) r = 1.0;
    double *Bj = Beta + (size_t)j * P;
    double *tb = work + (size_t)j * P;
    const double *sol = tmpsol + (size_t)j * P;
    for (int k = threadIdx.x; k < P; k += NT) {
        double b = Bj[k] * (1 - r) + sol[k] * r;
        tb[k] = b;
        Bj[k] = b + zeta[k];
    }
    __syncthreads();
    const double *Lj = Psi + (size_t)j * P * P;
    double acc = 0.0;
    for (int k = threadIdx.x; k < P; k += NT) {
        double t = 0.0;
        for (int i = k; i < P; i++) t += Lj[i + (size_t)k * P] * tb[i];
        acc += t * t;
    }
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
        double ph = (gq[j] - acc) / (double)N;
        if (isnan(ph)) ph = 1.0;
        if (ph < 0.0) ph = 1.0;
        if (ph < 0.01) ph = 0.01;
        if (ph > 1000.0) ph = 1000.0;
        phi[j] = ph;
    }
}

// ---------------------------------------------------------------- getLkhdPois
// ggplvm.cu:703-724: sum_nz (log W - delta t wt) y - sum_i W.
__global__ void __launch_bounds__(NT) k_lkhd_pois(int N, int Npad, const int *Yp,
        const int *Yi, const double *Yx, const double *W, const double *tits,
        const double *__restrict__ dtp, double wt, double *lkhds) {
    __shared__ double sh[4];
    const double dt = *dtp;
    const int j = blockIdx.x;
    const double *Wj = W + (size_t)j * Npad;
    double s = 0.0;
    for (int i = threadIdx.x; i < N; i += NT) s -= Wj[i];
    for (int l = Yp[j] + threadIdx.x; l < Yp[j + 1]; l += NT) {
        const int i = Yi[l];
        s += (log(Wj[i]) - dt * tits[i] * wt) * Yx[l];
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) lkhds[j] = s;
}

// ---------------------------------------------------------------- getLkhdj
// ggplvm.cu:744-770 from the factor diagonals, b_j, tBeta_j and the sums.
__global__ void __launch_bounds__(NT) k_lkhd_genes(int N, int P, const double *Psi,
        const double *DeltaChol, const double *bvec, const double *tBeta,
        const double *gq, const double *gs, const double *phi, double *lkhds) {
    __shared__ double sh[4];
    const int j = blockIdx.x;
    const double *Lj = Psi + (size_t)j * P * P;
    double ld = 0.0, bt = 0.0;
    for (int k = threadIdx.x; k < P; k += NT) {
        ld += log(DeltaChol[k + (size_t)k * P]) - log(Lj[k + (size_t)k * P]);
        bt += bvec[k + (size_t)j * P] * tBeta[k + (size_t)j * P];
    }
    ld = block_sum256(ld, sh);
    bt = block_sum256(bt, sh);
    if (threadIdx.x == 0) {
        const double ph = phi[j];
        double lk = -((double)N) / 2.0 * log(ph);
        lk -= gs[j] / 2.0;
        lk += ld;
        lk -= gq[j] / 2.0 / ph;
        lk += bt / 2.0 / ph;
        lkhds[j] = lk;
    }
}

// ---------------------------------------------------------------- collectLkhd
// ggplvm.cu:788-815; stage 1 over cells, stage 2 adds the gene sum.
__global__ void __launch_bounds__(NT) k_collect_cells(int N, const double *Oinvt,
        const double *tits, const double *Omega, const double *__restrict__ dtp, double wt,
        double igb, double igap1, double *part) {
    __shared__ double sh[4];
    const double wtdt = wt * *dtp;
    const int per = (N + RED_BLOCKS - 1) / RED_BLOCKS;
    const int lo = blockIdx.x * per, hi = min(N, lo + per);
    double s = 0.0;
    for (int i = lo + threadIdx.x; i < hi; i += NT)
        s += wtdt * Oinvt[i] * tits[i] + (igb / Omega[i] + igap1 * log(Omega[i]));
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(NT) k_collect_final(int J, const double *lkhds,
        const double *part, int npart, double *lkhd, int slot) {
    __shared__ double sh[4];
    double g = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < J; j += NT) g += lkhds[j];
    for (int b = threadIdx.x; b < npart; b += NT) c += part[b];
    g = block_sum256(g, sh);
    c = block_sum256(c, sh);
    if (threadIdx.x == 0) lkhd[slot] = g - c;
}

// ---------------------------------------------------------------- getGradLogDelta
// ggplvm.cu:210-253 per (gene, column) with the explicit inverse S_j.
// whitened: Sinv and tBeta are in the basis [covariates | Knm L^-T]; then the GP
// trace term tr(Psi^-1_GP K) = tr(Sinv_GP) and tb_GP^T K tb_GP = |tb_w,GP|^2, so
// each GP column carries its own diagonal entry and the column sums agree.
__global__ void __launch_bounds__(NT) k_grad_delta_gene(int P, int p, int M, int g0,
        const int *nh, const double *delta, const double *Ksym, const double *Sinv,
        const double *tBeta, const double *phi, double *gdt, int whitened) {
    __shared__ double sh[4];
    __shared__ double tb[2048];
    const int j = blockIdx.x;
    const double *S = Sinv + (size_t)j * P * P;
    const double *tbj = tBeta + (size_t)j * P;
    for (int k = threadIdx.x; k < P; k += NT) tb[k] = tbj[k];
    __syncthreads();
    const double ph = phi[j];
    // quadratic form tb_GP^T K tb_GP
    double qf = 0.0;
    for (int a = threadIdx.x; a < M; a += NT) {
        double t = 0.0;
        if (whitened) t = tb[g0 + a];
        else
            for (int b = 0; b < M; b++) t += Ksym[a + (size_t)b * M] * tb[g0 + b];
        qf += tb[g0 + a] * t;
    }
    qf = block_sum256(qf, sh);
    __shared__ double qfs;
    if (threadIdx.x == 0) qfs = qf;
    __syncthreads();
    for (int k = threadIdx.x; k < P; k += NT) {
        int l = 0;
        while (l < p - 1 && !(nh[l] <= k && k < nh[l + 1])) l++;
        const double dk = delta[l];
        double psi;
        if (k < g0 || whitened) {
            psi = S[k + (size_t)k * P];
        } else {
            const int kk = k - g0;
            psi = 0.0;
            for (int b = 0; b < M; b++) psi += S[(g0 + b) + (size_t)k * P] * Ksym[b + (size_t)kk * M];
        }
        double sum = 0.0;
        sum -= (dk - psi) / dk / dk / 2.0;
        if (k < g0) sum += tb[k] * tb[k] / dk / dk / ph / 2.0;
        else if (k == g0) sum += qfs / dk / dk / ph / 2.0;
        gdt[k + (size_t)j * P] = sum * dk;
    }
}

// collectGradLogDelta ggplvm.cu:256-281: group l of delta sums the per-gene
// terms of its columns and, for the GP group, the Titsias term over cells.
// Block (l, sp) sums genes [J sp / GDS, J (sp + 1) / GDS) and cells
// [N sp / GDS, N (sp + 1) / GDS) into part[l GDS + sp]; k_grad_delta_sum adds
// the GDS parts in order.
constexpr int GDS = 16;
__global__ void __launch_bounds__(NT) k_grad_delta_part(int N, int J, int P, int p,
        const int *nh, const double *gdt, const double *Oinvt, const double *tits,
        const double *delta, double wt, double *part) {
    __shared__ double sh[4];
    const int l = blockIdx.x, sp = blockIdx.y;
    const int c0 = nh[l], c1 = nh[l + 1];
    const int j0 = (int)((long long)J * sp / GDS), j1 = (int)((long long)J * (sp + 1) / GDS);
    double s = 0.0;
    for (int j = j0; j < j1; j++)
        for (int k = c0 + threadIdx.x; k < c1; k += NT) s += gdt[k + (size_t)j * P];
    if (l == p - 1) {
        const int i0 = (int)((long long)N * sp / GDS), i1 = (int)((long long)N * (sp + 1) / GDS);
        for (int i = i0 + threadIdx.x; i < i1; i += NT) s -= wt * Oinvt[i] * tits[i] * delta[p - 1];
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) part[l * GDS + sp] = s;
}
__global__ void k_grad_delta_sum(int p, const double *part, double *gd) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= p) return;
    double s = 0.0;
    for (int sp = 0; sp < GDS; sp++) s += part[l * GDS + sp];
    gd[l] = s;
}

// ---------------------------------------------------------------- tn / E
// ggplvm.cu:335-345 per-cell pieces: tn_ij = o nu - o z^T tBeta_j,
// E_i = sum_j tn^2 / W / phi / 2, TN_ij = tn_ij / phi_j.  The genes are cut
// into TN_SPLITS ranges (blockIdx.y) for occupancy; Epart holds the partial
// sums of E, added in order by k_tn_sum.
// A thread loads W, Nu, ZB and phi of TNU genes before using any of them
// (4 TNU loads in flight instead of 4); the sums still run in gene order.
constexpr int TN_SPLITS = GP_TN_SPLITS;
constexpr int TNU = 8;
__global__ void k_tn(int N, int Npad, int J, const double *__restrict__ W,
                     const double *__restrict__ Nu, const double *__restrict__ ZB,
                     const double *__restrict__ Omega, const double *__restrict__ phi,
                     double *__restrict__ TN, double *__restrict__ Epart) {
    const int i = blockIdx.x * NT + threadIdx.x;
    const int j0 = (int)((long long)J * blockIdx.y / TN_SPLITS);
    const int j1 = (int)((long long)J * (blockIdx.y + 1) / TN_SPLITS);
    if (i >= Npad) return;
    if (i >= N) {
        for (int j = j0; j < j1; j++) TN[i + (size_t)j * Npad] = 0.0;
        return;
    }
    const double om = Omega[i];
    double e = 0.0;
    int j = j0;
    for (; j + TNU <= j1; j += TNU) {
        double w[TNU], nu[TNU], zb[TNU], ph[TNU];
#pragma unroll
        for (int u = 0; u < TNU; u++) {
            const size_t a = i + (size_t)(j + u) * Npad;
            w[u] = W[a];
            nu[u] = Nu[a];
            zb[u] = ZB[a];
            ph[u] = phi[j + u];
        }
#pragma unroll
        for (int u = 0; u < TNU; u++) {
            const double o = w[u] / om;
            const double tn = o * nu[u] - o * zb[u];
            e += tn * tn / w[u] / ph[u] / 2.0;
            TN[i + (size_t)(j + u) * Npad] = tn / ph[u];
        }
    }
    for (; j < j1; j++) {
        const size_t a = i + (size_t)j * Npad;
        const double w = W[a], o = w / om;
        const double tn = o * Nu[a] - o * ZB[a];
        e += tn * tn / w / phi[j] / 2.0;
        TN[a] = tn / phi[j];
    }
    Epart[(size_t)blockIdx.y * Npad + i] = e;
}
__global__ void k_tn_sum(int N, int Npad, const double *Epart, double *E) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= N) return;
    double e = 0.0;
    for (int s = 0; s < TN_SPLITS; s++) e += Epart[(size_t)s * Npad + i];
    E[i] = e;
}

// FP32 form of the tn chain (mixed path): the same tn and E from an FP32
// product z^T tBeta (the Gram's FP32 whitened model matrix z32 = zw 2^s_c
// times the whitened beta-tilde scaled by 2^-s_c), TN written in FP32 for an
// FP32 C1a product.  tn = o nu - o zb and E are formed and summed in FP64.
__global__ void k_tn_prep32(int P, int M, int g0, int J, const double *__restrict__ tBw,
                            const double *__restrict__ zun, const double *__restrict__ tBeta,
                            float *__restrict__ B32, float *__restrict__ Bgp32) {
    const int idx = blockIdx.x * NT + threadIdx.x;
    if (idx < P * J) {
        const int c = idx % P;
        B32[idx] = (float)(tBw[idx] * zun[c]);
    }
    if (idx < M * J) {
        const int m = idx % M, j = idx / M;
        Bgp32[idx] = (float)tBeta[g0 + m + (size_t)j * P];
    }
}
__global__ void k_tn32(int N, int Npad, int J, const double *__restrict__ W,
                       const double *__restrict__ Nu, const float *__restrict__ ZB32,
                       const double *__restrict__ Omega, const double *__restrict__ phi,
                       float *__restrict__ TN32, double *__restrict__ Epart) {
    const int i = blockIdx.x * NT + threadIdx.x;
    const int j0 = (int)((long long)J * blockIdx.y / TN_SPLITS);
    const int j1 = (int)((long long)J * (blockIdx.y + 1) / TN_SPLITS);
    if (i >= Npad) return;
    if (i >= N) {
        for (int j = j0; j < j1; j++) TN32[i + (size_t)j * Npad] = 0.0f;
        return;
    }
    const double om = Omega[i];
    double e = 0.0;
    int j = j0;
    for (; j + TNU <= j1; j += TNU) {
        double w[TNU], nu[TNU], ph[TNU];
        float zb[TNU];
#pragma unroll
        for (int u = 0; u < TNU; u++) {
            const size_t a = i + (size_t)(j + u) * Npad;
            w[u] = W[a];
            nu[u] = Nu[a];
            zb[u] = ZB32[a];
            ph[u] = phi[j + u];
        }
#pragma unroll
        for (int u = 0; u < TNU; u++) {
            const double o = w[u] / om;
            const double tn = o * nu[u] - o * (double)zb[u];
            e += tn * tn / w[u] / ph[u] / 2.0;
            TN32[i + (size_t)(j + u) * Npad] = (float)(tn / ph[u]);
        }
    }
    for (; j < j1; j++) {
        const size_t a = i + (size_t)j * Npad;
        const double w = W[a], o = w / om;
        const double tn = o * Nu[a] - o * (double)ZB32[a];
        e += tn * tn / w / phi[j] / 2.0;
        TN32[a] = (float)(tn / phi[j]);
    }
    Epart[(size_t)blockIdx.y * Npad + i] = e;
}
__global__ void k_f32_to_f64(size_t n, const float *__restrict__ a, double *__restrict__ b) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx < n) b[idx] = (double)a[idx];
}

// ---------------------------------------------------------------- per-cell gradient// ---------------------------------------------------------------- per-cell gradient
// ggplvm.cu:346-375: gOmega, C1 and gX from U = sum_j o_ij Psi_j^{-1} z_i.
__global__ void k_cell_grad(int N, int Npad, int J, int M, int Q, int P, int g0, int ard,
        const double *Z, const double *U, const double *C1a, const double *A,
        const double *Omega, const double *Oinvt, const double *tits, const double *E,
        const double *X, const double *T, const double *rho, const double *__restrict__ dtp,
        double wt, double igb, double igap1, double *gO, double *gX, double *C1, double *ci1out) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= Npad) return;
    const double dt = *dtp;
    if (i >= N) {
        for (int k = 0; k < M; k++) C1[i + (size_t)k * Npad] = 0.0;
        return;
    }
    const double om = Omega[i], oii = Oinvt[i];
    double D = 0.0;
    for (int k = 0; k < P; k++) D += Z[i + (size_t)k * Npad] * U[i + (size_t)k * Npad];
    double sum = -((double)J) / om / 2.0 + D / om / 2.0 + E[i];
    sum += wt * dt * oii / om * tits[i];
    gO[i] = sum * om + igb / om - igap1;
    const double *Knm = Z + (size_t)g0 * Npad;
    double ci1 = 0.0, gx0 = 0.0;
    for (int k = 0; k < M; k++) {
        const size_t a = i + (size_t)k * Npad;
        double c = -U[i + (size_t)(g0 + k) * Npad] + C1a[a];
        c += wt * dt * oii * A[a];
        c *= Knm[a];
        C1[a] = c;
        ci1 += c;
        if (!ard) {
            const double t = (X[i] - T[k]) / 2.0;
            gx0 -= c * sin(t) * cos(t) / rho[0];
        }
    }
    if (!ard) gX[i] = gx0;
    // the other columns: gX[:, q] = (C1 T[:, q] - ci1 X[:, q]) / rho_q, the
    // product on rocBLAS and the rest in k_gx_finish
    ci1out[i] = ci1;
}
__global__ void k_gx_finish(int N, int q0, const double *ci1, const double *X, const double *rho,
                            double *gX) {
    const int i = blockIdx.x * NT + threadIdx.x;
    const int q = q0 + blockIdx.y;
    if (i >= N) return;
    const size_t a = i + (size_t)q * N;
    gX[a] = (gX[a] - ci1[i] * X[a]) / rho[q];
}

// ---------------------------------------------------------------- C2
__global__ void k_scale_rows(int N, int Npad, int M, const double *A, const double *o,
                             double *AO) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)Npad * M) return;
    const int i = (int)(idx % Npad);
    AO[idx] = (i < N) ? A[idx] * o[i] : 0.0;
}
__global__ void k_scale_cols(int P, int J, const double *B, const double *phi, double *Bo) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)P * J) return;
    Bo[idx] = B[idx] / phi[idx / P];
}
// Sum over genes of the GP block of Psi_j^{-1}.
__global__ void k_sum_sinv(int J, int P, int M, int g0, const double *Sinv, double *S2) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)M * M) return;
    const int a = (int)(idx % M), b = (int)(idx / M);
    const double *src = Sinv + (g0 + a) + (size_t)(g0 + b) * P;
    const size_t PP = (size_t)P * P;
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= J; j += 8) {          // 8 loads in flight, same order
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = src[(size_t)(j + u) * PP];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; j < J; j++) s += src[(size_t)j * PP];
    S2[idx] = s;
}
// ggplvm.cu:57-100 getC2
__global__ void k_c2(int M, int J, double wt, const double *__restrict__ dtp, const double *G2,
                     const double *B2, const double *S2, const double *Kinv, const double *K,
                     double *C2) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)M * M) return;
    const double dt = *dtp;
    const int j = (int)(idx % M), k = (int)(idx / M);
    if (j > k) return;
    const size_t a = j + (size_t)k * M;
    double c = -wt * dt * G2[a] / 2.0;
    c -= B2[a] / dt / 2.0;
    c -= S2[a] / dt / 2.0;
    c += ((double)J) * Kinv[a] / 2.0;
    c *= K[a];
    C2[a] = c;
    if (j < k) C2[k + (size_t)j * M] = c;
}

// ---------------------------------------------------------------- getGradTau / rho
// ggplvm.cu:103-145 and 165-205: one block per (inducing point m, dimension q).
__global__ void __launch_bounds__(NT) k_grad_tau_rho(int N, int Npad, int M, int Q, int ard,
        const double *C1, const double *C2, const double *X, const double *T,
        const double *rho, double *gT, double *grt) {
    __shared__ double sh[4];
    const int m = blockIdx.x, q = blockIdx.y;   // one block per (inducing point, dimension)
    const double *c1 = C1 + (size_t)m * Npad;
    {
        const double tm = T[m + (size_t)q * M];
        const double rq = rho[q];
        double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, r1 = 0.0, r2 = 0.0;
        const bool per = (q == 0 && !ard);
        for (int i = threadIdx.x; i < N; i += NT) {
            const double c = c1[i], x = X[i + (size_t)q * N];
            if (per) {
                const double t = (x - tm) / 2.0;
                s1 += c * sin(t) * cos(t);
                const double st = sin((x - tm) / 2.0);
                r1 += c * st * st / rq / rq;
            } else {
                s1 += c * x;
                s2 += c;
                const double d = x - tm;
                r1 += c * d * d / rq / rq / 2.0;
            }
        }
        for (int i = threadIdx.x; i < M; i += NT) {
            const double c = C2[m + (size_t)i * M], cc = C2[i + (size_t)m * M];
            const double ti = T[i + (size_t)q * M];
            if (per) {
                const double t = (ti - tm) / 2.0;
                s3 += c * sin(t) * cos(t);
                const double st = sin((ti - tm) / 2.0);
                r2 += cc * st * st / rq / rq;
            } else {
                s3 += c * ti;
                s4 += c;
                const double d = ti - tm;
                r2 += cc * d * d / rq / rq / 2.0;
            }
        }
        s1 = block_sum256(s1, sh);
        s2 = block_sum256(s2, sh);
        s3 = block_sum256(s3, sh);
        s4 = block_sum256(s4, sh);
        r1 = block_sum256(r1, sh);
        r2 = block_sum256(r2, sh);
        if (threadIdx.x == 0) {
            double g;
            if (per) g = s1 / rq + 2.0 * s3 / rq;
            else g = (s1 - s2 * tm) / rq + 2.0 * (s3 - s4 * tm) / rq;
            gT[m + (size_t)q * M] = g;
            grt[q + (size_t)m * Q] = (r1 + r2) * rq;
        }
    }
}
__global__ void k_collect_rho(int M, int Q, const double *grt, double *gr) {
    const int q = threadIdx.x;
    if (q >= Q) return;
    double s = 0.0;
    for (int k = 0; k < M; k++) s += grt[q + (size_t)k * Q];
    gr[q] = s;
}

// ---------------------------------------------------------------- L-BFGS
// util_cuda.cu:187-232 and ggplvm.cu:865-934 / 818-859.
__global__ void __launch_bounds__(NT) k_dot2(size_t n, const double *a, const double *b,
        const double *c, const double *d, double *part) {
    __shared__ double sh[4];
    const size_t per = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    const size_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    double s = 0.0, t = 0.0;
    for (size_t k = lo + threadIdx.x; k < hi; k += NT) {
        s += a[k] * b[k];
        t += c[k] * d[k];
    }
    s = block_sum256(s, sh);
    t = block_sum256(t, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s; part[2 * blockIdx.x + 1] = t; }
}
__global__ void __launch_bounds__(NT) k_dot2_final(const double *part, double *out) {
    __shared__ double sh[4];
    double s = 0.0, t = 0.0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += NT) { s += part[2 * b]; t += part[2 * b + 1]; }
    s = block_sum256(s, sh);
    t = block_sum256(t, sh);
    if (threadIdx.x == 0) { out[0] = s; out[1] = t; }
}
// mode 0: q -= (num/den) * y, store alpha in *alpha
// mode 1: z  = (num/den) * q
// mode 2: z += s * (alpha - num/den)
__global__ void k_lb_axpy(size_t n, int mode, const double *nd, double *alpha,
                          const double *x, double *y) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    const double r = nd[0] / nd[1];
    if (k == 0 && mode == 0) *alpha = r;
    if (k >= n) return;
    if (mode == 0) y[k] -= r * x[k];
    else if (mode == 1) y[k] = r * x[k];
    else y[k] += x[k] * (*alpha - r);
}
__global__ void k_lb_first(size_t n, const double *g, double *ss, double *g0) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    double v = -g[k] / 100000.0;
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    ss[k] = v;
    g0[k] = g[k];
}
__global__ void k_lb_diff(size_t n, const double *g, double *g0, double *xx) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    xx[k] = g[k] - g0[k];
    g0[k] = g[k];
}
__global__ void k_lb_clip(size_t n, const double *z, double *ss) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    double v = z[k];
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    ss[k] = -v;
}
// Parameter update (mode 0 = LBFGS with backups, mode 1 = Downdate from backups).
__global__ void k_lb_update(int N, int p, int M, int Q, int mode, double r, int rhoflag,
                            const double *ss, double *Omega, double *delta, double *X,
                            double *T, double *rho, double *SSslot, size_t tot) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    const size_t NQ = (size_t)N * Q, MQ = (size_t)M * Q;
    if (k < tot) SSslot[k] = r * ss[k];
    if (k < (size_t)N) {
        double v = (mode == 0) ? exp(log(Omega[k]) - ss[k]) : exp(log(Omega[k + N]) - r * ss[k]);
        if (mode == 0) Omega[k + N] = Omega[k];
        if (v > 1000.) v = 1000.;
        if (v < .001) v = .001;
        Omega[k] = v;
    }
    if (k < (size_t)p) {
        if (mode == 0) { delta[k + p] = delta[k]; delta[k] = exp(log(delta[k]) - ss[N + k]); }
        else delta[k] = exp(log(delta[k + p]) - r * ss[N + k]);
    }
    if (k < NQ) {
        if (mode == 0) { X[k + NQ] = X[k]; X[k] = X[k] - ss[N + p + k]; }
        else X[k] = X[k + NQ] - r * ss[N + p + k];
    }
    if (k < MQ) {
        if (mode == 0) { T[k + MQ] = T[k]; T[k] = T[k] - ss[N + p + NQ + k]; }
        else T[k] = T[k + MQ] - r * ss[N + p + NQ + k];
    }
}
// rho_0 update and mscale_d (util_cuda.cu:546-555) of rho_1..Q-1.
__global__ void k_lb_rho(int N, int p, int M, int Q, int mode, double r, int rhoflag,
                         const double *ss, double *rho) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const size_t off = (size_t)N + p + (size_t)N * Q + (size_t)M * Q;
    if (mode == 0) {
        rho[Q] = rho[0];
        rho[0] = exp(log(rho[0]) - ss[off] * (double)rhoflag);
    } else {
        rho[0] = exp(log(rho[Q]) - r * ss[off] * (double)rhoflag);
    }
    double mh = 0.0;
    for (int i = 1; i < Q; i++) mh += rho[i];
    mh /= (double)(Q - 1);
    for (int i = 1; i < Q; i++) rho[i] *= 10.0 / mh;
}

__global__ void k_add_zeta(size_t n, int P, const double *tB, const double *zeta, double *B) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    B[k] = tB[k] + zeta[k % P];
}

__global__ void k_fill(size_t n, double *p, double v) {
    const size_t k = (size_t)blockIdx.x * NT + threadIdx.x;
    if (k < n) p[k] = v;
}

}  // namespace

// ====================================================================== wrappers

#define LAUNCH_CHECK() GP_HIP(hipGetLastError())

namespace {
__global__ void k_fill_int(size_t n, int *p, int v) {
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i < n) p[i] = v;
}
}  // namespace

int gpk_fill_int(gp_ctx *c, int *p, size_t n, int v) {
    k_fill_int<<<nblk(n), NT, 0, c->stream>>>(n, p, v);
    LAUNCH_CHECK();
    return 0;
}

int gpk_fill(gp_ctx *c, double *p, size_t n, double v) {
    if (!n) return 0;
    k_fill<<<nblk(n), NT, 0, c->stream>>>(n, p, v);
    LAUNCH_CHECK();
    return 0;
}

int gpk_kuf(gp_ctx *c) {
    dim3 g(nblk(c->Npad), (c->M + 7) / 8);
    const size_t shm = (size_t)c->Q * NT * sizeof(double);      // Q <= 64: at most 128 KB
    if (shm > 65536 && !c->capturing)    // set before any recording (init_fit runs first)
        GP_HIP(hipFuncSetAttribute((const void *)k_kuf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    k_kuf<<<g, NT, shm, c->stream>>>(c->N, c->Npad, c->M, c->Q, c->X, c->T, c->rho,
                                   c->ard_only, c->Z + (size_t)c->g0 * c->Npad);
    LAUNCH_CHECK();
    k_kmm<<<nblk((size_t)c->M * c->M), NT, 0, c->stream>>>(c->M, c->Q, c->T, c->rho,
                                                           c->ard_only, c->K, c->Ksym, c->Kchol);
    LAUNCH_CHECK();
    return 0;
}

int gpk_expand_delta(gp_ctx *c) {
    k_expand_delta<<<nblk((size_t)c->P * c->P), NT, 0, c->stream>>>(
        c->P, c->p, c->M, c->g0, c->d_nh, c->delta, c->K, c->Kchol, c->DeltaInv, c->DeltaChol,
        c->DeltaW, c->DeltaWChol);
    LAUNCH_CHECK();
    return 0;
}

int gpk_oinvt(gp_ctx *c) {
    k_oinvt<<<nblk(c->N), NT, 0, c->stream>>>(c->N, c->Npad, c->J, c->W, c->Omega, c->Oinvt);
    LAUNCH_CHECK();
    return 0;
}

int gpk_titsias(gp_ctx *c) {
    const double one = 1.0, zero = 0.0;
    const double *Knm = c->Z + (size_t)c->g0 * c->Npad;
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad,
                          c->M, c->M, &one, Knm, c->Npad, c->Kinv, c->M, &zero, c->A, c->Npad));
    k_titsias<<<nblk(c->N), NT, 0, c->stream>>>(c->N, c->Npad, c->M, Knm, c->A, c->titsias);
    LAUNCH_CHECK();
    return 0;
}

int gpk_w_nu(gp_ctx *c) {
    // ZB = Z * Beta (eta for every cell and gene)
    const double one = 1.0, zero = 0.0;
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad,
                          c->J, c->P, &one, c->Z, c->Npad, c->Beta, c->P, &zero, c->ZB, c->Npad));
    double dt;
    GP_HIP(hipMemcpyAsync(&dt, c->delta + c->p - 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    dim3 g(nblk(c->Npad), c->J);
    k_w_nu<<<g, NT, 0, c->stream>>>(c->N, c->Npad, c->ZB, c->ls, c->titsias, dt, 0.1, c->W, c->Nu);
    LAUNCH_CHECK();
    k_nu_sparse<<<c->J, NT, 0, c->stream>>>(c->Npad, c->Yp, c->Yi, c->Yx, c->ZB, c->ls,
                                            c->titsias, dt, 0.1, c->Nu);
    LAUNCH_CHECK();
    return 0;
}

int gpk_r_sums(gp_ctx *c, bool w32) {
    if (w32 && (!c->w32 || !c->wsc || !c->wun)) { gp_set_error("FP32 weight buffers not allocated"); return 1; }
    k_r_sums_w32<<<c->J, NT, 0, c->stream>>>(c->N, c->Npad, c->W, c->Nu, c->Omega, c->R, c->gq, c->gs,
                                            w32 ? c->w32 : nullptr, w32 ? c->wsc : nullptr,
                                            w32 ? c->wun : nullptr);
    LAUNCH_CHECK();
    if (w32) c->w32_valid = 1;
    return 0;
}

int gpk_r_and_sums(gp_ctx *c, const double *Zsrc) {
    k_r_sums<<<c->J, NT, 0, c->stream>>>(c->N, c->Npad, c->W, c->Nu, c->Omega, c->R, c->gq, c->gs);
    LAUNCH_CHECK();
    // b_j = Z^T R_j  (util_cuda.cu:462-487 getZtOinvjnj for every gene)
    GP_TRY(gpk_gemm_tn_splitk(c, c->P, c->J, c->Npad, Zsrc, c->Npad, c->R, c->Npad, c->bvec, c->P));
    return 0;
}

namespace {
__global__ void k_sum_splits(size_t mn, int M, int ldc, int nsplit, const double *__restrict__ Wp,
                             double *__restrict__ C) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= mn) return;
    double s = 0.0;
    for (int q = 0; q < nsplit; q++) s += Wp[(size_t)q * mn + idx];
    C[(idx % M) + (idx / M) * (size_t)ldc] = s;
}
}  // namespace

int gpk_gemm_tn_splitk(gp_ctx *c, int M, int N, int K, const double *A, int lda, const double *B,
                       int ldb, double *C, int ldc) {
    const double one = 1.0, zero = 0.0;
    const long long tiles = (long long)((M + 127) / 128) * ((N + 127) / 128);
    int want = (int)((1024 + tiles - 1) / tiles);           // aim at ~1024 output tiles
    want = std::max(1, std::min(want, K / 512));
    if (want <= 1) {
        GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, M, N, K,
                              &one, A, lda, B, ldb, &zero, C, ldc));
        return 0;
    }
    const int Kc = ((K + want - 1) / want + 63) / 64 * 64;
    const int nfull = K / Kc, rem = K - nfull * Kc;
    const int nsplit = nfull + (rem > 0);
    const size_t mn = (size_t)M * N;
    if (c->skw_len < (size_t)nsplit * mn) {
        if (c->skw) (void)hipFree(c->skw);
        c->skw = nullptr;
        c->skw_len = (size_t)nsplit * mn;
        GP_HIP(hipMalloc((void **)&c->skw, c->skw_len * sizeof(double)));
    }
    GP_BLAS(rocblas_dgemm_strided_batched(c->blas, rocblas_operation_transpose,
                                          rocblas_operation_none, M, N, Kc, &one, A, lda, Kc, B,
                                          ldb, Kc, &zero, c->skw, M, (rocblas_stride)mn, nfull));
    if (rem > 0)
        GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, M, N,
                              rem, &one, A + (size_t)nfull * Kc, lda, B + (size_t)nfull * Kc, ldb,
                              &zero, c->skw + (size_t)nfull * mn, M));
    k_sum_splits<<<nblk(mn), NT, 0, c->stream>>>(mn, M, ldc, nsplit, c->skw, C);
    LAUNCH_CHECK();
    return 0;
}

int gpk_init_y0(gp_ctx *c) {
    dim3 g(nblk(c->Npad), c->J);
    k_y0_dense<<<g, NT, 0, c->stream>>>(c->N, c->Npad, c->ls, c->R);
    LAUNCH_CHECK();
    k_y0_sparse<<<c->J, NT, 0, c->stream>>>(c->Npad, c->Yp, c->Yi, c->Yx, c->ls, c->R);
    LAUNCH_CHECK();
    GP_TRY(gpk_gemm_tn_splitk(c, c->P, c->J, c->Npad, c->Z, c->Npad, c->R, c->Npad, c->bvec, c->P));
    return 0;
}

int gpk_beta_phi(gp_ctx *c, int jtr) { return gpk_beta_phi_in(c, jtr, c->Beta); }

int gpk_beta_phi_in(gp_ctx *c, int jtr, double *Beta) {
    // tBphi holds Psi_j^{-1} b_j on entry (solved by the caller)
    k_beta_phi<<<c->J, NT, 0, c->stream>>>(c->N, c->P, jtr, c->Psi, c->tBphi, c->zeta, c->gq,
                                           Beta, c->phi, c->pwork);
    LAUNCH_CHECK();
    return 0;
}

// Bw = T^-1 B with T = blockdiag(I, L^-T) (Kmm = L L^T): covariate rows
// copied, GP rows = L^T B_GP (L is the lower triangle of Kchol).
int gpk_beta_whiten(gp_ctx *c, const double *B, double *Bw) {
    const double one = 1.0;
    GP_HIP(hipMemcpy2DAsync(Bw, c->P * sizeof(double), B, c->P * sizeof(double),
                            c->g0 * sizeof(double), c->J, hipMemcpyDeviceToDevice, c->stream));
    GP_BLAS(rocblas_dtrmm(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, c->M, c->J, &one, c->Kchol, c->M, B + c->g0, c->P,
                          Bw + c->g0, c->P));
    return 0;
}

// b_w = T^T b with T = blockdiag(I, L^-T): GP rows L^-1 b_GP (a triangular
// solve with the Kmm factor), covariate rows unchanged; in place.
int gpk_b_whiten(gp_ctx *c) {
    const double one = 1.0;
    GP_BLAS(rocblas_dtrsm(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                          rocblas_diagonal_non_unit, c->M, c->J, &one, c->Kchol, c->M, c->bvec + c->g0,
                          c->P));
    return 0;
}

// B = T Bw: covariate rows copied, GP rows = L^-T Bw_GP.
int gpk_beta_unwhiten(gp_ctx *c, const double *Bw, double *B) {
    const double one = 1.0, zero = 0.0;
    GP_HIP(hipMemcpy2DAsync(B, c->P * sizeof(double), Bw, c->P * sizeof(double),
                            c->g0 * sizeof(double), c->J, hipMemcpyDeviceToDevice, c->stream));
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->M, c->J, c->M,
                          &one, c->Linv, c->M, Bw + c->g0, c->P, &zero, B + c->g0, c->P));
    return 0;
}

int gpk_lkhd_pois(gp_ctx *c) {
    k_lkhd_pois<<<c->J, NT, 0, c->stream>>>(c->N, c->Npad, c->Yp, c->Yi, c->Yx, c->W, c->titsias,
                                            c->delta + c->p - 1, 0.1, c->lkhds);
    LAUNCH_CHECK();
    return 0;
}

int gpk_lkhd_genes_with(gp_ctx *c, const double *tb, const double *Dchol) {
    k_lkhd_genes<<<c->J, NT, 0, c->stream>>>(c->N, c->P, c->Psi, Dchol, c->bvec, tb,
                                             c->gq, c->gs, c->phi, c->lkhds);
    LAUNCH_CHECK();
    return 0;
}

int gpk_collect(gp_ctx *c, int Nterm, int slot, double *obj) {
    int npart = 0;
    if (Nterm > 0) {
        k_collect_cells<<<RED_BLOCKS, NT, 0, c->stream>>>(Nterm, c->Oinvt, c->titsias, c->Omega,
                                                          c->delta + c->p - 1, 0.1, 100.0, 100.0,
                                                          c->red);
        LAUNCH_CHECK();
        npart = RED_BLOCKS;
    }
    k_collect_final<<<1, NT, 0, c->stream>>>(c->J, c->lkhds, c->red, npart, c->lkhd, slot);
    LAUNCH_CHECK();
    if (obj && !c->capturing) {
        GP_HIP(hipMemcpyAsync(obj, c->lkhd + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        GP_HIP(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int gpk_grad_delta(gp_ctx *c, int whitened) {
    if (c->P > 2048) { gp_set_error("P > 2048 unsupported"); return 1; }
    k_grad_delta_gene<<<c->J, NT, 0, c->stream>>>(c->P, c->p, c->M, c->g0, c->d_nh, c->delta,
                                                  c->Ksym, c->Sinv, whitened ? c->tBw : c->tBeta,
                                                  c->phi, c->gradDeltaTmp, whitened);
    LAUNCH_CHECK();
    if ((size_t)c->p * GDS > 4096) { gp_set_error("too many delta groups"); return 1; }
    // c->red is free here (the likelihood collection and the L-BFGS dots use
    // it within their own calls)
    k_grad_delta_part<<<dim3(c->p, GDS), NT, 0, c->stream>>>(c->N, c->J, c->P, c->p, c->d_nh,
                                                             c->gradDeltaTmp, c->Oinvt, c->titsias,
                                                             c->delta, 0.1, c->red);
    LAUNCH_CHECK();
    k_grad_delta_sum<<<(c->p + 63) / 64, 64, 0, c->stream>>>(c->p, c->red, c->grad + c->N);
    LAUNCH_CHECK();
    return 0;
}

// the tn chain with FP32 products (mixed path, whitened factor, the Gram's
// z32 current): ZB32 and TN32 occupy the FP64 ZB buffer's two halves
static int gpk_tn32(gp_ctx *c) {
    const size_t nj = (size_t)c->Npad * c->J;
    if (!c->tn_b32) {
        GP_HIP(hipMalloc((void **)&c->tn_b32, (size_t)(c->P + c->M) * c->J * sizeof(float)));
        GP_HIP(hipMalloc((void **)&c->tn_c1a32, (size_t)c->Npad * c->M * sizeof(float)));
    }
    float *B32 = c->tn_b32, *Bgp32 = c->tn_b32 + (size_t)c->P * c->J;
    float *ZB32 = reinterpret_cast<float *>(c->ZB), *TN32 = ZB32 + nj;
    k_tn_prep32<<<nblk((size_t)c->P * c->J), NT, 0, c->stream>>>(c->P, c->M, c->g0, c->J, c->tBw, c->zun,
                                                                c->tBeta, B32, Bgp32);
    LAUNCH_CHECK();
    const float one = 1.0f, zero = 0.0f;
    GP_BLAS(rocblas_sgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad, c->J, c->P,
                          &one, c->z32, c->Npad, B32, c->P, &zero, ZB32, c->Npad));
    k_tn32<<<dim3(nblk(c->Npad), TN_SPLITS), NT, 0, c->stream>>>(c->N, c->Npad, c->J, c->W, c->Nu, ZB32,
                                                                  c->Omega, c->phi, TN32, c->Epart);
    LAUNCH_CHECK();
    k_tn_sum<<<nblk(c->N), NT, 0, c->stream>>>(c->N, c->Npad, c->Epart, c->Ecell);
    LAUNCH_CHECK();
    // C1a = (TN / phi) tBeta_GP^T
    GP_BLAS(rocblas_sgemm(c->blas, rocblas_operation_none, rocblas_operation_transpose, c->Npad, c->M, c->J,
                          &one, TN32, c->Npad, Bgp32, c->M, &zero, c->tn_c1a32, c->Npad));
    const size_t nc = (size_t)c->Npad * c->M;
    k_f32_to_f64<<<nblk(nc), NT, 0, c->stream>>>(nc, c->tn_c1a32, c->C1a);
    LAUNCH_CHECK();
    return 0;
}

int gpk_tn(gp_ctx *c) {
    if (c->tn_f32 && c->precision == GP_PREC_MIXED && c->psi_whitened && c->zsplit_valid && c->z32)
        return gpk_tn32(c);
    // ZB = Z * tBeta
    const double one = 1.0, zero = 0.0;
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad,
                          c->J, c->P, &one, c->Z, c->Npad, c->tBeta, c->P, &zero, c->ZB, c->Npad));
    k_tn<<<dim3(nblk(c->Npad), TN_SPLITS), NT, 0, c->stream>>>(c->N, c->Npad, c->J, c->W, c->Nu,
                                                                c->ZB, c->Omega, c->phi, c->R,
                                                                c->Epart);
    LAUNCH_CHECK();
    k_tn_sum<<<nblk(c->N), NT, 0, c->stream>>>(c->N, c->Npad, c->Epart, c->Ecell);
    LAUNCH_CHECK();
    // C1a = (TN / phi) * tBeta_GP^T
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_transpose, c->Npad,
                          c->M, c->J, &one, c->R, c->Npad, c->tBeta + c->g0, c->P, &zero, c->C1a,
                          c->Npad));
    return 0;
}

int gpk_cell_grad(gp_ctx *c) {
    k_cell_grad<<<nblk(c->Npad), NT, 0, c->stream>>>(
        c->N, c->Npad, c->J, c->M, c->Q, c->P, c->g0, c->ard_only, c->Z, c->U, c->C1a, c->A,
        c->Omega, c->Oinvt, c->titsias, c->Ecell, c->X, c->T, c->rho, c->delta + c->p - 1, 0.1,
        100.0, 100.0,
        c->grad, c->grad + c->N + c->p, c->C1, c->ci1);
    LAUNCH_CHECK();
    const int q0 = c->ard_only ? 0 : 1;
    if (c->Q > q0) {
        const double one = 1.0, zero = 0.0;
        double *gX = c->grad + c->N + c->p;
        GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->N,
                              c->Q - q0, c->M, &one, c->C1, c->Npad, c->T + (size_t)q0 * c->M,
                              c->M, &zero, gX + (size_t)q0 * c->N, c->N));
        k_gx_finish<<<dim3(nblk(c->N), c->Q - q0), NT, 0, c->stream>>>(c->N, q0, c->ci1, c->X,
                                                                        c->rho, gX);
        LAUNCH_CHECK();
    }
    return 0;
}

int gpk_c2(gp_ctx *c, int whitened) {
    const double one = 1.0, zero = 0.0;
    k_sum_sinv<<<nblk((size_t)c->M * c->M), NT, 0, c->stream>>>(c->J, c->P, c->M, c->g0, c->Sinv,
                                                                 c->S2);
    LAUNCH_CHECK();
    if (whitened) {
        // sum_j Psi_j^{-1}[GP,GP] = L^-T (sum_j Sw_j[GP,GP]) L^-1  (G2 as scratch)
        GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->M, c->M,
                              c->M, &one, c->S2, c->M, c->Linv, c->M, &zero, c->G2, c->M));
        GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->M,
                              c->M, c->M, &one, c->Linv, c->M, c->G2, c->M, &zero, c->S2, c->M));
    }
    k_scale_rows<<<nblk((size_t)c->Npad * c->M), NT, 0, c->stream>>>(c->N, c->Npad, c->M, c->A,
                                                                      c->Oinvt, c->AO);
    LAUNCH_CHECK();
    GP_TRY(gpk_gemm_tn_splitk(c, c->M, c->M, c->Npad, c->A, c->Npad, c->AO, c->Npad, c->G2, c->M));
    k_scale_cols<<<nblk((size_t)c->P * c->J), NT, 0, c->stream>>>(c->P, c->J, c->tBeta, c->phi,
                                                                   c->tBphi);
    LAUNCH_CHECK();
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_transpose, c->M, c->M,
                          c->J, &one, c->tBphi + c->g0, c->P, c->tBeta + c->g0, c->P, &zero, c->B2,
                          c->M));
    k_c2<<<nblk((size_t)c->M * c->M), NT, 0, c->stream>>>(c->M, c->J, 0.1, c->delta + c->p - 1, c->G2, c->B2, c->S2,
                                                          c->Kinv, c->K, c->C2);
    LAUNCH_CHECK();
    return 0;
}

int gpk_grad_tau_rho(gp_ctx *c) {
    double *gT = c->grad + c->N + c->p + (size_t)c->N * c->Q;
    double *gr = gT + (size_t)c->M * c->Q;
    k_grad_tau_rho<<<dim3(c->M, c->Q), NT, 0, c->stream>>>(c->N, c->Npad, c->M, c->Q, c->ard_only, c->C1,
                                                c->C2, c->X, c->T, c->rho, gT, c->gradLogRhoTmp);
    LAUNCH_CHECK();
    k_collect_rho<<<1, 64, 0, c->stream>>>(c->M, c->Q, c->gradLogRhoTmp, gr);
    LAUNCH_CHECK();
    return 0;
}

static int lb_dots(gp_ctx *c, size_t n, const double *a, const double *b, const double *x,
                   const double *y, double *out) {
    k_dot2<<<RED_BLOCKS, NT, 0, c->stream>>>(n, a, b, x, y, c->red);
    LAUNCH_CHECK();
    k_dot2_final<<<1, NT, 0, c->stream>>>(c->red, out);
    LAUNCH_CHECK();
    return 0;
}

int gpk_lbfgs(gp_ctx *c, int itr, int rhoflag) {
    const size_t tot = (size_t)c->N + c->p + (size_t)c->N * c->Q + (size_t)c->M * c->Q + rhoflag;
    const int pt = (itr - 1) % GP_HIST, ptp1 = itr % GP_HIST;
    const unsigned nb = nblk(tot);
    if (itr == 0) {
        k_lb_first<<<nb, NT, 0, c->stream>>>(tot, c->grad, c->ss, c->grad0);
        LAUNCH_CHECK();
    } else {
        k_lb_diff<<<nb, NT, 0, c->stream>>>(tot, c->grad, c->grad0, c->XX + (size_t)pt * tot);
        LAUNCH_CHECK();
        double *q = c->lbq, *z = c->ss;
        GP_HIP(hipMemcpyAsync(q, c->grad, tot * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        const int lo = (itr > GP_HIST ? pt - GP_HIST + 1 : 0);
        double *nd = c->scal, *alpha = c->scal + 4;
        for (int i = pt; i >= lo; i--) {
            const int ii = (i + GP_HIST) % GP_HIST;
            const double *S = c->SS + (size_t)ii * tot, *Y = c->XX + (size_t)ii * tot;
            GP_TRY(lb_dots(c, tot, S, q, S, Y, nd));
            k_lb_axpy<<<nb, NT, 0, c->stream>>>(tot, 0, nd, alpha + ii, Y, q);
            LAUNCH_CHECK();
        }
        {
            const double *S = c->SS + (size_t)pt * tot, *Y = c->XX + (size_t)pt * tot;
            GP_TRY(lb_dots(c, tot, S, Y, Y, Y, nd));
            k_lb_axpy<<<nb, NT, 0, c->stream>>>(tot, 1, nd, nullptr, q, z);
            LAUNCH_CHECK();
        }
        for (int i = lo; i <= pt; i++) {
            const int ii = (i + GP_HIST) % GP_HIST;
            const double *S = c->SS + (size_t)ii * tot, *Y = c->XX + (size_t)ii * tot;
            GP_TRY(lb_dots(c, tot, Y, z, S, Y, nd));
            k_lb_axpy<<<nb, NT, 0, c->stream>>>(tot, 2, nd, alpha + ii, S, z);
            LAUNCH_CHECK();
        }
        k_lb_clip<<<nb, NT, 0, c->stream>>>(tot, z, c->ss);
        LAUNCH_CHECK();
    }
    const size_t nmax = std::max(tot, (size_t)c->N * c->Q);
    k_lb_update<<<nblk(nmax), NT, 0, c->stream>>>(c->N, c->p, c->M, c->Q, 0, 1.0, rhoflag, c->ss,
                                                   c->Omega, c->delta, c->X, c->T, c->rho,
                                                   c->SS + (size_t)ptp1 * tot, tot);
    LAUNCH_CHECK();
    k_lb_rho<<<1, 64, 0, c->stream>>>(c->N, c->p, c->M, c->Q, 0, 1.0, rhoflag, c->ss, c->rho);
    LAUNCH_CHECK();
    return 0;
}

int gpk_downdate(gp_ctx *c, int itr, double r, int rhoflag) {
    const size_t tot = (size_t)c->N + c->p + (size_t)c->N * c->Q + (size_t)c->M * c->Q + rhoflag;
    const int ptp1 = itr % GP_HIST;
    const size_t nmax = std::max(tot, (size_t)c->N * c->Q);
    k_lb_update<<<nblk(nmax), NT, 0, c->stream>>>(c->N, c->p, c->M, c->Q, 1, r, rhoflag, c->ss,
                                                   c->Omega, c->delta, c->X, c->T, c->rho,
                                                   c->SS + (size_t)ptp1 * tot, tot);
    LAUNCH_CHECK();
    k_lb_rho<<<1, 64, 0, c->stream>>>(c->N, c->p, c->M, c->Q, 1, r, rhoflag, c->ss, c->rho);
    LAUNCH_CHECK();
    return 0;
}

int gpk_beta_from_tilde(gp_ctx *c) {
    const size_t n = (size_t)c->P * c->J;
    k_add_zeta<<<nblk(n), NT, 0, c->stream>>>(n, c->P, c->tBeta, c->zeta, c->Beta);
    LAUNCH_CHECK();
    return 0;
}

// AT = A^T for one M x M column-major matrix (M^2 threads; 512 KB at M = 256)
__global__ void k_transpose_sq(int M, const double *__restrict__ A, double *__restrict__ AT) {
    const size_t idx = (size_t)blockIdx.x * NT + threadIdx.x;
    if (idx >= (size_t)M * M) return;
    const int i = (int)(idx % M), j = (int)(idx / M);
    AT[(size_t)i * M + j] = A[idx];
}

// Whitened basis for the mixed path: L^{-1} of the Kmm factor and the model
// matrix with its GP block replaced by Knm L^-T (rows of norm <= 1).
int gpk_whiten(gp_ctx *c) {
    const double one = 1.0, zero = 0.0;
    // c->Linv = L^-1 of the current Kmm factor (set with Kinv by the kernel-matrix step)
    GP_HIP(hipMemcpyAsync(c->Zw, c->Z, (size_t)c->Npad * c->g0 * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    const double *Knm = c->Z + (size_t)c->g0 * c->Npad;
    // Knm L^-T as a plain (NN) product on an explicit transpose of L^-1: rocBLAS's
    // transposed-B kernel for this tall shape (Npad x M x M) ran ~8x slower alone
    // (2.07 against 0.24 ms at Npad = 100096, M = 256).
    const size_t mm = (size_t)c->M * c->M;
    k_transpose_sq<<<nblk(mm), NT, 0, c->stream>>>(c->M, c->Linv, c->LinvT);
    LAUNCH_CHECK();
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad,
                          c->M, c->M, &one, Knm, c->Npad, c->LinvT, c->M, &zero,
                          c->Zw + (size_t)c->g0 * c->Npad, c->Npad));
    return 0;
}

// Back to the reference basis: tBeta = T^-T tBw and U = Uw T^-1 with
// T = blockdiag(I, L):  GP rows of tBeta = L^-T tBw_GP, GP columns of U = Uw_GP L^-1.
int gpk_unwhiten_beta(gp_ctx *c) {
    const double one = 1.0, zero = 0.0;
    GP_HIP(hipMemcpy2DAsync(c->tBeta, c->P * sizeof(double), c->tBw, c->P * sizeof(double),
                            c->g0 * sizeof(double), c->J, hipMemcpyDeviceToDevice, c->stream));
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, c->M, c->J,
                          c->M, &one, c->Linv, c->M, c->tBw + c->g0, c->P, &zero,
                          c->tBeta + c->g0, c->P));
    return 0;
}

int gpk_unwhiten_u(gp_ctx *c) {
    const double one = 1.0, zero = 0.0;
    GP_HIP(hipMemcpyAsync(c->U, c->Uw, (size_t)c->Npad * c->g0 * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    GP_BLAS(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, c->Npad, c->M,
                          c->M, &one, c->Uw + (size_t)c->g0 * c->Npad, c->Npad, c->Linv, c->M,
                          &zero, c->U + (size_t)c->g0 * c->Npad, c->Npad));
    return 0;
}
