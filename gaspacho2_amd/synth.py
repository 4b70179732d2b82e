"""Synthetic single-cell count data in the reference's file formats.

Generative model (the workloads' recipe): latent X ~ N(0, I_Q); generating
inducing points ~ N(0, I_Q); SE-ARD lengthscales l_q ~ U(0.5, 2) (rho = l^2);
per-gene f_g = K_xz a_g with a_g ~ N(0, 0.5^2); intercepts b_g ~ N(-1, 1);
size factors s_n ~ LogNormal(0, 0.5^2); counts y ~ Poisson(s_n exp(f_g + b_g)).
Counts are stored CSC by gene (SparseMat.cu:37-52) with double values.

The PCA initial latent follows init.cu:304-357 (getNu0 / getNuTNu / getX):
nu0 = log((y + 0.5) / s) - Z beta, X = nu0 V / sqrt(lambda) * sqrt(N) with
the top-Q eigenpairs of nu0^T nu0, largest first.  Here beta is the
intercept-only fit (per-gene mean of log((y + 0.5) / s)).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Optional

import os

import numpy as np


@dataclass
class Dataset:
    N: int
    J: int
    Q: int
    M: int
    Yp: np.ndarray           # int32 [J+1]
    Yi: np.ndarray           # int32 [L]
    Yx: np.ndarray           # float64 [L]
    ls: np.ndarray           # float64 [N]
    Zcov: np.ndarray         # float64 N x ncol (Fortran), intercept first
    nh: np.ndarray           # int32 [p+1]
    X0: np.ndarray           # float64 N x Q (Fortran)
    c: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    v: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    body: bytes = b""
    X_true: Optional[np.ndarray] = None

    @property
    def L(self) -> int:
        return int(self.Yp[-1])

    @property
    def P(self) -> int:
        return int(self.nh[-1])

    def zero_fraction(self) -> float:
        return 1.0 - self.L / float(self.N * self.J)


def _se_ard(X: np.ndarray, Z: np.ndarray, rho: np.ndarray) -> np.ndarray:
    d = np.zeros((X.shape[0], Z.shape[0]))
    for q in range(X.shape[1]):
        d += (X[:, q:q + 1] - Z[None, :, q]) ** 2 / (2.0 * rho[q])
    return np.exp(-d)


def make_dataset(N: int, G: int, Q: int, M: int, seed: int = 0, n_numeric: int = 0,
                 n_levels: int = 0, gen_M: Optional[int] = None, pca: bool = True,
                 block: int = 256, model_seed: Optional[int] = None, progress=None) -> Dataset:
    """Counts from the generative model above; optional covariates
    (n_numeric N(0,1) columns and one categorical variable with n_levels).

    model_seed: draw the per-gene functions (generating inducing points,
    lengthscales, a_g, b_g) from their own stream, so several data seeds share
    one set of genes (new cells, size factors and counts per seed).
    progress, if given, is called with (genes done, G) after each gene block."""
    rng = np.random.default_rng(seed)
    Mg = gen_M or M
    X = rng.standard_normal((N, Q))
    if model_seed is None:
        Zg = rng.standard_normal((Mg, Q))
        rho = rng.uniform(0.5, 2.0, Q) ** 2
        Kxz = _se_ard(X, Zg, rho)
        a = rng.normal(0.0, 0.5, (Mg, G))
        b = rng.normal(-1.0, 1.0, G)
    else:
        mrng = np.random.default_rng([model_seed, 7919])
        Zg = mrng.standard_normal((Mg, Q))
        rho = mrng.uniform(0.5, 2.0, Q) ** 2
        Kxz = _se_ard(X, Zg, rho)
        a = mrng.normal(0.0, 0.5, (Mg, G))
        b = mrng.normal(-1.0, 1.0, G)
    s = np.exp(rng.normal(0.0, 0.5, N))
    # covariates
    cov_cols = [np.ones(N)]
    c_list, v_list, body = [], [], b""
    if n_numeric:
        num = rng.standard_normal((N, n_numeric))
        for k in range(n_numeric):
            cov_cols.append(num[:, k])
        c_list.append(1)
        v_list.append(n_numeric)
        body += np.asfortranarray(num).tobytes(order="F")
    if n_levels:
        cat = rng.integers(0, n_levels, N).astype(np.int32)
        for lev in range(n_levels):
            cov_cols.append((cat == lev).astype(np.float64))
        c_list.append(n_levels)
        v_list.append(1)
        body += cat.tobytes()
    Zcov = np.asfortranarray(np.stack(cov_cols, axis=1))
    beta_cov = rng.normal(0.0, 0.2, (Zcov.shape[1], G)) if Zcov.shape[1] > 1 else None
    Yp = np.zeros(G + 1, dtype=np.int64)
    Yi_parts, Yx_parts = [], []
    nu0_gram = None
    for g0 in range(0, G, block):
        g1 = min(G, g0 + block)
        eta = Kxz @ a[:, g0:g1] + b[g0:g1]
        if beta_cov is not None:
            eta += Zcov[:, 1:] @ beta_cov[1:, g0:g1]
        lam = s[:, None] * np.exp(eta)
        y = rng.poisson(lam)
        for k in range(g1 - g0):
            nz = np.flatnonzero(y[:, k])
            Yi_parts.append(nz.astype(np.int32))
            Yx_parts.append(y[nz, k].astype(np.float64))
            Yp[g0 + k + 1] = Yp[g0 + k] + nz.size
        if progress is not None:
            progress(g1, G)
    Yi = np.concatenate(Yi_parts) if Yi_parts else np.zeros(0, np.int32)
    Yx = np.concatenate(Yx_parts) if Yx_parts else np.zeros(0)
    ncol = Zcov.shape[1]
    nh = np.array([0, 1] + ([] if not n_numeric else [1 + n_numeric]) +
                  ([] if not n_levels else [1 + n_numeric + n_levels]), dtype=np.int32)
    nh = np.concatenate([nh, [ncol + M]]).astype(np.int32)
    ds = Dataset(N=N, J=G, Q=Q, M=M, Yp=Yp.astype(np.int32), Yi=Yi, Yx=Yx, ls=s,
                 Zcov=Zcov, nh=nh, X0=np.zeros((N, Q), order="F"),
                 c=np.array(c_list, np.int32), v=np.array(v_list, np.int32), body=body,
                 X_true=X)
    ds.X0 = pca_init(ds) if pca else np.asfortranarray(X)
    return ds


def dense_counts(ds: Dataset, genes: Optional[slice] = None) -> np.ndarray:
    g = range(ds.J) if genes is None else range(*genes.indices(ds.J))
    out = np.zeros((ds.N, len(g)))
    for k, j in enumerate(g):
        lo, hi = ds.Yp[j], ds.Yp[j + 1]
        out[ds.Yi[lo:hi], k] = ds.Yx[lo:hi]
    return out


def pca_init(ds: Dataset, max_cells: int = 20000, seed: int = 1) -> np.ndarray:
    """init.cu:304-357 PCA initial latent (see module docstring).

    nu0 = S + u 1^T - 1 m^T with S sparse (log((y+0.5)/0.5) at non-zeros),
    u_i = log(0.5 / s_i) and m_j the per-gene mean; the eigenvectors come from
    a random cell subsample, the projection uses every cell."""
    import scipy.sparse as sp
    N, J, Q = ds.N, ds.J, ds.Q
    vals = np.log(ds.Yx + 0.5) - np.log(0.5)
    S = sp.csc_matrix((vals, ds.Yi, ds.Yp), shape=(N, J)).tocsr()
    u = np.log(0.5) - np.log(ds.ls)
    m = (np.asarray(S.sum(axis=0)).ravel() + u.sum()) / N
    rng = np.random.default_rng(seed)
    rows = np.arange(N) if N <= max_cells else np.sort(rng.choice(N, max_cells, replace=False))
    nu = S[rows].toarray() + u[rows, None] - m[None, :]
    G = nu.T @ nu
    lam, V = np.linalg.eigh(G)
    lam, V = lam[::-1][:Q], V[:, ::-1][:, :Q]
    V = V * np.where(V[np.abs(V).argmax(axis=0), np.arange(Q)] < 0, -1.0, 1.0)   # sign convention
    proj = np.asarray(S @ V) + u[:, None] * V.sum(axis=0)[None, :] - (m @ V)[None, :]
    X = proj / np.sqrt(lam) * np.sqrt(rows.size)
    return np.asfortranarray(X)


# ----------------------------------------------------------------------------
# reference file formats
def write_sparsemat(path: str, ds: Dataset) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<iii", ds.N, ds.J, ds.L))
        f.write(ds.Yi.astype("<i4").tobytes())
        f.write(ds.Yp.astype("<i4").tobytes())
        f.write(ds.Yx.astype("<f8").tobytes())


def write_dataframe(path: str, ds: Dataset) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", ds.N, len(ds.c)))
        f.write(ds.c.astype("<i4").tobytes())
        f.write(ds.v.astype("<i4").tobytes())
        f.write(ds.body)


def write_doubles(path: str, x: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(np.asarray(x, dtype="<f8").ravel(order="F").tobytes())


@dataclass
class BatchedData:
    """Per-gene cell subsets with their own covariates (the uneven batched
    GLM workload): gene g uses cells cell[off[g]:off[g+1]] (ascending), p_g =
    ncov[g] design columns before the GP block (intercept + p_g - 1 FP32
    N(0,1) covariates, column-major per gene in cov), and non-zero counts at
    subset positions ypos[yoff[g]:yoff[g+1]] with values yval."""
    off: np.ndarray
    cell: np.ndarray
    ncov: np.ndarray
    cov: np.ndarray
    yoff: np.ndarray
    ypos: np.ndarray
    yval: np.ndarray


def _batched_block(ds, Kxz, sizes, ncov, g0, g1, seed):
    """Genes g0..g1-1 of make_batched, from their own stream."""
    rng = np.random.default_rng([seed, 104729, g0])
    N, Mg = ds.N, Kxz.shape[1]
    A = rng.normal(0.0, 0.5, (Mg, g1 - g0))
    F = Kxz @ A                                  # latent part of eta, N x block
    out = []
    for g in range(g0, g1):
        n, pc = int(sizes[g]), int(ncov[g]) - 1
        if 2 * n < N:     # sorted sample without replacement, no full permutation
            c = np.unique(rng.integers(0, N, int(n * 1.3) + 16))
            while c.size < n:
                c = np.unique(np.concatenate([c, rng.integers(0, N, n)]))
            c = np.sort(rng.choice(c, n, replace=False))
        else:
            c = np.sort(rng.choice(N, n, replace=False))
        c = c.astype(np.int32)
        C = rng.standard_normal((pc, n), dtype=np.float32)      # column-major n x pc
        gam = rng.normal(0.0, 0.2, pc)
        b = rng.normal(-1.0, 1.0)
        eta = F[c, g - g0] + gam @ C + b
        y = rng.poisson(ds.ls[c] * np.exp(eta))
        nz = np.flatnonzero(y)
        out.append((c, C.ravel(), nz.astype(np.int32), y[nz].astype(np.float64)))
    return out


def make_batched(ds: Dataset, G: int, nmin: int, nmax: int, pmin: int = 2, pmax: int = 16,
                 seed: int = 0, gene_block: int = 256, threads: Optional[int] = None) -> BatchedData:
    """Subset sizes log-uniform in [nmin, nmax], p_g ~ U{pmin..pmax}; counts
    y ~ Poisson(s_i exp(f_g(x_i) + c_i . gamma_g + b_g)) with f_g from the
    generating SE-ARD model of ds (its true latent X).  Blocks of gene_block
    genes draw their subsets, covariates, coefficients and counts from their
    own streams (so the data do not depend on the thread count) and the
    latent part of eta of a block is one product Kxz (N x 64) by its
    coefficients."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng([seed, 104729])
    N, Q = ds.N, ds.Q
    X = ds.X_true if ds.X_true is not None else np.asarray(ds.X0)
    Mg = 64
    Zg = rng.standard_normal((Mg, Q))
    rho = rng.uniform(0.5, 2.0, Q) ** 2
    Kxz = _se_ard(X, Zg, rho)                       # N x Mg, once
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), G)).astype(np.int64)
    sizes = np.clip(sizes, 1, N)
    ncov = rng.integers(pmin, pmax + 1, G).astype(np.int32)
    off = np.zeros(G + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    blocks = [(g0, min(G, g0 + gene_block)) for g0 in range(0, G, gene_block)]
    nt = threads or min(16, os.cpu_count() or 1)
    # single-threaded BLAS inside each worker: a block's product then does
    # not depend on how many threads the library would have split it over
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1, user_api="blas"), ThreadPoolExecutor(max_workers=nt) as ex:
        parts = list(ex.map(lambda b: _batched_block(ds, Kxz, sizes, ncov, b[0], b[1], seed), blocks))
    genes = [x for part in parts for x in part]
    yoff = np.zeros(G + 1, np.int64)
    yoff[1:] = np.cumsum([x[2].size for x in genes])
    return BatchedData(off=off.astype(np.int32), cell=np.concatenate([x[0] for x in genes]), ncov=ncov,
                       cov=np.concatenate([x[1] for x in genes]) if genes else np.zeros(0, np.float32),
                       yoff=yoff.astype(np.int32), ypos=np.concatenate([x[2] for x in genes]),
                       yval=np.concatenate([x[3] for x in genes]))
