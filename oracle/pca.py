"""CPU oracle of the PCA initial latent (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of gp_pca_init
(gaspacho2_amd/csrc/gp_pca.hip).  The product never imports it.

Restates the unused helpers getNu0 / getNuTNu / getX of the reference
(src/init.cu:304-357):
  getNu0   nu0[i, j] = log((y_ij + 0.5) / s_i) - sum_k Z[i, k] beta[k, j]
           (the merge over the gene's stored counts, init.cu:305-326)
  getNuTNu upper triangle of nu0^T nu0                      (init.cu:329-340)
  getX     X[i, M-k-1] = sum_j nu0[i, j] evec[j, k] / sqrt(eval[k]) * sqrt(N)
           with the top eigenpairs in ascending order        (init.cu:343-357)
beta is the least-squares fit of log((y + 0.5) / s) on Z (the reference
leaves it to the caller; with dependent columns the residual nu0 is the
projection residual onto their span, whatever beta is chosen); each eigenvector is signed so that its entry of
largest magnitude is positive.

Pinned in tests/test_oracle_pins.py by what the mathematics fixes: X^T X =
N I, X spans the top left singular vectors of nu0 (numpy SVD), and nu0 is
orthogonal to the covariate columns.
"""
from __future__ import annotations

import numpy as np


def nu0(N, Yp, Yi, Yx, ls, Z):
    """getNu0 with a least-squares beta; dense N x J."""
    J = len(Yp) - 1
    Y = np.log(0.5 / ls)[:, None] * np.ones((1, J))
    for j in range(J):
        lo, hi = Yp[j], Yp[j + 1]
        Y[Yi[lo:hi], j] = np.log((Yx[lo:hi] + 0.5) / ls[Yi[lo:hi]])
    Z = np.asarray(Z, dtype=np.float64).reshape(N, -1, order="F")
    beta = np.linalg.lstsq(Z, Y, rcond=None)[0]       # minimum norm if Z is rank deficient
    return Y - Z @ beta


def pca_init(N, Yp, Yi, Yx, ls, Z, Q):
    """(X as N x Q, eigenvalues largest first)."""
    nu = nu0(N, Yp, Yi, Yx, ls, Z)
    G = nu.T @ nu                                   # getNuTNu (full symmetric)
    lam, V = np.linalg.eigh(G)                      # ascending
    lam_top, V_top = lam[-Q:], V[:, -Q:]            # the top Q, ascending
    M = Q
    X = np.zeros((N, Q))
    for k in range(M):                              # getX: column M-k-1 <- pair k
        v = V_top[:, k]
        am = int(np.argmax(np.abs(v)))
        v = v if v[am] >= 0 else -v
        X[:, M - k - 1] = nu @ v / np.sqrt(lam_top[k]) * np.sqrt(N)
    return X, lam_top[::-1].copy()
