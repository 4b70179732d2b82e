"""Encoder warm start of the latent X (no counterpart in the reference, which
reads the starting X from an external PCA file, src/ggplvm.cu:1088-1095).

Pairs (counts, fitted X) from earlier fits train the GPU encoder
(gaspacho2_amd.lib.Encoder); a new data set then starts its fit from the
encoder's output instead of the PCA latent.  The fit itself is unchanged and
remains the judge: its final objective is compared with the PCA-started fit.

Two training targets:
  "residual": the fitted X minus that data set's own PCA latent, i.e. the
              correction the L-BFGS phase made, expressed in the PCA frame of
              the same data set; warm start = PCA latent + encoder mean;
  "aligned":  the fitted X itself, rotated (orthogonal Procrustes) into the
              frame of the first data set through the prediction of an encoder
              trained on that data set alone (the fitted latent space is only
              defined up to such a rotation); warm start = encoder mean.
"""
from __future__ import annotations

import time
from typing import List, Sequence, Tuple

import numpy as np

from . import lib

TARGETS = ("residual", "aligned")


def fitted_latent(ctx: lib.GPContext, N: int, Q: int) -> np.ndarray:
    """The context's current X as an N x Q array."""
    return ctx.get("X")[:N * Q].reshape(Q, N).T.copy()


def fit_from(ds, X0: np.ndarray, T0: np.ndarray, itrmax: int, jtr_cap: int,
             precision: int = lib.PREC_MIXED, ard_only: bool = False, verbose: bool = False):
    """gp_fit from the starting latent X0; returns (X, objective, seconds,
    trace) with trace = (objectives, seconds, kinds) of gp_fit_trace
    (verbose: the library prints every objective as the fit runs)."""
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                      np.asfortranarray(X0), T0, ard_only=ard_only, precision=precision)
    try:
        t = time.perf_counter()
        obj = g.fit(itrmax, jtr_cap, verbose=verbose)
        dt = time.perf_counter() - t
        return fitted_latent(g, ds.N, ds.Q), obj, dt, g.fit_trace()
    finally:
        g.close()


def steps_to_reach(trace, target: float, rtol: float = 1e-6) -> Tuple[int, float]:
    """(L-BFGS steps, seconds) after which a fit trace first reaches the
    objective target within rtol, or (None, None).  Only the last outer
    iteration's L-BFGS phase is searched: every outer iteration refits the
    GLM before its quasi-likelihood start, so objectives of different outer
    iterations are not on one scale (the steps count from the fit's start)."""
    obj, sec, kind = trace
    kind = np.asarray(kind)
    last = int(np.flatnonzero(kind == 1)[-1]) if np.any(kind == 1) else 0
    n = 0
    for i, (v, t, k) in enumerate(zip(obj, sec, kind)):
        if k == 2:
            n += 1
            if i > last and v >= target - rtol * abs(target):
                return n, float(t)
    return None, None


def _add(enc: lib.Encoder, ds, X: np.ndarray) -> None:
    enc.add(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls, X)


def train_encoder(train: Sequence[Tuple[object, np.ndarray]], target: str = "residual",
                  epochs: int = 30, lr: float = 1e-3, batch: int = 1024, H1: int = 512,
                  H2: int = 256, seed: int = 2, log=None) -> Tuple[lib.Encoder, np.ndarray]:
    """Train an encoder on [(data set, fitted X)]; every data set has the same
    genes.  Returns (encoder, per-epoch losses)."""
    if target not in TARGETS:
        raise ValueError(f"target must be one of {TARGETS}")
    ds0 = train[0][0]
    F, Q = ds0.J, ds0.Q
    ncell = sum(ds.N for ds, _ in train)
    rots = [np.eye(Q)] * len(train)
    if target == "aligned":
        enc0 = lib.Encoder(F, Q, H1=H1, H2=H2, batch=batch, capacity=ds0.N, seed=seed - 1)
        try:
            _add(enc0, ds0, train[0][1])
            enc0.train(epochs, lr=lr)
            for k in range(1, len(train)):
                ds, X = train[k]
                mu0, _ = enc0.predict(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls)
                U, _, Vt = np.linalg.svd(X.T @ mu0)
                rots[k] = U @ Vt
                if log:
                    log(f"aligned: data set {k}: rms(X R - mu0) {np.sqrt(np.mean((X @ rots[k] - mu0) ** 2)):.3f}")
        finally:
            enc0.close()
    enc = lib.Encoder(F, Q, H1=H1, H2=H2, batch=batch, capacity=ncell, seed=seed)
    for k, (ds, X) in enumerate(train):
        _add(enc, ds, X @ rots[k] if target == "aligned" else X - ds.X0)
    losses = enc.train(epochs, lr=lr)
    return enc, losses


def warm_start(enc: lib.Encoder, ds, target: str = "residual") -> np.ndarray:
    """Starting latent for the data set ds from a trained encoder."""
    mu, _ = enc.predict(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls)
    if not np.all(np.isfinite(mu)):
        raise lib.GPError("encoder produced a non-finite latent")
    return ds.X0 + mu if target == "residual" else mu
