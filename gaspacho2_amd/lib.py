"""ctypes binding of libgaspacho2.so (the C ABI declared in include/gaspacho2.h).

The binding mirrors the reference driver's launch groups (src/ggplvm.cu main):
``init_fit`` (ggplvm.cu:1111-1137), ``irls_step`` (1156-1181), ``ql_start``
(1205-1211), ``elbo`` (1259-1278), ``grad`` (1215-1240), ``lbfgs_step`` (1251),
``downdate`` (1289), ``end_outer`` (1319-1324) and ``fit`` (1146-1330).
There is no CPU fallback: if the library is not built, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgaspacho2.so")

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP64_WHITENED = 2

_lib = None


class GPError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libgaspacho2.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GPError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "gp_version": (C.c_int, []),
        "gp_last_error": (C.c_char_p, []),
        "gp_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip,
                                ip, ip, dp, C.c_longlong, dp, dp, dp, dp, C.c_int, C.c_int]),
        "gp_destroy": (C.c_int, [vp]),
        "gp_set_precision": (C.c_int, [vp, C.c_int]),
        "gp_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "gp_get_option": (C.c_int, [vp, C.c_char_p, ip]),
        "gp_autotune": (C.c_int, [vp, C.c_int]),
        "gp_eval_graph": (C.c_int, [vp, C.c_int, dp]),
        "gp_score_tests": (C.c_int, [vp, C.c_int, dp, C.c_int, ip, ip, dp]),
        "gp_bfit_create": (C.c_int, [C.POINTER(vp), vp, C.c_int, ip, ip, ip, C.POINTER(C.c_float),
                                     ip, ip, dp]),
        "gp_bfit_destroy": (C.c_int, [vp]),
        "gp_bfit_init": (C.c_int, [vp]),
        "gp_bfit_step": (C.c_int, [vp, dp]),
        "gp_bfit_run": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, ip, dp]),
        "gp_bfit_get": (C.c_int, [vp, C.c_char_p, dp, C.c_size_t]),
        "gp_bfit_dims": (C.c_int, [vp, ip, ip]),
        "gp_bfit_autotune": (C.c_int, [vp, C.c_int, ip]),
        "gp_bfit_set_grid": (C.c_int, [vp, C.c_int]),
        "gp_score_tests_dosage": (C.c_int, [vp, C.c_int, C.POINTER(C.c_ubyte), C.c_int, ip, ip, dp]),
        "gp_score_tests_interaction": (C.c_int, [vp, C.c_int, C.POINTER(C.c_ubyte), dp, C.c_int, ip,
                                                 ip, dp, dp]),
        "gp_score_tests_dosage_noalloc": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, vp, vp]),
        "gp_score_tests_donor": (C.c_int, [vp, C.c_int, ip, C.c_int, C.POINTER(C.c_ubyte), dp, C.c_int,
                                           ip, ip, dp, dp]),
        "gp_pca_init": (C.c_int, [C.c_int, C.c_int, ip, ip, dp, dp, C.c_int, dp, C.c_int, dp, dp]),
        "gp_enc_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_longlong, C.c_ulonglong]),
        "gp_enc_destroy": (C.c_int, [vp]),
        "gp_enc_dims": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "gp_enc_add": (C.c_int, [vp, C.c_int, C.c_int, ip, ip, dp, dp, dp]),
        "gp_enc_loss_grad": (C.c_int, [vp, C.c_int, dp]),
        "gp_enc_adam_step": (C.c_int, [vp, C.c_double]),
        "gp_enc_train": (C.c_int, [vp, C.c_int, C.c_double, dp]),
        "gp_enc_predict": (C.c_int, [vp, C.c_int, C.c_int, ip, ip, dp, dp, dp, dp]),
        "gp_enc_get": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_float), C.c_size_t]),
        "gp_enc_set": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_float), C.c_size_t]),
        "gp_buffer_len": (C.c_size_t, [vp, C.c_char_p]),
        "gp_get_buffer": (C.c_int, [vp, C.c_char_p, dp, C.c_size_t]),
        "gp_set_buffer": (C.c_int, [vp, C.c_char_p, dp, C.c_size_t]),
        "gp_init_fit": (C.c_int, [vp]),
        "gp_irls_step": (C.c_int, [vp, C.c_int, dp]),
        "gp_ql_start": (C.c_int, [vp, dp]),
        "gp_elbo": (C.c_int, [vp, C.c_int, dp]),
        "gp_grad": (C.c_int, [vp]),
        "gp_lbfgs_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "gp_downdate": (C.c_int, [vp, C.c_int, C.c_double, C.c_int]),
        "gp_end_outer": (C.c_int, [vp]),
        "gp_fit": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, dp]),
        "gp_fit_trace": (C.c_int, [vp, dp, dp, ip, C.c_int, ip]),
        "gp_fit_trace_steps": (C.c_int, [vp, ip, dp, C.c_int, ip]),
        "gp_timing_enable": (C.c_int, [vp, C.c_int]),
        "gp_timing_reset": (C.c_int, [vp]),
        "gp_timing_get": (C.c_int, [vp, C.c_char_p, dp, ip]),
        "gp_init_ta_with_per": (C.c_int, [C.c_int, C.c_int, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def init_ta_with_per(M: int, Q: int) -> np.ndarray:
    """Initial inducing points (util.c:73-87); M x Q column-major as a flat array."""
    lib = load()
    T = np.zeros(M * Q, dtype=np.float64)
    if lib.gp_init_ta_with_per(M, Q, _dp(T)) != 0:
        raise GPError(lib.gp_last_error().decode())
    return T


def pca_init(N: int, Yp, Yi, Yx, ls, Zcov, Q: int):
    """PCA initial latent on the GPU (gp_pca_init, init.cu:304-357): returns
    (X as N x Q, eigenvalues largest first).  Zcov: N x p0, intercept first."""
    lib = load()
    Yp = np.ascontiguousarray(Yp, dtype=np.int32)
    Yi = np.ascontiguousarray(Yi, dtype=np.int32)
    Yx = np.ascontiguousarray(Yx, dtype=np.float64)
    ls = np.ascontiguousarray(ls, dtype=np.float64)
    Z = np.asfortranarray(np.asarray(Zcov, dtype=np.float64).reshape(N, -1, order="F"))
    zf = Z.ravel(order="F")
    X = np.zeros(N * Q)
    ev = np.zeros(Q)
    st = lib.gp_pca_init(int(N), len(Yp) - 1, _ip(Yp), _ip(Yi), _dp(Yx), _dp(ls), Z.shape[1], _dp(zf), int(Q),
                         _dp(X), _dp(ev))
    if st != 0:
        raise GPError(f"status {st}: {lib.gp_last_error().decode()}")
    return X.reshape(Q, N).T.copy(), ev


class GPContext:
    """One fit on one GPU (a gp_ctx)."""

    def __init__(self, N: int, J: int, M: int, Q: int, nh, Yp, Yi, Yx, Zcov, ls, X0, T0,
                 ard_only: bool = False, precision: int = PREC_FP64):
        self._lib = load()
        self.N, self.J, self.M, self.Q = int(N), int(J), int(M), int(Q)
        self.nh = np.ascontiguousarray(nh, dtype=np.int32)
        self.p = len(self.nh) - 1
        self.P = int(self.nh[-1])
        self._keep = [np.ascontiguousarray(Yp, dtype=np.int32),
                      np.ascontiguousarray(Yi, dtype=np.int32),
                      np.ascontiguousarray(Yx, dtype=np.float64),
                      np.asfortranarray(np.asarray(Zcov, dtype=np.float64).reshape(self.N, -1, order="F")).ravel(order="F"),
                      np.ascontiguousarray(ls, dtype=np.float64),
                      np.ascontiguousarray(X0, dtype=np.float64).ravel(order="F") if np.ndim(X0) == 1 else np.asfortranarray(X0, dtype=np.float64).ravel(order="F"),
                      np.ascontiguousarray(T0, dtype=np.float64).ravel(order="F") if np.ndim(T0) == 1 else np.asfortranarray(T0, dtype=np.float64).ravel(order="F")]
        Ypa, Yia, Yxa, Zc, lsa, Xa, Ta = self._keep
        h = C.c_void_p()
        st = self._lib.gp_create(C.byref(h), self.N, self.J, self.M, self.Q, self.p, _ip(self.nh),
                                 _ip(Ypa), _ip(Yia), _dp(Yxa), len(Yxa), _dp(Zc), _dp(lsa),
                                 _dp(Xa), _dp(Ta), int(bool(ard_only)), int(precision))
        self._check(st)
        self._h = h
        self.last_nonpd = 0

    # -- plumbing
    def _check(self, st: int) -> int:
        if st > 0:
            raise GPError(f"status {st}: {self._lib.gp_last_error().decode()}")
        self.last_nonpd = -st if st < 0 else 0
        return st

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- buffers
    def buffer_len(self, name: str) -> int:
        return int(self._lib.gp_buffer_len(self._h, name.encode()))

    def get(self, name: str) -> np.ndarray:
        n = self.buffer_len(name)
        if n == 0:
            raise GPError(f"unknown buffer {name}")
        out = np.empty(n, dtype=np.float64)
        self._check(self._lib.gp_get_buffer(self._h, name.encode(), _dp(out), n))
        return out

    def set(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(self._lib.gp_set_buffer(self._h, name.encode(), _dp(a), a.size))

    def set_precision(self, precision: int) -> None:
        self._check(self._lib.gp_set_precision(self._h, int(precision)))

    def set_option(self, name: str, value: int) -> None:
        self._check(self._lib.gp_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int()
        self._check(self._lib.gp_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    def autotune(self, verbose: bool = False) -> dict:
        """Pick the mixed Gram form and segments and the contraction form and
        gene groups by timing them on the current state."""
        self._check(self._lib.gp_autotune(self._h, int(bool(verbose))))
        return {k: self.get_option(k) for k in ("gram_waves", "gram_splits", "gc_form", "gc_groups")}

    # -- launch groups
    def init_fit(self) -> None:
        self._check(self._lib.gp_init_fit(self._h))

    def irls_step(self, jtr: int) -> float:
        o = C.c_double()
        self._check(self._lib.gp_irls_step(self._h, int(jtr), C.byref(o)))
        return o.value

    def ql_start(self) -> float:
        o = C.c_double()
        self._check(self._lib.gp_ql_start(self._h, C.byref(o)))
        return o.value

    def elbo(self, slot: int = 0) -> float:
        o = C.c_double()
        self._check(self._lib.gp_elbo(self._h, int(slot), C.byref(o)))
        return o.value

    def eval_graph(self, slot: int = 0) -> float:
        """ELBO (returned) and gradient (device grad buffer) by replaying the recorded unit."""
        o = C.c_double()
        self._check(self._lib.gp_eval_graph(self._h, int(slot), C.byref(o)))
        return o.value

    def grad(self) -> None:
        self._check(self._lib.gp_grad(self._h))

    def lbfgs_step(self, jtr: int, rhoflag: int) -> None:
        self._check(self._lib.gp_lbfgs_step(self._h, int(jtr), int(rhoflag)))

    def downdate(self, jtr: int, sr: float, rhoflag: int) -> None:
        self._check(self._lib.gp_downdate(self._h, int(jtr), float(sr), int(rhoflag)))

    def end_outer(self) -> None:
        self._check(self._lib.gp_end_outer(self._h))

    def fit(self, itrmax: int, jtr_cap: int = 0, verbose: bool = False) -> float:
        o = C.c_double()
        self._check(self._lib.gp_fit(self._h, int(itrmax), int(jtr_cap), int(bool(verbose)),
                                     None, None, C.byref(o)))
        return o.value

    def fit_trace(self):
        """(objective, seconds, kind) after every step of the last fit; kind 0
        IRLS, 1 quasi-likelihood start, 2 L-BFGS."""
        n = C.c_int()
        self._check(self._lib.gp_fit_trace(self._h, None, None, None, 0, C.byref(n)))
        o, t, k = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value, np.int32)
        self._check(self._lib.gp_fit_trace(self._h, _dp(o), _dp(t), _ip(k), n.value, C.byref(n)))
        return o, t, k

    def fit_trace_steps(self):
        """(entries x 4) int array (kind, itr, jtr, ktr) and the final
        line-search sr of every entry of the last fit (the oracle's
        fit_trace layout)."""
        o, _, k = self.fit_trace()
        n = C.c_int()
        step, sr = np.zeros(3 * len(o), np.int32), np.zeros(len(o))
        self._check(self._lib.gp_fit_trace_steps(self._h, _ip(step), _dp(sr), len(o), C.byref(n)))
        info = np.concatenate([k[:, None], step.reshape(-1, 3)], axis=1)
        return o, sr, info

    # -- score tests (ggplvm_de.cu:491-532)
    def score_tests(self, G, gene, snp) -> np.ndarray:
        """Statistics of tests (gene[t], snp[t]); G is N x S (float: doubles,
        uint8: 0/1/2 dosages), column-major or a 2-D array."""
        G = np.asarray(G)
        gene = np.ascontiguousarray(gene, dtype=np.int32)
        snp = np.ascontiguousarray(snp, dtype=np.int32)
        out = np.zeros(len(gene))
        if G.dtype == np.uint8:
            Gf = np.asfortranarray(G.reshape(self.N, -1, order="F"))
            flat = Gf.ravel(order="F")
            st = self._lib.gp_score_tests_dosage(self._h, Gf.shape[1],
                                                 flat.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                                 len(gene), _ip(gene), _ip(snp), _dp(out))
        else:
            Gf = np.asfortranarray(G.astype(np.float64).reshape(self.N, -1, order="F"))
            flat = Gf.ravel(order="F")
            st = self._lib.gp_score_tests(self._h, Gf.shape[1], _dp(flat), len(gene), _ip(gene),
                                          _ip(snp), _dp(out))
        self._check(st)
        return out

    def score_tests_interaction(self, dosage, factor, gene, snp):
        """(statistics, two-sided p-values) of tests (gene[t], snp[t]) on the
        interaction columns g_l * factor - mean(g_l * factor) (factor None:
        the dosage columns themselves); dosage is N x S uint8."""
        D = np.asfortranarray(np.asarray(dosage, dtype=np.uint8).reshape(self.N, -1, order="F"))
        flat = D.ravel(order="F")
        gene = np.ascontiguousarray(gene, dtype=np.int32)
        snp = np.ascontiguousarray(snp, dtype=np.int32)
        stat, pval = np.zeros(len(gene)), np.zeros(len(gene))
        f = None if factor is None else np.ascontiguousarray(factor, dtype=np.float64)
        st = self._lib.gp_score_tests_interaction(self._h, D.shape[1],
                                                  flat.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                                  None if f is None else _dp(f), len(gene), _ip(gene),
                                                  _ip(snp), _dp(stat), _dp(pval))
        self._check(st)
        return stat, pval

    def score_tests_donor(self, donor, dosage, factor, gene, snp):
        """(statistics, two-sided p-values) with genotypes per donor: cell i
        has donor[i]'s genotypes, dosage is ndonor x S uint8; factor None
        or a per-cell interaction factor (see score_tests_interaction)."""
        donor = np.ascontiguousarray(donor, dtype=np.int32)
        D = np.asfortranarray(np.asarray(dosage, dtype=np.uint8))
        nd = D.shape[0]
        flat = D.ravel(order="F")
        gene = np.ascontiguousarray(gene, dtype=np.int32)
        snp = np.ascontiguousarray(snp, dtype=np.int32)
        stat, pval = np.zeros(len(gene)), np.zeros(len(gene))
        f = None if factor is None else np.ascontiguousarray(factor, dtype=np.float64)
        st = self._lib.gp_score_tests_donor(self._h, nd, _ip(donor), D.shape[1],
                                            flat.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                            None if f is None else _dp(f), len(gene), _ip(gene),
                                            _ip(snp), _dp(stat), _dp(pval))
        self._check(st)
        return stat, pval

    def score_tests_device(self, S: int, dosage_ptr: int, T: int, gene_ptr: int, snp_ptr: int,
                           stat_ptr: int) -> None:
        """gp_score_tests_dosage_noalloc: every argument is a device address
        (N x S dosage bytes, T gene and SNP ints, T statistics); nothing is
        allocated."""
        self._check(self._lib.gp_score_tests_dosage_noalloc(self._h, int(S), C.c_void_p(dosage_ptr),
                                                            int(T), C.c_void_p(gene_ptr),
                                                            C.c_void_p(snp_ptr), C.c_void_p(stat_ptr)))

    # -- timing
    def timing(self, on: bool = True) -> None:
        self._check(self._lib.gp_timing_enable(self._h, int(bool(on))))

    def timing_reset(self) -> None:
        self._check(self._lib.gp_timing_reset(self._h))

    def timing_get(self, name: str):
        ms = C.c_double()
        cnt = C.c_int()
        self._check(self._lib.gp_timing_get(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class BatchedFit:
    """Per-gene Poisson GLM / Laplace fits on cell subsets (gp_bfit_*), on the
    latent space of an initialised GPContext."""

    def __init__(self, ctx: GPContext, off, cell, ncov, cov, yoff, ypos, yval):
        self._lib = load()
        self.ctx = ctx
        self._keep = [np.ascontiguousarray(off, dtype=np.int32), np.ascontiguousarray(cell, dtype=np.int32),
                      np.ascontiguousarray(ncov, dtype=np.int32), np.ascontiguousarray(cov, dtype=np.float32),
                      np.ascontiguousarray(yoff, dtype=np.int32), np.ascontiguousarray(ypos, dtype=np.int32),
                      np.ascontiguousarray(yval, dtype=np.float64)]
        o, c, nc, cv, yo, yp, yv = self._keep
        h = C.c_void_p()
        st = self._lib.gp_bfit_create(C.byref(h), ctx._h, len(o) - 1, _ip(o), _ip(c), _ip(nc),
                                      cv.ctypes.data_as(C.POINTER(C.c_float)), _ip(yo), _ip(yp), _dp(yv))
        ctx._check(st)
        self._h = h
        J, PB = C.c_int(), C.c_int()
        self._lib.gp_bfit_dims(h, C.byref(J), C.byref(PB))
        self.J, self.PB = J.value, PB.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gp_bfit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self.ctx._check(self._lib.gp_bfit_init(self._h))

    def step(self) -> float:
        o = C.c_double()
        self.ctx._check(self._lib.gp_bfit_step(self._h, C.byref(o)))
        return o.value

    def autotune(self, verbose: bool = False) -> int:
        """Persistent workgroups per CU picked by timing the pass (1-4)."""
        m = C.c_int()
        self.ctx._check(self._lib.gp_bfit_autotune(self._h, int(bool(verbose)), C.byref(m)))
        return m.value

    def set_grid(self, grid_mult: int) -> None:
        self.ctx._check(self._lib.gp_bfit_set_grid(self._h, int(grid_mult)))

    def run(self, maxit: int, tol: float = 1e-7, graph: bool = True):
        it, o = C.c_int(), C.c_double()
        self.ctx._check(self._lib.gp_bfit_run(self._h, int(maxit), float(tol), int(bool(graph)),
                                              C.byref(it), C.byref(o)))
        return it.value, o.value

    def get(self, name: str) -> np.ndarray:
        n = {"Beta": self.J * self.PB, "phi": self.J, "lkhd": self.J, "Psi": self.J * self.PB * self.PB}[name]
        out = np.empty(n)
        self.ctx._check(self._lib.gp_bfit_get(self._h, name.encode(), _dp(out), n))
        return out


class Encoder:
    """Amortization encoder (gp_enc_*): MLP from log1p(y / s) expression
    features to a Gaussian over the latent position, trained on the GPU on
    (counts, fitted X) pairs and used to warm-start X.

    Parameters live in one FP32 array in the padded layout
    [W1 (H1 x Fp) | b1 (H1) | W2 (H2 x H1) | b2 (H2) | W3 (O x H2) | b3 (O)],
    Fp = F rounded up to 64, O = 64 (rows >= 2Q are zero)."""

    def __init__(self, F: int, Q: int, H1: int = 512, H2: int = 256, batch: int = 1024,
                 capacity: int = 0, seed: int = 0):
        self._lib = load()
        h = C.c_void_p()
        self._check(self._lib.gp_enc_create(C.byref(h), int(F), int(H1), int(H2), int(Q), int(batch),
                                            int(capacity), C.c_ulonglong(int(seed))))
        self._h = h
        d = (C.c_longlong * 12)()
        self._check(self._lib.gp_enc_dims(h, d))
        (self.F, self.Fp, self.H1, self.H2, self.Q, self.O, self.batch) = [int(x) for x in d[:7]]
        self.n_params = int(d[8])

    def _check(self, st: int) -> None:
        if st != 0:
            raise GPError(f"status {st}: {self._lib.gp_last_error().decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gp_enc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_cells(self) -> int:
        d = (C.c_longlong * 12)()
        self._check(self._lib.gp_enc_dims(self._h, d))
        return int(d[7])

    @staticmethod
    def _counts(N, Yp, Yi, Yx, ls):
        Yp = np.ascontiguousarray(Yp, dtype=np.int32)
        Yi = np.ascontiguousarray(Yi, dtype=np.int32)
        Yx = np.ascontiguousarray(Yx, dtype=np.float64)
        ls = np.ascontiguousarray(ls, dtype=np.float64)
        if len(ls) != N:
            raise GPError("size factors must have N entries")
        return Yp, Yi, Yx, ls

    def add(self, N: int, Yp, Yi, Yx, ls, X) -> None:
        """Append N training cells: gene-major CSR counts (J = F genes), size
        factors, and target latents X (N x Q, column-major flat or 2-D)."""
        Yp, Yi, Yx, ls = self._counts(N, Yp, Yi, Yx, ls)
        X = np.asarray(X, dtype=np.float64)
        X = np.ascontiguousarray(X.T if X.ndim == 2 else X.reshape(self.Q, N)).reshape(-1)
        self._check(self._lib.gp_enc_add(self._h, int(N), len(Yp) - 1, _ip(Yp), _ip(Yi), _dp(Yx), _dp(ls),
                                         _dp(X)))

    def loss_grad(self, n: int) -> float:
        """Loss and gradients (left on the device) of the first n loaded cells."""
        o = C.c_double()
        self._check(self._lib.gp_enc_loss_grad(self._h, int(n), C.byref(o)))
        return o.value

    def adam_step(self, lr: float) -> None:
        self._check(self._lib.gp_enc_adam_step(self._h, float(lr)))

    def train(self, epochs: int, lr: float = 1e-3) -> np.ndarray:
        out = np.zeros(max(int(epochs), 1))
        self._check(self._lib.gp_enc_train(self._h, int(epochs), float(lr), _dp(out)))
        return out[:int(epochs)]

    def predict(self, N: int, Yp, Yi, Yx, ls):
        """(mu, log-variance), each N x Q."""
        Yp, Yi, Yx, ls = self._counts(N, Yp, Yi, Yx, ls)
        mu = np.zeros(N * self.Q)
        lv = np.zeros(N * self.Q)
        self._check(self._lib.gp_enc_predict(self._h, int(N), len(Yp) - 1, _ip(Yp), _ip(Yi), _dp(Yx), _dp(ls),
                                             _dp(mu), _dp(lv)))
        return mu.reshape(self.Q, N).T.copy(), lv.reshape(self.Q, N).T.copy()

    def get(self, name: str) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._lib.gp_enc_get(self._h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_float)),
                                         self.n_params))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self._lib.gp_enc_set(self._h, b"params", p.ctypes.data_as(C.POINTER(C.c_float)), p.size))

    def unpack(self, flat: np.ndarray) -> dict:
        """Split a parameter-layout array into W1, b1, W2, b2, W3, b3."""
        F, Fp, H1, H2, O = self.F, self.Fp, self.H1, self.H2, self.O
        o = 0
        out = {}
        for name, shape in (("W1", (H1, Fp)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)),
                            ("W3", (O, H2)), ("b3", (O,))):
            n = int(np.prod(shape))
            out[name] = flat[o:o + n].reshape(shape)
            o += n
        return out
