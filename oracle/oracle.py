"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
The method names mirror gaspacho2_amd.lib.GPContext so that a test can drive
both through the same launch-group sequence.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    sig = {
        "orc_choldc": (C.c_int, [dp, C.c_int]),
        "orc_cholsl": (None, [dp, C.c_int, dp, dp]),
        "orc_getRb": (None, [dp, dp, C.c_int]),
        "orc_quadform": (C.c_double, [dp, dp, C.c_int]),
        "orc_mscale": (None, [C.c_int, dp, C.c_double]),
        "orc_lbfgs_two_loop": (None, [dp, dp, dp, C.c_int, C.c_int, C.c_int, dp, dp]),
        "orc_kernel_mats": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp]),
        "orc_kinv_kmn": (None, [C.c_int, C.c_int, dp, dp, dp, dp]),
        "orc_expand_delta": (None, [C.c_int, ip, C.c_int, dp, dp, dp]),
        "orc_dataframe_layout": (C.c_int, [C.c_int, ip, ip, C.c_int, ip]),
        "orc_dataframe_expand": (None, [C.c_int, C.c_int, ip, ip, C.c_char_p, ip, dp]),
        "orc_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, ip, dp, C.c_int,
                            dp, dp, dp, dp, C.c_int, C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_grad_len": (C.c_size_t, [vp]),
        "orc_buf": (dp, [vp, C.c_char_p, C.POINTER(C.c_size_t)]),
        "orc_init_fit": (None, [vp]),
        "orc_irls_step": (C.c_double, [vp, C.c_int, C.c_int]),
        "orc_ql_start": (C.c_double, [vp, C.c_int]),
        "orc_elbo": (C.c_double, [vp, C.c_int, C.c_int]),
        "orc_grad": (None, [vp]),
        "orc_lbfgs": (None, [vp, C.c_int, C.c_int]),
        "orc_downdate": (None, [vp, C.c_int, C.c_double, C.c_int]),
        "orc_beta_from_tilde_beta": (None, [vp]),
        "orc_expand_delta_k": (None, [vp]),
        "orc_fit": (C.c_double, [vp, C.c_int, C.c_int, C.c_int]),
        "orc_fit_trace": (C.c_int, [vp, dp, dp, ip, C.c_int]),
        "orc_init_ta_with_per": (None, [C.c_int, C.c_int, dp]),
        "orc_mt19937_kat": (C.c_uint32, [C.c_uint32, C.c_int]),
        # individual kernels, for pin tests
        "orc_get_kernel_mats": (None, [vp]), "orc_chol_kmm": (C.c_int, [vp]),
        "orc_get_kinv_kmn": (None, [vp]), "orc_chol_delta_inv": (C.c_int, [vp]),
        "orc_get_w_for_beta": (None, [vp]), "orc_get_psi": (None, [vp]),
        "orc_chol_psi": (C.c_int, [vp]), "orc_get_beta_phi": (None, [vp, C.c_int]),
        "orc_get_lkhd_pois": (None, [vp]), "orc_collect_lkhd": (C.c_double, [vp, C.c_int, C.c_int]),
        "orc_get_oinvt": (None, [vp]), "orc_get_lkhd": (None, [vp]),
        "orc_get_tilde_beta": (None, [vp]), "orc_grad_log_delta": (None, [vp]),
        "orc_grad_log_omega": (None, [vp]), "orc_solve_psi": (None, [vp]),
        "orc_get_c2": (None, [vp]), "orc_grad_tau": (None, [vp]), "orc_grad_log_rho": (None, [vp]),
        "orc_score_stats": (None, [vp, C.c_int, dp, C.c_int, ip, ip, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def init_ta_with_per(M: int, Q: int) -> np.ndarray:
    T = np.zeros(M * Q)
    load().orc_init_ta_with_per(M, Q, _dp(T))
    return T


class OracleState:
    """orc_state with the launch-group methods of GPContext."""

    def __init__(self, N, J, M, Q, nh, Yp, Yi, Yx, Zcov, ls, X0, T0, ard_only=False, nthreads=0):
        self._lib = load()
        self.N, self.J, self.M, self.Q = int(N), int(J), int(M), int(Q)
        self.nh = np.ascontiguousarray(nh, dtype=np.int32)
        self.p = len(self.nh) - 1
        self.P = int(self.nh[-1])
        keep = [np.ascontiguousarray(Yp, dtype=np.int32), np.ascontiguousarray(Yi, dtype=np.int32),
                np.ascontiguousarray(Yx, dtype=np.float64),
                np.asarray(Zcov, dtype=np.float64).reshape(self.N, -1, order="F").ravel(order="F").copy(),
                np.ascontiguousarray(ls, dtype=np.float64),
                np.asarray(X0, dtype=np.float64).reshape(self.N, self.Q, order="F").ravel(order="F").copy(),
                np.asarray(T0, dtype=np.float64).reshape(self.M, self.Q, order="F").ravel(order="F").copy()]
        self._keep = keep
        Ypa, Yia, Yxa, Zc, lsa, Xa, Ta = keep
        self._h = self._lib.orc_create(self.N, self.J, self.M, self.Q, self.p, _ip(self.nh), _ip(Ypa),
                                       _ip(Yia), _dp(Yxa), len(Yxa), _dp(Zc), _dp(lsa), _dp(Xa), _dp(Ta),
                                       int(bool(ard_only)), int(nthreads))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self, name: str) -> np.ndarray:
        n = C.c_size_t()
        ptr = self._lib.orc_buf(self._h, name.encode(), C.byref(n))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(n.value,))

    def get(self, name: str) -> np.ndarray:
        return self.view(name).copy()

    def set(self, name: str, arr) -> None:
        v = self.view(name)
        v[:] = np.asarray(arr, dtype=np.float64).ravel()

    def call(self, fn: str, *args):
        return getattr(self._lib, fn)(self._h, *args)

    # launch groups
    def init_fit(self):
        self._lib.orc_init_fit(self._h)

    def irls_step(self, jtr: int) -> float:
        return self._lib.orc_irls_step(self._h, 0, int(jtr))

    def ql_start(self) -> float:
        return self._lib.orc_ql_start(self._h, 0)

    def elbo(self, slot: int = 0) -> float:
        return self._lib.orc_elbo(self._h, 0, int(slot))

    def grad(self):
        self._lib.orc_grad(self._h)

    def lbfgs_step(self, jtr: int, rhoflag: int):
        self._lib.orc_lbfgs(self._h, int(jtr), int(rhoflag))

    def downdate(self, jtr: int, sr: float, rhoflag: int):
        self._lib.orc_downdate(self._h, int(jtr), float(sr), int(rhoflag))

    def end_outer(self):
        self._lib.orc_expand_delta_k(self._h)
        self._lib.orc_beta_from_tilde_beta(self._h)

    def score_stats(self, G, gene, snp) -> np.ndarray:
        """ggplvm_de.cu:491-532 for tests (gene[t], snp[t]); G is N x S."""
        Gf = np.asfortranarray(np.asarray(G, dtype=np.float64).reshape(self.N, -1, order="F"))
        gene = np.ascontiguousarray(gene, dtype=np.int32)
        snp = np.ascontiguousarray(snp, dtype=np.int32)
        out = np.zeros(len(gene))
        self._lib.orc_score_stats(self._h, Gf.shape[1], _dp(Gf.ravel(order="F")), len(gene),
                                  _ip(gene), _ip(snp), _dp(out))
        return out

    def fit(self, itrmax: int, jtr_cap: int = 0, verbose: bool = False) -> float:
        return self._lib.orc_fit(self._h, int(itrmax), int(jtr_cap), int(bool(verbose)))

    def fit_trace(self):
        """Trace of the last fit: objectives, final line-search sr, and an
        (entries x 4) int array of (kind, itr, jtr, ktr); kind 0 = IRLS step,
        1 = quasi-likelihood start, 2 = accepted L-BFGS step."""
        n = self._lib.orc_fit_trace(self._h, None, None, None, 0)
        obj, sr = np.zeros(n), np.zeros(n)
        info = np.zeros(4 * n, dtype=np.int32)
        self._lib.orc_fit_trace(self._h, _dp(obj), _dp(sr), _ip(info), n)
        return obj, sr, info.reshape(n, 4)
