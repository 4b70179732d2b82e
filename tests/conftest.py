import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def small_ds():
    from gaspacho2_amd import synth
    return synth.make_dataset(300, 20, 3, 8, seed=0, n_numeric=2, n_levels=3)


@pytest.fixture(scope="session")
def tiny_ds():
    from gaspacho2_amd import synth
    return synth.make_dataset(120, 6, 3, 5, seed=3, n_numeric=1, n_levels=2)


def make_oracle(ds, T0=None, ard_only=False):
    from oracle.oracle import OracleState, init_ta_with_per
    if T0 is None:
        T0 = init_ta_with_per(ds.M, ds.Q)
    return OracleState(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                       ds.X0, T0, ard_only=ard_only)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def fit_jtrmax(itr, phase, cap=0):
    """JTRMAX of ggplvm.cu:1152, 1213 (shared by both phases, reset each outer
    iteration): 200 at itr 0, else 100, and 10000 for the L-BFGS phase of
    itr 30; a positive cap bounds every inner loop."""
    j = 200 if itr == 0 else 100
    if phase == 2 and itr == 30:
        j = 10000
    return min(j, cap) if cap > 0 else j


def check_fit_schedule(obj, info, sr, itrmax, cap):
    """Assert that a fit trace (objective, (kind, itr, jtr, ktr), final sr per
    entry) follows the reference schedule ggplvm.cu:1146-1330; returns counts
    of the rules the trace reached."""
    kind, itr, jtr, ktr = np.asarray(info).T
    assert np.all(np.diff(itr) >= 0) and itr[-1] == itrmax - 1
    seen = dict(irls_converged=0, irls_cap=0, lbfgs_converged=0, lbfgs_cap=0, downdates=0,
                long_phase_steps=0)
    for t in range(itrmax):
        sel = np.flatnonzero(itr == t)
        k = kind[sel]
        irls, ql, lb = sel[k == 0], sel[k == 1], sel[k == 2]
        assert len(ql) == 1 and irls.max() < ql[0] < lb.min(), t
        assert np.array_equal(jtr[irls], np.arange(len(irls))), t
        lk = obj[irls]
        jmax = fit_jtrmax(t, 0, cap)
        conv = [j > 0 and abs(lk[j - 1] - lk[j]) / abs(lk[j]) < 1e-7 for j in range(len(lk))]
        stop = [conv[j] or j == jmax - 1 for j in range(len(lk))]
        assert stop[-1] and not any(stop[:-1]), t
        seen["irls_converged" if conv[-1] else "irls_cap"] += 1
        assert np.array_equal(jtr[lb], np.arange(len(lb))), t
        seq = [obj[ql[0]]] + list(obj[lb])
        for m, e in enumerate(lb):
            assert 1 <= ktr[e] <= 7, (t, m)
            assert seq[m + 1] > seq[m] or ktr[e] == 7, (t, m)
            n = ktr[e] - 1
            allowed = [0.1 ** a * 0.01 ** (n - a) for a in range(n + 1)]
            assert min(abs(sr[e] - v) for v in allowed) <= 1e-15, (t, m, sr[e])
            seen["downdates"] += n
        jmax = fit_jtrmax(t, 2, cap)
        conv = [m > 0 and abs(seq[m + 1] - seq[m]) < abs(seq[m + 1]) * 1e-8 for m in range(len(lb))]
        stop = [conv[m] or m == jmax - 1 for m in range(len(lb))]
        assert stop[-1] and not any(stop[:-1]), t
        seen["lbfgs_converged" if conv[-1] else "lbfgs_cap"] += 1
        if t == 30:
            seen["long_phase_steps"] = len(lb)
    return seen
