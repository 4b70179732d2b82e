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
