"""Shared fixtures.  Tests marked ``gpu`` need an MI355X (they call libpmg through its
C ABI); everything else runs on the CPU against the FP64 oracle."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libpmg.so")


def _have_gpu() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def params(**kw):
    from to_merge_master_parmoon_amd import _abi
    base = dict(sc_gmres_restart=30, sc_lin_red_factor_saddle=1e-10, sc_lin_res_norm_min_saddle=0.0,
                sc_lin_maxit_saddle=200)
    base.update(kw)
    return _abi.default_params(**base)


def unpin_pressure(x, n_udof):
    """x with the mean of the P1disc constants (every 4th pressure entry) removed: with
    Dirichlet data on the whole boundary the pressure is fixed only up to a constant."""
    y = np.array(x, dtype=np.float64, copy=True)
    y[3 * n_udof::4] -= y[3 * n_udof::4].mean()
    return y


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


@pytest.fixture(scope="session")
def oseen8():
    from to_merge_master_parmoon_amd import problem
    return problem.oseen3d(8)


@pytest.fixture(scope="session")
def oseen16():
    from to_merge_master_parmoon_amd import problem
    return problem.oseen3d(16)
