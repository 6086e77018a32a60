import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
import oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def orc():
    return oracle


@pytest.fixture(scope="session")
def ctr():
    return __graft_entry__.load()


@pytest.fixture(scope="session")
def dev(ctr):
    import torch
    ctr.require_device()
    return torch.device("cuda:0")


def _wide(a):
    """float64, or complex128 when the input is complex (so that |got - ref|
    includes the imaginary parts)."""
    a = np.asarray(a)
    return a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)


def rel_max(got, ref):
    got, ref = _wide(got), _wide(ref)
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


def rel_rmse(got, ref):
    got, ref = _wide(got), _wide(ref)
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


def error_metrics(got, ref):
    """rel RMSE, max-abs, max-rel (over max |ref|) and PSNR = 20 log10(max |ref|
    / RMSE) of `got` against `ref`."""
    import math
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = got - ref
    rmse = math.sqrt(float(np.mean(d * d)))
    peak = float(np.abs(ref).max())
    return {
        "rel_rmse": float(np.linalg.norm(d) / np.linalg.norm(ref)),
        "max_abs": float(np.abs(d).max()),
        "max_rel": float(np.abs(d).max()) / peak,
        "psnr_db": 20.0 * math.log10(peak / rmse) if rmse > 0 else float("inf"),
    }


def format_metrics(m):
    return (f"rel RMSE {m['rel_rmse']:.2e}, max-abs {m['max_abs']:.2e}, "
            f"max-rel {m['max_rel']:.2e}, PSNR {m['psnr_db']:.1f} dB")
