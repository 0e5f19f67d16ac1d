import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref built from the reference sources")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def have_ref():
    from oracle import oracle as O
    return O.ref_available()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def fields_equal(a, b):
    """Bitwise equality: every element that is not NaN has the same bit
    pattern in both arrays (so +0 and -0 differ), and both arrays have NaN at
    the same positions.  NaN payloads are not compared: the GPU's default NaN
    and the host's differ in sign bit, and no finite result depends on them."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if not np.issubdtype(a.dtype, np.floating):
        return np.array_equal(a, b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    view = np.dtype(f"u{a.dtype.itemsize}")
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return np.array_equal(a.view(view)[~na], b.view(view)[~nb])


def all_finite(a):
    """Every element finite: a parity claim on a solution is only a claim
    when the compared values are numbers."""
    return bool(np.isfinite(np.asarray(a)).all())
