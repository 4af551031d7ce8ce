import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("pgfem-3d_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")


def rel_inf(a, b):
    """max |a-b| / max |b| (infinity-norm relative error; 0 when both are zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    return 0.0 if num == 0.0 else num / max(den, 1e-300)


def rel_cols(a, b, floor=1e-6):
    """Largest error of a [comp][n] array relative to each column's own size: max over columns q
    of max_c |a[c,q] - b[c,q]| / max(max_c |b[c,q]|, floor * max |b|).  Holds each QP (or
    element) to its own magnitude rather than the array's largest entry; the floor only keeps
    columns that are zero to rounding from dividing by zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.ndim == 1:
        a, b = a[None], b[None]
    big = np.abs(b).max() if b.size else 0.0
    if big == 0.0:
        return float(np.abs(a).max()) if a.size else 0.0
    scale = np.maximum(np.abs(b).max(axis=0), floor * big)
    return float((np.abs(a - b).max(axis=0) / scale).max())


def rel_csr_rows(a, b, Ap, floor=1e-6):
    """rel_cols for CSR values: each row's error relative to that row's largest entry."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    Ap = np.asarray(Ap, dtype=np.int64)
    full = np.diff(Ap) > 0  # empty rows hold nothing to compare
    if not full.any():
        return 0.0
    rows = np.flatnonzero(full)
    big = np.abs(b).max()
    worst = 0.0
    step = max(1, len(rows) * (1 << 24) // max(len(b), 1))  # about 16M entries per chunk
    for k in range(0, len(rows), step):
        rs = rows[k:k + step]
        lo, hi = Ap[rs[0]], Ap[rs[-1] + 1]
        starts = Ap[rs] - lo
        row_max = np.maximum.reduceat(np.abs(b[lo:hi]), starts)
        row_err = np.maximum.reduceat(np.abs(a[lo:hi] - b[lo:hi]), starts)
        scale = np.maximum(np.maximum(row_max, floor * big), 1e-300)
        worst = max(worst, float((row_err / scale).max()))
    return worst
