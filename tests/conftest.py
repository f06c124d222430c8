import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gvv-differentiable-cuda-renderer_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libgvv_amd.so")


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
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def rel_err(x, ref):
    """Max-norm relative error max|x - ref| / max|ref| (0 when both are all zero)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max() if ref.size else 0.0
    diff = np.abs(x - ref).max() if ref.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def elem_rel_stats(x, ref, floor=1e-3):
    """Elementwise |x-ref|/|ref| over entries with |ref| > floor*max|ref|: (max, median)."""
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    if ref.size == 0 or np.abs(ref).max() == 0:
        return 0.0, 0.0
    m = np.abs(ref) > floor * np.abs(ref).max()
    r = np.abs(x[m] - ref[m]) / np.abs(ref[m])
    return float(r.max()), float(np.median(r))


def pytest_terminal_summary(terminalreporter):
    """Accuracy of every GPU-vs-oracle comparison at the full workload sizes (cases whose name
    starts with 'full'), one line per output tensor."""
    pu = sys.modules.get("parity_util")
    recs = [r for r in getattr(pu, "ACCURACY", []) if r["case"].startswith("full")]
    if not recs:
        return
    tr = terminalreporter
    tr.write_sep("-", "accuracy against the oracle at full size (relative errors; face: mismatching ids)")
    for r in recs:
        if r["output"] == "face":
            tr.write_line(f"{r['case']:42s} face    mismatches {int(r['maxnorm'])} of {r['n']}")
        else:
            tr.write_line(f"{r['case']:42s} {r['output']:7s} maxnorm {r['maxnorm']:.1e} elem max {r['elem_max']:.1e} "
                          f"p99.9 {r['elem_p999']:.1e} median {r['elem_median']:.1e}")
