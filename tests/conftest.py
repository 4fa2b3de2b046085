import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    """The FP64 CPU oracle (built here if the shared object is missing)."""
    from oracle import Oracle, HERE
    if not os.path.exists(os.path.join(HERE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return Oracle()


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def sea_from_golden(g, case):
    """Rebuild a sea dictionary stored by tests/golden/make_golden.py."""
    d = {}
    for key in ("omega", "A", "k", "phase", "theta", "ramptime", "starttime", "duration"):
        if f"{case}__{key}" in g:
            d[key] = g[f"{case}__{key}"]
    for key in ("depth", "g", "rho", "swl", "mtheta", "tofmax", "x0", "y0", "bandwidth", "cutoff"):
        d[key] = float(g[f"{case}__{key}"])
    d["order"] = int(g[f"{case}__order"])
    return d


def scaled_err(a, b):
    """max |a-b| / max(max|b|, rms(b), tiny): the scale-normalised error."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.max(np.abs(b)) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if b.size else 0.0


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False
