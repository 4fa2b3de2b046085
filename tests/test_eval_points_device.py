"""wave_EvalPointsDevice keeps its work on the device (z clamp, batch,
elevation sum and ramps in device passes) and returns bit for bit what
wave_EvalPoints returns for the same points, under the reference programs'
waveinput.dat files: direct wave type with time and x ramps, the s-grid
(also with a batch large enough for cell-ordered queries) and the Smit
vertical theory."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from tests.conftest import ROOT, golden

LIB = os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm.so")
PROG = os.path.join(ROOT, "tests", "points_device_prog.py")

pytestmark = pytest.mark.gpu


def _points(g, n, seed):
    a = np.asarray(g["args"])[np.asarray(g["op"]) <= 5]
    lo, hi = a.min(0), a.max(0)
    rng = np.random.default_rng(seed)
    # a margin around the programs' points: clamps below the seabed, points
    # above the surface, outside the ramps and the grid
    x = rng.uniform(lo[0] - 5.0, hi[0] + 5.0, n)
    y = rng.uniform(lo[1] - 5.0, hi[1] + 5.0, n)
    z = rng.uniform(lo[2] - 10.0, hi[2] + 2.0, n)
    t = rng.uniform(lo[3], hi[3] + 1.0, n)
    return dict(x=x, y=y, z=z, t=t)


@pytest.mark.parametrize("n", [3000, 70000])
@pytest.mark.parametrize("fixture", ["abi_direct.npz", "abi_grid.npz", "abi_smit.npz"])
def test_eval_points_device_matches_host(fixture, n):
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        pp, op = os.path.join(wd, "pts.npz"), os.path.join(wd, "out.npz")
        np.savez(pp, **_points(g, n, 5 + n))
        r = subprocess.run([sys.executable, PROG, LIB, wd, pp, op], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        o = np.load(op)
    assert list(o["rc"]) == [0, 0, 0]
    host, dev = o["host"], o["dev"]
    assert np.all(np.isfinite(host))
    for c, name in enumerate(("eta", "u", "v", "w", "p")):
        assert np.array_equal(host[c], dev[c]), (fixture, name, float(np.max(np.abs(host[c] - dev[c]))))
