"""Prefetch micro-batcher of the per-point API: after wave_PrefetchPoints,
per-point calls with the same arguments return exactly the uncached values,
without a GPU round trip per call."""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from tests.conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm.so")

PROGRAM = r'''
import ctypes as C, json, os, sys, time
import numpy as np
os.chdir(sys.argv[2])
L = C.CDLL(sys.argv[1]); d = C.c_double; P = C.POINTER(d)
for f in ("wave_VeloX", "wave_VeloY", "wave_VeloZ", "wave_DynPres"):
    getattr(L, f).restype = d; getattr(L, f).argtypes = [d, d, d, d]
L.wave_SurfElev.restype = d; L.wave_SurfElev.argtypes = [d, d, d]
L.wave_Kinematics.restype = P; L.wave_Kinematics.argtypes = [d, d, d, d]
L.wave_PrefetchPoints.restype = C.c_int
L.wave_PrefetchPoints.argtypes = [C.c_int64, P, P, P, P, C.c_uint32]
L.wave_PrefetchStats.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
assert L.wave_Initialize() == 0
wt = int(sys.argv[3])
rng = np.random.default_rng(5)
n = 200
x, y = rng.uniform(0, 60, n), rng.uniform(-5, 5, n)
z, t = rng.uniform(-8, -0.5, n), np.full(n, 2.25)
def calls():
    out = np.zeros((n, 8))
    for q in range(n):
        a = (x[q], y[q], z[q], t[q])
        out[q, 0] = L.wave_SurfElev(x[q], y[q], t[q])
        out[q, 1] = L.wave_VeloX(*a); out[q, 2] = L.wave_VeloY(*a); out[q, 3] = L.wave_VeloZ(*a)
        out[q, 4] = L.wave_DynPres(*a) if wt == 1 else 0.0
        k = L.wave_Kinematics(*a); out[q, 5:8] = [k[0], k[1], k[3]]
    return out
calls()
t0 = time.perf_counter(); A = calls(); t_unc = time.perf_counter() - t0
ptr = lambda v: np.ascontiguousarray(v).ctypes.data_as(P)
assert L.wave_PrefetchPoints(n, ptr(x), ptr(y), ptr(z), ptr(t), 0x1F) == 0
t0 = time.perf_counter(); B = calls(); t_c = time.perf_counter() - t0
h, m = C.c_int64(), C.c_int64(); L.wave_PrefetchStats(C.byref(h), C.byref(m))
print(json.dumps({"maxdiff": float(np.max(np.abs(A - B))), "scale": float(np.max(np.abs(A))),
                  "hits": h.value, "misses": m.value, "us_per_call_uncached": t_unc / (7 * n) * 1e6,
                  "us_per_call_cached": t_c / (7 * n) * 1e6}))
'''


@pytest.mark.gpu
@pytest.mark.parametrize("fixture,wt", [("abi_direct.npz", 1), ("abi_grid.npz", 5)])
def test_prefetched_points_equal_uncached(fixture, wt):
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        r = subprocess.run([sys.executable, "-c", PROGRAM, LIB, wd, str(wt)], capture_output=True, text=True,
                           timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        res = json.loads(r.stdout.strip().splitlines()[-1])
        print(fixture, res, flush=True)
        assert res["maxdiff"] == 0.0 and res["scale"] > 0
        assert res["hits"] >= 6 * 200
        assert res["us_per_call_cached"] < res["us_per_call_uncached"]
