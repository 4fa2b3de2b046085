"""Per-point API caches.

wave_PrefetchPoints: after an explicit prefetch, per-point calls with the
same arguments return the kept values without a GPU round trip per call:
bit-equal to the uncached values on the s-grid (wave type 5), equal to
rounding level in direct mode (wave type 1), where the batch sums the same
terms in another order (segments, and for FP32 batches of 2048 points or more
the eta2 quadratic form).

Automatic record-and-replay on the s-grid: the points of one time value are
evaluated in one batch at the next time value, and per-point calls from 16
host threads are then served from the kept values, bit-equal to uncached
calls."""
import json
import os
import subprocess
import sys
import tempfile

import pytest

from tests.conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm.so")
DRV = os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm_hostdrv.so")

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
L.wave_PrefetchAuto(0)
wt, n, xmax = int(sys.argv[3]), int(sys.argv[4]), float(sys.argv[5])
rng = np.random.default_rng(5)
x, y = rng.uniform(0, xmax, n), rng.uniform(-5, 5, n)
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
scale = np.maximum(np.max(np.abs(A), axis=0), 1e-300)
print(json.dumps({"maxdiff": float(np.max(np.abs(A - B))), "scaled": float(np.max(np.abs(A - B) / scale)),
                  "scale": float(np.max(np.abs(A))), "hits": h.value, "misses": m.value,
                  "us_per_call_uncached": t_unc / (7 * n) * 1e6, "us_per_call_cached": t_c / (7 * n) * 1e6}))
'''


def _run(fixture, args, env=None):
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        r = subprocess.run([sys.executable, "-c", PROGRAM, LIB, wd] + [str(a) for a in args], capture_output=True,
                           text=True, timeout=600, env=dict(os.environ, **(env or {})))
        assert r.returncode == 0, r.stderr[-2000:]
        return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.gpu
@pytest.mark.parametrize("fixture,wt", [("abi_direct.npz", 1), ("abi_grid.npz", 5)])
def test_prefetched_points_equal_uncached(fixture, wt):
    res = _run(fixture, [wt, 200, 60.0])
    print(fixture, res, flush=True)
    if wt == 5:
        assert res["maxdiff"] == 0.0
    else:
        assert res["scaled"] <= 1e-12
    assert res["scale"] > 0 and res["hits"] >= 6 * 200
    assert res["us_per_call_cached"] < res["us_per_call_uncached"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_prefetch_direct_large_batch(precision, tol):
    """Direct mode, 4096 points (FP32 batches of this size take eta2 from the
    MFMA quadratic form, single points the pair loop): kept values equal the
    uncached ones to the stated bound."""
    res = _run("abi_direct.npz", [1, 4096, 120.0], env={"CFDWM_PRECISION": precision})
    print(precision, res, flush=True)
    assert res["scaled"] <= tol and res["hits"] >= 6 * 4096


AUTO = r'''
import ctypes as C, json, os, sys
import numpy as np
os.chdir(sys.argv[2])
L = C.CDLL(sys.argv[1]); H = C.CDLL(sys.argv[3])
d = C.c_double; P = C.POINTER(d); I = C.c_int64
H.cfdwm_host_sweep.argtypes = [C.c_int, I, P, P, P, d, P, P]
L.wave_PrefetchAutoStats.argtypes = [C.POINTER(I)] * 4
L.wave_force_update.argtypes = [d]
assert L.wave_Initialize() == 0
n = int(sys.argv[4])
rng = np.random.default_rng(8)
x, y, z = rng.uniform(0, 110, n), rng.uniform(-18, 18, n), rng.uniform(-25, -0.5, n)
ptr = lambda v: v.ctypes.data_as(P)
def sweep(t, threads=16):
    out = np.zeros(4 * n); sec = d()
    assert H.cfdwm_host_sweep(threads, n, ptr(x), ptr(y), ptr(z), t, ptr(out), C.byref(sec)) == 0
    return out, sec.value
res = {}
sweep(0.10)                        # generation 1: every point queried on its own, remembered
A, res["s_replayed_step"] = sweep(0.20)      # generation 2: one batched replay, then hits
L.wave_PrefetchAuto(0)
B, res["s_uncached"] = sweep(0.20)           # the same calls without the cache
L.wave_PrefetchAuto(1)
L.wave_force_update(0.80)          # the grid advances: a new generation at the same points
C2, _ = sweep(0.80)
st = [I() for _ in range(4)]; L.wave_PrefetchAutoStats(*[C.byref(v) for v in st])
res["hits"], res["misses"], res["replays"], res["replayed"] = [v.value for v in st]
L.wave_PrefetchAuto(0)
D, _ = sweep(0.80)
res["equal"] = bool(np.array_equal(A, B)) and bool(np.array_equal(C2, D))
res["us_per_call_cached"] = res["s_replayed_step"] / (4 * n) * 1e6
res["us_per_call_uncached"] = res["s_uncached"] / (4 * n) * 1e6
print(json.dumps(res))
'''


@pytest.mark.gpu
def test_auto_prefetch_sgrid_threads():
    """Wave type 5, 2000 points, 16 host threads: the second time value is
    served from one replayed batch (bit-equal to uncached calls, also after
    wave_force_update), at a fraction of the uncached per-call cost."""
    g = golden("abi_grid.npz")
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        r = subprocess.run([sys.executable, "-c", AUTO, LIB, wd, DRV, "2000"], capture_output=True, text=True,
                           timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res, flush=True)
    assert res["equal"]
    assert res["replays"] >= 2 and res["replayed"] >= 2 * 2 * 2000
    assert res["hits"] >= 2 * 4 * 2000
    assert res["us_per_call_cached"] < res["us_per_call_uncached"]
