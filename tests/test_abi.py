"""The C-ABI boundary: libcfdwm loads and exports every symbol include/cfdwm.h
declares (CPU), and the reference's own call programs give the reference's
answers through it (GPU)."""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from tests.conftest import ROOT, golden, scaled_err

HEADER = os.path.join(ROOT, "include", "cfdwm.h")
LIB = os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"CFDWM_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


def test_header_lists_reference_api():
    from cfdwavemaker_amd._lib import REFERENCE_API, BATCH_API
    decl = set(declared_symbols())
    assert set(REFERENCE_API) <= decl
    assert set(BATCH_API) <= decl


def test_library_exports_every_declared_symbol():
    from cfdwavemaker_amd._lib import load
    assert os.path.exists(LIB), "build libcfdwm.so first (make -C cfdwavemaker_amd/csrc)"
    L = load(LIB)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def run_program(fixture, files=False):
    """Run a fixture's call program through libcfdwm in a process of its own;
    with files, also the text of every file the run wrote (including files
    written at process exit)."""
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        pp = os.path.join(wd, "prog.npz")
        op = os.path.join(wd, "out.npz")
        np.savez(pp, op=g["op"], args=g["args"])
        r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "golden", "ref_abi.py"), LIB, wd, pp, op],
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        o = np.load(op)
        if not files:
            return g, int(o["rc"]), o["out"]
        written = {}
        for root, _, names in os.walk(wd):
            for nm in names:
                if nm not in ("waveinput.dat", "prog.npz", "out.npz"):
                    p = os.path.join(root, nm)
                    written[os.path.relpath(p, wd)] = open(p).read()
        return g, int(o["rc"]), o["out"], written


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["abi_direct.npz", "abi_grid.npz", "abi_smit.npz", "abi_t005.npz"])
def test_reference_programs_through_abi(fixture):
    g, rc, out = run_program(fixture)
    assert rc == int(g["rc"]) == 0
    ref = g["out"]
    ops = g["op"]
    for op in np.unique(ops):
        sel = ops == op
        for c in range(4):
            err = scaled_err(out[sel, c], ref[sel, c])
            print(f"{fixture} op {op} col {c} err {err:.2e}", flush=True)
            assert err <= 1e-9, (fixture, int(op), c, err)
    if fixture == "abi_t005.npz":
        # the reference's own targetwave.dat (tests/t005_focused_wave)
        assert np.max(np.abs(out[:, 0] - g["targetwave"][:, 1])) <= 1e-9 * np.max(np.abs(g["targetwave"][:, 1]))


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["abi_direct.npz", "abi_grid.npz"])
def test_reference_style_wrapper(fixture):
    """The same call programs through cfdwavemaker_amd.cfdwavemaker.CFDwavemaker,
    the class and method names of the reference's examples/python/cfdwavemaker.py:
    its answers equal the compiled reference's (1e-9 scale-normalised), and its
    batched eval_points equals its per-point calls (rounding level: wave type 1
    batches sum in another order)."""
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(str(g["waveinput"]))
        pp, op_ = os.path.join(wd, "prog.npz"), os.path.join(wd, "out.npz")
        np.savez(pp, op=g["op"], args=g["args"])
        r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "wrapper_prog.py"), wd, pp, op_],
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        o = dict(np.load(op_))
    assert int(o["rc"]) == 0 and int(o["initialized"]) == 1
    out, ref, ops = o["out"], g["out"], g["op"]
    for op in np.unique(ops):
        sel = ops == op
        for c in range(4):
            err = scaled_err(out[sel, c], ref[sel, c])
            assert err <= 1e-9, (fixture, int(op), c, err)
    pts, batch = o["pts"], o["batch"]
    col = {0: 0, 1: 1, 2: 2, 3: 3}
    for k, q in enumerate(pts):
        opq = int(ops[q])
        got = batch[k] if opq == 5 else batch[k, col[opq]:col[opq] + 1]
        want = out[q] if opq == 5 else out[q, :1]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max())), (q, got, want)


@pytest.mark.gpu
def test_missing_waveinput_returns_minus_one():
    code = ("import ctypes,os,sys; os.chdir(sys.argv[2]); L=ctypes.CDLL(sys.argv[1]); "
            "sys.exit(0 if L.wave_Initialize()==-1 and L.CFDwavemaker_is_initialized()==0 else 3)")
    with tempfile.TemporaryDirectory() as wd:
        sub = os.path.join(wd, "a", "b")
        os.makedirs(sub)
        r = subprocess.run([sys.executable, "-c", code, LIB, sub], capture_output=True, timeout=120)
        assert r.returncode == 0, r.stderr
