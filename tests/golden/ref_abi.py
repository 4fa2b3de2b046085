"""Drive a CFDwavemaker-compatible C ABI with a small "program" of calls.

Usage: python ref_abi.py <libpath> <workdir> <program.npz> <out.npz>

The library reads ``waveinput.dat`` from ``workdir`` (wave_Initialize), then
executes the program: ``op`` (int array) and ``args`` (n x 4 doubles).  Each
call writes up to four doubles into ``out`` (n x 4).  Op codes:

  0 SurfElev(x, y, t)            1 VeloX(x, y, z, t)      2 VeloY      3 VeloZ
  4 DynPres(x, y, z, t)          5 Kinematics(x, y, z, t) -> 4 values
  6 force_update(t)              7 phase_speed(opt)       8 mean_period(opt)
  9 mean_length(opt)            10 water_depth()

The same driver runs the compiled reference (tests/golden/make_golden.py) and
this project's libcfdwm (tests/test_abi.py), each in its own process
because both keep one process-global wave field.
"""
import ctypes as C
import os
import sys

import numpy as np


def run(libpath, workdir, prog):
    libpath = os.path.abspath(libpath)
    os.chdir(workdir)
    L = C.CDLL(libpath)
    d = C.c_double
    L.wave_Initialize.restype = C.c_int
    for f in ("wave_VeloX", "wave_VeloY", "wave_VeloZ", "wave_DynPres"):
        getattr(L, f).restype = d
        getattr(L, f).argtypes = [d, d, d, d]
    L.wave_SurfElev.restype = d
    L.wave_SurfElev.argtypes = [d, d, d]
    L.wave_Kinematics.restype = C.POINTER(d)
    L.wave_Kinematics.argtypes = [d, d, d, d]
    L.wave_force_update.restype = None
    L.wave_force_update.argtypes = [d]
    for f in ("wave_phase_speed", "wave_mean_period", "wave_mean_length"):
        getattr(L, f).restype = d
        getattr(L, f).argtypes = [C.c_int]
    L.wave_water_depth.restype = d
    rc = L.wave_Initialize()
    op, a = prog["op"], prog["args"]
    out = np.zeros((len(op), 4))
    for q in range(len(op)):
        o = int(op[q])
        x, y, z, t = a[q]
        if o == 0:
            out[q, 0] = L.wave_SurfElev(x, y, t)
        elif o in (1, 2, 3, 4):
            f = (L.wave_VeloX, L.wave_VeloY, L.wave_VeloZ, L.wave_DynPres)[o - 1]
            out[q, 0] = f(x, y, z, t)
        elif o == 5:
            p = L.wave_Kinematics(x, y, z, t)
            out[q, :] = [p[0], p[1], p[2], p[3]]
        elif o == 6:
            L.wave_force_update(t)
        elif o in (7, 8, 9):
            f = (L.wave_phase_speed, L.wave_mean_period, L.wave_mean_length)[o - 7]
            out[q, 0] = f(int(x))
        elif o == 10:
            out[q, 0] = L.wave_water_depth()
    return rc, out


if __name__ == "__main__":
    lib, wd, prog_path, out_path = sys.argv[1:5]
    prog = dict(np.load(prog_path))
    rc, out = run(lib, wd, prog)
    np.savez(out_path, rc=rc, out=out)
