"""Reference-style Python binding of libcfdwm.

Same class, method names and argument meaning as the reference's ctypes
wrapper (examples/python/cfdwavemaker.py:41-217), pointed at this library,
plus ``eval_points`` for the batched entry.  ``waveinput.dat`` is read from
the working directory by ``initialize()`` (wave_Initialize).
"""
from __future__ import annotations

import os

import numpy as np

from ._lib import check, load

OUT_ETA, OUT_U, OUT_V, OUT_W, OUT_P, FP32 = 0x1, 0x2, 0x4, 0x8, 0x10, 0x100


class CFDwavemaker:
    def __init__(self, libpath=None, workdir=None):
        if workdir is not None:
            os.chdir(workdir)
        self.lib = load(libpath)

    # lifecycle
    def initialize(self):
        return self.lib.wave_Initialize()

    def cleanup(self):
        return self.lib.wave_Cleanup()

    def is_initialized(self):
        return bool(self.lib.CFDwavemaker_is_initialized())

    def set_mpi_comm(self, fortran_comm):
        self.lib.wave_set_mpi_comm(int(fortran_comm))

    # point queries
    def velocity_x(self, x, y, z, t):
        return self.lib.wave_VeloX(x, y, z, t)

    def velocity_y(self, x, y, z, t):
        return self.lib.wave_VeloY(x, y, z, t)

    def velocity_z(self, x, y, z, t):
        return self.lib.wave_VeloZ(x, y, z, t)

    def velocity(self, x, y, z, t):
        return (self.lib.wave_VeloX(x, y, z, t), self.lib.wave_VeloY(x, y, z, t),
                self.lib.wave_VeloZ(x, y, z, t))

    def dynamic_pressure(self, x, y, z, t):
        return self.lib.wave_DynPres(x, y, z, t)

    def nonhydrostatic_pressure(self, x, y, z, t):
        return self.lib.wave_nhPres(x, y, z, t)

    def surface_elevation(self, x, y, t):
        return self.lib.wave_SurfElev(x, y, t)

    def volume_fraction(self, x, y, z, t, delta_cellsize):
        return self.lib.wave_VFrac(x, y, z, t, delta_cellsize)

    def seabed(self, x, y):
        return self.lib.wave_Seabed(x, y)

    def piston_position(self, t):
        return self.lib.wave_piston_position(t)

    def kinematics(self, x, y, z, t, n=4):
        ptr = self.lib.wave_Kinematics(x, y, z, t)
        return [ptr[i] for i in range(n)]

    # grid / probes / helpers
    def force_update(self, t):
        self.lib.wave_force_update(t)

    def update_probes(self, t):
        self.lib.update_probes(t)

    def phase_speed(self, opt=1):
        return self.lib.wave_phase_speed(opt)

    def mean_period(self, opt=1):
        return self.lib.wave_mean_period(opt)

    def mean_length(self, opt=1):
        return self.lib.wave_mean_length(opt)

    def water_depth(self):
        return self.lib.wave_water_depth()

    # batched entry (no counterpart in the reference)
    def eval_points(self, x, y, z, t, outputs=("eta", "u", "v", "w", "p"), fp32=False):
        """Evaluate arrays of points in one launch; returns a dict of arrays."""
        x, y, z, t = (np.ascontiguousarray(np.ravel(a), dtype=np.float64)
                      for a in np.broadcast_arrays(x, y, z, t))
        n = x.size
        res = {k: np.zeros(n) for k in outputs}
        import ctypes as C
        P = C.POINTER(C.c_double)
        ptr = lambda a: a.ctypes.data_as(P)   # noqa: E731
        args = [ptr(res[k]) if k in res else None for k in ("eta", "u", "v", "w", "p")]
        rc = self.lib.wave_EvalPoints(n, ptr(x), ptr(y), ptr(z), ptr(t), *args, FP32 if fp32 else 0)
        check(rc, "wave_EvalPoints")
        return res
