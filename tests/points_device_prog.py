"""Evaluate one point set through wave_EvalPoints (host arrays) and
wave_EvalPointsDevice (device arrays) after wave_Initialize, in a process of
its own (the library keeps one process-global wave field).

Usage: python points_device_prog.py <libpath> <workdir> <points.npz> <out.npz>
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main(lib, wd, pts_path, out_path):
    import torch                                      # one HIP runtime: torch's, loaded first
    sys.path.insert(0, ROOT)
    from cfdwavemaker_amd._lib import load
    P = np.load(pts_path)
    x, y, z, t = (np.ascontiguousarray(P[k], dtype=np.float64) for k in ("x", "y", "z", "t"))
    os.chdir(wd)
    L = load(os.path.abspath(lib))
    rc0 = L.wave_Initialize()
    n = x.size
    flags = 0x1F
    host = [np.zeros(n) for _ in range(5)]
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc1 = L.wave_EvalPoints(n, ptr(x), ptr(y), ptr(z), ptr(t), *[ptr(h) for h in host], flags)
    dev = torch.device("cuda:0")
    X, Y, Z, T = (torch.tensor(a, dtype=torch.float64, device=dev) for a in (x, y, z, t))
    D = [torch.full((n,), float("nan"), dtype=torch.float64, device=dev) for _ in range(5)]
    torch.cuda.synchronize()
    rc2 = L.wave_EvalPointsDevice(n, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), T.data_ptr(),
                                  *[d.data_ptr() for d in D], flags)
    torch.cuda.synchronize()
    np.savez(out_path, rc=np.array([rc0, rc1, rc2]), host=np.array(host),
             dev=np.array([d.cpu().numpy() for d in D]))


if __name__ == "__main__":
    main(*sys.argv[1:5])
