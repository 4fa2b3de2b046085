"""Run a golden call program (see tests/golden/ref_abi.py for the op codes)
through the reference-style Python wrapper cfdwavemaker_amd.cfdwavemaker
(the class and method names of examples/python/cfdwavemaker.py), in a process
of its own: the wave field is process-global.

Usage: python wrapper_prog.py <workdir> <program.npz> <out.npz>

Besides the program's outputs it writes the batched eval_points of every
point op of the program (before any force_update), for the test to compare
with the per-point calls."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from cfdwavemaker_amd.cfdwavemaker import CFDwavemaker  # noqa: E402


def main(wd, prog_path, out_path):
    prog = dict(np.load(prog_path))
    wm = CFDwavemaker(workdir=wd)
    rc = wm.initialize()
    op, a = prog["op"], prog["args"]
    out = np.zeros((len(op), 4))
    first_update = next((q for q in range(len(op)) if int(op[q]) == 6), len(op))
    pts = [q for q in range(first_update) if int(op[q]) in (0, 1, 2, 3, 5)]
    batch = wm.eval_points(a[pts, 0], a[pts, 1], a[pts, 2], a[pts, 3], outputs=("eta", "u", "v", "w"))
    for q in range(len(op)):
        o = int(op[q])
        x, y, z, t = a[q]
        if o == 0:
            out[q, 0] = wm.surface_elevation(x, y, t)
        elif o in (1, 2, 3, 4):
            f = (wm.velocity_x, wm.velocity_y, wm.velocity_z, wm.dynamic_pressure)[o - 1]
            out[q, 0] = f(x, y, z, t)
        elif o == 5:
            out[q, :] = wm.kinematics(x, y, z, t)
        elif o == 6:
            wm.force_update(t)
        elif o in (7, 8, 9):
            f = (wm.phase_speed, wm.mean_period, wm.mean_length)[o - 7]
            out[q, 0] = f(int(x))
        elif o == 10:
            out[q, 0] = wm.water_depth()
    ok = wm.is_initialized()
    wm.cleanup()
    np.savez(out_path, rc=rc, out=out, pts=np.array(pts, dtype=np.int64),
             batch=np.stack([batch[k] for k in ("eta", "u", "v", "w")], 1), initialized=int(ok))


if __name__ == "__main__":
    main(*sys.argv[1:4])
