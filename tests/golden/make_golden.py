"""Write golden vectors from the compiled reference CFDwavemaker.

Run in the development container only (it needs the reference tree to build
oracle/_ref/):   make -C oracle ref && python tests/golden/make_golden.py

Outputs (small npz fixtures, committed):
  points.npz  every Irregular evaluation (eta1 ... uvw2) at seeded points for
              four sea states that exercise mtheta/tofmax/x0/swl, short-crested
              duplicates (k_i == k_m), bandwidth/cutoff filters, individual
              ramps, shallow and deep water;
  linear.npz  the linear (order 0) long-crested workload the reference runs on the
              CPU: eta, u, w, dp on 64 x 32 points of a vertical section at t = 0;
  pairs.npz   the reference pair cache (Irregular::rebuild_pair_cache) of those seas;
  smit.npz    the Smit et al. members u_smit/v_smit/w_smit/dp_smit of those seas
              (orders 2 and 1) and their pair-cache H coefficients;
  grid.npz    lsGridSpline fields after initialize_kinematics and three
              wave_force_update steps (plain and ignore_at_init), plus spline
              queries at each step;
  lsgrid.npz  the legacy two-slot lsGrid (src/lsgrid.cpp) of the cases in
              cases.LSGRID_CASES: both slots of ETA/UX/UY/UZ after the fill and
              after each update, queries and CheckTime probes at every stage,
              and the reference-written .vtu of case A (both slots);
  abi_*.npz   a waveinput.dat and a program of C-ABI calls (ref_abi.py) with
              the reference's answers: direct mode with ramps, s-grid mode
              with updates, direct mode with vertical_theory smit, and the
              reference's t005 focused-wave input.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import KINDS, GRID_FIELDS, RefHarness, REF_DIR  # noqa: E402
from cfdwavemaker_amd.seastate import make_sea, write_waveinput  # noqa: E402
from tests.golden.cases import sea_cases, points_for, GRID_CASE, LSGRID_CASES, lsgrid_sea, lsgrid_points  # noqa: E402

REF_SRC = "/root/reference"


def make_points(R):
    out = {}
    for name, sea in sea_cases().items():
        R.setup(sea)
        t, x, y, z = points_for(sea, seed=ord(name))
        for key in ("omega", "A", "k", "phase", "theta", "ramptime", "starttime", "duration"):
            if sea.get(key) is not None:
                out[f"{name}__{key}"] = np.asarray(sea[key], float)
        for key in ("depth", "g", "rho", "swl", "mtheta", "tofmax", "x0", "y0", "bandwidth", "cutoff", "order"):
            out[f"{name}__{key}"] = np.float64(sea[key])
        out[f"{name}__pts"] = np.stack([t, x, y, z])
        for kind in KINDS:
            out[f"{name}__{kind}"] = R.eval(kind, t, x, y, z)
    np.savez_compressed(os.path.join(HERE, "points.npz"), **out)


def make_linear(R):
    """The linear long-crested workload of the CPU reference (order 0, FP64):
    N=256 JONSWAP sea (Hs 3 m, Tp 10 s, depth 50 m), 64 x on [0, 500] m x 32 z
    on [-50 m, eta(x)] at t = 0; eta, u, w and dp through Irregular's dispatch."""
    from cfdwavemaker_amd.seastate import config_sea
    sea = config_sea(1).as_dict()
    R.setup(sea)
    xs = np.linspace(0.0, 500.0, 64)
    eta = R.eval("eta", np.zeros(64), xs, np.zeros(64))
    z = -50.0 + (eta[:, None] + 50.0) * np.linspace(0.0, 1.0, 32)[None, :]
    x = np.repeat(xs, 32)
    t, y = np.zeros_like(x), np.zeros_like(x)
    out = {}
    for key in ("omega", "A", "k", "phase", "theta"):
        out[f"L__{key}"] = np.asarray(sea[key], float)
    for key in ("depth", "g", "rho", "swl", "mtheta", "tofmax", "x0", "y0", "bandwidth", "cutoff", "order"):
        out[f"L__{key}"] = np.float64(sea[key])
    out["L__pts"] = np.stack([t, x, y, z.ravel()])
    for kind in ("eta", "u", "w", "dp"):
        out[f"L__{kind}"] = R.eval(kind, t, x, y, z.ravel())
    np.savez_compressed(os.path.join(HERE, "linear.npz"), **out)


def make_pairs(R):
    out = {}
    for name, sea in sea_cases().items():
        R.setup(sea)
        im, fl, v = R.pairs()
        out[f"{name}__im"], out[f"{name}__flags"], out[f"{name}__vals"] = im, fl, v
        out[f"{name}__kmax"] = np.float64(R.kmax())
    np.savez_compressed(os.path.join(HERE, "pairs.npz"), **out)


def make_smit(R):
    """Smit et al. s-coordinate members (vertical_theory smit) of the same seas,
    orders 2 and 1, at the same points, with the pair cache's H coefficients."""
    out = {}
    for name, sea in sea_cases().items():
        for order in (2, 1):
            sd = dict(sea, order=order)
            R.setup(sd)
            R.set_vertical_theory(1)
            t, x, y, z = points_for(sd, seed=ord(name))
            key = f"{name}{order}"
            out[f"{key}__pts"] = np.stack([t, x, y, z])
            out[f"{key}__H"] = R.pairs_h()
            for kind in ("u_smit", "v_smit", "w_smit", "dp_smit"):
                out[f"{key}__{kind}"] = R.eval(kind, t, x, y, z)
    np.savez_compressed(os.path.join(HERE, "smit.npz"), **out)


def make_grid(R):
    sea = sea_cases()["B"]
    R.setup(sea)
    out = {}
    g = GRID_CASE
    for ign in (0, 1):
        R.grid_init(g["domain"], g["nx"], g["ny"], g["nl"], g["t0"], g["dt"], g["tan_a"],
                    g["tan_b"], g["depth"], 0.0, (30.0, 70.0, -25.0, 25.0), 1, ign)
        for step in range(4):
            if step:
                R.grid_force_update(g["t0"] + step * g["dt"] + 0.05)
            st = R.grid_state()
            for f in GRID_FIELDS:
                out[f"ign{ign}__s{step}__{f}"] = R.grid_export(f)
            rng = np.random.default_rng(100 + 10 * ign + step)
            n = 64
            q = np.stack([st["t0"] + rng.uniform(0, g["dt"], n), rng.uniform(-5, 125, n),
                          rng.uniform(-25, 25, n), rng.uniform(-g["depth"] - 2, 3.0, n)])
            out[f"ign{ign}__s{step}__q"] = q
            out[f"ign{ign}__s{step}__qout"] = R.grid_query(*q)
            out[f"ign{ign}__s{step}__state"] = np.array([st["tstep"], st["t0"], st["dx"], st["dy"], st["ds"]])
    np.savez_compressed(os.path.join(HERE, "grid.npz"), **out)


def make_lsgrid(R):
    """Legacy lsGrid through the compiled reference (ref_harness.cpp)."""
    out = {}
    for name, c in LSGRID_CASES.items():
        R.setup(lsgrid_sea(c["sea"]))
        R.lsgrid_init(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"],
                      c["depth"], c["swl"], c["ignore"], c["ignore_at_init"], c["init_only"])
        for stage, tt in enumerate((None,) + tuple(c["updates"])):
            if tt is not None:
                R.lsgrid_update(tt)
            t0 = R.lsgrid_t0()
            key = f"{name}__s{stage}"
            out[key + "__t0"] = np.float64(t0)
            fields = ("ETA",) if name == "vtu" else ("ETA", "UX", "UY", "UZ")
            for f in fields:
                out[key + "__" + f] = R.lsgrid_export(f)
            q = np.stack(lsgrid_points(c, t0, seed=1000 + 10 * stage + len(name)))
            out[key + "__q"] = q
            out[key + "__qout"] = R.lsgrid_query(*q)
            probes = t0 + c["dt"] * np.array([0.0, 0.5, 1.0, 1.0 + 1e-9, 1.5, 3.0])
            out[key + "__ct_t"] = probes
            out[key + "__ct"] = np.array([R.lsgrid_check_time(v) for v in probes], dtype=np.int8)
            if name == "A" and stage == 2:
                with tempfile.TemporaryDirectory() as td:
                    for last in (0, 1):
                        path = os.path.join(td, f"k{last}.vtu")
                        R.lsgrid_write_vtu(path, last)
                        with open(path, "rb") as fh:
                            out[f"{key}__vtu{last}"] = np.frombuffer(fh.read(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "lsgrid.npz"), **out)


def run_abi(lib, waveinput_text, prog, files=False):
    """Run a call program in a process of its own; with files, also return the
    text of every file the library wrote into the run directory
    (spectral_components.dat, .vtu dumps, including those written at exit)."""
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "waveinput.dat"), "w") as f:
            f.write(waveinput_text)
        pp = os.path.join(wd, "prog.npz")
        op = os.path.join(wd, "out.npz")
        np.savez(pp, **prog)
        subprocess.run([sys.executable, os.path.join(HERE, "ref_abi.py"), lib, wd, pp, op],
                       check=True, stdout=subprocess.DEVNULL)
        r = np.load(op)
        if not files:
            return int(r["rc"]), r["out"]
        written = {}
        for root, _, names in os.walk(wd):
            for nm in names:
                if nm in ("waveinput.dat", "prog.npz", "out.npz"):
                    continue
                p = os.path.join(root, nm)
                written[os.path.relpath(p, wd)] = open(p).read()
        return int(r["rc"]), r["out"], written


def file_fields(written):
    """npz entries for the files a run wrote: file__<name> = text."""
    return {"file__" + k: np.str_(v) for k, v in written.items()}


def waveinput_spreading(sea_rows, dirs, depth, extra_general=(), irregular_keys=(), lsgrid=None, tail=""):
    """waveinput.dat in the reference's ndir > 0 form (src/CFDwavemaker.cpp:
    729-797): nfreq rows (omega, A, k, phase), then ndir rows (theta, D)."""
    L = ["@v310", "[wave type]", "irregular", "", "[general input data]", f"depth {depth!r}", "mtheta 0.0"]
    L += list(extra_general) + ["", "[wave input irregular]"] + list(irregular_keys)
    L += [f"nfreq {len(sea_rows)}", f"ndir {len(dirs)}"]
    L += [" ".join(repr(float(v)) for v in r) for r in sea_rows]
    L += [" ".join(repr(float(v)) for v in d) for d in dirs]
    L.append("")
    if lsgrid:
        L += lsgrid
    return "\n".join(L) + "\n" + tail


def prog_direct(seed, n=60, depth=30.0):
    rng = np.random.default_rng(seed)
    ops, args = [], []
    for q in range(n):
        x, y = rng.uniform(-50, 300), rng.uniform(-30, 30)
        z, t = rng.uniform(-depth - 1, 2.0), rng.uniform(-1.0, 25.0)
        for o in (0, 1, 2, 3, 4):
            ops.append(o)
            args.append((x, y, z, t))
    for o, opt in ((7, 1), (7, 2), (8, 1), (8, 2), (9, 1), (9, 2), (10, 0)):
        ops.append(o)
        args.append((opt, 0, 0, 0))
    return dict(op=np.array(ops, np.int32), args=np.array(args, float))


def prog_grid(seed, lsg, depth, nsteps=3, nq=40):
    rng = np.random.default_rng(seed)
    ops, args = [], []
    b = lsg["bounds"]
    for step in range(nsteps + 1):
        t0 = lsg["t0"] + step * lsg["dt"]
        if step:
            ops.append(6)
            args.append((0, 0, 0, t0 + 0.01))
        for q in range(nq):
            x, y = rng.uniform(b[0] - 5, b[1] + 5), rng.uniform(b[2] - 5, b[3] + 5)
            z, t = rng.uniform(-depth - 1, 1.0), t0 + rng.uniform(0.0, lsg["dt"])
            for o in (0, 1, 2, 3, 5):
                ops.append(o)
                args.append((x, y, z, t))
    return dict(op=np.array(ops, np.int32), args=np.array(args, float))


def make_abi():
    lib = os.path.join(REF_DIR, "libCFDwavemaker_ref.so")
    # 1. direct mode (wavetype 1) with ramps, filters, swl, reference point
    sea = make_sea(24, 3.0, 10.0, 30.0, seed=21)
    sea.swl, sea.mtheta, sea.tofmax, sea.x0, sea.y0 = 0.25, 15.0, 5.0, 10.0, 2.0
    with tempfile.NamedTemporaryFile("r", suffix=".dat") as tf:
        write_waveinput(tf.name, sea, ramps={"time_rampup": (0.0, 4.0), "x_rampdown": (250.0, 280.0)},
                        bandwidth="0.6", cutoff=2.2, amplify=1.3)
        text = open(tf.name).read()
    prog = prog_direct(21)
    rc, out = run_abi(lib, text, prog)
    np.savez_compressed(os.path.join(HERE, "abi_direct.npz"), waveinput=text, rc=rc, out=out, **prog)
    # 2. s-grid mode (wavetype 5): short-crested, three updates
    sea = make_sea(8, 4.0, 11.0, 40.0, seed=22, ndir=4)
    lsg = dict(bounds=(0.0, 150.0, -30.0, 30.0), nx=9, ny=5, nl=6, t0=0.0, dt=0.5, tan_a=0.7, tan_b=1.5)
    with tempfile.NamedTemporaryFile("r", suffix=".dat") as tf:
        write_waveinput(tf.name, sea, lsgrid=lsg, extrap=2)
        text = open(tf.name).read()
    prog = prog_grid(22, lsg, sea.depth)
    rc, out = run_abi(lib, text, prog)
    np.savez_compressed(os.path.join(HERE, "abi_grid.npz"), waveinput=text, rc=rc, out=out, **prog)
    # 3. direct mode with the Smit et al. vertical theory (experimental in the
    #    reference): VeloX/Y/Z, DynPres and SurfElev through the dispatch
    sea = make_sea(16, 2.5, 9.0, 25.0, seed=23, ndir=2)
    sea.swl, sea.tofmax, sea.x0 = 0.2, 4.0, 5.0
    with tempfile.NamedTemporaryFile("r", suffix=".dat") as tf:
        write_waveinput(tf.name, sea, ramps={"time_rampup": (0.0, 3.0)}, bandwidth="0.8",
                        vertical_theory="smit")
        text = open(tf.name).read()
    prog = prog_direct(23, n=40, depth=sea.depth)
    rc, out = run_abi(lib, text, prog)
    np.savez_compressed(os.path.join(HERE, "abi_smit.npz"), waveinput=text, rc=rc, out=out, **prog)
    # 4. the reference's own t005 focused-wave input (data file of its test
    #    suite), sampled where its targetwave.dat was written (x=10, y=0)
    t005 = os.path.join(REF_SRC, "tests", "t005_focused_wave")
    text = open(os.path.join(t005, "waveinput.dat")).read()
    tw = np.loadtxt(os.path.join(t005, "targetwave.dat"))
    sel = np.arange(0, tw.shape[0], 16)
    ops = np.zeros(sel.size, np.int32)
    args = np.zeros((sel.size, 4))
    args[:, 0] = 10.0
    args[:, 3] = tw[sel, 0]
    prog = dict(op=ops, args=args)
    rc, out = run_abi(lib, text, prog)
    np.savez_compressed(os.path.join(HERE, "abi_t005.npz"), waveinput=text, rc=rc, out=out,
                        targetwave=tw[sel], **prog)


def make_abi_inputs():
    """Input-format and output-file coverage of the waveinput.dat loader and
    the s-grid dumps, each a call program with the reference's answers and the
    text of the files the reference wrote (spectral_components.dat, .vtu):
      abi_t008     the reference's tests/t008_3D_irregular_sea/waveinput_linear.dat
                   (312 components, bandwidth auto, s-grid with ignore_at_init);
      abi_example  the reference's examples/python/linux/waveinput.dat (ndir 1
                   spreading form, normalize, amplify, ramps; direct mode);
      abi_ndir     a 10 x 5 frequency x direction spreading input (direct mode);
      abi_iramp    individual_ramp rows (direct mode, start times on both sides of 0);
      abi_vtk      an s-grid with [lsgrid2vtk]: dumps at init and after each
                   update, the end-of-run dump of the following slot, and a
                   storage directory that has to be created."""
    from cfdwavemaker_amd.seastate import wavenumber
    lib = os.path.join(REF_DIR, "libCFDwavemaker_ref.so")
    # t008, without its ignore_subdomain / ignore_at_init lines: the reference
    # allocates the s-grid arrays uninitialised (lsgrid_spline.cpp:35-61) and
    # ignore_at_init leaves the ignored columns unwritten, so its queries there
    # (and in every cell of this 5 x 5 grid, which all touch an ignored
    # column) read whatever the heap held; the ignore paths are pinned by the
    # harness fixtures (grid.npz), which zero the arrays first
    text = open(os.path.join(REF_SRC, "tests", "t008_3D_irregular_sea", "waveinput_linear.dat")).read()
    text = "\n".join(ln for ln in text.splitlines() if not ln.strip().startswith(("ignore_subdomain", "ignore_at_init")))
    text += "\n"
    lsg = dict(bounds=(-100.0, 100.0, -100.0, 100.0), t0=0.0, dt=0.5)
    prog = prog_grid(31, lsg, 40.0, nsteps=2, nq=20)
    rc, out, wr = run_abi(lib, text, prog, files=True)
    np.savez_compressed(os.path.join(HERE, "abi_t008.npz"), waveinput=text, rc=rc, out=out, **prog,
                        **file_fields(wr))
    # python example input (tank scale: depth 1.2 m)
    text = open(os.path.join(REF_SRC, "examples", "python", "linux", "waveinput.dat")).read()
    rng = np.random.default_rng(32)
    ops, args = [], []
    for q in range(40):
        x, y, z, t = rng.uniform(-2, 8), rng.uniform(-1, 1), rng.uniform(-1.3, 0.05), rng.uniform(-0.2, 20.0)
        for o in (0, 1, 2, 3, 4):
            ops.append(o)
            args.append((x, y, z, t))
    prog = dict(op=np.array(ops, np.int32), args=np.array(args, float))
    rc, out, wr = run_abi(lib, text, prog, files=True)
    np.savez_compressed(os.path.join(HERE, "abi_example.npz"), waveinput=text, rc=rc, out=out, **prog,
                        **file_fields(wr))
    # synthetic spreading input: 10 frequencies x 5 directions, unsorted rows
    rng = np.random.default_rng(33)
    w = rng.permutation(np.linspace(0.45, 1.6, 10))
    k = wavenumber(w, 30.0)
    A = 0.4 * np.exp(-((w - 0.8) / 0.3) ** 2)
    ph = rng.uniform(0, 2 * np.pi, 10)
    th = np.linspace(-0.9, 0.9, 5)
    D = np.cos(th) ** 2
    D /= np.sqrt(np.sum(D * D))
    text = waveinput_spreading(np.stack([w, A, k, ph], 1), np.stack([th, D], 1), 30.0,
                               extra_general=["swl 0.1", "gravity 9.81", "rho 1025.0"],
                               irregular_keys=["bandwidth 0.7"])
    prog = prog_direct(33, n=40, depth=30.0)
    rc, out, wr = run_abi(lib, text, prog, files=True)
    np.savez_compressed(os.path.join(HERE, "abi_ndir.npz"), waveinput=text, rc=rc, out=out, **prog,
                        **file_fields(wr))
    # individual ramps
    sea = make_sea(18, 2.0, 8.0, 25.0, seed=34)
    rng = np.random.default_rng(34)
    sea.ramptime = rng.uniform(0.5, 4.0, 18)
    sea.starttime = rng.uniform(-3.0, 6.0, 18)
    sea.duration = rng.uniform(8.0, 30.0, 18)
    with tempfile.NamedTemporaryFile("r", suffix=".dat") as tf:
        write_waveinput(tf.name, sea, individual_ramp=True)
        text = open(tf.name).read()
    prog = prog_direct(34, n=40, depth=25.0)
    rc, out, wr = run_abi(lib, text, prog, files=True)
    np.savez_compressed(os.path.join(HERE, "abi_iramp.npz"), waveinput=text, rc=rc, out=out, **prog,
                        **file_fields(wr))
    # [lsgrid2vtk]
    sea = make_sea(8, 3.0, 9.0, 30.0, seed=35, ndir=3)
    lsg = dict(bounds=(0.0, 100.0, -20.0, 20.0), nx=6, ny=4, nl=5, t0=0.0, dt=0.5, tan_a=0.7, tan_b=1.5)
    with tempfile.NamedTemporaryFile("r", suffix=".dat") as tf:
        write_waveinput(tf.name, sea, lsgrid=lsg, extrap=2)
        text = open(tf.name).read()
    text += "[lsgrid2vtk]\nstorage_path ./vtkout/\nfilename waves_\nvtk_timelabel Time\n"
    prog = prog_grid(35, lsg, sea.depth, nsteps=2, nq=10)
    rc, out, wr = run_abi(lib, text, prog, files=True)
    np.savez_compressed(os.path.join(HERE, "abi_vtk.npz"), waveinput=text, rc=rc, out=out, **prog,
                        **file_fields(wr))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if sys.argv[1:] == ["inputs"]:
        make_abi_inputs()
        print("input-format fixtures written to", HERE)
        return
    if sys.argv[1:] == ["lsgrid"]:
        make_lsgrid(RefHarness())
        print("legacy-grid fixtures written to", HERE)
        return
    R = RefHarness()
    make_points(R)
    make_linear(R)
    make_pairs(R)
    make_smit(R)
    make_grid(R)
    make_lsgrid(R)
    make_abi()
    make_abi_inputs()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
