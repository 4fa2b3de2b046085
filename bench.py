"""Benchmark: second-order kinematics points/s at N=1024 components on one MI355X.

Workload (the N=1024 short-crested configuration): 64 frequencies x 16
directions (cos^2s, s=4) = 1024 components, Hs 8 m, Tp 14 s, depth 100 m;
s-grid 256 x 256 columns x 16 sigma levels x 4 time levels over 1 km x 1 km,
then 1e7 cubic-spline queries at uniform random points in the 1 km x 1 km x
(-100 m, eta) box at t = 0.25 s.

One step = one full initialize_kinematics of the s-grid (column pass: eta1 +
eta2 + Taylor gradients; level pass: first + second order + Taylor at every
level; FD gradients) plus the 1e7 queries, all on the GPU with inputs resident
in HBM.  value = grid points (4 x 256 x 256 x 16) per second of step time.

usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--precision bf16|fp32|fp64]

Precision: bf16 (default) = FP32 sums with FP64 phase reduction whose MFMA
quadratic forms take one bf16 product per term; fp32 = the same with three
bf16 products per term (hi/lo split); fp64 = FP64 throughout.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "bf16": 157.3e12, "fp64": 78.6e12}   # MI355X vector peaks (FLOP/s)
PEAK_UNPACKED = {"fp32": 78.6e12, "bf16": 78.6e12, "fp64": 78.6e12}   # without packed FP32 instructions
PEAK_MFMA_BF16 = 2.5e15                        # MI355X dense bf16 matrix peak (FLOP/s)
PEAK_SCLK_MHZ = 2400.0                         # its engine clock: 256 CUs x 4 x 1024 FLOP/clk x 2.4 GHz
PEAK_HBM = 8.0e12                              # MI355X HBM3E (B/s)
# algorithmic work of the secondary kernels: a spline query moves 12 B in + 16 B
# out + its share of one pass over the packed grid (128 MB over 1e7 queries);
# the column linear sums take 4 N (3 nl + 4) FLOP per column; the surrogate's
# 29-256-256-256-65 network 2 (29*256 + 3*256^2 + 256*65) FLOP per column
QUERY_BYTES = 41.0
MLP_FLOP_PER_COLUMN = 2 * (29 * 256 + 3 * 256 * 256 + 256 * 65)
NX = NY = 256
NL = 16
DOMAIN = (0.0, 1000.0, 0.0, 1000.0)
T0, DT, TAN_A, TAN_B, DEPTH, HS = 0.0, 0.5, 0.7, 1.5, 100.0, 8.0


def _grid_fill_seconds(ncols: int):
    """One timed s-grid fill (initialize_kinematics) of an ncols x ncols subset
    of the N=1024 grid (same spacing) by the compiled reference's OpenMP code,
    or by the oracle's column body where the reference is not present."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = (DOMAIN[1] - DOMAIN[0]) / (NX - 1)
    dom = (0.0, dx * (ncols - 1), 0.0, dx * (ncols - 1))
    try:
        from oracle import RefHarness
        R = RefHarness()
        R.setup(s)
        return "reference", R.grid_init(dom, ncols, ncols, NL, T0, DT, TAN_A, TAN_B, DEPTH)
    except FileNotFoundError:
        from oracle import Oracle
        O = Oracle().sea(s)
        xs, ys = np.meshgrid(np.arange(ncols) * dx, np.arange(ncols) * dx, indexing="ij")
        tot = 0.0
        for t in (T0 - DT, T0, T0 + DT, T0 + 2 * DT):
            sec, *_ = O.column_fill(xs.ravel(), ys.ravel(), t, NL, TAN_A, TAN_B, DEPTH)
            tot += sec
        return "port", tot


def _lsgrid_fill_seconds(ncols: int):
    """One timed legacy-grid fill (initialize_surface_elevation +
    initialize_kinematics, two time levels) of an ncols x ncols subset of the
    N=1024 grid by the compiled reference's OpenMP code (src/lsgrid.cpp)."""
    from cfdwavemaker_amd.seastate import config_sea
    from oracle import RefHarness
    s = config_sea(3).as_dict()
    dx = (DOMAIN[1] - DOMAIN[0]) / (NX - 1)
    dom = (0.0, dx * (ncols - 1), 0.0, dx * (ncols - 1))
    R = RefHarness()
    R.setup(s)
    return "reference", R.lsgrid_init(dom, ncols, ncols, NL, T0, DT, TAN_A, TAN_B, DEPTH)


def _direct_seconds(ncolumns: int):
    """One timed direct-mode evaluation by the compiled reference (Irregular
    members in an OpenMP parallel loop over points): eta at ncolumns (x, t)
    columns of the N=512 long-crested workload and u, v, w, dp at its 64 z
    per column (the points of the GPU direct leg's first columns)."""
    import time as _t
    from cfdwavemaker_amd.seastate import config_sea
    from oracle import RefHarness
    d = config_sea(2).as_dict()
    R = RefHarness()
    R.setup(d)
    nz = 64
    xs = np.linspace(0.0, 500.0, 256)[:ncolumns]
    ts = np.arange(ncolumns, dtype=float) % 16
    ys = np.zeros(ncolumns)
    t0 = _t.perf_counter()
    eta = R.eval("eta", ts, xs, ys)
    frac = np.linspace(0.0, 1.0, nz)
    z = (-30.0 + (eta[:, None] + 30.0) * frac[None, :]).ravel()
    rep = lambda a: np.repeat(a, nz)
    for kind in ("u", "v", "w", "dp"):
        R.eval(kind, rep(ts), rep(xs), rep(ys), z)
    return "reference", _t.perf_counter() - t0, ncolumns * nz


def cpu_worker(spec: str, out_path: str) -> None:
    """Timing protocol of one CPU baseline: one warm-up run, then five runs
    at each of two sample sizes; per size the median.  spec: "grid:a,b"
    (ncols x ncols s-grid subsets), "lsgrid:a,b" (the same for the legacy
    two-slot grid) or "direct:a,b" (columns of 64 points)."""
    mode, sizes = spec.split(":")
    sizes = [int(v) for v in sizes.split(",")]
    res = {"mode": mode, "runs": {}}
    if mode == "grid":
        kind, _ = _grid_fill_seconds(sizes[0])                  # warm-up
        for n in sizes:
            secs = [_grid_fill_seconds(n)[1] for _ in range(5)]
            res["runs"][str(n)] = {"points": 4 * n * n * NL, "seconds": secs}
    elif mode == "lsgrid":
        kind, _ = _lsgrid_fill_seconds(sizes[0])                # warm-up
        for n in sizes:
            secs = [_lsgrid_fill_seconds(n)[1] for _ in range(5)]
            res["runs"][str(n)] = {"points": 2 * n * n * NL, "seconds": secs}
    else:
        kind, _, _ = _direct_seconds(sizes[0])                 # warm-up
        for n in sizes:
            runs = [_direct_seconds(n) for _ in range(5)]
            res["runs"][str(n)] = {"points": runs[0][2], "seconds": [r[1] for r in runs]}
    res["kind"] = kind
    with open(out_path, "w") as f:
        json.dump(res, f)


def host_cpu() -> str:
    """CPU model and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def cpu_baseline(spec: str, threads: int, what: str):
    """Run one CPU baseline (cpu_worker protocol) in a child process with
    `threads` OpenMP threads; returns the cpu_baseline object: the larger
    sample's median rate, both medians and the linearity of the two."""
    import tempfile
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "cpu.json")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", spec, out],
                           cwd=td, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                           timeout=900)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        with open(out) as f:
            c = json.load(f)
    rates = {}
    for n, run in c["runs"].items():
        med = float(np.median(run["seconds"]))
        rates[n] = {"points": run["points"], "median_s": med, "runs_s": run["seconds"],
                    "points_per_s": run["points"] / med}
    keys = list(rates)
    lin = rates[keys[-1]]["points_per_s"] / rates[keys[0]]["points_per_s"]
    return {"value": rates[keys[-1]]["points_per_s"], "unit": "points/s", "cores": threads, "kind": c["kind"],
            "host": host_cpu(), "sample": what, "protocol": "1 warm-up, median of 5 runs at each of 2 sizes",
            "sizes": rates, "linearity": lin}


# PMC calibration of the committed profiles: a 256 MiB device copy reads
# 131,088 KiB by FETCH_SIZE and a 256 MiB fill writes 262,144 KiB by WRITE_SIZE
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
PROFILE_BASE = os.path.join(ROOT, "profiles", "sgrid_fill_query_N1024_256x256x16x4_")


def profiled_traffic(kernel: str, precision: str = "bf16"):
    """HBM-side bytes per launch of the largest launch of `kernel` from the
    committed FETCH_SIZE / WRITE_SIZE summaries (rocprofv3 --pmc, KiB per
    dispatch, calibrated as above); None if they do not list it."""
    tot = 0.0
    for counter, scale in (("fetch_size", FETCH_SCALE), ("write_size", WRITE_SCALE)):
        best = None
        try:
            with open(PROFILE_BASE + precision + "_" + counter + ".md") as f:
                for line in f:
                    c = [x.strip() for x in line.strip().strip("|").split("|")]
                    if len(c) == 6 and kernel in c[0] and c[2] == counter.upper():
                        g, kib = int(c[1]), float(c[4])
                        if best is None or g > best[0]:
                            best = (g, kib)
        except OSError:
            return None
        if best is None:
            return None
        tot += best[1] * 1024.0 * scale
    return tot


def rocprof_summary(db_path: str, command: str) -> str:
    """Markdown summary of a rocprofv3 results database: per kernel and launch
    size the dispatch count and durations, and the mean of any PMC counter per
    dispatch (FETCH_SIZE / WRITE_SIZE in KiB, instruction counters in
    instructions, busy counters in cycles summed over the units)."""
    import sqlite3
    con = sqlite3.connect(db_path)
    lines = [f"command: `{command}`", ""]
    rows = con.execute("select name, grid_x, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name, grid_x order by sum(duration) desc").fetchall()
    if rows:
        lines += ["| kernel | grid_x | calls | total ms | avg ms | min ms | max ms |", "|---|---|---|---|---|---|---|"]
        for nm, gx, n, tot, avg, mn, mx in rows:
            if tot < 1e4:
                continue
            short = nm.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            lines.append(f"| `{short}` | {gx} | {n} | {tot / 1e6:.3f} | {avg / 1e6:.3f} | {mn / 1e6:.3f} | {mx / 1e6:.3f} |")
    pmc = con.execute("select kernel_name, grid_size, counter_name, count(*), avg(value), avg(duration) "
                      "from counters_collection group by kernel_name, grid_size, counter_name "
                      "order by avg(duration) desc").fetchall()
    if pmc:
        lines += ["", "| kernel | grid_size | counter | dispatches | mean per dispatch | mean ms |", "|---|---|---|---|---|---|"]
        for nm, gs, cn, n, v, d in pmc:
            short = nm.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            lines.append(f"| `{short}` | {gs} | {cn} | {n} | {v:.1f} | {d / 1e6:.3f} |")
    return "\n".join(lines) + "\n"


def surrogate_leg(steps: int, warmup: int, prec_name: str = "bf16") -> dict:
    """Learned second-order surrogate on the 512 x 512 column x 16 level
    workload (N = 1024), hybrid network: features and the exact eta2 quadratic
    form -> network (velocity and pressure profiles) -> assembly on the GPU;
    the residual check against the exact kernels on a sample of columns
    (elevation gate 1e-3 Hs; velocity and pressure errors stated); and the
    exact GPU path on the same columns (the s-grid fill at the same precision,
    per time level) as the fallback and the comparison."""
    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd import surrogate as S
    from cfdwavemaker_amd.seastate import config_sea

    sea = config_sea(5)
    d = sea.as_dict()
    prec = {"bf16": engine.BF16, "fp32": engine.FP32}[prec_name]
    es = engine.Sea(d, precision=prec, with_direct=True)
    sur = S.Surrogate.load()
    n = 512
    g = np.linspace(DOMAIN[0], DOMAIN[1], n)
    xx, yy = (q.ravel() for q in np.meshgrid(g, g, indexing="ij"))
    xd = torch.tensor(xx, dtype=torch.float64, device="cuda")
    yd = torch.tensor(yy, dtype=torch.float64, device="cuda")
    t = 0.25
    for _ in range(max(warmup, 1)):
        sur.fields(es, sea, xd, yd, t)
    engine.sync()
    engine.prof_enable(True)
    engine.prof_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out, _, Yp = sur.fields(es, sea, xd, yd, t)
    engine.sync()
    t_sur = (time.perf_counter() - t0) / steps
    prof = engine.prof_read(reset=True)
    engine.prof_enable(False)
    pts = n * n * S.NLEV
    # the network alone on the same features (bf16 MFMA GEMMs with fused bias + GELU)
    X, _, _ = S.linear_fields(es, xd, yd, t, float(sea.depth), S.peak_wavenumber(sea))
    Xs = X[:, sur.cols].contiguous()
    sur.mlp.forward(Xs)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(steps):
        sur.mlp.forward(Xs)
    torch.cuda.synchronize()
    t_mlp = (time.perf_counter() - t3) / steps
    mlp_flop = 2 * (sur.mlp.nin * 256 + 2 * 256 * 256 + 256 * 65) * n * n
    eta = out[:, 0].cpu().numpy().astype(float)
    y2 = Yp.cpu().numpy().astype(float)
    e2net = sur.eta2_net.cpu().numpy().astype(float) if sur.eta2_net is not None else None
    sample = np.arange(0, n * n, 256)
    gate = 1e-3 * HS
    t1 = time.perf_counter()
    res = sur.residual_check(es, xx, yy, t, eta, y2, sample, gate, sea.depth, e2net)
    t_chk = time.perf_counter() - t1
    ok = res["pass"]
    # exact GPU path on the same columns: s-grid fill (4 time levels per fill)
    grid = es.grid(DOMAIN, n, n, S.NLEV, T0, DT, TAN_A, TAN_B, DEPTH)
    engine.sync()
    t2 = time.perf_counter()
    for _ in range(steps):
        grid.refill()
    engine.sync()
    t_exact = (time.perf_counter() - t2) / steps / 4.0
    eff = pts / (t_sur + (0.0 if ok else t_exact))
    res.update({"eta_err_over_hs": res["eta_max_err_m"] / HS, "gate": "eta error <= 1e-3 Hs vs exact",
                "ms": t_chk * 1e3})
    return {"workload": "N=1024 short-crested sea, 512x512 columns x 16 levels, one time level",
            "mode": sur.mode, "features": len(sur.cols), "precision": prec_name,
            "points": pts, "surrogate_points_per_s": pts / t_sur, "ms_surrogate": t_sur * 1e3,
            "kernels_ms": {k: v[0] / max(v[1], 1) for k, v in prof.items()},
            "mlp": {"bound": "mfma", "flop_per_column": mlp_flop / (n * n), "ms": t_mlp * 1e3,
                    "achieved": mlp_flop / t_mlp / 1e12, "peak": PEAK_MFMA_BF16 / 1e12, "unit": "TFLOP/s",
                    "frac": mlp_flop / t_mlp / PEAK_MFMA_BF16},
            "residual_check": res,
            "exact_points_per_s": pts / t_exact, "ms_exact_per_time_level": t_exact * 1e3,
            "effective_points_per_s": eff, "fallback_to_exact": not ok,
            "speedup_over_exact": (pts / t_sur) / (pts / t_exact)}


def batched_leg(nstates: int, prec_name: str, err_cols: int = 4) -> dict:
    """Batched workload: nstates of the 512 random sea states (N 128-2048, Hs
    1-8 m, Tp 6-16 s, depth 20-200 m, half long-crested, half 16-direction),
    each on its own 64 x 64 x 8 s-grid over [0, 20 lambda_p]^2 with 4 time
    levels, refilled in one call: per group of states (bounded by a device
    memory budget) the form tables are built, the Z operands generated, the
    forms of every state of the group evaluated in one persistent launch per
    product mode (tile order, grouping and grid size from the tuner), the
    column / level / gradient / packing passes run, and the tables released,
    all inside the timed call.  Per state, the error of an err_cols x err_cols
    column subsample (all levels and time levels) against an FP64 fill of the
    same columns."""
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd import surrogate as S
    from cfdwavemaker_amd.seastate import config_sea

    prec = {"bf16": engine.BF16, "fp32": engine.FP32, "fp64": engine.FP64}[prec_name]
    os.environ.setdefault("CFDWM_QUIET", "1")
    seas, grids = [], []
    t0 = time.perf_counter()
    for st in range(nstates):
        sea = config_sea(4, st)
        lam = 2 * np.pi / S.peak_wavenumber(sea)
        es = engine.Sea(sea.as_dict(), precision=prec, with_direct=False)
        g = es.grid((0.0, 20 * lam, 0.0, 20 * lam), 64, 64, 8, 0.0, 0.5, 0.7, 1.5, sea.depth, defer_fill=True)
        seas.append((st, sea, es, lam))
        grids.append(g)
        if (st + 1) % 128 == 0:
            print(f"[bench] batched: {st + 1}/{nstates} seas set up", file=sys.stderr, flush=True)
    t_setup = time.perf_counter() - t0
    # first call: the tuner times its candidate launch shapes on the first
    # group of each product mode (and the call fills every grid)
    t1 = time.perf_counter()
    tune = engine.refill_batched(grids, release=True)
    engine.sync()
    t_first = time.perf_counter() - t1
    print(f"[bench] batched: first call (tuning) {t_first:.2f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    stats = engine.refill_batched(grids, release=True)
    engine.sync()
    el = time.perf_counter() - t1
    print(f"[bench] batched: timed call {el:.2f} s", file=sys.stderr, flush=True)
    pts = nstates * 64 * 64 * 8 * 4
    # per-state error: FP64 fill of err_cols^2 columns that coincide with grid
    # columns 0, 21, 42, 63 (for err_cols 4) of the 64 x 64 grid
    step = 63 // (err_cols - 1)
    idx = np.arange(err_cols) * step
    errs = []
    t2 = time.perf_counter()
    for k, (st, sea, es, lam) in enumerate(seas):
        dx = 20 * lam / 63
        e64 = es.clone(engine.FP64)
        g64 = e64.grid((0.0, idx[-1] * dx, 0.0, idx[-1] * dx), err_cols, err_cols, 8, 0.0, 0.5, 0.7, 1.5, sea.depth)
        e = {"state": st, "components": int(len(sea.omega))}
        for f in ("ETA", "UX", "UY", "UZ"):
            ref = g64.export(f)
            got = grids[k].export(f)
            if f == "ETA":
                got = got.reshape(4, 64, 64)[:, idx][:, :, idx].ravel()
                e["eta_err_over_hs"] = float(np.max(np.abs(got - ref)) / sea.hs)
            else:
                got = got.reshape(4, 64, 64, 8)[:, idx][:, :, idx].ravel()
                e[f + "_scaled"] = float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))
        errs.append(e)
        g64.close()
        e64.close()
    t_err = time.perf_counter() - t2
    worst = {key: max(e[key] for e in errs) for key in ("eta_err_over_hs", "UX_scaled", "UY_scaled", "UZ_scaled")}
    return {"workload": f"{nstates} of the 512 random sea states (N 128-2048), 64x64x8 s-grid x 4 time levels each, "
                        f"{prec_name}; form tables built inside the timed call",
            "states": nstates, "points": pts, "points_per_s": pts / el, "ms_refill_all": el * 1e3,
            "setup_s": t_setup, "first_call_s": t_first, "tables_ms": stats["tables_ms"],
            "forms_ms": stats["forms_ms"],
            "forms_exec_tflops": stats["mfma_flop"] / (stats["forms_ms"] * 1e-3) / 1e12 if stats["forms_ms"] else None,
            "groups": stats["groups"], "tiles": stats["tiles"],
            "launch": {"blocks": stats["blocks"], "order": stats["order"], "group_forms": stats["group_forms"],
                       "group_ptiles": stats["group_ptiles"], "tuned_on_first_call": bool(tune["tuned"])},
            "per_state_err_vs_fp64": {"columns_per_state": err_cols * err_cols, "worst": worst, "states": errs,
                                      "seconds": t_err}}


def abi_worker(out_path: str) -> None:
    """The batched C-ABI entries on the s-grid of the N=1024 workload, in a
    process of their own (the ABI keeps one process-global wave field):
    waveinput.dat with the same sea and [lsgrid] block, FP32, then 1e6 points
    (eta at z = 0 and u, v, w at z) through wave_EvalPoints (host arrays,
    PCIe copies included) and wave_EvalPointsDevice (device arrays)."""
    import ctypes as C
    import tempfile
    import torch
    from cfdwavemaker_amd._lib import load
    from cfdwavemaker_amd.seastate import config_sea, write_waveinput
    n, reps = 1_000_000, 5
    rng = np.random.default_rng(7)
    x = rng.uniform(DOMAIN[0], DOMAIN[1], n)
    y = rng.uniform(DOMAIN[2], DOMAIN[3], n)
    z = rng.uniform(-DEPTH, 0.0, n)
    t = np.full(n, T0 + 0.25)
    res = {}
    with tempfile.TemporaryDirectory() as wd:
        write_waveinput(os.path.join(wd, "waveinput.dat"), config_sea(3),
                        lsgrid=dict(bounds=DOMAIN, nx=NX, ny=NY, nl=NL, t0=T0, dt=DT, tan_a=TAN_A, tan_b=TAN_B))
        os.environ["CFDWM_PRECISION"] = "fp32"
        L = load(os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm.so"))
        cwd = os.getcwd()
        os.chdir(wd)
        t0 = time.perf_counter()
        rc = L.wave_Initialize()
        res["initialize_s"] = time.perf_counter() - t0
        os.chdir(cwd)
        if rc != 0:
            raise RuntimeError(f"wave_Initialize returned {rc}")
    flags = 0x1 | 0x2 | 0x4 | 0x8
    host = [np.zeros(n) for _ in range(4)]
    P = C.POINTER(C.c_double)
    ptr = lambda a: a.ctypes.data_as(P)

    def run_host():
        return L.wave_EvalPoints(n, ptr(x), ptr(y), ptr(z), ptr(t), *[ptr(h) for h in host], None, flags)
    dev = torch.device("cuda:0")
    X, Y, Z, T = (torch.tensor(a, dtype=torch.float64, device=dev) for a in (x, y, z, t))
    D = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4)]

    def run_dev():
        return L.wave_EvalPointsDevice(n, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), T.data_ptr(),
                                       *[d.data_ptr() for d in D], None, flags)
    for name, fn in (("host_arrays", run_host), ("device_arrays", run_dev)):
        if fn() != 0:
            raise RuntimeError(f"{name}: batched entry failed")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / reps
        res[name] = {"points_per_s": n / el, "ms": el * 1e3}
    res["bitwise_equal"] = bool(all(np.array_equal(h, d.cpu().numpy()) for h, d in zip(host, D)))
    res["workload"] = (f"wave_Initialize on the {NX}x{NY}x{NL} s-grid of the N=1024 sea (FP32), then "
                       f"{n} points: eta at z = 0 and u, v, w at z, FP64 arrays")
    # per-point calls as a CFD host makes them: 16 OpenMP threads, each
    # calling wave_SurfElev and wave_VeloX/Y/Z per point; the first time value
    # queries every point on its own (and remembers them), the next one is
    # served from one replayed batch; then the same with the cache off
    H = C.CDLL(os.path.join(ROOT, "cfdwavemaker_amd", "libcfdwm_hostdrv.so"))
    I64 = C.c_int64
    H.cfdwm_host_sweep.argtypes = [C.c_int, I64, P, P, P, C.c_double, P, P]
    L.wave_PrefetchAuto.argtypes = [C.c_int]
    L.wave_PrefetchAutoStats.argtypes = [C.POINTER(I64)] * 4
    npp, threads = 20_000, 16
    px, py, pz = x[:npp].copy(), y[:npp].copy(), z[:npp].copy()

    def sweep(t):
        out = np.zeros(4 * npp)
        sec = C.c_double()
        if H.cfdwm_host_sweep(threads, npp, ptr(px), ptr(py), ptr(pz), t, ptr(out), C.byref(sec)) != 0:
            raise RuntimeError("host sweep failed")
        return out, sec.value
    L.wave_PrefetchAuto(1)
    _, s_first = sweep(T0 + 0.25)
    A, s_cached = sweep(T0 + 0.30)
    st = [I64() for _ in range(4)]
    L.wave_PrefetchAutoStats(*[C.byref(v) for v in st])
    L.wave_PrefetchAuto(0)
    B, s_unc = sweep(T0 + 0.30)
    calls = 4 * npp
    res["per_point_threads"] = {
        "workload": f"{npp} points x (wave_SurfElev + wave_VeloX/Y/Z) from {threads} host threads per time value",
        "us_per_call_first_time_value": s_first / calls * 1e6,
        "us_per_call_replayed": s_cached / calls * 1e6,
        "us_per_call_uncached": s_unc / calls * 1e6,
        "calls_per_s_replayed": calls / s_cached, "bitwise_equal_to_uncached": bool(np.array_equal(A, B)),
        "cache": dict(zip(("hits", "misses", "replays", "replayed_points"), (v.value for v in st)))}
    with open(out_path, "w") as f:
        json.dump(res, f)


def abi_leg() -> dict:
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "abi.json")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--abi-worker", out],
                           stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, timeout=600)
        if r.returncode != 0:
            return {"error": r.stderr[-500:]}
        with open(out) as f:
            return json.load(f)


def update_leg(engine, grid, updates: int) -> dict:
    """Time stepping of the filled grid (CheckTime / update_lsgrid): each
    force_update advances the grid by one time level and computes that new
    level (256 x 256 columns x 16 levels: column and level passes, FD
    gradients, packing); per-update time and grid points per second."""
    grid.refill()
    n = grid.force_update(T0 + 1.5 * DT)                # warm-up: one update
    engine.sync()
    engine.prof_enable(True)
    engine.prof_read(reset=True)
    t0 = time.perf_counter()
    done = 0
    for k in range(updates):
        done += grid.force_update(T0 + (k + 2.5) * DT)
    engine.sync()
    el = time.perf_counter() - t0
    prof = engine.prof_read(reset=True)
    engine.prof_enable(False)
    if n != 1 or done != updates:
        raise RuntimeError(f"update leg: expected one time level per call, got {n} and {done}/{updates}")
    return {"workload": f"{updates} updates of the {NX}x{NY}x{NL} s-grid, one new time level each",
            "ms_per_update": el / updates * 1e3, "points_per_s": NX * NY * NL * updates / el,
            "kernels_ms": {k: v[0] / max(v[1], 1) for k, v in prof.items()}}


def lsgrid_leg(engine, sea, nq: int, steps: int) -> dict:
    """The legacy two-slot grid (lsGrid) on the same sea: a full fill (a
    re-initialising update: 2 time levels of 256 x 256 columns x 16 levels),
    one-step updates (1 time level) and nq tri/bilinear queries with FP64
    coordinates resident on the device, in cell order (k_lsquery: 32 B in +
    32 B out per query plus its share of one pass over the 25.7 MB of slots;
    queries_per_s counts the ordering pass and sort too)."""
    import torch
    g = sea.lsgrid(DOMAIN, NX, NY, NL, T0, DT, TAN_A, TAN_B, DEPTH)
    engine.sync()
    engine.prof_enable(True)
    engine.prof_read(reset=True)
    t_fill = 0.0
    for k in range(steps):
        tj = T0 + (10 * (k + 1)) * DT + 0.01
        c0 = time.perf_counter()
        if g.update(tj) != 1:
            raise RuntimeError("lsgrid leg: expected a re-initialisation")
        engine.sync()
        t_fill += time.perf_counter() - c0
    t_step = 0.0
    for k in range(steps):
        c0 = time.perf_counter()
        if g.update(g.t0 + 1.5 * DT) != 0:
            raise RuntimeError("lsgrid leg: expected one step")
        engine.sync()
        t_step += time.perf_counter() - c0
    fill_prof = engine.prof_read(reset=True)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device="cpu").manual_seed(5)
    t0 = g.t0
    qt = (t0 + torch.rand(nq, generator=gen, dtype=torch.float64) * DT).to(dev)
    qx = (torch.rand(nq, generator=gen, dtype=torch.float64) * 1000.0).to(dev)
    qy = (torch.rand(nq, generator=gen, dtype=torch.float64) * 1000.0).to(dev)
    qz = torch.zeros(nq, dtype=torch.float64, device=dev)
    out = torch.empty((nq, 4), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    g.query_device(nq, qt.data_ptr(), qx.data_ptr(), qy.data_ptr(), qz.data_ptr(), out.data_ptr())
    u = torch.rand(nq, generator=gen, dtype=torch.float64).to(dev)
    qz.copy_(-DEPTH + u * (out[:, 0] + DEPTH))
    torch.cuda.synchronize()
    engine.prof_read(reset=True)
    c0 = time.perf_counter()
    for _ in range(steps):
        g.query_device(nq, qt.data_ptr(), qx.data_ptr(), qy.data_ptr(), qz.data_ptr(), out.data_ptr())
    t_q = (time.perf_counter() - c0) / steps
    prof = engine.prof_read(reset=True)
    engine.prof_enable(False)
    q_ms = prof.get("k_lsquery", (float("nan"), 1))
    q_avg = q_ms[0] / max(q_ms[1], 1) * 1e-3
    o_ms = prof.get("k_lsquery_sort", (0.0, 1))
    o_avg = o_ms[0] / max(o_ms[1], 1) * 1e-3
    slot_bytes = 2 * (NX * NY * NL * 3 + NX * NY) * (8 if sea.precision == engine.FP64 else 4)
    alg = 64.0 * nq + slot_bytes
    g.close()
    return {"workload": f"legacy lsGrid of the N=1024 sea, {NX}x{NY} columns x {NL} levels x 2 time levels; "
                        f"{nq} tri/bilinear queries with FP64 coordinates on the device",
            "fill_ms": t_fill / steps * 1e3, "fill_points_per_s": 2 * NX * NY * NL * steps / t_fill,
            "step_ms": t_step / steps * 1e3, "step_points_per_s": NX * NY * NL * steps / t_step,
            "fill_kernels_ms": {k: v[0] / max(v[1], 1) for k, v in fill_prof.items()},
            "query_ms": t_q * 1e3, "ordering_ms": o_avg * 1e3, "query_kernel_ms": q_avg * 1e3,
            "queries_per_s": nq / (q_avg + o_avg),
            "roofline": {"bound": "hbm", "kernel": "k_lsquery", "bytes_per_launch": alg,
                         "byte_rule": "64 B per query (t, x, y, z in, eta, u, v, w out, FP64) + one pass over the "
                                      "two slots", "avg_launch_ms": q_avg * 1e3,
                         "achieved": alg / q_avg / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                         "frac": alg / q_avg / PEAK_HBM}}


def direct_leg(steps: int, smit: bool = False) -> dict:
    """Direct (per-point semantics) workload: N=512 long-crested sea, 256 x on
    [0, 500] m x 16 t on [0, 15] s columns, 64 z per column on [-30, eta]:
    eta1 + eta2 per column (FP32: eta2 as the bf16 MFMA quadratic form) and
    u, v, w, dp (unfiltered second-order sums; with smit, the Smit et al.
    s-coordinate members, top level one spacing below the surface) at the
    262,144 points, FP32 and FP64, errors of FP32 vs FP64."""
    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd.seastate import config_sea

    d = config_sea(2).as_dict()
    nx, nt, nz = 256, 16, 64
    xs = np.linspace(0.0, 500.0, nx)
    ts = np.arange(nt, dtype=float)
    X, T = (q.ravel() for q in np.meshgrid(xs, ts, indexing="ij"))
    dev = torch.device("cuda:0")
    xc = torch.tensor(X, dtype=torch.float64, device=dev)
    yc = torch.zeros_like(xc)
    tc = torch.tensor(T, dtype=torch.float64, device=dev)
    res, fields = {}, {}
    for prec, name in ((engine.FP32, "fp32"), (engine.FP64, "fp64")):
        es = engine.Sea(d, precision=prec, with_direct=not smit)
        if smit:
            es.set_vertical_theory(1)

        def run():
            c = es.eval_fields_device(tc, xc, yc, want=("eta1", "eta2"))
            eta = c["eta1"] + c["eta2"]
            # Smit velocities jump to 0 above eta + swl: keep the top level one
            # spacing below the surface, so that FP32 and FP64 elevations do not
            # put the same point on opposite sides of it
            top = 1.0 - 1.0 / nz if smit else 1.0
            frac = torch.linspace(0.0, top, nz, dtype=torch.float64, device=dev)
            z = (-30.0 + (eta[:, None] + 30.0) * frac[None, :]).reshape(-1)
            rep = lambda a: a[:, None].expand(-1, nz).reshape(-1).contiguous()
            f = es.eval_fields_device(rep(tc), rep(xc), rep(yc), z, want=("u", "v", "w", "p"))
            return eta, f

        run()
        engine.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            eta, f = run()
        engine.sync()
        el = (time.perf_counter() - t0) / steps
        res[name] = {"points_per_s": nx * nt * nz / el, "ms": el * 1e3}
        fields[name] = (eta.cpu().numpy(), {k: v.cpu().numpy() for k, v in f.items()})
        es.close()
    e32, f32 = fields["fp32"]
    e64, f64 = fields["fp64"]
    err = {"eta_max_err_over_hs": float(np.max(np.abs(e32 - e64)) / 6.0)}
    for k in ("u", "w", "p"):
        err[k + "_scaled_err"] = float(np.max(np.abs(f32[k] - f64[k])) / np.max(np.abs(f64[k])))
    what = ("Smit et al. s-coordinate u, v, w, dp (vertical_theory smit)" if smit
            else "u, v, w, dp (unfiltered second order)")
    return {"workload": "N=512 long-crested sea, 256 x 16 (x, t) columns x 64 z: eta1+eta2 per column, "
                        f"{what} per point", "points": nx * nt * nz,
            "fp32": res["fp32"], "fp64": res["fp64"], "fp32_vs_fp64": err,
            "eta2": "fp32: bf16 MFMA quadratic form (hi/lo split, FP32 accumulation)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32", "fp64"])
    ap.add_argument("--queries", type=int, default=10_000_000)
    ap.add_argument("--cpu-cols", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-worker", nargs=2, default=None)
    ap.add_argument("--abi-worker", default=None)
    ap.add_argument("--no-abi", action="store_true", help="skip the batched C-ABI leg")
    ap.add_argument("--rocprof-summary", nargs=2, metavar=("DB", "COMMAND"), default=None,
                    help="print a markdown summary of a rocprofv3 results database and exit")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct per-point workload leg")
    ap.add_argument("--no-lsgrid", action="store_true", help="skip the legacy two-slot grid leg")
    ap.add_argument("--updates", type=int, default=5,
                    help="time-stepping updates of the grid to time after the fill (0: skip)")
    ap.add_argument("--batched-states", type=int, default=512,
                    help="sea states of the batched leg (512 = the full batched workload; 0 skips it)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one plain 256 MiB device copy first (reference for PMC byte counters)")
    a = ap.parse_args()
    if a.rocprof_summary:
        print(rocprof_summary(*a.rocprof_summary), end="")
        return
    if a.cpu_worker:
        cpu_worker(a.cpu_worker[0], a.cpu_worker[1])
        return
    if a.abi_worker:
        abi_worker(a.abi_worker)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: the s-grid fill would shard by columns across GPUs, "
              f"but only N=1 is measured here; got --gpus {a.gpus}", file=sys.stderr)
        sys.exit(2)

    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd.seastate import config_sea

    prec = {"bf16": engine.BF16, "fp32": engine.FP32, "fp64": engine.FP64}[a.precision]
    s = config_sea(3).as_dict()
    print(f"[bench] building N={len(s['omega'])} sea ({a.precision})", file=sys.stderr, flush=True)
    sea = engine.Sea(s, precision=prec, with_direct=False)
    pair_terms, diag_terms, _ = sea.terms()
    T = pair_terms + diag_terms
    t_build = time.perf_counter()
    grid = sea.grid(DOMAIN, NX, NY, NL, T0, DT, TAN_A, TAN_B, DEPTH)
    print(f"[bench] first fill {time.perf_counter() - t_build:.2f} s", file=sys.stderr, flush=True)

    # query points: uniform in the box, z uniform between the seabed and the local surface
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(3)
    nq = a.queries
    qx = (torch.rand(nq, generator=g, dtype=torch.float32) * 1000.0).to(dev)
    qy = (torch.rand(nq, generator=g, dtype=torch.float32) * 1000.0).to(dev)
    qz = torch.zeros(nq, dtype=torch.float32, device=dev)
    # results {eta, u, v, w} per query (the layout of wave_Kinematics)
    out4 = torch.empty(4 * nq, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    tq = 0.25

    def query():
        grid.kinematics_device_f32(nq, tq, qx.data_ptr(), qy.data_ptr(), qz.data_ptr(), out4.data_ptr())

    query()
    engine.sync()
    u = torch.rand(nq, generator=g, dtype=torch.float32).to(dev)
    qz.copy_(-DEPTH + u * (out4[0::4] + DEPTH))
    torch.cuda.synchronize()

    def step():
        grid.refill()
        query()
        engine.sync()

    if a.calib_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(a.warmup):
        step()
    engine.prof_enable(True)
    engine.prof_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step()
        print(f"[bench] step {k + 1}/{a.steps} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    prof = engine.prof_read(reset=True)
    engine.prof_enable(False)

    points = 4 * NX * NY * NL
    value = points * a.steps / el
    # queries: cell ordering (key pass + radix sort) plus the spline kernel
    q_ms, q_n = prof.get("k_query", (float("nan"), 1))
    q_ms += prof.get("k_query_sort", (0.0, 1))[0]
    qfi = grid.qform_info()
    if qfi is not None:
        # dominant kernel: the second-order velocity pair sums as bf16 MFMA
        # quadratic forms (the launch k_qf_forms times; with one bf16 product
        # per term, k_qf_forms_eta times the separate three-product eta2 launch).
        # Algorithmic work: 2 (2N)^2 = 8 N^2 FLOP per column-slot and surface
        # form, over the 4 x 256 x 256 column-slots of a fill: u, v, w (24 N^2)
        # for the velocity launch, eta2 + u, v, w (32 N^2) for the forms step
        # (the 40 N^2 of five forms counts a pressure form the s-grid does not
        # keep).  The kernel executes more: every form of the depth bases
        # (2 R + R_w velocity forms), nlv bf16 products per term and the pair
        # triangle in tiles (exec_flop_per_launch)
        ncomp = len(s["omega"])
        cs = 4 * NX * NY
        f_ms, f_n = prof.get("k_qf_forms", (float("nan"), 1))
        e_ms, e_n = prof.get("k_qf_forms_eta", (0.0, 0))
        per_fill = max(f_n // max(a.steps, 1), 1)
        avg = f_ms / max(f_n, 1) * 1e-3
        fill_exec = qfi["mfma_flop"] / per_fill
        split = qfi["products"] == 1
        R = qfi["rank"]
        nvf = qfi["nform"] - 1                        # velocity forms: 2 R (u, v) + R_w (w)
        alg_launch = (24.0 if split else 32.0) * ncomp * ncomp * cs / per_fill
        # the eta2 form runs three products per term, each velocity form one
        flop_launch = fill_exec * nvf / (nvf + 3.0) if split else fill_exec
        roof = {"bound": "mfma", "kernel": "k_qf_forms_il<1> (velocity forms)" if split else "k_qf_forms_il<3>",
                "achieved": alg_launch / avg / 1e12,
                "peak": PEAK_MFMA_BF16 / 1e12, "unit": "TFLOP/s", "frac": alg_launch / avg / PEAK_MFMA_BF16,
                "traffic": profiled_traffic("k_qf_forms_il<1>" if split else "k_qf_forms_il<3>", a.precision),
                "traffic_source": "profiles/*_fetch_size.md + *_write_size.md (rocprofv3 --pmc, calibrated)",
                "avg_launch_ms": avg * 1e3, "flop_per_launch": alg_launch,
                "flop_rule": ("24 N^2 per column-slot (u, v, w forms x 2 (2N)^2)" if split
                              else "32 N^2 per column-slot (eta2, u, v, w forms x 2 (2N)^2)"),
                "exec_flop_per_launch": flop_launch, "exec_achieved": flop_launch / avg / 1e12,
                "exec_frac": flop_launch / avg / PEAK_MFMA_BF16, "exec_over_alg": flop_launch / alg_launch,
                "forms": qfi["nform"], "bf16_products": qfi["products"],
                "depth_basis_ranks": {"cosh_type_u_v": R, "sinh_type_w": nvf - 2 * R},
                "depth_basis_err": qfi["basis_err"], "depth_basis_tol": qfi["basis_tol"]}
        if split and e_n:
            e_avg = e_ms / e_n * 1e-3
            step_alg = 32.0 * ncomp * ncomp * cs
            roof["forms_step"] = {"kernels": "k_qf_forms_il<3> (eta2) + k_qf_forms_il<1> (velocities)",
                                  "ms": (avg + e_avg) * 1e3, "eta2_launch_ms": e_avg * 1e3,
                                  "flop": step_alg, "achieved": step_alg / (avg + e_avg) / 1e12,
                                  "frac": step_alg / (avg + e_avg) / PEAK_MFMA_BF16,
                                  "exec_flop": fill_exec, "exec_frac": fill_exec / (avg + e_avg) / PEAK_MFMA_BF16}
        # shader clock during the forms launch, from one more fill (untimed)
        # with the in-kernel clock measurement on; the dense peak assumes the
        # peak engine clock, and MFMA power with these operands lowers it
        grid.qform_clock(True)
        grid.refill()
        engine.sync()
        sclk = grid.qform_clock(False)
        if sclk:
            roof["sclk_mhz"] = sclk
            roof["peak_sclk_mhz"] = PEAK_SCLK_MHZ
            roof["frac_at_sclk"] = roof["frac"] * PEAK_SCLK_MHZ / sclk
            roof["exec_frac_at_sclk"] = roof["exec_frac"] * PEAK_SCLK_MHZ / sclk
    else:
        lev_ms, lev_n = prof.get("k_levels_grid", (float("nan"), 1))
        lev_avg = lev_ms / max(lev_n, 1) * 1e-3
        flop_per_launch = 16.0 * T * points
        roof = {"bound": "valu", "kernel": "k_levels_grid", "achieved": flop_per_launch / lev_avg / 1e12,
                "peak": PEAK[a.precision] / 1e12, "unit": "TFLOP/s",
                "frac": flop_per_launch / lev_avg / PEAK[a.precision], "traffic": None,
                "avg_launch_ms": lev_avg * 1e3, "flop_per_launch": flop_per_launch}

    # rooflines of the other kernels of the step
    lev = prof.get("k_levels_grid", (float("nan"), 1))
    col = prof.get("k_columns_grid", (float("nan"), 1))
    t_lin = (lev[0] / max(lev[1], 1) + col[0] / max(col[1], 1)) * 1e-3
    lin_flop = 4.0 * len(s["omega"]) * (3 * NL + 4) * 4 * NX * NY
    t_q = q_ms / max(q_n, 1) * 1e-3
    rooflines = {
        "spline_query": {"kernels": "k_query_sort + k_query", "bound": "hbm", "bytes_per_query": QUERY_BYTES,
                         "achieved": QUERY_BYTES * nq / t_q / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                         "frac": QUERY_BYTES * nq / t_q / PEAK_HBM},
        "column_linear_sums": {"kernels": "k_columns_grid + k_levels_grid", "bound": "valu",
                               "flop_per_column": 4.0 * len(s["omega"]) * (3 * NL + 4),
                               "achieved": lin_flop / t_lin / 1e12, "peak": PEAK[a.precision] / 1e12,
                               "unit": "TFLOP/s", "frac": lin_flop / t_lin / PEAK[a.precision],
                               # the FP32 peak counts packed FP32 FMAs, which this build does not
                               # emit; one FMA per lane and cycle gives half of it
                               "peak_unpacked": PEAK_UNPACKED[a.precision] / 1e12,
                               "frac_unpacked": lin_flop / t_lin / PEAK_UNPACKED[a.precision]},
    }
    upd = update_leg(engine, grid, a.updates) if a.updates > 0 else None
    lsg = None if a.no_lsgrid else lsgrid_leg(engine, sea, nq, a.steps)

    # accuracy gate: FP32 (or FP64) grid vs an FP64 grid on an 8 x 8 block of the same columns
    dx = (DOMAIN[1] - DOMAIN[0]) / (NX - 1)
    sub = (0.0, 7 * dx, 0.0, 7 * dx)
    g64 = engine.Sea(s, precision=engine.FP64, with_direct=False).grid(sub, 8, 8, NL, T0, DT, TAN_A, TAN_B, DEPTH)
    gxx = sea.grid(sub, 8, 8, NL, T0, DT, TAN_A, TAN_B, DEPTH)
    eta_err = float(np.max(np.abs(gxx.export("ETA") - g64.export("ETA"))))
    uref = g64.export("UX")
    u_err = float(np.max(np.abs(gxx.export("UX") - uref)) / np.max(np.abs(uref)))

    sur = None if a.no_surrogate else surrogate_leg(a.steps, a.warmup, a.precision if a.precision != "fp64" else "fp32")
    bat = batched_leg(a.batched_states, a.precision if a.precision != "fp64" else "fp32") \
        if a.batched_states > 0 else None
    dirl = None if a.no_direct else direct_leg(a.steps)
    smitl = None if a.no_direct else direct_leg(a.steps, smit=True)
    abil = None if a.no_abi else abi_leg()

    cpu = None
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baselines ({threads} threads)", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(f"grid:{a.cpu_cols // 2},{a.cpu_cols}", threads,
                               f"s-grid fills of {a.cpu_cols // 2}^2 and {a.cpu_cols}^2 columns x {NL} levels x 4 "
                               f"time levels of the same N=1024 sea and spacing (value: the larger)")
        except Exception as e:  # reported, not fatal
            cpu = {"value": None, "error": str(e)[-300:]}
        if lsg is not None:
            try:
                lsg["cpu_baseline"] = cpu_baseline("lsgrid:4,8", threads,
                                                   "legacy-grid fills of 4^2 and 8^2 columns x 16 levels x 2 time "
                                                   "levels of the same N=1024 sea and spacing (value: the larger)")
                lsg["fill_over_cpu"] = lsg["fill_points_per_s"] / lsg["cpu_baseline"]["value"]
            except Exception as e:  # reported, not fatal
                lsg["cpu_baseline"] = {"value": None, "error": str(e)[-300:]}
        if dirl is not None:
            try:
                dirl["cpu_baseline"] = cpu_baseline("direct:4,8", threads,
                                                    "eta at 4 and 8 (x, t) columns and u, v, w, dp at their 64 z "
                                                    "each, N=512 long-crested (value: the larger)")
                dirl["fp32_over_cpu"] = dirl["fp32"]["points_per_s"] / dirl["cpu_baseline"]["value"]
            except Exception as e:  # reported, not fatal
                dirl["cpu_baseline"] = {"value": None, "error": str(e)[-300:]}

    line = {
        "metric": "2nd-order kinematics points/s (N=1024), exact GPU",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16/f32", "fp32": "f32", "fp64": "f64"}[a.precision],
        "data": "synthetic (JONSWAP + cos^2s spreading, seeded)",
        "config": {"workload": "N=1024 short-crested sea, s-grid 256x256 columns x 16 levels x 4 time "
                               "levels + 1e7 spline queries", "components": len(s["omega"]),
                   "pair_terms": pair_terms, "diag_terms": diag_terms,
                   "points_per_step": points, "queries_per_step": nq, "precision": a.precision},
        "roofline": roof,
        "rooflines_other": rooflines,
        "kernels_ms": {k: v[0] / max(v[1], 1) for k, v in prof.items()},
        "query_rate": nq / (q_ms / max(q_n, 1) * 1e-3) if q_ms == q_ms else None,
        "accuracy": {"eta_max_err_m": eta_err, "eta_err_over_hs": eta_err / HS, "ux_scaled_err": u_err,
                     "gate": "eta error <= 1e-3 Hs vs FP64", "pass": eta_err <= 1e-3 * HS},
        "cpu_baseline": cpu,
        "surrogate": sur,
        "batched": bat,
        "direct": dirl,
        "direct_smit": smitl,
        "update": upd,
        "legacy_lsgrid": lsg,
        "abi_batched": abil,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
