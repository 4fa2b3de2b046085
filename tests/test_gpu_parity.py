"""GPU parity: the HIP path (through libcfdwm's C ABI) against the reference's
golden vectors and the FP64 oracle.

FP64 mode bar: scale-normalised error |err| / max|ref| <= 1e-9 (the stated
contract is 1e-6; the profile ratio form and GPU libm differ from the
reference only at rounding level).  FP32 mode bars are stated per test.
"""
import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden, scaled_err
from tests.golden.cases import GRID_CASE

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
CASES = ["A", "B", "C", "D", "E"]
KINDS = ["eta1", "eta2", "eta", "u1", "v1", "w1", "dp1", "phi1_dxdz", "phi1_dydz", "phi1_dzdz",
         "dp1_dz", "u2", "v2", "w2", "dp2", "u", "v", "w", "dp", "uvw2_u", "uvw2_v", "uvw2_w"]


@pytest.fixture(scope="module")
def eng():
    from cfdwavemaker_amd import engine
    engine.load()
    return engine


@pytest.mark.parametrize("case", CASES)
def test_point_kinds_fp64_vs_golden(eng, case):
    g = golden("points.npz")
    sea = eng.Sea(sea_from_golden(g, case), precision=eng.FP64)
    t, x, y, z = g[f"{case}__pts"]
    for kind in KINDS:
        got = sea.eval(kind, t, x, y, z)
        ref = g[f"{case}__{kind}"]
        err = scaled_err(got, ref)
        print(f"case {case} {kind:10s} scaled err {err:.2e}", flush=True)
        assert err <= FP64_TOL, (case, kind, err)


def test_linear_section_vs_golden(eng):
    """The reference's CPU linear case: N=256, order 0, a 64 x 32 vertical
    section at t = 0 (eta, u, w, dp); FP64 to the golden vectors, FP32 to the
    FP32 bound."""
    g = golden("linear.npz")
    t, x, y, z = g["L__pts"]
    for prec, tol in ((eng.FP64, FP64_TOL), (eng.FP32, 1e-5)):
        sea = eng.Sea(sea_from_golden(g, "L"), precision=prec)
        for kind in ("eta", "u", "w", "dp"):
            err = scaled_err(sea.eval(kind, t, x, y, z), g[f"L__{kind}"])
            print(f"linear prec {prec} {kind:4s} scaled err {err:.2e}", flush=True)
            assert err <= tol, (prec, kind, err)


@pytest.mark.parametrize("ign", [0, 1])
def test_sgrid_fp64_vs_golden(eng, ign):
    g = golden("grid.npz")
    p = golden("points.npz")
    sea = eng.Sea(sea_from_golden(p, "B"), precision=eng.FP64, with_direct=False)
    c = GRID_CASE
    grid = sea.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"],
                    c["depth"], 0.0, (30.0, 70.0, -25.0, 25.0), ign)
    for step in range(4):
        if step:
            grid.force_update(c["t0"] + step * c["dt"] + 0.05)
        st = grid.state()
        ref_state = g[f"ign{ign}__s{step}__state"]
        assert st["tstep"] == int(ref_state[0]) and abs(st["t0"] - ref_state[1]) < 1e-15
        for f in eng.GRID_FIELDS:
            ref = g[f"ign{ign}__s{step}__{f}"]
            got = grid.export(f)
            if f == "IGNORE":
                assert np.array_equal(got, ref)
                continue
            err = scaled_err(got, ref)
            assert err <= FP64_TOL, (step, f, err)
        q = g[f"ign{ign}__s{step}__q"]
        got = grid.query(*q)
        ref = g[f"ign{ign}__s{step}__qout"]
        for col in range(4):
            err = scaled_err(got[:, col], ref[:, col])
            print(f"ign {ign} step {step} query col {col} err {err:.2e}", flush=True)
            assert err <= FP64_TOL, (step, col, err)


def test_fp32_point_kinds_error_bound(eng, oracle):
    """FP32 mode (FP64 phase reduction, FP64 totals): stated bound 1e-5 of the field scale."""
    g = golden("points.npz")
    for case in ("A", "B", "E"):
        sea = eng.Sea(sea_from_golden(g, case), precision=eng.FP32)
        t, x, y, z = g[f"{case}__pts"]
        for kind in ("eta", "u", "w", "dp", "uvw2_u", "uvw2_w"):
            err = scaled_err(sea.eval(kind, t, x, y, z), g[f"{case}__{kind}"])
            print(f"fp32 case {case} {kind} err {err:.2e}", flush=True)
            assert err <= 1e-5, (case, kind, err)


def test_config2_subset_fp64_vs_oracle(eng, oracle):
    """Direct second-order mode at the N=512 long-crested workload (subset of points)."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(2).as_dict()
    sea = eng.Sea(s, precision=eng.FP64)
    osea = oracle.sea(s)
    rng = np.random.default_rng(2)
    n = 24
    t = rng.choice(np.arange(16.0), n)
    x = rng.uniform(0, 500, n)
    y = np.zeros(n)
    eta = osea.eval("eta", t, x, y)
    z = rng.uniform(-30.0, 1.0, n)
    for kind in ("eta", "u", "w", "dp"):
        ref = eta if kind == "eta" else osea.eval(kind, t, x, y, z)
        got = sea.eval(kind, t, x, y, z)
        err = scaled_err(got, ref)
        print(f"cfg2 {kind} err {err:.2e}", flush=True)
        assert err <= FP64_TOL, (kind, err)


def test_config3_columns_fp64_fp32_vs_oracle(eng, oracle):
    """s-grid column fill at the N=1024 short-crested workload: a 3x3-column
    grid (16 levels, 4 time levels) against the oracle's column body."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    osea = oracle.sea(s)
    dom = (0.0, 1000.0 * 2 / 255, 0.0, 1000.0 * 2 / 255)   # the first 3x3 columns of the 256^2 grid
    ref = osea.grid(dom, 3, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    fields = ["ETA", "UX", "UY", "UZ"]
    R = {f: ref.export(f) for f in fields}
    hs = 8.0
    for prec, tol in ((eng.FP64, FP64_TOL), (eng.FP32, 1e-5)):
        sea = eng.Sea(s, precision=prec, with_direct=False)
        grid = sea.grid(dom, 3, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
        for f in fields:
            err = scaled_err(grid.export(f), R[f])
            print(f"cfg3 prec {prec} {f} err {err:.2e}", flush=True)
            assert err <= tol, (prec, f, err)
        # elevation gate of the benchmark: |eta error| <= 1e-3 Hs
        assert np.max(np.abs(grid.export("ETA") - R["ETA"])) <= 1e-3 * hs


def test_split_k_matches_single_segment(eng, monkeypatch):
    """The pair sum split over segments (split-K) equals the single-segment sum."""
    g = golden("points.npz")
    s = sea_from_golden(g, "B")
    t, x, y, z = g["B__pts"]
    out = {}
    for nseg in ("1", "7", "64"):
        monkeypatch.setenv("CFDWM_NSEG", nseg)
        sea = eng.Sea(s, precision=eng.FP64)
        out[nseg] = {k: sea.eval(k, t, x, y, z) for k in ("eta", "u", "dp", "uvw2_w")}
        grid = sea.grid((0.0, 120.0, -20.0, 20.0), 7, 5, 6, 0.0, 0.5, 0.7, 1.5, 30.0)
        out[nseg]["UX"] = grid.export("UX")
        out[nseg]["ETA"] = grid.export("ETA")
    for k in out["1"]:
        for nseg in ("7", "64"):
            err = scaled_err(out[nseg][k], out["1"][k])
            assert err <= 1e-12, (k, nseg, err)


def test_eta2_quadratic_form_matches_pair_loop(eng, oracle, monkeypatch):
    """FP32 s-grid elevation with the filtered eta2 pair sum on the matrix
    cores (bf16 hi/lo split, FP32 accumulation) against the FP32 pair loop and
    the FP64 oracle: stated bound 1e-5 of the elevation scale."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dom = (0.0, 1000.0 * 20 / 255, 0.0, 1000.0 * 12 / 255)   # 21 x 13 columns of the 256^2 grid
    ref = oracle.sea(s).grid(dom, 21, 13, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    eta_ref = ref.export("ETA")
    out = {}
    for qf in ("1", "0"):
        monkeypatch.setenv("CFDWM_QFORM", qf)
        grid = eng.Sea(s, precision=eng.FP32, with_direct=False).grid(dom, 21, 13, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
        out[qf] = grid.export("ETA")
        err = scaled_err(out[qf], eta_ref)
        print(f"qform {qf} ETA err vs oracle {err:.2e}", flush=True)
        assert err <= 1e-5, (qf, err)
    assert scaled_err(out["1"], out["0"]) <= 1e-5


def test_points_eta2_quadratic_form(eng, oracle, monkeypatch):
    """FP32 point batches (>= 2048 points) take the filtered eta2 pair sum from
    the MFMA quadratic form; against the FP64 oracle and the FP32 pair loop."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(2).as_dict()
    rng = np.random.default_rng(21)
    n = 4096
    t, x, y = rng.uniform(0, 15, n), rng.uniform(0, 500, n), np.zeros(n)
    osea = oracle.sea(s)
    sub = np.arange(0, n, 16)
    ref = osea.eval("eta2", t[sub], x[sub], y[sub])
    out = {}
    for qf in ("1", "0"):
        monkeypatch.setenv("CFDWM_QFORM", qf)
        sea = eng.Sea(s, precision=eng.FP32, with_direct=False)
        out[qf] = sea.eval("eta2", t, x, y)
        err = scaled_err(out[qf][sub], ref)
        print(f"points eta2 qform {qf} err vs oracle {err:.2e}", flush=True)
        assert err <= 1e-5, (qf, err)
    assert scaled_err(out["1"], out["0"]) <= 1e-5
