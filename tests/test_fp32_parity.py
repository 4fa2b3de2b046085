"""FP32 and BF16 modes against the reference's golden vectors and the FP64
oracle: the s-grid query paths, the reference's C-ABI call programs and the
one-slot grid update.

Stated bounds (scale-normalised error |err| / max|ref|):
  FP32 (FP64 phase reduction, three-product bf16 quadratic forms): 2e-5 for
      grid fields, queries and ABI outputs;
  BF16 (one bf16 product per quadratic-form term): the elevation gate
      |eta err| <= 1e-3 Hs, and 3e-3 of the field scale for velocities (the
      rounding of one bf16 product: 0.8e-3 to 2.4e-3 on these blocks,
      depending on the depth basis; with three products the same fills
      stay within 2e-5, so the basis itself contributes no more).
"""
import os

import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden, scaled_err
from tests.golden.cases import GRID_CASE
from tests.test_abi import run_program

pytestmark = pytest.mark.gpu

FP32_TOL = 2e-5
BF16_VEL_TOL = 3e-3


@pytest.fixture(scope="module")
def eng():
    from cfdwavemaker_amd import engine
    engine.load()
    return engine


def _grid(eng, prec, ign):
    p = golden("points.npz")
    sea = eng.Sea(sea_from_golden(p, "B"), precision=prec, with_direct=False)
    c = GRID_CASE
    grid = sea.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"],
                    c["depth"], 0.0, (30.0, 70.0, -25.0, 25.0), ign)
    return sea, grid


@pytest.mark.parametrize("ign", [0, 1])
def test_fp32_grid_fields_and_queries_vs_golden(eng, ign):
    """FP32 s-grid (fill, three updates, FD gradients, packing) and its spline
    query with FP64 coordinates (k_query<float, double>) against the compiled
    reference's grid arrays and get_kinematics_at_point outputs."""
    g = golden("grid.npz")
    sea, grid = _grid(eng, eng.FP32, ign)
    c = GRID_CASE
    for step in range(4):
        if step:
            grid.force_update(c["t0"] + step * c["dt"] + 0.05)
        for f in ("ETA", "UX", "UY", "UZ", "ETAdt", "Udt", "Udx", "Wds"):
            err = scaled_err(grid.export(f), g[f"ign{ign}__s{step}__{f}"])
            assert err <= FP32_TOL, (step, f, err)
        got = grid.query(*g[f"ign{ign}__s{step}__q"])
        ref = g[f"ign{ign}__s{step}__qout"]
        for col in range(4):
            err = scaled_err(got[:, col], ref[:, col])
            print(f"fp32 ign {ign} step {step} query col {col} err {err:.2e}", flush=True)
            assert err <= FP32_TOL, (step, col, err)


@pytest.mark.parametrize("prec_name", ["fp32", "bf16"])
def test_device_f32_query_vs_oracle(eng, oracle, prec_name):
    """The FP32 device query (k_query<float, float>, FP32 coordinates and
    outputs) on an FP32 or BF16 grid of the N=1024 sea's first 24 x 20
    columns, against the FP64 oracle grid queried at the same points: a batch
    under the cell-ordering threshold and one above it (65,536 queries)."""
    import torch
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    dom = (0.0, 23 * dx, 0.0, 19 * dx)
    og = oracle.sea(s).grid(dom, 24, 20, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    prec = eng.FP32 if prec_name == "fp32" else eng.BF16
    grid = eng.Sea(s, precision=prec, with_direct=False).grid(dom, 24, 20, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    rng = np.random.default_rng(31)
    t = 0.3
    for n in (5000, 70000):
        x = rng.uniform(0.0, dom[1], n).astype(np.float32)
        y = rng.uniform(0.0, dom[3], n).astype(np.float32)
        # z between the seabed and the local surface of the oracle grid
        eta = og.query(np.full(n, t), x.astype(float), y.astype(float), np.zeros(n))[:, 0]
        z = (-100.0 + rng.uniform(0.0, 1.0, n) * (eta + 100.0)).astype(np.float32)
        ref = og.query(np.full(n, t), x.astype(float), y.astype(float), z.astype(float))
        dev = torch.device("cuda:0")
        X, Y, Z = (torch.tensor(a, device=dev) for a in (x, y, z))
        O = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(4)]
        grid.query_device_f32(n, t, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), *[o.data_ptr() for o in O])
        eng.sync()
        got = np.stack([o.cpu().numpy().astype(float) for o in O], axis=1)
        e_eta = float(np.max(np.abs(got[:, 0] - ref[:, 0])))
        print(f"{prec_name} n {n} eta max err {e_eta:.3e} m", flush=True)
        assert e_eta <= 1e-3 * 8.0
        for col in range(4):
            err = scaled_err(got[:, col], ref[:, col])
            print(f"{prec_name} n {n} col {col} scaled err {err:.2e}", flush=True)
            assert err <= (FP32_TOL if prec_name == "fp32" else BF16_VEL_TOL), (n, col, err)


@pytest.mark.parametrize("prec_name", ["fp32", "bf16"])
@pytest.mark.parametrize("fixture", ["abi_direct.npz", "abi_grid.npz", "abi_smit.npz"])
def test_reference_programs_fp32(monkeypatch, fixture, prec_name):
    """The reference's C-ABI call programs (wave types 1 and 5, ramps, swl,
    Smit) with CFDWM_PRECISION=fp32 / bf16 against the reference's outputs."""
    monkeypatch.setenv("CFDWM_PRECISION", prec_name)
    g, rc, out = run_program(fixture)
    assert rc == int(g["rc"]) == 0
    ref = g["out"]
    ops = g["op"]
    tol = FP32_TOL if prec_name == "fp32" else BF16_VEL_TOL
    for op in np.unique(ops):
        sel = ops == op
        for c in range(4):
            err = scaled_err(out[sel, c], ref[sel, c])
            print(f"{prec_name} {fixture} op {op} col {c} err {err:.2e}", flush=True)
            assert err <= tol, (fixture, int(op), c, err)


@pytest.mark.parametrize("prec_name", ["fp32", "bf16"])
def test_force_update_one_slot_vs_oracle(eng, oracle, prec_name):
    """The one-slot update (CheckTime / update, the forms run on one time
    level) of a 4 x 3 column block of the N=1024 sea, three updates, against
    the oracle's grid after the same updates."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    dom = (10 * dx, 13 * dx, 40 * dx, 42 * dx)
    og = oracle.sea(s).grid(dom, 4, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    prec = eng.FP32 if prec_name == "fp32" else eng.BF16
    grid = eng.Sea(s, precision=prec, with_direct=False).grid(dom, 4, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    for k in range(1, 4):
        t = k * 0.5 + 0.05
        assert grid.force_update(t) == 1
        og.force_update(t)
        for f in ("ETA", "UX", "UY", "UZ", "ETAdt", "Udt"):
            ref = og.export(f)
            got = grid.export(f)
            if f == "ETA":
                e = float(np.max(np.abs(got - ref)))
                assert e <= 1e-3 * 8.0, (k, e)
            err = scaled_err(got, ref)
            print(f"{prec_name} update {k} {f} scaled err {err:.2e}", flush=True)
            assert err <= (FP32_TOL if prec_name == "fp32" else BF16_VEL_TOL), (k, f, err)


def test_bf16_mode_reports_one_product(eng):
    """A BF16 sea's fill runs the one-product forms; FP32 the three-product ones."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    for prec, products in ((eng.BF16, 1), (eng.FP32, 3)):
        g = eng.Sea(s, precision=prec, with_direct=False).grid((0, 3 * dx, 0, 3 * dx), 4, 4, 8, 0.0, 0.5,
                                                                0.7, 1.5, 100.0)
        info = g.qform_info()
        assert info is not None and info["products"] == products, info


def test_bf16_side_stream_fill_equals_serial_fill(eng, monkeypatch):
    """The velocity forms of a one-chunk bf16 fill run on a second stream
    beside the column pass and the level pass joins them: fills and updates
    equal those of the serial order (CFDWM_QF_OVERLAP=0) bit for bit."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    dom = (50 * dx, 65 * dx, 20 * dx, 35 * dx)
    out = {}
    for ov in ("1", "0"):
        monkeypatch.setenv("CFDWM_QF_OVERLAP", ov)
        g = eng.Sea(s, precision=eng.BF16, with_direct=False).grid(dom, 16, 16, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
        assert g.qform_info() is not None
        res = [g.export(f) for f in ("ETA", "UX", "UY", "UZ", "Udt")]
        g.force_update(0.55)
        res += [g.export(f) for f in ("ETA", "UX", "UY", "UZ", "Udt")]
        out[ov] = res
    for a, b in zip(out["1"], out["0"]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_segmented_level_pass_matches_one_segment(eng, monkeypatch):
    """A bf16 fill of 64 x 64 columns x 16 levels x 4 time levels splits its
    column and level passes over component segments (partial records reduced
    in segment order; the level pass's records carry the linear-term basis
    sums, evaluated at the levels once per column by the reduction).  With
    CFDWM_NSEG=1 the same fill runs in one segment.  The two agree to the FP32
    rounding of the partial sums."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    n, nl = 64, 16
    out = {}
    for nseg in ("1", "default"):
        if nseg == "1":
            monkeypatch.setenv("CFDWM_NSEG", "1")
        else:
            monkeypatch.delenv("CFDWM_NSEG", raising=False)
        g = eng.Sea(s, precision=eng.BF16, with_direct=False).grid((0.0, (n - 1) * dx, 0.0, (n - 1) * dx), n, n, nl,
                                                                    0.0, 0.5, 0.7, 1.5, 100.0)
        assert g.qform_info() is not None
        out[nseg] = {f: g.export(f) for f in ("ETA", "UX", "UY", "UZ")}
        del g
    for f in ("ETA", "UX", "UY", "UZ"):
        err = scaled_err(out["default"][f], out["1"][f])
        print(f"segments vs one segment {f} {err:.2e}", flush=True)
        assert err <= 1e-5, (f, err)


@pytest.mark.parametrize("prec_name,lin_tol,bound,n", [("bf16", "1e-5", 1e-4, 256), ("fp32", "1e-6", 2e-5, 256),
                                                       ("bf16", "1e-5", 1e-4, 12)])
def test_linear_term_bases_vs_per_level_profiles(eng, oracle, monkeypatch, prec_name, lin_tol, bound, n):
    """The level pass's first-order and diagonal terms through the
    linear-term profile bases (default in the bf16 mode, CFDWM_LIN_TOL) against
    the per-level profiles (CFDWM_LIN_TOL=0) on the same n x n x 16 x 4 fill
    (n = 12: the level pass splits the components over segments and sums
    their partial records): the difference stays within a small multiple of
    the bases' tolerance of the field scale, and the FP32 fill with them stays
    within the FP32 bound of the oracle on a corner block of columns."""
    from cfdwavemaker_amd.seastate import config_sea
    s = config_sea(3).as_dict()
    dx = 1000.0 / 255
    nl = 16
    prec = eng.FP32 if prec_name == "fp32" else eng.BF16
    out = {}
    for tol in (lin_tol, "0"):
        monkeypatch.setenv("CFDWM_LIN_TOL", tol)
        g = eng.Sea(s, precision=prec, with_direct=False).grid((0.0, (n - 1) * dx, 0.0, (n - 1) * dx), n, n, nl, 0.0,
                                                                0.5, 0.7, 1.5, 100.0)
        assert g.qform_info() is not None
        out[tol] = {f: g.export(f) for f in ("ETA", "UX", "UY", "UZ")}
        del g
    for f in ("UX", "UY", "UZ"):
        err = scaled_err(out[lin_tol][f], out["0"][f])
        print(f"{prec_name} lin_tol {lin_tol} {f} bases vs per-level {err:.2e}", flush=True)
        assert err <= bound, (f, err)
    assert np.array_equal(out[lin_tol]["ETA"], out["0"]["ETA"])
    if prec_name == "fp32":
        ref = oracle.sea(s).grid((0.0, 3 * dx, 0.0, 3 * dx), 4, 4, nl, 0.0, 0.5, 0.7, 1.5, 100.0)
        for f in ("UX", "UY", "UZ"):
            got = out[lin_tol][f].reshape(4, n, n, nl)[:, :4, :4, :]
            assert scaled_err(got, ref.export(f).reshape(4, 4, 4, nl)) <= FP32_TOL, f
