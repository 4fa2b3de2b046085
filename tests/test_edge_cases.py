"""Edge cases of the batched paths: empty batches, batch sizes around the
thresholds where the implementation changes path (2,048 points: the eta2 pair
sum as the MFMA form in FP32; 65,536 queries: cell order), single points,
and an empty batched refill."""
import numpy as np
import pytest

from tests.conftest import golden, scaled_err, sea_from_golden
from tests.golden.cases import GRID_CASE

pytestmark = pytest.mark.gpu


def _sea(prec, case="B"):
    from cfdwavemaker_amd import engine as eng
    return eng.Sea(sea_from_golden(golden("points.npz"), case), precision=prec, with_direct=True)


def test_empty_batches():
    import torch
    from cfdwavemaker_amd import engine as eng
    for prec in (eng.FP64, eng.FP32):
        s = _sea(prec)
        e = np.zeros(0)
        for kind in ("eta1", "eta2", "eta", "u", "v", "w", "dp"):
            out = s.eval(kind, e, e, e, e if kind not in ("eta1", "eta2", "eta") else None)
            assert out.shape == (0,)
        c = GRID_CASE
        g = s.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"], c["depth"])
        assert g.query(e, e, e, e).shape == (0, 4)
        buf = torch.zeros(4, dtype=torch.float32, device="cuda")
        g.query_device_f32(0, 0.0, *([buf.data_ptr()] * 7))
        eng.sync()
        f = s.eval_fields_device(*(torch.zeros(0, dtype=torch.float64, device="cuda") for _ in range(4)))
        assert all(v.numel() == 0 for v in f.values())
    st = eng.refill_batched([])
    assert st["batched"] == 0 and st["single"] == 0


@pytest.mark.parametrize("n", [1, 255, 257, 2047, 2048, 2049, 4097])
def test_point_batch_sizes_fp32_vs_fp64(n):
    """FP32 point batches below and above the MFMA-form threshold agree with
    the FP64 evaluation (golden-pinned) to the FP32 bound."""
    from cfdwavemaker_amd import engine as eng
    rng = np.random.default_rng(n)
    x = rng.uniform(-50, 150, n)
    y = rng.uniform(-20, 20, n)
    t = rng.uniform(0, 30, n)
    z = rng.uniform(-25, 0.5, n)
    s64, s32 = _sea(eng.FP64), _sea(eng.FP32)
    for kind in ("eta2", "eta", "u", "w"):
        zz = None if kind.startswith("eta") else z
        a, b = s32.eval(kind, t, x, y, zz), s64.eval(kind, t, x, y, zz)
        ref = s64.eval(kind, *(np.resize(q, 4096) for q in (t, x, y)), None if zz is None else np.resize(zz, 4096))
        # scale: the field's magnitude over a fixed 4096-point sample (a single
        # point's own value can be arbitrarily small)
        err = float(np.max(np.abs(a - b)) / max(np.max(np.abs(ref)), 1e-300))
        assert err <= 1e-5, (kind, n, err)


@pytest.mark.parametrize("n", [65535, 65536, 65537])
def test_query_sizes_around_cell_order_threshold(n):
    """Batches just below, at and above the cell-order threshold return, bit
    for bit, what the same queries return in small batches."""
    from cfdwavemaker_amd import engine as eng
    s = _sea(eng.FP32)
    c = GRID_CASE
    g = s.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"], c["depth"])
    rng = np.random.default_rng(n)
    x0, x1, y0, y1 = c["domain"]
    pts = (rng.uniform(c["t0"], c["t0"] + c["dt"], n), rng.uniform(x0, x1, n), rng.uniform(y0, y1, n),
           rng.uniform(-c["depth"], 0.5, n))
    big = g.query(*pts)
    small = np.concatenate([g.query(*(q[a:a + 10000] for q in pts)) for a in range(0, n, 10000)])
    assert np.array_equal(big, small)


def test_fill_in_several_chunks_matches_single_chunk():
    """FP32 fills whose Z operands exceed one chunk (2 GiB: 131,072 column
    slots at N = 2048) are processed in point chunks; the shared columns of a
    200 x 200 grid (two chunks) and a 40 x 40 grid (one) agree."""
    from cfdwavemaker_amd import engine as eng
    from cfdwavemaker_amd.seastate import make_sea
    s = eng.Sea(make_sea(2048, 4.0, 10.0, 50.0, seed=21).as_dict(), precision=eng.FP32, with_direct=False)
    dx = 4.0
    big = s.grid((0.0, 199 * dx, 0.0, 199 * dx), 200, 200, 4, 0.0, 0.5, 0.7, 1.5, 50.0)
    assert big.qform_info() is not None
    small = s.grid((0.0, 39 * dx, 0.0, 39 * dx), 40, 40, 4, 0.0, 0.5, 0.7, 1.5, 50.0)
    eb = big.export("ETA").reshape(4, 200, 200)[:, :40, :40]
    es = small.export("ETA").reshape(4, 40, 40)
    assert float(np.max(np.abs(eb - es))) <= 1e-6 * float(np.max(np.abs(es)))
    ub = big.export("UX").reshape(4, 200, 200, 4)[:, :40, :40, :]
    us = small.export("UX").reshape(4, 40, 40, 4)
    assert float(np.max(np.abs(ub - us))) <= 1e-6 * float(np.max(np.abs(us)))


def test_point_batch_eta2_in_several_chunks():
    """FP32 point batches take eta2 from the MFMA form in chunks of 131,072
    points at N = 2048: one batch of 150,000 points (two chunks) agrees with
    three batches of 50,000 (one chunk each)."""
    from cfdwavemaker_amd import engine as eng
    from cfdwavemaker_amd.seastate import make_sea
    s = eng.Sea(make_sea(2048, 4.0, 10.0, 50.0, seed=22).as_dict(), precision=eng.FP32, with_direct=False)
    rng = np.random.default_rng(22)
    n = 150_000
    t, x, y = rng.uniform(0, 20, n), rng.uniform(0, 500, n), rng.uniform(-100, 100, n)
    big = s.eval("eta2", t, x, y)
    small = np.concatenate([s.eval("eta2", t[a:a + 50000], x[a:a + 50000], y[a:a + 50000])
                            for a in range(0, n, 50000)])
    assert float(np.max(np.abs(big - small))) <= 1e-6 * float(np.max(np.abs(small)))


def test_forms_clock_measurement_leaves_fill_unchanged():
    """The shader-clock measurement of the form launches (a diagnostic) leaves
    the fill bit for bit unchanged and reports a clock in the GPU's range;
    fills after it is switched off are not measured."""
    from cfdwavemaker_amd import engine as eng
    from cfdwavemaker_amd.seastate import make_sea
    s = eng.Sea(make_sea(256, 4.0, 10.0, 50.0, seed=5).as_dict(), precision=eng.FP32, with_direct=False)
    g = s.grid((0.0, 124.0, 0.0, 124.0), 32, 32, 4, 0.0, 0.5, 0.7, 1.5, 50.0)
    assert g.qform_info() is not None
    names = ("ETA", "UX", "UZ")
    ref = [g.export(k) for k in names]
    assert g.qform_clock(True) is None            # the fill at creation was not measured
    g.refill()
    assert all(np.array_equal(a, g.export(k)) for a, k in zip(ref, names))
    mhz = g.qform_clock(False)
    assert mhz is not None and 100.0 < mhz < 3500.0, mhz
    g.refill()
    assert g.qform_clock(False) is None
    assert all(np.array_equal(a, g.export(k)) for a, k in zip(ref, names))
