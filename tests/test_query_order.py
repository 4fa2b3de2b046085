"""Cell-ordered spline queries: a batch large enough to be sorted by cell
(65,536 queries or more) returns bit for bit what the same queries return in
the caller's order (batches below that size, whose results the golden query
vectors pin), for the host FP64 entry and the device FP32 entry, in both
grid precisions."""
import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden
from tests.golden.cases import GRID_CASE

pytestmark = pytest.mark.gpu

N_BIG, N_CHUNK = 200_000, 50_000
# the golden grid (35 columns) and a 64 x 48 column grid (12-bit sort keys)
SHAPES = {"golden": GRID_CASE,
          "64x48": dict(GRID_CASE, domain=(0.0, 630.0, -235.0, 235.0), nx=64, ny=48, nl=8)}


def _grid(prec, shape):
    from cfdwavemaker_amd import engine as eng
    c = SHAPES[shape]
    sea = eng.Sea(sea_from_golden(golden("points.npz"), "B"), precision=prec, with_direct=False)
    g = sea.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"], c["tan_b"], c["depth"])
    return sea, g


def _points(seed, shape):
    c = SHAPES[shape]
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = c["domain"]
    # a margin outside the domain exercises the edge clamps
    x = rng.uniform(x0 - 5.0, x1 + 5.0, N_BIG)
    y = rng.uniform(y0 - 5.0, y1 + 5.0, N_BIG)
    z = rng.uniform(-c["depth"], 1.0, N_BIG)
    t = rng.uniform(c["t0"], c["t0"] + c["dt"], N_BIG)
    return t, x, y, z


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_queries_cell_order_matches_caller_order(precision, shape):
    from cfdwavemaker_amd import engine as eng
    prec = eng.FP64 if precision == "fp64" else eng.FP32
    sea, g = _grid(prec, shape)
    t, x, y, z = _points(11, shape)
    big = g.query(t, x, y, z)
    parts = [g.query(t[a:a + N_CHUNK], x[a:a + N_CHUNK], y[a:a + N_CHUNK], z[a:a + N_CHUNK])
             for a in range(0, N_BIG, N_CHUNK)]
    small = np.concatenate(parts)
    assert np.all(np.isfinite(big))
    assert np.array_equal(big, small)


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_f32_queries_cell_order_matches_caller_order(precision, shape):
    import torch
    from cfdwavemaker_amd import engine as eng
    prec = eng.FP64 if precision == "fp64" else eng.FP32
    sea, g = _grid(prec, shape)
    t, x, y, z = _points(12, shape)
    tq = float(t[0])
    dev = torch.device("cuda:0")
    X, Y, Z = (torch.tensor(a, dtype=torch.float32, device=dev) for a in (x, y, z))
    big = [torch.empty(N_BIG, dtype=torch.float32, device=dev) for _ in range(4)]
    small = [torch.empty(N_BIG, dtype=torch.float32, device=dev) for _ in range(4)]
    g.query_device_f32(N_BIG, tq, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), *[o.data_ptr() for o in big])
    eng.sync()
    for a in range(0, N_BIG, N_CHUNK):
        g.query_device_f32(N_CHUNK, tq, X[a:].data_ptr(), Y[a:].data_ptr(), Z[a:].data_ptr(),
                           *[o[a:].data_ptr() for o in small])
    eng.sync()
    for b, s in zip(big, small):
        b, s = b.cpu().numpy(), s.cpu().numpy()
        assert np.all(np.isfinite(b))
        assert np.array_equal(b, s)
