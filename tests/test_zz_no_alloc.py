"""The allocation-free device query (cfdwm_grid_query_device_f32_noalloc)
against the allocating one it is a variant of (cfdwm_grid_query_device_f32):
two calls in a row through the C ABI, the second with a smaller batch, both
large enough to take the cell-ordered path whose buffers live in the fixed
device array; bit-identical outputs required.  Also: a batch under the
ordering threshold, and a batch too large for the array is refused before
any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def last_error():
    from cfdwavemaker_amd._lib import last_error as le
    return le()


def _lib():
    from cfdwavemaker_amd._lib import load
    L = load()
    L.cfdwm_grid_query_device_f32_noalloc.restype = C.c_int
    L.cfdwm_grid_query_device_f32_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_double] + [C.c_void_p] * 7
    return L


def _queries(torch, n, seed, extent):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = (torch.rand(n, generator=g, dtype=torch.float32) * extent).cuda()
    y = (torch.rand(n, generator=g, dtype=torch.float32) * extent).cuda()
    z = (-40.0 + torch.rand(n, generator=g, dtype=torch.float32) * 40.0).cuda()
    return x, y, z


def test_noalloc_query_matches_allocating_query():
    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd.seastate import config_sea
    L = _lib()
    sea = engine.Sea(config_sea(3).as_dict(), precision=engine.FP32, with_direct=False)
    dx = 1000.0 / 255
    extent = 47 * dx
    grid = sea.grid((0.0, extent, 0.0, extent), 48, 48, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    t = 0.3
    for n, seed in ((300_000, 1), (120_000, 2), (5_000, 3)):
        print(f"noalloc query n={n}", flush=True)
        x, y, z = _queries(torch, n, seed, extent)
        ref = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(4)]
        got = [torch.full((n,), float("nan"), dtype=torch.float32, device="cuda") for _ in range(4)]
        grid.query_device_f32(n, t, x.data_ptr(), y.data_ptr(), z.data_ptr(), *[o.data_ptr() for o in ref])
        engine.sync()
        rc = L.cfdwm_grid_query_device_f32_noalloc(grid.h, n, t, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                                                   *[o.data_ptr() for o in got])
        assert rc == 0, last_error()
        engine.sync()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), (n, float((a - b).abs().max()))


def test_noalloc_query_refuses_oversized_batch():
    import torch
    from cfdwavemaker_amd import engine
    from cfdwavemaker_amd.seastate import config_sea
    L = _lib()
    sea = engine.Sea(config_sea(3).as_dict(), precision=engine.FP32, with_direct=False)
    grid = sea.grid((0.0, 100.0, 0.0, 100.0), 8, 8, 8, 0.0, 0.5, 0.7, 1.5, 100.0)
    n = 8_000_000
    buf = torch.zeros(n, dtype=torch.float32, device="cuda")
    p = buf.data_ptr()
    rc = L.cfdwm_grid_query_device_f32_noalloc(grid.h, n, 0.3, p, p, p, p, p, p, p)
    assert rc == -1
    msg = last_error()
    assert "exceed the maximum" in msg, msg
