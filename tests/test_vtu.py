"""s-grid .vtu export: XML structure, node positions from the sigma map,
velocities of the query's base slot, cell connectivity."""
import os
import tempfile
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden
from tests.golden.cases import GRID_CASE

pytestmark = pytest.mark.gpu


def _arrays(root):
    out = {}
    for da in root.iter("DataArray"):
        name = da.get("Name") or "Points"
        out[name] = np.array(da.text.split(), dtype=float)
    return out


@pytest.mark.parametrize("ny", [5, 1])
def test_vtu_export(ny):
    from cfdwavemaker_amd import engine as eng
    from cfdwavemaker_amd import surrogate as S
    c = GRID_CASE
    sea = eng.Sea(sea_from_golden(golden("points.npz"), "B"), precision=eng.FP64, with_direct=False)
    nx, nl = c["nx"], c["nl"]
    g = sea.grid(c["domain"], nx, ny, nl, c["t0"], c["dt"], c["tan_a"], c["tan_b"], c["depth"])
    g.force_update(c["t0"] + c["dt"] + 0.05)                   # tstep 1: base slot (tstep + 1) % 4 = 2
    st = g.state()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "k.vtu")
        g.write_vtu(path, time_label="TimeValue")
        root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    npts, ncell = int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))
    assert npts == nx * ny * nl
    A = _arrays(root)
    assert abs(A["TimeValue"][0] - st["t0"]) < 1e-6
    tt = (st["tstep"] + 1) % 4
    nv = nx * ny * nl
    vel = A["Velocity"].reshape(-1, 3)
    # the reference's number formats: velocities %g (6 significant digits),
    # positions %12.4f
    for k, f in enumerate(("UX", "UY", "UZ")):
        ref = g.export(f)[tt * nv:(tt + 1) * nv]
        assert np.allclose(vel[:, k], ref, rtol=6e-6, atol=1e-12 * max(1.0, np.abs(ref).max()))
    pts = A["Points"].reshape(-1, 3)
    eta = g.export("ETA")[tt * nx * ny:(tt + 1) * nx * ny]
    sig = S.sigma_levels(nl, c["tan_a"], c["tan_b"])
    z = (eta[:, None] + sig[None, :] * (c["depth"] + eta[:, None])).ravel()
    assert np.allclose(pts[:, 2], z, rtol=0, atol=5.1e-5)
    conn = A["connectivity"].astype(int)
    if ny > 1:
        assert ncell == (nx - 1) * (ny - 1) * (nl - 1)
        assert np.all(A["types"] == 12)
        i, j, m = 1, 2, 3
        n = lambda a, b, cc: (a * ny + b) * nl + cc
        cell = ((i * (ny - 1) + j) * (nl - 1) + m)
        assert list(conn[8 * cell:8 * cell + 8]) == [n(i, j, m), n(i + 1, j, m), n(i + 1, j + 1, m), n(i, j + 1, m),
                                                     n(i, j, m + 1), n(i + 1, j, m + 1), n(i + 1, j + 1, m + 1),
                                                     n(i, j + 1, m + 1)]
    else:
        assert ncell == (nx - 1) * (nl - 1) and np.all(A["types"] == 9)
    assert A["offsets"][-1] == conn.size
