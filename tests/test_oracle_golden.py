"""Pin the FP64 oracle against golden vectors written from the compiled reference.

The oracle restates the reference in the same expression order, so agreement
is at rounding level (bit-exact in practice on this libm); the bound used here
is 1e-12 of the field scale.
"""
import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden, scaled_err
from oracle import KINDS, GRID_FIELDS

CASES = ["A", "B", "C", "D", "E"]
TOL = 1e-12


@pytest.mark.parametrize("case", CASES)
def test_point_functions(oracle, case):
    g = golden("points.npz")
    sea = oracle.sea(sea_from_golden(g, case))
    t, x, y, z = g[f"{case}__pts"]
    for kind in KINDS:
        ref = g[f"{case}__{kind}"]
        got = sea.eval(kind, t, x, y, z)
        assert scaled_err(got, ref) <= TOL, (case, kind, scaled_err(got, ref))


def test_linear_section(oracle):
    """The reference's CPU linear case (order 0: no clamp, u1(z - swl))."""
    g = golden("linear.npz")
    sea = oracle.sea(sea_from_golden(g, "L"))
    t, x, y, z = g["L__pts"]
    for kind in ("eta", "u", "w", "dp"):
        err = scaled_err(sea.eval(kind, t, x, y, z), g[f"L__{kind}"])
        assert err <= TOL, (kind, err)


@pytest.mark.parametrize("case", CASES)
def test_pair_cache(oracle, case):
    g = golden("pairs.npz")
    p = golden("points.npz")
    sea = oracle.sea(sea_from_golden(p, case))
    im, fl, v = sea.pairs()
    assert np.array_equal(im, g[f"{case}__im"])
    assert np.array_equal(fl, g[f"{case}__flags"])
    ref = g[f"{case}__vals"]
    for c in range(7):
        assert scaled_err(v[:, c], ref[:, c]) <= TOL, c
    assert sea.kmax() == float(g[f"{case}__kmax"])


@pytest.mark.parametrize("ign", [0, 1])
def test_sgrid_fill_update_query(oracle, ign):
    from tests.golden.cases import GRID_CASE
    g = golden("grid.npz")
    p = golden("points.npz")
    sea = oracle.sea(sea_from_golden(p, "B"))
    c = GRID_CASE
    grid = sea.grid(c["domain"], c["nx"], c["ny"], c["nl"], c["t0"], c["dt"], c["tan_a"],
                    c["tan_b"], c["depth"], 0.0, (30.0, 70.0, -25.0, 25.0), ign)
    for step in range(4):
        if step:
            grid.force_update(c["t0"] + step * c["dt"] + 0.05)
        st = grid.state()
        ref_state = g[f"ign{ign}__s{step}__state"]
        assert st["tstep"] == int(ref_state[0]) and st["t0"] == ref_state[1]
        for f in GRID_FIELDS:
            ref = g[f"ign{ign}__s{step}__{f}"]
            got = grid.export(f)
            if f == "IGNORE":
                assert np.array_equal(got, ref)
                continue
            assert scaled_err(got, ref) <= TOL, (step, f)
        q = g[f"ign{ign}__s{step}__q"]
        got = grid.query(*q)
        ref = g[f"ign{ign}__s{step}__qout"]
        for col in range(4):
            assert scaled_err(got[:, col], ref[:, col]) <= TOL, (step, col)
