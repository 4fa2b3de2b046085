"""Pin the oracle's Smit et al. (2017) s-coordinate members (vertical_theory
smit, src/Irregular.cpp:1623-2027) against the compiled reference, and check
their linear limit against Airy theory.

Golden vectors: tests/golden/smit.npz, written from the compiled reference by
tests/golden/make_golden.py (five seas, orders 2 and 1).
"""
import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden, scaled_err
from oracle import SMIT_KINDS

CASES = ["A", "B", "C", "D", "E"]
TOL = 1e-12


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("order", [2, 1])
def test_smit_members_vs_reference(oracle, case, order):
    g = golden("smit.npz")
    d = sea_from_golden(golden("points.npz"), case)
    d["order"] = order
    sea = oracle.sea(d)
    sea.set_vertical_theory(1)
    key = f"{case}{order}"
    t, x, y, z = g[f"{key}__pts"]
    h = sea.pairs_h()
    assert h.shape == g[f"{key}__H"].shape
    assert scaled_err(h, g[f"{key}__H"]) <= TOL
    for kind, disp in zip(SMIT_KINDS, ("u", "v", "w", "dp")):
        ref = g[f"{key}__{kind}"]
        assert np.all(np.isfinite(ref))
        assert scaled_err(sea.eval(kind, t, x, y, z), ref) <= TOL, (key, kind)
        # the dispatching member routes to the Smit member (Irregular.cpp:25, :53, :77, :143)
        assert scaled_err(sea.eval(disp, t, x, y, z), ref) <= TOL, (key, disp)


def test_smit_linear_limit_is_airy(oracle):
    """As the amplitudes scale by eps -> 0 the stretch height and every
    second-order term vanish at O(eps): u_smit / eps tends to Airy's linear
    velocity u1 at unit amplitude (cosh k(z+h)/cosh kh profiles), with an
    O(eps) relative remainder.  A wrong stretched profile, a dropped ramp or
    a second-order term of the wrong order fails this."""
    d = sea_from_golden(golden("points.npz"), "A")
    d.update(swl=0.0, ramptime=None, starttime=None, duration=None)
    rng = np.random.default_rng(5)
    n = 24
    t = rng.uniform(0.0, 20.0, n)
    x = rng.uniform(-50.0, 50.0, n)
    y = rng.uniform(-20.0, 20.0, n)
    z = rng.uniform(-0.9 * d["depth"], -0.1 * d["depth"], n)
    base = oracle.sea(d)
    lin = {c: base.eval(c, t, x, y, z) for c in ("u1", "v1", "w1", "dp1")}
    errs = []
    for eps in (1e-3, 1e-4):
        de = dict(d, A=np.asarray(d["A"]) * eps)
        s = oracle.sea(de)
        s.set_vertical_theory(1)
        e = max(scaled_err(s.eval(k, t, x, y, z) / eps, lin[c])
                for k, c in zip(SMIT_KINDS, ("u1", "v1", "w1", "dp1")))
        errs.append(e)
    assert errs[0] < 5e-3 and errs[1] < 5e-4, errs
    assert errs[1] < 0.2 * errs[0], errs          # the remainder is O(eps)
