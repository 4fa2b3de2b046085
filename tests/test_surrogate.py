"""Learned second-order surrogate: linear-field feature kernel against the
oracle, the committed network's gain over linear theory on held-out sea
states, and the inference pipeline (features -> network -> assembly ->
residual check)."""
import os

import numpy as np
import pytest

from tests.conftest import golden, sea_from_golden

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cfdwavemaker_amd",
                       "surrogate_weights.npz")


def test_weights_file_shapes():
    z = np.load(WEIGHTS, allow_pickle=False)
    nin, width, nh, nout = (int(v) for v in z["dims"])
    assert (nin, nout) == (29, 65) and nh == 3 and width == 256
    n = 0
    d = nin
    for l in range(nh + 1):
        o = width if l < nh else nout
        n += d * o + o
        d = o
    assert z["params"].size == n and np.all(np.isfinite(z["params"]))
    assert np.all(z["in_std"] > 0) and np.all(z["out_std"] > 0)


@pytest.mark.gpu
def test_linear_features_vs_oracle(oracle):
    import torch
    from cfdwavemaker_amd import engine, surrogate as S
    g = golden("points.npz")
    s = sea_from_golden(g, "B")
    h = float(s["depth"])
    es = engine.Sea(s, precision=engine.FP64, with_direct=False)
    osea = oracle.sea(s)
    rng = np.random.default_rng(11)
    n = 64
    x, y, t = rng.uniform(0, 200, n), rng.uniform(-50, 50, n), 3.7
    kp = S.peak_wavenumber(s)
    X, Lin = S.linear_fields(es, torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda"), t, h, kp)
    engine.sync()
    X, Lin = X.cpu().numpy().astype(float), Lin.cpu().numpy().astype(float)
    e1 = osea.eval("eta1", t, x, y)
    sc = np.abs(e1).max()
    assert np.abs(X[:, 0] - e1).max() <= 1e-6 * sc
    d = 1e-3
    sx = (osea.eval("eta1", t, x + d, y) - osea.eval("eta1", t, x - d, y)) / (2 * d)
    sy = (osea.eval("eta1", t, x, y + d) - osea.eval("eta1", t, x, y - d)) / (2 * d)
    sl = max(np.abs(sx).max(), np.abs(sy).max())
    assert np.abs(X[:, 1] - sx).max() <= 1e-5 * sl and np.abs(X[:, 2] - sy).max() <= 1e-5 * sl
    sig = S.sigma_levels()
    z = e1[:, None] + sig[None, :] * (h + e1[:, None])
    kinds = (("u1", "phi1_dxdz"), ("v1", "phi1_dydz"), ("w1", "phi1_dzdz"), ("dp1", "dp1_dz"))
    for c, (k, gk) in enumerate(kinds):
        ref = np.zeros_like(z)
        for l in range(z.shape[1]):
            zl = z[:, l]
            below = osea.eval(k, t, x, y, np.minimum(zl, 0.0))
            grad = osea.eval(gk, t, x, y, np.zeros(n))
            ref[:, l] = below + np.maximum(zl, 0.0) * grad
        got = Lin[:, 1 + c::4]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"{k} levels scaled err {err:.2e}", flush=True)
        assert err <= 2e-6, (k, err)
        if c < 3:
            feat = X[:, 3 + c:27:3]
            assert np.allclose(feat, got[:, ::2])


@pytest.mark.gpu
def test_network_gain_over_linear_theory():
    """Held-out sea states (never in the training set): the network's
    second-order prediction beats the method without it (linear theory,
    correction zero) for eta2 and the velocities."""
    import torch
    from cfdwavemaker_amd import surrogate as S
    from cfdwavemaker_amd.mlp import Mlp
    m, _ = Mlp.load_file(WEIGHTS)
    X, Y = S.build_dataset(range(20000, 20003), ncol=24)
    P = m.forward(torch.tensor(X, device="cuda")).cpu().numpy()
    for name, sl in (("eta2", slice(0, 1)), ("u2", slice(1, None, 4)), ("w2", slice(3, None, 4))):
        e = np.sqrt(np.mean((P[:, sl] - Y[:, sl]) ** 2))
        e0 = np.sqrt(np.mean(Y[:, sl] ** 2))            # error without the network
        print(f"{name}: rel rms error with network {e / e0:.3f} (without: 1.000)", flush=True)
        assert e < 0.8 * e0, (name, e / e0)
    m.close()


@pytest.mark.gpu
def test_inference_pipeline_and_residual_check():
    import torch
    from cfdwavemaker_amd import engine, surrogate as S
    from cfdwavemaker_amd.seastate import config_sea
    sea = config_sea(5)
    d = sea.as_dict()
    es = engine.Sea(d, precision=engine.FP32, with_direct=False)
    sur = S.Surrogate.load()
    n = 32
    g = np.linspace(0, 300, n)
    x, y = (a.ravel() for a in np.meshgrid(g, g, indexing="ij"))
    out, Lin, Yp = sur.fields(es, sea, torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda"), 0.25)
    engine.sync()
    out, Lin, Yp = (a.cpu().numpy() for a in (out, Lin, Yp))
    assert np.all(np.isfinite(out)) and np.allclose(out, Lin + Yp, atol=1e-6 * np.abs(out).max())
    err, ok = sur.residual_check(es, x, y, 0.25, out[:, 0], np.arange(0, n * n, 7), 1e-3 * sea.hs)
    print(f"surrogate eta max err on sample {err:.4f} m (gate {1e-3 * sea.hs} m): {'pass' if ok else 'fallback'}")
    assert np.isfinite(err) and ok == (err <= 1e-3 * sea.hs)
