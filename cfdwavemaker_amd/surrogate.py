"""Learned second-order surrogate: a per-column MLP that predicts the
second-order correction (eta2 and u2, v2, w2, dp2 at the column's levels)
from linear-field features, with the exact kernels as fallback and judge.

The reference has no learned part; what the network replaces is the pair sums
of src/Irregular.cpp (Irregular::eta2 with the pair cache, and the direct
u2/v2/w2/dp2 members).  Columns are evaluated at the levels
z_l = eta1 + sigma_l (h + eta1) of the lsGridSpline sigma map (tan-stretched,
src/lsgrid_spline.cpp:114-119) with the linear elevation eta1.

Features (29): eta1, d eta1/dx, d eta1/dy, (u1, v1, w1) at the 8 even levels,
the depth h and the spectral peak wavenumber kp.  Outputs (65): eta2, then per
level (u2, v2, w2, dp2).  Every GPU step (features, network, assembly) runs in
libcfdwm's HIP kernels; this module orchestrates them.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import engine
from ._lib import check
from .mlp import Mlp
from .seastate import SeaState, config_sea

NLEV = 16
FSTRIDE = 2
NIN = 3 + 3 * (NLEV // FSTRIDE) + 2
NOUT = 1 + 4 * NLEV
TAN_A, TAN_B = 0.7, 1.5
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_weights.npz")


def sigma_levels(nl=NLEV, tan_a=TAN_A, tan_b=TAN_B):
    """sigma of each level (lsgrid_spline.cpp:114-119 sigmaS2cart), -1 at the bed, 0 at the surface."""
    s = np.arange(nl) / (nl - 1)
    return -np.power(np.tan((1.0 - s) * tan_a), tan_b) / np.power(np.tan(tan_a), tan_b)


def peak_wavenumber(sea: SeaState | dict) -> float:
    d = sea.as_dict() if isinstance(sea, SeaState) else sea
    return float(np.asarray(d["k"])[int(np.argmax(np.asarray(d["A"])))])


def sea_scales(sea: SeaState | dict):
    """Per-sea scales that make features and targets dimensionless: with
    a = Hs / 2 (Hs = 4 sqrt(sum A^2 / 2)), the peak wavenumber kp and
    omega_p = sqrt(g kp tanh(kp h)):  features / [a, kp a, kp a, (omega_p a) x 24,
    1 / kp, 1 / a] and targets / [kp a^2, (omega_p kp a^2) x 3 and rho g kp a^2
    per level].  Returns (sx[29], sy[65])."""
    d = sea.as_dict() if isinstance(sea, SeaState) else sea
    A = np.asarray(d["A"], float)
    a = 2.0 * np.sqrt(np.sum(A * A) / 2.0)
    kp = peak_wavenumber(d)
    g, rho, h = float(d.get("g", 9.81)), float(d.get("rho", 1025.0)), float(d["depth"])
    wp = np.sqrt(g * kp * np.tanh(kp * h))
    sx = np.concatenate([[a, kp * a, kp * a], np.full(3 * (NLEV // FSTRIDE), wp * a), [1.0 / kp, 1.0 / a]])
    lev = np.tile([wp * kp * a * a] * 3 + [rho * g * kp * a * a], NLEV)
    sy = np.concatenate([[kp * a * a], lev])
    return sx, sy


def linear_fields(esea: "engine.Sea", x, y, t, h, kp, sigma=None):
    """Features X [n, 29] and linear fields Lin [n, 65] (float32 cuda tensors)
    for columns x, y (float64 cuda tensors) at time t."""
    import torch
    sigma = sigma_levels() if sigma is None else np.asarray(sigma, float)
    n = x.shape[0]
    X = torch.empty((n, NIN), dtype=torch.float32, device=x.device)
    Lin = torch.empty((n, NOUT), dtype=torch.float32, device=x.device)
    sg = np.ascontiguousarray(sigma, dtype=np.float64)
    L = esea.L
    check(L.cfdwm_surrogate_linear(esea.h, n, x.data_ptr(), y.data_ptr(), float(t), float(h), len(sg),
                                   sg.ctypes.data_as(C.POINTER(C.c_double)), FSTRIDE, float(kp),
                                   X.data_ptr(), Lin.data_ptr()), "cfdwm_surrogate_linear")
    return X, Lin


def exact_second_order(esea: "engine.Sea", x, y, t, eta1, h, sigma=None):
    """Exact second-order parts (numpy [n, 65]): eta2 (pair cache), then per
    level u2, v2, w2, dp2 (direct members) at z_l = eta1 + sigma_l (h + eta1)."""
    sigma = sigma_levels() if sigma is None else np.asarray(sigma, float)
    x, y, eta1 = (np.asarray(a, float) for a in (x, y, eta1))
    n = x.size
    Y = np.zeros((n, NOUT))
    Y[:, 0] = esea.eval("eta2", t, x, y)
    z = eta1[:, None] + sigma[None, :] * (h + eta1[:, None])          # n x nlev
    xx = np.repeat(x, len(sigma))
    yy = np.repeat(y, len(sigma))
    zz = z.ravel()
    for c, kind in enumerate(("u2", "v2", "w2", "dp2")):
        Y[:, 1 + c::4] = esea.eval(kind, t, xx, yy, zz).reshape(n, len(sigma))
    return Y


def columns_for(sea: SeaState, ncol: int, seed: int):
    """ncol x ncol columns over two peak wavelengths, random offset."""
    lam = 2.0 * np.pi / peak_wavenumber(sea)
    rng = np.random.default_rng(seed)
    x0, y0 = rng.uniform(0, 4 * lam, 2)
    g = np.linspace(0.0, 2.0 * lam, ncol)
    X, Y = np.meshgrid(x0 + g, y0 + g, indexing="ij")
    return X.ravel(), Y.ravel()


def build_dataset(states, ncol=64, precision=engine.FP32, log=None):
    """Features and exact targets over the random sea states of the batched
    workload (N 128-2048, Hs 1-8 m, Tp 6-16 s, depth 20-200 m, long- and
    short-crested), ncol x ncol columns each, made dimensionless with each
    sea's scales.  Returns float32 numpy X, Y."""
    import torch
    Xs, Ys = [], []
    for st in states:
        sea = config_sea(4, st)
        d = sea.as_dict()
        es = engine.Sea(d, precision=precision, with_direct=True)
        x, y = columns_for(sea, ncol, 7000 + st)
        t = float(np.random.default_rng(9000 + st).uniform(0.0, 600.0))
        xd = torch.tensor(x, dtype=torch.float64, device="cuda")
        yd = torch.tensor(y, dtype=torch.float64, device="cuda")
        X, Lin = linear_fields(es, xd, yd, t, sea.depth, peak_wavenumber(sea))
        engine.sync()
        Xh = X.cpu().numpy()
        Yh = exact_second_order(es, x, y, t, Xh[:, 0].astype(np.float64), sea.depth)
        sx, sy = sea_scales(d)
        Xs.append((Xh / sx).astype(np.float32))
        Ys.append((Yh / sy).astype(np.float32))
        es.close()
        if log:
            log(f"state {st}: N={len(d['omega'])} hs={sea.hs:.2f} tp={sea.tp:.2f} h={sea.depth:.1f}")
    return np.concatenate(Xs), np.concatenate(Ys)


def train(X, Y, steps=4000, batch=4096, lr=2e-3, width=256, nhidden=3, seed=0, report=200):
    """Train the MLP on the GPU with the library's kernels; returns (Mlp, loss curve)."""
    import torch
    m = Mlp(X.shape[1], width, nhidden, Y.shape[1], seed=seed)
    sd_in = X.std(0)
    sd_in[sd_in < 1e-12] = 1.0
    sd_out = Y.std(0)
    sd_out[sd_out < 1e-12] = 1.0
    m.set_norm(X.mean(0), sd_in, Y.mean(0), sd_out)
    Xd = torch.tensor(X, dtype=torch.float32, device="cuda")
    Yd = torch.tensor(Y, dtype=torch.float32, device="cuda")
    curve = m.train(Xd, Yd, steps=steps, batch=batch, lr=lr, seed=seed + 1, report=report)
    return m, curve


class Surrogate:
    """Trained network + the inference path with a residual check against the
    exact elevation on a sample of columns."""

    def __init__(self, mlp: Mlp):
        self.mlp = mlp
        self.base_norm = mlp.get_norm()        # dimensionless standardisation
        self._scaled_for = None

    @classmethod
    def load(cls, path=WEIGHTS):
        m, _ = Mlp.load_file(path)
        return cls(m)

    def set_sea(self, sea):
        """Fold the sea's scales into the network's standardisation."""
        sx, sy = sea_scales(sea)
        mi, si, mo, so = self.base_norm
        self.mlp.set_norm(mi * sx, si * sx, mo * sy, so * sy)

    def fields(self, esea, sea, x, y, t):
        """Surrogate total fields [n, 65] (eta, then u, v, w, p per level), the
        linear part and the predicted second-order part (float32 cuda tensors)
        for columns x, y (float64 cuda tensors) at time t."""
        import torch
        d = sea.as_dict() if isinstance(sea, SeaState) else sea
        key = id(sea)
        if self._scaled_for != key:
            self.set_sea(d)
            self._scaled_for = key
        X, Lin = linear_fields(esea, x, y, t, float(d["depth"]), peak_wavenumber(d))
        Yp = self.mlp.forward(X)
        out = torch.empty_like(Lin)
        check(esea.L.cfdwm_surrogate_add(Lin.numel(), Lin.data_ptr(), Yp.data_ptr(), out.data_ptr()),
              "cfdwm_surrogate_add")
        return out, Lin, Yp

    @staticmethod
    def residual_check(esea, x, y, t, eta_pred, sample, gate):
        """max |eta_surrogate - eta_exact| over the sampled columns (host
        arrays) and whether it is within the gate."""
        xs, ys = np.asarray(x)[sample], np.asarray(y)[sample]
        ex = esea.eval("eta1", t, xs, ys) + esea.eval("eta2", t, xs, ys)
        err = float(np.max(np.abs(np.asarray(eta_pred)[sample] - ex)))
        return err, err <= gate
