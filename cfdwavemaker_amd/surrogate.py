"""Learned second-order surrogate: a per-column MLP that predicts the
second-order correction (u2, v2, w2, dp2 at the column's levels, and its own
estimate of eta2) from linear-field features, with the exact kernels as
fallback and judge.

The reference has no learned part; what the network replaces is the pair sums
of src/Irregular.cpp (Irregular::eta2 with the pair cache, :223-293, and the
direct u2/v2/w2/dp2 members, :450-803).  Columns are evaluated at the levels
z_l = eta1 + sigma_l (h + eta1) of the lsGridSpline sigma map (tan-stretched,
src/lsgrid_spline.cpp:114-119) with the linear elevation eta1.

Two modes:
  hybrid  -- the elevation's second-order part is the exact filtered eta2,
             evaluated for every column as one MFMA quadratic form (three
             bf16 products per term, about 1/15 of the matrix work of an exact
             s-grid fill); the network predicts the velocity and pressure
             profiles, the other 14/15, and its eta2 output serves as a
             monitor.  The elevation therefore meets the 1e-3 Hs gate with the
             exact form's error.
  learned -- the network predicts eta2 too (the first round's surrogate).

Feature columns (computed on the GPU for every column; a network uses a
subset, stored with its weights):
  0..28   eta1, d eta1/dx, d eta1/dy, (u1, v1, w1) at the 8 even levels, h, kp
  29..31  the quadrature (Hilbert pair) of eta1, sum A r sin(theta), and its
          x and y slopes
  32      the exact eta2 of the column (hybrid networks)
Outputs (65): eta2, then per level (u2, v2, w2, dp2).  Every GPU step
(features, eta2 form, network, assembly) runs in libcfdwm's HIP kernels; this
module orchestrates them.
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
NBASE = 3 + 3 * (NLEV // FSTRIDE) + 2          # 29 local linear features
COL_QUAD = NBASE                                # 3 quadrature features
COL_ETA2 = NBASE + 3                            # exact eta2
NFEAT = NBASE + 4                               # 33 feature columns
NIN = NBASE
NOUT = 1 + 4 * NLEV
TAN_A, TAN_B = 0.7, 1.5
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_weights.npz")

FEATURE_SETS = {
    "local": list(range(NBASE)),
    "local+quad": list(range(NBASE + 3)),
    "local+quad+eta2": list(range(NFEAT)),
}


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
    1 / kp, 1 / a, a, kp a, kp a, kp a^2] and targets / [kp a^2,
    (omega_p kp a^2) x 3 and rho g kp a^2 per level].  Returns (sx[33], sy[65])."""
    d = sea.as_dict() if isinstance(sea, SeaState) else sea
    A = np.asarray(d["A"], float)
    a = 2.0 * np.sqrt(np.sum(A * A) / 2.0)
    kp = peak_wavenumber(d)
    g, rho, h = float(d.get("g", 9.81)), float(d.get("rho", 1025.0)), float(d["depth"])
    wp = np.sqrt(g * kp * np.tanh(kp * h))
    sx = np.concatenate([[a, kp * a, kp * a], np.full(3 * (NLEV // FSTRIDE), wp * a), [1.0 / kp, 1.0 / a],
                         [a, kp * a, kp * a], [kp * a * a]])
    lev = np.tile([wp * kp * a * a] * 3 + [rho * g * kp * a * a], NLEV)
    sy = np.concatenate([[kp * a * a], lev])
    return sx, sy


def linear_fields(esea: "engine.Sea", x, y, t, h, kp, sigma=None, eta2=True):
    """Features X [n, 33] and linear fields Lin [n, 65] (float32 cuda tensors)
    for columns x, y (float64 cuda tensors) at time t; with eta2, column 32
    holds the exact eta2 (filtered pair sum, as the MFMA quadratic form for
    FP32 seas and batches of 2048 columns or more) and the float64 eta2 tensor
    is returned as well (else None)."""
    import torch
    sigma = sigma_levels() if sigma is None else np.asarray(sigma, float)
    n = x.shape[0]
    # the library's kernels run on its own stream: every tensor they touch is
    # created before the first launch and read by torch only after a sync
    X = torch.empty((n, NFEAT), dtype=torch.float32, device=x.device)
    Lin = torch.empty((n, NOUT), dtype=torch.float32, device=x.device)
    tt = torch.full((n,), float(t), dtype=torch.float64, device=x.device) if eta2 else None
    torch.cuda.synchronize()
    sg = np.ascontiguousarray(sigma, dtype=np.float64)
    L = esea.L
    check(L.cfdwm_surrogate_features(esea.h, n, x.data_ptr(), y.data_ptr(), float(t), float(h), len(sg),
                                     sg.ctypes.data_as(C.POINTER(C.c_double)), FSTRIDE, float(kp), NFEAT,
                                     X.data_ptr(), Lin.data_ptr()), "cfdwm_surrogate_features")
    e2 = None
    if eta2:
        e2 = esea.eval_fields_device(tt, x, y, want=("eta2",))["eta2"]
        check(L.cfdwm_surrogate_set_column(n, e2.data_ptr(), X.data_ptr(), NFEAT, COL_ETA2),
              "cfdwm_surrogate_set_column")
    engine.sync()
    return X, Lin, e2


def exact_second_order(esea: "engine.Sea", x, y, t, eta1, h, sigma=None):
    """Exact second-order parts (numpy [n, 65]): eta2 (pair cache), then per
    level the pair-cache velocity sums of the s-grid fill (Irregular::uvw2,
    src/Irregular.cpp:975-1102) and the direct pressure dp2 (:739-803) at
    z_l = eta1 + sigma_l (h + eta1), evaluated at min(z_l, 0) as the
    reference's dispatch and column update do (above the still-water level
    the second-order part is its value at z = 0; only the first-order part
    is extended by its gradient)."""
    sigma = sigma_levels() if sigma is None else np.asarray(sigma, float)
    x, y, eta1 = (np.asarray(a, float) for a in (x, y, eta1))
    n = x.size
    Y = np.zeros((n, NOUT))
    Y[:, 0] = esea.eval("eta2", t, x, y)
    z = eta1[:, None] + sigma[None, :] * (h + eta1[:, None])          # n x nlev
    xx = np.repeat(x, len(sigma))
    yy = np.repeat(y, len(sigma))
    zz = np.minimum(z.ravel(), 0.0)
    for c, kind in enumerate(("uvw2_u", "uvw2_v", "uvw2_w", "dp2")):
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
    """All 33 feature columns and the exact targets over the random sea states
    of the batched workload (N 128-2048, Hs 1-8 m, Tp 6-16 s, depth 20-200 m,
    long- and short-crested), ncol x ncol columns each, made dimensionless
    with each sea's scales.  Returns float32 numpy X [rows, 33], Y [rows, 65]."""
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
        X, Lin, _ = linear_fields(es, xd, yd, t, sea.depth, peak_wavenumber(sea))
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
    """Trained network + the inference path (features, exact eta2 form for
    hybrid networks, network, assembly) + a residual check against the exact
    kernels on a sample of columns."""

    def __init__(self, mlp: Mlp, cols=None, mode="learned"):
        self.mlp = mlp
        self.cols = list(range(NBASE)) if cols is None else [int(c) for c in cols]
        if len(self.cols) != mlp.nin:
            raise ValueError("feature columns and network input width differ")
        self.mode = mode
        self.base_norm = mlp.get_norm()        # dimensionless standardisation
        self._scaled_for = None
        self._cols_t = None

    @classmethod
    def load(cls, path=WEIGHTS):
        m, z = Mlp.load_file(path)
        cols = z["feature_cols"] if "feature_cols" in z.files else None
        mode = str(z["mode"]) if "mode" in z.files else "learned"
        return cls(m, cols, mode)

    def save(self, path, **extra):
        self.mlp.save(path, feature_cols=np.asarray(self.cols, dtype=np.int32), mode=np.str_(self.mode), **extra)

    @property
    def needs_eta2(self):
        return self.mode == "hybrid" or COL_ETA2 in self.cols

    def set_sea(self, sea):
        """Fold the sea's scales into the network's standardisation."""
        sx, sy = sea_scales(sea)
        sx = sx[self.cols]
        mi, si, mo, so = self.base_norm
        self.mlp.set_norm(mi * sx, si * sx, mo * sy, so * sy)

    def fields(self, esea, sea, x, y, t):
        """Surrogate total fields [n, 65] (eta, then u, v, w, p per level), the
        linear part and the second-order part (float32 cuda tensors) for
        columns x, y (float64 cuda tensors) at time t.  Hybrid networks take
        the elevation's second-order part from the exact eta2 form."""
        import torch
        d = sea.as_dict() if isinstance(sea, SeaState) else sea
        key = id(sea)
        if self._scaled_for != key:
            self.set_sea(d)
            self._scaled_for = key
        X, Lin, e2 = linear_fields(esea, x, y, t, float(d["depth"]), peak_wavenumber(d), eta2=self.needs_eta2)
        if self.cols == list(range(NFEAT)):
            Xs = X
        else:
            if self._cols_t is None or self._cols_t.device != X.device:
                self._cols_t = torch.tensor(self.cols, dtype=torch.long, device=X.device)
            Xs = X.index_select(1, self._cols_t).contiguous()
            torch.cuda.synchronize()
        Yp = self.mlp.forward(Xs)
        engine.sync()
        self.eta2_net = None
        if self.mode == "hybrid":
            self.eta2_net = Yp[:, 0].clone()   # the network's own estimate (monitor)
        out = torch.empty_like(Lin)
        torch.cuda.synchronize()
        if self.mode == "hybrid":
            # exact second-order elevation into the network's eta2 output
            check(esea.L.cfdwm_surrogate_set_column(Yp.shape[0], e2.data_ptr(), Yp.data_ptr(), NOUT, 0),
                  "cfdwm_surrogate_set_column")
        check(esea.L.cfdwm_surrogate_add(Lin.numel(), Lin.data_ptr(), Yp.data_ptr(), out.data_ptr()),
              "cfdwm_surrogate_add")
        engine.sync()
        return out, Lin, Yp

    @staticmethod
    def residual_check(esea, x, y, t, eta_host, y2_host, sample, gate, h, eta2_net_host=None):
        """Against the exact kernels on the sampled columns (host arrays):
        the elevation error max |eta_surrogate - (eta1 + eta2)_exact| and
        whether it is within the gate; per second-order field (u2, v2, w2,
        dp2 at the linear levels) the max and relative RMS error of the
        predicted second-order part; for hybrid networks, the error of the
        network's own eta2 estimate (not used for the output)."""
        xs, ys = np.asarray(x)[sample], np.asarray(y)[sample]
        e1 = esea.eval("eta1", t, xs, ys)
        ex = exact_second_order(esea, xs, ys, t, e1, h)
        eta_err = float(np.max(np.abs(np.asarray(eta_host, float)[sample] - (e1 + ex[:, 0]))))
        y2 = np.asarray(y2_host, float)[sample]
        res = {"columns": int(np.size(sample)), "eta_max_err_m": eta_err, "pass": bool(eta_err <= gate)}
        for c, name in enumerate(("u2", "v2", "w2", "dp2")):
            d = y2[:, 1 + c::4] - ex[:, 1 + c::4]
            ref = ex[:, 1 + c::4]
            rr = float(np.sqrt(np.mean(ref ** 2)))
            res[name] = {"max_abs_err": float(np.max(np.abs(d))), "max_abs_exact": float(np.max(np.abs(ref))),
                         "rel_rms_err": float(np.sqrt(np.mean(d ** 2)) / rr) if rr > 0 else 0.0}
        if eta2_net_host is not None:
            res["eta2_net_max_err_m"] = float(np.max(np.abs(np.asarray(eta2_net_host, float)[sample] - ex[:, 0])))
        return res
