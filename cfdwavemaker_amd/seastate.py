"""Synthetic sea states and ``waveinput.dat`` writer.

CFDwavemaker reads its wave components from ``waveinput.dat`` (section
``[wave input irregular]``, src/CFDwavemaker.cpp:495-799); the components are
normally produced by an external spectral tool.  This module generates them
from a JONSWAP spectrum with optional cos^(2s) directional spreading, and
writes a ``waveinput.dat`` the reference parser and this library both read.

Conventions (the frozen benchmark definitions):
  * omega equally spaced on [wmin, wmax];
  * JONSWAP S(w) = alpha g^2 w^-5 exp(-1.25 (wp/w)^4) gamma^r,
    r = exp(-(w-wp)^2 / (2 sigma^2 wp^2)), sigma = 0.07 (w <= wp) / 0.09;
  * A_i = sqrt(2 S(w_i) dw), rescaled so that 4 sqrt(sum A^2 / 2) = Hs;
  * k from the linear dispersion relation w^2 = g k tanh(k h) (Newton);
  * phases uniform on [0, 2 pi) from numpy.random.default_rng(seed);
  * directional seas use the reference's ndir > 0 convention
    (src/CFDwavemaker.cpp:706-797): every frequency row is repeated for each
    direction with amplitude A(w) D_j and the frequency's phase, where
    D_j = sqrt(D(theta_j) dtheta) normalised so that sum D_j^2 = 1, and the
    rows are stably sorted by omega.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81


def wavenumber(omega, depth, g=G):
    """Solve w^2 = g k tanh(k h) for k by Newton iteration (to 1e-15)."""
    w = np.asarray(omega, dtype=np.float64)
    k = w * w / g
    k = np.where(depth * k < 1.0, w / np.sqrt(g * depth), k)
    for _ in range(100):
        th = np.tanh(k * depth)
        f = g * k * th - w * w
        df = g * th + g * k * depth * (1.0 - th * th)
        dk = f / df
        k = k - dk
        if np.all(np.abs(dk) <= 1e-15 * np.abs(k)):
            break
    return k


def jonswap(omega, hs, tp, gamma=3.3, g=G):
    w = np.asarray(omega, dtype=np.float64)
    wp = 2.0 * math.pi / tp
    sigma = np.where(w <= wp, 0.07, 0.09)
    r = np.exp(-((w - wp) ** 2) / (2.0 * sigma ** 2 * wp ** 2))
    return g * g * w ** -5 * np.exp(-1.25 * (wp / w) ** 4) * gamma ** r


@dataclass
class SeaState:
    """Component arrays in reference order (stably sorted by omega)."""
    omega: np.ndarray
    A: np.ndarray
    k: np.ndarray
    phase: np.ndarray
    theta: np.ndarray
    depth: float
    g: float = G
    rho: float = 1025.0
    swl: float = 0.0
    mtheta: float = 0.0
    tofmax: float = 0.0
    x0: float = 0.0
    y0: float = 0.0
    bandwidth: float = 100.0
    cutoff: float = 100000.0
    order: int = 2
    hs: float = 0.0
    tp: float = 0.0
    nfreq: int = 0
    ndir: int = 1
    ramptime: np.ndarray | None = None
    starttime: np.ndarray | None = None
    duration: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.omega.size)

    def as_dict(self) -> dict:
        d = dict(omega=self.omega, A=self.A, k=self.k, phase=self.phase, theta=self.theta,
                 depth=self.depth, g=self.g, rho=self.rho, swl=self.swl, mtheta=self.mtheta,
                 tofmax=self.tofmax, x0=self.x0, y0=self.y0, bandwidth=self.bandwidth,
                 cutoff=self.cutoff, order=self.order, normalize=0, ampl=1.0, bw_auto=0)
        if self.ramptime is not None:
            d.update(ramptime=self.ramptime, starttime=self.starttime, duration=self.duration)
        return d


def make_sea(nfreq, hs, tp, depth, gamma=3.3, wmin=0.3, wmax=2.5, seed=0, ndir=1,
             spread_s=4.0, mean_dir=0.0, g=G, rho=1025.0) -> SeaState:
    """JONSWAP sea, long-crested (ndir=1) or cos^(2s)-spread over (-pi/2, pi/2)."""
    w = np.linspace(wmin, wmax, nfreq)
    dw = (wmax - wmin) / (nfreq - 1) if nfreq > 1 else 1.0
    S = jonswap(w, hs, tp, gamma, g)
    A = np.sqrt(2.0 * S * dw)
    A *= hs / (4.0 * math.sqrt(np.sum(A * A) / 2.0))
    k = wavenumber(w, depth, g)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0.0, 2.0 * math.pi, nfreq)
    if ndir <= 1:
        om, AA, kk, pp, th = w, A, k, ph, np.zeros(nfreq)
    else:
        dth = math.pi / ndir
        thj = -math.pi / 2 + dth * (np.arange(ndir) + 0.5) + mean_dir
        D = np.cos(thj - mean_dir) ** (2.0 * spread_s)
        Dj = np.sqrt(D * dth)
        Dj /= math.sqrt(np.sum(Dj * Dj))
        om = np.repeat(w, ndir)
        kk = np.repeat(k, ndir)
        pp = np.repeat(ph, ndir)
        AA = (A[:, None] * Dj[None, :]).ravel()
        th = np.tile(thj, nfreq)
        order = np.argsort(om, kind="stable")
        om, AA, kk, pp, th = om[order], AA[order], kk[order], pp[order], th[order]
    return SeaState(omega=om.astype(np.float64), A=AA, k=kk, phase=pp, theta=th,
                    depth=float(depth), g=g, rho=rho, hs=hs, tp=tp, nfreq=nfreq, ndir=max(ndir, 1))


# ------------------------------------------------------------------ configs
def config_sea(cfg: int, state: int = 0) -> SeaState:
    """Sea states of the benchmark configurations (seeds = config index)."""
    if cfg == 1:
        s = make_sea(256, 3.0, 10.0, 50.0, seed=1)
        s.order = 0
        return s
    if cfg == 2:
        return make_sea(512, 6.0, 12.0, 30.0, seed=2)
    if cfg in (3, 5):
        return make_sea(64, 8.0, 14.0, 100.0, seed=3, ndir=16)
    if cfg == 4:
        rng = np.random.default_rng(4000 + state)
        n = int(rng.integers(128, 2049))
        hs = float(rng.uniform(1.0, 8.0))
        tp = float(rng.uniform(6.0, 16.0))
        h = float(rng.uniform(20.0, 200.0))
        if state % 2 == 0:
            return make_sea(n, hs, tp, h, seed=4000 + state)
        nf = max(n // 16, 2)
        return make_sea(nf, hs, tp, h, seed=4000 + state, ndir=16)
    raise ValueError(cfg)


# ------------------------------------------------------------ waveinput.dat
def write_waveinput(path, sea: SeaState, lsgrid: dict | None = None, ramps: dict | None = None,
                    extrap: int | None = None, bandwidth=None, cutoff=None, normalize=None,
                    amplify=None, individual_ramp=False, vertical_theory=None, version="@v310") -> None:
    """Write a ``waveinput.dat`` that the reference parser accepts.

    Components are written one row per component with ndir 0 (already expanded
    and sorted), which the parser re-sorts stably by omega (a no-op here).
    """
    L = [version, "[wave type]", "irregular", "", "[general input data]",
         f"depth {sea.depth!r}", f"mtheta {sea.mtheta!r}"]
    if normalize is not None:
        L.append(f"normalize {int(normalize)}")
    if amplify is not None:
        L.append(f"amplify {amplify!r}")
    L += [f"swl {sea.swl!r}", f"gravity {sea.g!r}", f"rho {sea.rho!r}", "",
          "[wave reference point]", f"time {sea.tofmax!r}", f"x {sea.x0!r}", f"y {sea.y0!r}", ""]
    if ramps:
        L.append("[ramps]")
        for key in ("time_rampup", "time_rampdown", "x_rampup", "x_rampdown", "y_rampup", "y_rampdown"):
            if key in ramps:
                a, b = ramps[key]
                L.append(f"{key} {a!r} {b!r}")
        L.append("")
    L.append("[wave input irregular]")
    if extrap is not None:
        L.append(f"kinematics_surf_extrap {int(extrap)}")
    if bandwidth is not None:
        L.append(f"bandwidth {bandwidth}")
    if cutoff is not None:
        L.append(f"cutoff {cutoff!r}")
    if individual_ramp:
        L.append("individual_ramp true")
    if vertical_theory is not None:
        L.append(f"vertical_theory {vertical_theory}")
    L += [f"nfreq {sea.n}", "ndir 0"]
    for i in range(sea.n):
        row = [sea.omega[i], sea.A[i], sea.k[i], sea.phase[i], sea.theta[i]]
        if individual_ramp:
            row += [sea.ramptime[i], sea.starttime[i], sea.duration[i]]
        L.append(" ".join(repr(float(v)) for v in row))
    L.append("")
    if lsgrid:
        b = lsgrid["bounds"]
        L += ["[lsgrid]", f"bounds {b[0]!r} {b[1]!r} {b[2]!r} {b[3]!r}",
              f"nx {int(lsgrid['nx'])}", f"ny {int(lsgrid['ny'])}", f"nl {int(lsgrid['nl'])}",
              f"t0 {lsgrid.get('t0', 0.0)!r}", f"dt {lsgrid.get('dt', 0.5)!r}",
              f"stretch_params {lsgrid.get('tan_a', 0.7)!r} {lsgrid.get('tan_b', 1.5)!r}"]
        if "ignore_subdomain" in lsgrid:
            q = lsgrid["ignore_subdomain"]
            L.append(f"ignore_subdomain {q[0]!r} {q[1]!r} {q[2]!r} {q[3]!r}")
        if lsgrid.get("ignore_at_init"):
            L.append("ignore_at_init 1")
        L.append("")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
