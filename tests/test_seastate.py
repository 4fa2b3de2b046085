"""The synthetic sea states behind the benchmark and the fixtures: linear
dispersion, the Hs normalisation, the omega grid, the directional weights of
the reference's ndir > 0 convention (every frequency repeated per direction
with amplitude A(w) D_j, sum_j D_j^2 = 1, the frequency's phase), seeded
phases, and the frozen configurations."""
import math

import numpy as np
import pytest

from cfdwavemaker_amd.seastate import G, config_sea, make_sea, wavenumber


def test_wavenumber_solves_linear_dispersion():
    rng = np.random.default_rng(1)
    w = rng.uniform(0.05, 5.0, 500)
    for h in (0.6, 5.0, 30.0, 100.0, 3000.0):
        k = wavenumber(w, h)
        assert np.all(np.abs(G * k * np.tanh(k * h) - w * w) <= 1e-12 * w * w)
    # limits: deep water w^2 / g, shallow water w / sqrt(g h)
    assert np.allclose(wavenumber(np.array([2.0]), 5000.0), 4.0 / G, rtol=1e-14)
    w0 = np.array([1e-3])
    assert np.allclose(wavenumber(w0, 2.0), w0 / math.sqrt(G * 2.0), rtol=1e-6)


@pytest.mark.parametrize("cfg,n,hs,depth", [(1, 256, 3.0, 50.0), (2, 512, 6.0, 30.0), (3, 1024, 8.0, 100.0)])
def test_config_seas(cfg, n, hs, depth):
    s = config_sea(cfg)
    assert s.n == n and s.depth == depth
    assert abs(4.0 * math.sqrt(np.sum(s.A ** 2) / 2.0) - hs) <= 1e-12 * hs
    w = np.unique(s.omega)
    assert w[0] == pytest.approx(0.3, abs=1e-15) and w[-1] == pytest.approx(2.5, abs=1e-15)
    assert np.allclose(np.diff(w), (2.5 - 0.3) / (w.size - 1), rtol=1e-12)
    assert np.all(np.diff(s.omega) >= 0.0)                    # sorted by omega (stable)
    assert np.all((s.phase >= 0.0) & (s.phase < 2 * math.pi))
    assert np.allclose(s.k, wavenumber(s.omega, depth), rtol=0, atol=0)


def test_directional_weights_follow_the_reference_convention():
    s = config_sea(3)                                         # 64 frequencies x 16 directions
    lc = make_sea(64, 8.0, 14.0, 100.0, seed=3)               # same spectrum, long-crested
    A2 = (s.A ** 2).reshape(64, 16)
    assert np.allclose(A2.sum(1), lc.A ** 2, rtol=1e-13)      # sum_j D_j^2 = 1 per frequency
    ph = s.phase.reshape(64, 16)
    assert np.all(ph == ph[:, :1])                            # the frequency's phase for every direction
    assert np.allclose(ph[:, 0], lc.phase)
    th = s.theta.reshape(64, 16)
    assert np.all(np.abs(th) < math.pi / 2)
    D2 = A2 / A2.sum(1, keepdims=True)
    # cos^8 spreading: D_j^2 proportional to cos^8(theta_j)
    assert np.allclose(D2, np.cos(th) ** 8 / np.sum(np.cos(th) ** 8, 1, keepdims=True), rtol=1e-12)


def test_phases_are_seeded():
    a, b = make_sea(32, 2.0, 8.0, 20.0, seed=5), make_sea(32, 2.0, 8.0, 20.0, seed=5)
    c = make_sea(32, 2.0, 8.0, 20.0, seed=6)
    assert np.array_equal(a.phase, b.phase) and not np.array_equal(a.phase, c.phase)
    assert np.array_equal(a.phase, np.random.default_rng(5).uniform(0.0, 2 * math.pi, 32))


@pytest.mark.parametrize("state", [0, 1, 2, 7, 100, 511])
def test_batched_states(state):
    s = config_sea(4, state)
    assert 1.0 <= s.hs <= 8.0 and 6.0 <= s.tp <= 16.0 and 20.0 <= s.depth <= 200.0
    assert abs(4.0 * math.sqrt(np.sum(s.A ** 2) / 2.0) - s.hs) <= 1e-12 * s.hs
    if state % 2 == 0:
        assert s.ndir == 1 and 128 <= s.n <= 2048
    else:
        assert s.ndir == 16 and s.n % 16 == 0 and 8 * 16 <= s.n <= 2048
    assert np.array_equal(config_sea(4, state).phase, s.phase)  # reproducible
