"""Separable depth basis of the pair-term profiles (host code of libcfdwm; no
GPU needed).  The approximation e^{-kappa t} ~ sum_q c_q(kappa) e^{-mu_q t},
c = E^{-1} [e^{-kappa tau_s}], E[s][q] = e^{-mu_q tau_s}, is rebuilt here from
the returned skeleton points and checked against numpy's exp on off-grid
samples; the interpolation property (exact at the skeleton depths and
wavenumbers) is checked as well."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def eng():
    from cfdwavemaker_amd import engine
    engine.load()
    return engine


def _approx(b, kap, t):
    mu, tau = b["mu"], b["tau"]
    E = np.exp(-np.outer(tau, mu))
    c = np.linalg.solve(E, np.exp(-np.outer(tau, kap)))          # rank x nk
    return (np.exp(-np.outer(t, mu)) @ c).T                        # nk x nt


@pytest.mark.parametrize("kmax,tmax,tol", [(1.4, 200.0, 1e-6), (0.9, 60.0, 1e-7), (3.0, 40.0, 1e-5)])
def test_basis_error_bound(eng, kmax, tmax, tol):
    b = eng.depth_basis(kmax, tmax, tol)
    assert b["ok"] and b["err"] <= tol
    rng = np.random.default_rng(7)
    kap = rng.uniform(0, kmax, 400)
    t = rng.uniform(0, tmax, 500)
    err = np.abs(_approx(b, kap, t) - np.exp(-np.outer(kap, t))).max()
    assert err <= 3 * tol, err
    # coefficients stay O(1) (no cancellation between basis terms)
    E = np.exp(-np.outer(b["tau"], b["mu"]))
    c = np.linalg.solve(E, np.exp(-np.outer(b["tau"], kap)))
    assert np.abs(c).max() < 4.0


def test_basis_interpolates_at_skeleton_points(eng):
    b = eng.depth_basis(1.4, 200.0, 1e-6)
    kap = b["mu"]
    got = _approx(b, kap, b["tau"])
    assert np.allclose(got, np.exp(-np.outer(kap, b["tau"])), atol=1e-9)


def test_rank_grows_with_tolerance(eng):
    r = [eng.depth_basis(1.4, 200.0, tol)["rank"] for tol in (1e-4, 1e-6, 1e-8)]
    assert r[0] < r[1] < r[2] <= 32
