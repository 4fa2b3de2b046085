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


def test_unreachable_tolerance_reports_failure(eng):
    """A tolerance that no rank <= 32 reaches is reported as a failure (negative
    rank from cfdwm_depth_basis), with the measured error above the tolerance:
    the s-grid fills then take the pair loop instead of the forms."""
    b = eng.depth_basis(1.4, 200.0, 1e-9)
    assert not b["ok"]
    assert b["err"] > 1e-9


@pytest.mark.gpu
def test_tight_basis_tolerance_falls_back_to_pair_loop(eng, oracle, monkeypatch):
    """An FP32 fill whose depth basis cannot meet CFDWM_QF_TOL runs the pair
    loop (no forms) and stays within the FP32 bound against the oracle."""
    from cfdwavemaker_amd.seastate import config_sea
    from tests.conftest import scaled_err
    monkeypatch.setenv("CFDWM_QF_TOL", "1e-9")
    s = config_sea(3).as_dict()
    dom = (0.0, 1000.0 * 2 / 255, 0.0, 1000.0 * 2 / 255)
    ref = oracle.sea(s).grid(dom, 3, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    grid = eng.Sea(s, precision=eng.FP32, with_direct=False).grid(dom, 3, 3, 16, 0.0, 0.5, 0.7, 1.5, 100.0)
    assert grid.qform_info() is None
    assert grid.qform_reason() == "no depth basis within the tolerance"
    for f in ("ETA", "UX", "UY", "UZ"):
        err = scaled_err(grid.export(f), ref.export(f))
        assert err <= 1e-5, (f, err)
