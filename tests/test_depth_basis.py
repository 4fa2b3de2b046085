"""Separable depth bases of the pair-term profiles (host code of libcfdwm; no
GPU needed).  The exponential basis e^{-kappa t} ~ sum_q c_q(kappa) e^{-mu_q t},
c = E^{-1} [e^{-kappa tau_s}], E[s][q] = e^{-mu_q tau_s}, and the profile bases
the fills use, P_kappa(z) = e^{kappa z} +- e^{-kappa (z + 2h)} ~ sum_q
c_q(kappa) P_{mu_q}(z), c = F^{-1} [P_kappa(z_s)], are rebuilt here from the
returned skeleton points and checked against numpy's exp on off-grid samples;
the interpolation property (exact at the skeleton depths and wavenumbers) is
checked as well."""
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


def _profile(kap, z, h, sign):
    """e^{kz} + sign e^{-k(z+2h)}, kap x z."""
    return np.exp(np.outer(kap, z)) + sign * np.exp(-np.outer(kap, z + 2 * h))


def _papprox(b, kap, z, h, sign):
    F = _profile(b["mu"], b["zs"], h, sign).T                       # F[s][q] = P_{mu_q}(z_s)
    c = np.linalg.solve(F, _profile(kap, b["zs"], h, sign).T)       # rank x nk
    return c.T @ _profile(b["mu"], z, h, sign), c                   # nk x nz


@pytest.mark.parametrize("kmax,h,sign,tol", [(1.27, 100.0, 1, 1e-4), (1.27, 100.0, -1, 1e-4),
                                             (1.27, 100.0, 1, 1e-6), (1.27, 100.0, -1, 1e-6),
                                             (0.3, 40.0, 1, 1e-6), (0.3, 40.0, -1, 1e-7)])
def test_profile_basis_error_bound(eng, kmax, h, sign, tol):
    """The profile basis rebuilt with numpy from the returned skeleton meets the
    tolerance on random off-grid (kappa, z) in [0, kmax] x [-2h, 0], and the
    rounding of its coefficients is amplified by less than 4."""
    b = eng.profile_basis(kmax, h, sign, tol)
    assert b["ok"] and b["err"] <= tol
    assert np.all(b["zs"] <= 0) and np.all(b["zs"] >= -h) and np.all(b["mu"] <= kmax)
    rng = np.random.default_rng(11)
    kap = np.concatenate([rng.uniform(0, kmax, 400), [0.0, kmax]])
    z = np.concatenate([rng.uniform(-2 * h, 0, 500), [0.0, -2 * h, -h]])
    got, c = _papprox(b, kap, z, h, sign)
    err = np.abs(got - _profile(kap, z, h, sign)).max()
    assert err <= 3 * tol, err
    # rounding amplification: independent errors e_q in the coefficients give
    # |sum_q e_q P_{mu_q}(z)| of order sqrt(sum_q c_q^2 P_{mu_q}(z)^2) |e|; the
    # construction keeps it a small multiple of the profile's own scale
    pm = _profile(b["mu"], z, h, sign)                              # rank x nz
    amp = np.sqrt((c ** 2).T @ (pm ** 2)).max(axis=1)
    scale = np.maximum(np.abs(_profile(kap, z, h, sign)).max(axis=1), 0.05)
    assert (amp / scale).max() < 4.0, (amp / scale).max()


def test_profile_basis_interpolates_at_skeleton_points(eng):
    """Exact at the skeleton wavenumbers (c is a unit vector there) and at the
    skeleton depths (the interpolation conditions)."""
    h = 100.0
    for sign in (1, -1):
        b = eng.profile_basis(1.27, h, sign, 1e-6)
        got, c = _papprox(b, b["mu"], b["zs"], h, sign)
        assert np.allclose(c, np.eye(b["rank"]), atol=1e-8)
        kap = np.linspace(0.01, 1.27, 37)
        got, _ = _papprox(b, kap, b["zs"], h, sign)
        assert np.allclose(got, _profile(kap, b["zs"], h, sign), rtol=0, atol=1e-9 * 2)


def test_profile_basis_scale_invariant_and_smaller_than_exponential(eng):
    """The basis depends on kmax h only (mu h and zs / h agree across scales;
    a power-of-two rescaling keeps every product kappa z bit for bit, so the
    pivots agree too), and one profile family needs no more terms than the
    exponential basis of e^{-kappa t}, t in [0, 2h], that it replaces."""
    a = eng.profile_basis(1.27, 100.0, 1, 1e-6)
    b = eng.profile_basis(5.08, 25.0, 1, 1e-6)
    assert eng.profile_basis(12.7, 10.0, 1, 1e-6)["rank"] == a["rank"]
    assert a["rank"] == b["rank"]
    assert np.allclose(a["mu"] * 100.0, b["mu"] * 25.0, rtol=1e-12)
    assert np.allclose(a["zs"] / 100.0, b["zs"] / 25.0, rtol=1e-12)
    for tol in (1e-4, 1e-6):
        e = eng.depth_basis(1.27, 200.0, tol)["rank"]
        for sign in (1, -1):
            assert eng.profile_basis(1.27, 100.0, sign, tol)["rank"] <= e


def test_profile_basis_rank_and_bad_arguments(eng):
    r = [eng.profile_basis(1.27, 100.0, 1, tol)["rank"] for tol in (1e-3, 1e-5, 1e-7)]
    assert r[0] < r[1] < r[2] <= 32
    b = eng.profile_basis(1.27, 100.0, -1, 1e-14, rmax=8)
    assert not b["ok"] and b["err"] > 1e-14 and b["rank"] == 8
    for args in ((0.0, 100.0, 1), (1.0, 0.0, 1), (1.0, 100.0, 0)):
        with pytest.raises(ValueError):
            eng.profile_basis(*args)


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
