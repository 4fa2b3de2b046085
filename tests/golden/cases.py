"""Sea states, points and grid shape shared by the golden-vector writer and the tests."""
import numpy as np

from cfdwavemaker_amd.seastate import make_sea


def sea_cases():
    cases = {}
    a = make_sea(24, 3.0, 10.0, 30.0, seed=11).as_dict()
    a.update(swl=0.3, mtheta=10.0, tofmax=3.0, x0=1.5, y0=-2.0)
    cases["A"] = a
    b = make_sea(8, 3.0, 9.0, 30.0, seed=12, ndir=4).as_dict()
    cases["B"] = b
    c = make_sea(20, 2.0, 8.0, 40.0, seed=13).as_dict()
    rng = np.random.default_rng(13)
    c.update(bandwidth=0.5, cutoff=1.8, ramptime=rng.uniform(0.5, 3.0, 20),
             starttime=rng.uniform(-1.0, 2.0, 20), duration=rng.uniform(5.0, 20.0, 20))
    cases["C"] = c
    d = make_sea(16, 1.0, 7.0, 5.0, seed=14).as_dict()
    cases["D"] = d
    e = make_sea(16, 4.0, 12.0, 500.0, seed=15, wmin=0.4, wmax=1.6).as_dict()
    cases["E"] = e
    return cases


def points_for(sea, n=40, seed=0):
    rng = np.random.default_rng(seed)
    h = sea["depth"]
    t = rng.uniform(-2.0, 30.0, n)
    t[:3] = [0.0, -1.0, 1e-13]
    x = rng.uniform(-100.0, 400.0, n)
    y = rng.uniform(-50.0, 50.0, n)
    z = rng.uniform(-h, 0.2 * h, n)
    z[3:6] = [-h, 0.0, sea.get("swl", 0.0) + 0.5]
    return t, x, y, z


GRID_CASE = dict(domain=(0.0, 120.0, -20.0, 20.0), nx=7, ny=5, nl=6, t0=0.0, dt=0.5,
                 tan_a=0.7, tan_b=1.5, depth=30.0)


# The legacy lsGrid's unit-test sea (unittests/lsgrid_tests.cpp:7-57): three
# long-crested components set by hand, deep-water wavenumbers, depth 300 m.
REF_PI = 3.1415926535897            # the reference's PI (src/lsgrid.h:1-3)


def unit_sea():
    w = np.array([2 * REF_PI / 10.0, 2 * REF_PI / 5.0, 2 * REF_PI / 4.0])
    return dict(omega=w, A=np.array([3.1, 2.1, 1.1]), k=w ** 2 / 9.81, phase=np.zeros(3),
                theta=np.array([REF_PI / 8.0, 0.0, 0.0]), depth=300.0, g=9.81, rho=1025.0, swl=0.0,
                mtheta=10.0, tofmax=0.0, x0=0.0, y0=0.0, bandwidth=100.0, cutoff=100000.0, order=2)


NO_IGNORE = (1e12, -1e12, 1e12, -1e12)
# Legacy two-slot grid cases: sea, grid, update times (one step forward or a
# re-initialisation jump), init_only and ignore settings.
LSGRID_CASES = {
    # lsgrid_tests.cpp:64-83 (test1): nx 1, ny 3, nl 20, tan_b 2, t0 4
    "unit": dict(sea="unit", domain=(0.0, 30.0, 0.0, 20.0), nx=1, ny=3, nl=20, t0=4.0, dt=0.5,
                 tan_a=7 * REF_PI / 18, tan_b=2.0, depth=300.0, swl=0.0, ignore=NO_IGNORE, ignore_at_init=0,
                 init_only=0, updates=(4.6, 5.2, 9.7, 10.1)),
    # lsgrid_tests.cpp:126-145 (vtueksporter): 61 x 31 x 10, tan_b 1.5, t0 4
    "vtu": dict(sea="unit", domain=(0.0, 300.0, 0.0, 100.0), nx=61, ny=31, nl=10, t0=4.0, dt=0.5,
                tan_a=7 * REF_PI / 18, tan_b=1.5, depth=300.0, swl=0.0, ignore=NO_IGNORE, ignore_at_init=0,
                init_only=0, updates=()),
    # swl, mtheta, tofmax, focus point, an ignore region from the start
    "A": dict(sea="A", domain=(0.0, 120.0, -20.0, 20.0), nx=7, ny=5, nl=6, t0=0.0, dt=0.5, tan_a=0.7,
              tan_b=1.5, depth=30.0, swl=0.3, ignore=(30.0, 70.0, -5.0, 30.0), ignore_at_init=1, init_only=0,
              updates=(0.6, 1.1, 3.0, 3.4)),
    # init_only: both slots share one array; the ignore region still holds in updates
    "A_io": dict(sea="A", domain=(0.0, 120.0, -20.0, 20.0), nx=7, ny=5, nl=6, t0=0.0, dt=0.5, tan_a=0.7,
                 tan_b=1.5, depth=30.0, swl=0.3, ignore=(30.0, 70.0, -5.0, 30.0), ignore_at_init=0, init_only=1,
                 updates=(0.6, 1.1, 3.0)),
    # individual frequency ramps and pair filters, a single row (ny 1)
    "C": dict(sea="C", domain=(-10.0, 60.0, 0.0, 0.0), nx=9, ny=1, nl=8, t0=0.5, dt=0.25, tan_a=0.9,
              tan_b=1.2, depth=40.0, swl=0.0, ignore=NO_IGNORE, ignore_at_init=0, init_only=0,
              updates=(0.8, 1.0, 2.5)),
}


def lsgrid_sea(name):
    return unit_sea() if name == "unit" else sea_cases()[name]


def lsgrid_points(c, t0, seed, n=120):
    """Query points of a legacy-grid case at one stage: inside and outside the
    domain, above the surface and below the seabed, t before, inside and after
    [t0, t0 + dt], and the grid nodes' columns at the surface and the seabed."""
    rng = np.random.default_rng(seed)
    d, h, dt = c["domain"], c["depth"], c["dt"]
    wx, wy = d[1] - d[0], d[3] - d[2]
    t = t0 + rng.uniform(-0.3 * dt, 1.3 * dt, n)
    x = rng.uniform(d[0] - 0.1 * wx - 1.0, d[1] + 0.1 * wx + 1.0, n)
    y = rng.uniform(d[2] - 0.1 * wy - 1.0, d[3] + 0.1 * wy + 1.0, n)
    z = rng.uniform(-h - 2.0, 3.0, n) + c["swl"]
    k = n // 4
    t[:k] = t0
    x[:k] = d[0] + (d[1] - d[0]) / max(1, c["nx"] - 1) * rng.integers(0, c["nx"], k)
    y[:k] = d[2] + (d[3] - d[2]) / max(1, c["ny"] - 1) * rng.integers(0, c["ny"], k)
    z[: k // 2] = -h
    return t, x, y, z
