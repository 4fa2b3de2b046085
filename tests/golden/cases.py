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
