"""Object interface over libcfdwm: sea states and s-grids resident on the GPU.

``Sea`` mirrors the reference's ``Irregular`` (src/Irregular.h:47-193):
``eval(kind, t, x, y, z)`` evaluates the member of that name (eta1, eta2, u1,
..., uvw2) for arrays of points.  ``Grid`` mirrors ``lsGridSpline``
(src/lsgrid_spline.h:15-129): construction runs initialize_kinematics,
``force_update`` is wave_force_update, ``query`` is get_kinematics_at_point.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import GridDescC, SeaParams, check, load

KINDS = ["eta1", "eta2", "eta", "u1", "v1", "w1", "dp1", "phi1_dxdz", "phi1_dydz",
         "phi1_dzdz", "dp1_dz", "u2", "v2", "w2", "dp2", "u", "v", "w", "dp",
         "uvw2_u", "uvw2_v", "uvw2_w"]
KIND = {k: i for i, k in enumerate(KINDS)}
# the Smit et al. s-coordinate members (cfdwm_eval kinds 23-26)
KIND.update({"u_smit": 23, "v_smit": 24, "w_smit": 25, "dp_smit": 26})
GRID_FIELDS = ["ETA", "ETAdx", "ETAdy", "ETAdt", "UX", "UY", "UZ", "Udt", "Vdt", "Wdt",
               "Udx", "Udy", "Uds", "Vdx", "Vdy", "Vds", "Wdx", "Wdy", "Wds", "IGNORE"]
# precision of a sea: FP64; FP32 (FP64 phase reduction, three-product bf16
# quadratic forms); BF16 (FP32 whose quadratic forms take one bf16 product)
FP64, FP32, BF16 = 0, 1, 2


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class Sea:
    """A sea state (component arrays sorted by omega) on the GPU."""

    def __init__(self, sea: dict, precision: int = FP64, with_direct: bool = True):
        self.L = load()
        self.precision = precision
        n = len(sea["omega"])
        arrs = [_f64(sea[k]) for k in ("omega", "A", "k", "phase", "theta")]
        ramps = [None if sea.get(k) is None else _f64(sea[k]) for k in ("ramptime", "starttime", "duration")]
        p = SeaParams()
        p.depth = float(sea.get("depth", 300.0)); p.g = float(sea.get("g", 9.81))
        p.rho = float(sea.get("rho", 1.0)); p.swl = float(sea.get("swl", 0.0))
        p.mtheta = float(sea.get("mtheta", 0.0)); p.tofmax = float(sea.get("tofmax", 0.0))
        p.x0 = float(sea.get("x0", 0.0)); p.y0 = float(sea.get("y0", 0.0))
        p.bandwidth = float(sea.get("bandwidth", 100.0)); p.cutoff = float(sea.get("cutoff", 100000.0))
        p.ampl = float(sea.get("ampl", 1.0)); p.order = int(sea.get("order", 2))
        p.normalize = int(sea.get("normalize", 0)); p.bw_auto = int(sea.get("bw_auto", 0))
        p.precision = int(precision); p.with_direct = int(bool(with_direct))
        h = C.c_void_p()
        check(self.L.cfdwm_sea_create(C.byref(h), n, *[_p(a) for a in arrs], *[_p(a) for a in ramps],
                                      C.byref(p)), "cfdwm_sea_create")
        self.h = h
        self.n = n
        self.depth = p.depth

    def close(self):
        if getattr(self, "h", None):
            self.L.cfdwm_sea_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_vertical_theory(self, vt: int):
        """1: u, v, w, p of every point entry follow the Smit et al. members
        (vertical_theory smit); 0: Sharma & Dean."""
        check(self.L.cfdwm_sea_set_vertical_theory(self.h, int(vt)), "cfdwm_sea_set_vertical_theory")

    def terms(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.L.cfdwm_sea_terms(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def eval(self, kind, t, x, y, z=None):
        t, x, y = np.broadcast_arrays(np.asarray(t, float), np.asarray(x, float), np.asarray(y, float))
        shape = t.shape
        t, x, y = (_f64(a.ravel()) for a in (t, x, y))
        zz = None if z is None else _f64(np.broadcast_to(np.asarray(z, float), shape).ravel())
        out = np.zeros(t.size)
        k = KIND[kind] if isinstance(kind, str) else int(kind)
        check(self.L.cfdwm_eval(self.h, k, t.size, _p(t), _p(x), _p(y), _p(zz), _p(out)), f"cfdwm_eval({kind})")
        return out.reshape(shape)

    def eval_fields_device(self, t, x, y, z=None, want=("eta1", "eta2", "u", "v", "w", "p")):
        """Direct-mode fields in one pass on float64 cuda tensors (no ramps,
        no z clamp); returns a dict of float64 cuda tensors."""
        import torch
        n = t.shape[0]
        out = {k: torch.empty(n, dtype=torch.float64, device=t.device) for k in want}
        ptr = lambda k: out[k].data_ptr() if k in out else None
        check(self.L.cfdwm_eval_fields(self.h, n, t.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       None if z is None else z.data_ptr(), ptr("eta1"), ptr("eta2"), ptr("u"),
                                       ptr("v"), ptr("w"), ptr("p"), 1), "cfdwm_eval_fields")
        return out

    def grid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
             domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, defer_fill=False) -> "Grid":
        return Grid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl, domain_ignore,
                    ignore_at_init, defer_fill)

    def lsgrid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
               domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, init_only=False) -> "LsGrid":
        """The legacy two-slot grid (lsGrid): filled at t0 and t0 + dt."""
        return LsGrid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl, domain_ignore,
                      ignore_at_init, init_only)

    def clone(self, precision: int) -> "Sea":
        """The same sea at another precision (host coefficient tables copied)."""
        h = C.c_void_p()
        check(self.L.cfdwm_sea_clone(C.byref(h), self.h, int(precision)), "cfdwm_sea_clone")
        other = Sea.__new__(Sea)
        other.L, other.precision, other.h, other.n, other.depth = self.L, precision, h, self.n, self.depth
        return other


class Grid:
    def __init__(self, sea: Sea, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
                 domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, defer_fill=False):
        self.sea, self.L = sea, sea.L
        d = GridDescC()
        for q in range(4):
            d.domain[q] = float(domain[q])
            d.domain_ignore[q] = float(domain_ignore[q])
        d.nx, d.ny, d.nl = int(nx), int(ny), int(nl)
        d.t0, d.dt, d.tan_a, d.tan_b = float(t0), float(dt), float(tan_a), float(tan_b)
        d.water_depth, d.swl, d.ignore_at_init = float(water_depth), float(swl), int(ignore_at_init)
        d.defer_fill = int(bool(defer_fill))
        h = C.c_void_p()
        check(self.L.cfdwm_grid_create(C.byref(h), sea.h, C.byref(d)), "cfdwm_grid_create")
        self.h = h
        self.nx, self.ny, self.nl = nx, ny, nl

    def close(self):
        if getattr(self, "h", None):
            self.L.cfdwm_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def refill(self):
        check(self.L.cfdwm_grid_refill(self.h), "cfdwm_grid_refill")

    def force_update(self, t):
        n = C.c_int()
        check(self.L.cfdwm_grid_force_update(self.h, float(t), C.byref(n)), "cfdwm_grid_force_update")
        return n.value

    def state(self):
        ts = C.c_int()
        v = [C.c_double() for _ in range(4)]
        self.L.cfdwm_grid_state(self.h, C.byref(ts), *[C.byref(q) for q in v])
        return dict(tstep=ts.value, t0=v[0].value, dx=v[1].value, dy=v[2].value, ds=v[3].value)

    def export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else int(which)
        n = self.L.cfdwm_grid_export(self.h, w, None)
        out = np.zeros(n)
        got = self.L.cfdwm_grid_export(self.h, w, _p(out))
        if got != n:
            raise RuntimeError("cfdwm_grid_export failed")
        return out

    def query(self, t, x, y, z):
        t, x, y, z = (_f64(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        check(self.L.cfdwm_grid_query(self.h, t.size, _p(t), _p(x), _p(y), _p(z), _p(out)), "cfdwm_grid_query")
        return out

    def write_vtu(self, path, last=False, time_label="TimeValue"):
        """ASCII .vtu of the query's base slot (last: the following slot)."""
        check(self.L.cfdwm_grid_write_vtu(self.h, str(path).encode(), int(bool(last)), time_label.encode()),
              "cfdwm_grid_write_vtu")

    def release_workspace(self):
        """Free the sea's form tables and this grid's operand workspace."""
        check(self.L.cfdwm_grid_release_workspace(self.h), "cfdwm_grid_release_workspace")

    def qform_info(self):
        """Quadratic forms of the last fill: None when the pair loop ran."""
        nf, rk = C.c_int(), C.c_int()
        fl, er = C.c_double(), C.c_double()
        used = self.L.cfdwm_grid_qform_info(self.h, C.byref(nf), C.byref(rk), C.byref(fl), C.byref(er))
        if used != 1:
            return None
        pr, tol, why = C.c_int(), C.c_double(), C.c_int()
        self.L.cfdwm_grid_qform_mode(self.h, C.byref(pr), C.byref(tol), C.byref(why))
        return dict(nform=nf.value, rank=rk.value, mfma_flop=fl.value, basis_err=er.value,
                    products=pr.value, basis_tol=tol.value)

    QFORM_REASONS = ("used", "fp64", "switched off", "no second-order pairs", "frequency ramps differ",
                     "grid deeper than the depth basis", "no depth basis within the tolerance")

    def qform_reason(self):
        """Why the last fill did (\"used\") or did not use the MFMA forms."""
        pr, tol, why = C.c_int(), C.c_double(), C.c_int()
        self.L.cfdwm_grid_qform_mode(self.h, C.byref(pr), C.byref(tol), C.byref(why))
        return self.QFORM_REASONS[why.value]

    def qform_clock(self, enable):
        """Mean shader clock (MHz) of the last fill's form launches, or None
        if that fill was not measured; then measure later fills or not."""
        mhz = C.c_double()
        r = self.L.cfdwm_grid_qform_clock(self.h, 1 if enable else 0, C.byref(mhz))
        return mhz.value if r == 0 else None

    def kinematics_device_f32(self, n, t, x_ptr, y_ptr, z_ptr, out4_ptr):
        """Device pointers; out4 receives {eta, u, v, w} per query (the layout
        of wave_Kinematics), enqueued on the library stream."""
        check(self.L.cfdwm_grid_kinematics_device_f32(self.h, int(n), float(t), x_ptr, y_ptr, z_ptr, out4_ptr),
              "cfdwm_grid_kinematics_device_f32")

    def query_device_f32(self, n, t, x_ptr, y_ptr, z_ptr, eta_ptr, u_ptr, v_ptr, w_ptr):
        """Device pointers (e.g. torch tensors' data_ptr()), enqueued on the library stream."""
        check(self.L.cfdwm_grid_query_device_f32(self.h, int(n), float(t), x_ptr, y_ptr, z_ptr,
                                                 eta_ptr, u_ptr, v_ptr, w_ptr), "cfdwm_grid_query_device_f32")


def _grid_desc(domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl, domain_ignore, ignore_at_init,
               defer_fill=False):
    d = GridDescC()
    for q in range(4):
        d.domain[q] = float(domain[q])
        d.domain_ignore[q] = float(domain_ignore[q])
    d.nx, d.ny, d.nl = int(nx), int(ny), int(nl)
    d.t0, d.dt, d.tan_a, d.tan_b = float(t0), float(dt), float(tan_a), float(tan_b)
    d.water_depth, d.swl, d.ignore_at_init = float(water_depth), float(swl), int(ignore_at_init)
    d.defer_fill = int(bool(defer_fill))
    return d


class LsGrid:
    """The legacy two-slot s-grid (lsGrid, src/lsgrid.h:15-114) on the GPU.

    Method names follow the reference class: eta(t, x, y) and u/v/w(t, x, y, z)
    (lsgrid.cpp:360-376), CheckTime (:378-386), update (:749-827),
    export_vtu (:1354-1552); query() evaluates all four at once.  Slot 0
    holds t0, slot 1 t0 + dt; export() gives slot 0 then slot 1.
    """

    def __init__(self, sea: Sea, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
                 domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, init_only=False):
        self.sea, self.L = sea, sea.L
        d = _grid_desc(domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl, domain_ignore, ignore_at_init)
        h = C.c_void_p()
        check(self.L.cfdwm_lsgrid_create(C.byref(h), sea.h, C.byref(d), int(bool(init_only))),
              "cfdwm_lsgrid_create")
        self.h = h
        self.nx, self.ny, self.nl = nx, ny, nl

    def close(self):
        if getattr(self, "h", None):
            rc = self.L.cfdwm_lsgrid_destroy(self.h)
            self.h = None
            check(rc, "cfdwm_lsgrid_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def CheckTime(self, t) -> bool:
        r = self.L.cfdwm_lsgrid_check_time(self.h, float(t))
        check(min(r, 0), "cfdwm_lsgrid_check_time")
        return r == 1

    def update(self, t_target) -> int:
        """1 when the grid was re-initialised (a jump of more than one step), 0 after one step."""
        r = C.c_int()
        check(self.L.cfdwm_lsgrid_update(self.h, float(t_target), C.byref(r)), "cfdwm_lsgrid_update")
        return r.value

    def state(self):
        t0, uc = C.c_double(), C.c_int()
        check(self.L.cfdwm_lsgrid_state(self.h, C.byref(t0), C.byref(uc)), "cfdwm_lsgrid_state")
        return dict(t0=t0.value, update_count=uc.value)

    @property
    def t0(self):
        return self.state()["t0"]

    def export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else int(which)
        n = self.L.cfdwm_lsgrid_export(self.h, w, None)
        if n < 0:
            check(int(n), "cfdwm_lsgrid_export")
        out = np.zeros(n)
        got = self.L.cfdwm_lsgrid_export(self.h, w, _p(out))
        if got != n:
            raise RuntimeError("cfdwm_lsgrid_export failed")
        return out

    def query(self, t, x, y, z):
        t, x, y, z = (_f64(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        check(self.L.cfdwm_lsgrid_query(self.h, t.size, _p(t), _p(x), _p(y), _p(z), _p(out)),
              "cfdwm_lsgrid_query")
        return out

    def query_device(self, n, t_ptr, x_ptr, y_ptr, z_ptr, out4_ptr):
        """Device pointers of doubles; out4 receives {eta, u, v, w} per point."""
        check(self.L.cfdwm_lsgrid_query_device(self.h, int(n), t_ptr, x_ptr, y_ptr, z_ptr, out4_ptr),
              "cfdwm_lsgrid_query_device")

    def eta(self, t, x, y):
        return self.query(t, x, y, 0.0)[:, 0]

    def u(self, t, x, y, z):
        return self.query(t, x, y, z)[:, 1]

    def v(self, t, x, y, z):
        return self.query(t, x, y, z)[:, 2]

    def w(self, t, x, y, z):
        return self.query(t, x, y, z)[:, 3]

    def write_vtu(self, path, last=False, time_label="TimeValue"):
        check(self.L.cfdwm_lsgrid_write_vtu(self.h, str(path).encode(), int(bool(last)), time_label.encode()),
              "cfdwm_lsgrid_write_vtu")

    def set_vtk(self, on=True, directory="./", prefix="kinematics", time_label="TimeValue"):
        check(self.L.cfdwm_lsgrid_set_vtk(self.h, int(bool(on)), directory.encode(), prefix.encode(),
                                          time_label.encode()), "cfdwm_lsgrid_set_vtk")


def sync():
    check(load().cfdwm_sync(), "cfdwm_sync")


def prof_enable(on=True):
    check(load().cfdwm_prof_enable(int(on)), "cfdwm_prof_enable")


def prof_read(reset=True):
    L = load()
    cap = 64
    names = C.create_string_buffer(64 * cap)
    ms = np.zeros(cap)
    cnt = np.zeros(cap, dtype=np.int64)
    n = L.cfdwm_prof_read(names, _p(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), cap, int(reset))
    out = {}
    for q in range(n):
        nm = names.raw[64 * q:64 * q + 64].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[q]), int(cnt[q]))
    return out


def depth_basis(kmax, tmax, tol=1e-6, rmax=32):
    """Skeleton-exponential depth basis (host only): dict(rank, mu, tau, err, ok)."""
    L = load()
    mu = np.zeros(rmax)
    tau = np.zeros(rmax)
    err = C.c_double()
    r = L.cfdwm_depth_basis(float(kmax), float(tmax), float(tol), int(rmax), _p(mu), _p(tau), C.byref(err))
    if r == 0:
        raise ValueError("cfdwm_depth_basis: invalid arguments")
    k = abs(r)
    return dict(rank=k, mu=mu[:k].copy(), tau=tau[:k].copy(), err=err.value, ok=r > 0)


def profile_basis(kmax, h, sign, tol=1e-6, rmax=32):
    """Depth basis of the pair-term profiles e^{kz} + sign e^{-k(z+2h)} on
    z in [-2h, 0], skeleton depths in [-h, 0] (host only): dict(rank, mu, zs,
    err, ok)."""
    L = load()
    mu = np.zeros(rmax)
    zs = np.zeros(rmax)
    err = C.c_double()
    r = L.cfdwm_profile_basis(float(kmax), float(h), int(sign), float(tol), int(rmax), _p(mu), _p(zs),
                              C.byref(err))
    if r == 0:
        raise ValueError("cfdwm_profile_basis: invalid arguments")
    k = abs(r)
    return dict(rank=k, mu=mu[:k].copy(), zs=zs[:k].copy(), err=err.value, ok=r > 0)


def refill_batched(grids, budget_gb=96.0, release=True):
    """Refill many grids with their second-order forms in persistent launches
    over all of them (autotuned launch shape); returns the batch statistics."""
    from ._lib import BatchStatsC
    L = load()
    arr = (C.c_void_p * len(grids))(*[g.h for g in grids])
    st = BatchStatsC()
    check(L.cfdwm_grids_refill_batched(arr, len(grids), float(budget_gb), int(bool(release)), C.byref(st)),
          "cfdwm_grids_refill_batched")
    return {f: getattr(st, f) for f, _ in BatchStatsC._fields_}
