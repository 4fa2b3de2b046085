"""CPU oracle (test infrastructure only).

ctypes bindings for ``liboracle.so`` (the FP64 C restatement of the reference
algorithm, ``cfdwm_oracle.c``) and, where it has been built, for the compiled
reference under ``_ref/`` (``libcfdwm_refh.so``: a driver around the
reference's ``Irregular``/``lsGridSpline`` classes; ``libCFDwavemaker_ref.so``:
the reference's own C ABI).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
use this package, always as the checker.  The product (``cfdwavemaker_amd``)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_ubyte)

# evaluation kinds (cfdwm_oracle.h / ref_harness.cpp share the numbering):
# the Sharma-Dean members, then the Smit et al. members
KINDS = ["eta1", "eta2", "eta", "u1", "v1", "w1", "dp1", "phi1_dxdz", "phi1_dydz",
         "phi1_dzdz", "dp1_dz", "u2", "v2", "w2", "dp2", "u", "v", "w", "dp",
         "uvw2_u", "uvw2_v", "uvw2_w", "uvw2_p"]
SMIT_KINDS = ["u_smit", "v_smit", "w_smit", "dp_smit"]
KIND = {k: i for i, k in enumerate(KINDS + SMIT_KINDS)}
GRID_FIELDS = ["ETA", "ETAdx", "ETAdy", "ETAdt", "UX", "UY", "UZ", "Udt", "Vdt", "Wdt",
               "Udx", "Udy", "Uds", "Vdx", "Vdy", "Vds", "Wdx", "Wdy", "Wds", "IGNORE"]


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    target = "all" if with_ref is None else ("ref" if with_ref else "oracle")
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)
    if with_ref is None or with_ref:
        pass


def _arr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _sea_args(sea):
    """Positional arguments shared by or_sea_create and refh_setup."""
    n = len(sea["omega"])
    keep = [_arr(sea[k]) for k in ("omega", "A", "k", "phase", "theta")]
    rt = _arr(sea.get("ramptime"))
    ts = _arr(sea.get("starttime"))
    du = _arr(sea.get("duration"))
    keep += [rt, ts, du]
    args = [n] + [_ptr(a) for a in keep] + [
        float(sea.get("depth", 300.0)), float(sea.get("g", 9.81)), float(sea.get("rho", 1.0)),
        float(sea.get("swl", 0.0)), float(sea.get("mtheta", 0.0)), float(sea.get("tofmax", 0.0)),
        float(sea.get("x0", 0.0)), float(sea.get("y0", 0.0)), float(sea.get("bandwidth", 100.0)),
        float(sea.get("cutoff", 100000.0)), int(sea.get("order", 2)), int(sea.get("normalize", 0)),
        float(sea.get("ampl", 1.0)), int(sea.get("bw_auto", 0))]
    return args, keep


_SEA_ARGTYPES = [C.c_int] + [_dp] * 8 + [C.c_double] * 10 + [C.c_int, C.c_int, C.c_double, C.c_int]


class Oracle:
    """FP64 C restatement (liboracle.so)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        L = self.lib = C.CDLL(path)
        L.or_sea_create.restype = C.c_void_p
        L.or_sea_create.argtypes = _SEA_ARGTYPES
        L.or_sea_destroy.argtypes = [C.c_void_p]
        L.or_sea_components.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_sea_kmax.restype = C.c_double
        L.or_sea_kmax.argtypes = [C.c_void_p]
        L.or_pair_count.restype = C.c_long
        L.or_pair_count.argtypes = [C.c_void_p]
        L.or_pair_export.argtypes = [C.c_void_p, _ip, _bp, _dp]
        L.or_sea_set_vertical_theory.argtypes = [C.c_void_p, C.c_int]
        L.or_pair_export_h.argtypes = [C.c_void_p, _dp]
        L.or_eval.argtypes = [C.c_void_p, C.c_int, C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.or_freq_ramp.restype = C.c_double
        L.or_freq_ramp.argtypes = [C.c_double] * 4
        L.or_ramp.restype = C.c_double
        L.or_ramp.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_double]
        L.or_spectral_moments.argtypes = [C.c_void_p, _dp]
        for f in ("or_mean_period", "or_phase_speed", "or_mean_length"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p, C.c_int]
        L.or_grid_create.restype = C.c_void_p
        L.or_grid_create.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     _dp, C.c_int]
        L.or_grid_destroy.argtypes = [C.c_void_p]
        L.or_grid_force_update.restype = C.c_int
        L.or_grid_force_update.argtypes = [C.c_void_p, C.c_double]
        L.or_grid_state.argtypes = [C.c_void_p, _ip, _dp, _dp, _dp, _dp]
        L.or_grid_export.restype = C.c_long
        L.or_grid_export.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_grid_query.argtypes = [C.c_void_p, C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.or_sigmaS2cart.restype = C.c_double
        L.or_sigmaS2cart.argtypes = [C.c_double] * 6
        L.or_cart2sigmaS.restype = C.c_double
        L.or_cart2sigmaS.argtypes = [C.c_double] * 6
        L.or_lsgrid_create.restype = C.c_void_p
        L.or_lsgrid_create.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       _dp, C.c_int, C.c_int]
        L.or_lsgrid_destroy.argtypes = [C.c_void_p]
        L.or_lsgrid_from_fields.restype = C.c_void_p
        L.or_lsgrid_from_fields.argtypes = [_dp, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [_dp] * 4
        L.or_lsgrid_check_time.restype = C.c_int
        L.or_lsgrid_check_time.argtypes = [C.c_void_p, C.c_double]
        L.or_lsgrid_update.restype = C.c_int
        L.or_lsgrid_update.argtypes = [C.c_void_p, C.c_double]
        L.or_lsgrid_t0.restype = C.c_double
        L.or_lsgrid_t0.argtypes = [C.c_void_p]
        L.or_lsgrid_export.restype = C.c_long
        L.or_lsgrid_export.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_lsgrid_query.argtypes = [C.c_void_p, C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.or_column_fill.restype = C.c_double
        L.or_column_fill.argtypes = [C.c_void_p, C.c_long, _dp, _dp, C.c_double, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.c_double,
                                     _dp, _dp, _dp, _dp]

    # -- sea ---------------------------------------------------------------
    def sea(self, sea: dict) -> "OracleSea":
        args, keep = _sea_args(sea)
        h = self.lib.or_sea_create(*args)
        return OracleSea(self, h, len(sea["omega"]))

    def freq_ramp(self, t, rt, ts, du):
        return self.lib.or_freq_ramp(t, rt, ts, du)

    def ramp(self, enabled, flags, se12, t, x, y):
        se = _arr(np.asarray(se12, dtype=np.float64))
        return self.lib.or_ramp(int(enabled), int(flags), _ptr(se), t, x, y)

    def sigmaS2cart(self, s, eta, depth, swl, a, b):
        return self.lib.or_sigmaS2cart(s, eta, depth, swl, a, b)

    def lsgrid_from_fields(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl, eta, u, v, w):
        """Legacy-grid interpolation over given slots (eta [2 nx ny], u/v/w [2 nx ny nl])."""
        d = _arr(np.asarray(domain, float))
        keep = [_arr(np.ravel(a)) for a in (eta, u, v, w)]
        h = self.lib.or_lsgrid_from_fields(_ptr(d), nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl,
                                           *[_ptr(a) for a in keep])
        return OracleLsGrid(self, h)

    def cart2sigmaS(self, z, eta, depth, swl, a, b):
        return self.lib.or_cart2sigmaS(z, eta, depth, swl, a, b)


class OracleSea:
    def __init__(self, owner: Oracle, handle, n):
        self.o, self.h, self.n = owner, handle, n

    def __del__(self):
        try:
            self.o.lib.or_sea_destroy(self.h)
        except Exception:
            pass

    def components(self):
        A = np.zeros(self.n)
        ph = np.zeros(self.n)
        bw = C.c_double()
        self.o.lib.or_sea_components(self.h, _ptr(A), _ptr(ph), C.byref(bw))
        return A, ph, bw.value

    def kmax(self):
        return self.o.lib.or_sea_kmax(self.h)

    def pairs(self):
        n = self.o.lib.or_pair_count(self.h)
        im = np.zeros((n, 2), dtype=np.int32)
        fl = np.zeros(n, dtype=np.uint8)
        v = np.zeros((n, 7))
        self.o.lib.or_pair_export(self.h, im.ctypes.data_as(_ip), fl.ctypes.data_as(_bp), _ptr(v))
        return im, fl, v

    def set_vertical_theory(self, vt):
        """1: u/v/w/dp dispatch to the Smit members (Irregular.cpp:25, :53, :77, :143)."""
        self.o.lib.or_sea_set_vertical_theory(self.h, int(vt))

    def pairs_h(self):
        """Smit H_nm_plus, H_nm_minus of each pair-cache entry."""
        n = self.o.lib.or_pair_count(self.h)
        h = np.zeros((n, 2))
        self.o.lib.or_pair_export_h(self.h, _ptr(h))
        return h

    def eval(self, kind, t, x, y, z=None):
        t, x, y = np.broadcast_arrays(np.asarray(t, float), np.asarray(x, float), np.asarray(y, float))
        shape = t.shape
        t, x, y = (_arr(a.ravel()) for a in (t, x, y))
        zz = None if z is None else _arr(np.broadcast_to(np.asarray(z, float), shape).ravel())
        out = np.zeros(t.size)
        k = KIND[kind] if isinstance(kind, str) else int(kind)
        self.o.lib.or_eval(self.h, k, t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(zz), _ptr(out))
        return out.reshape(shape)

    def spectral_moments(self):
        m = np.zeros(3)
        self.o.lib.or_spectral_moments(self.h, _ptr(m))
        return m

    def grid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
             domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0):
        d = _arr(np.asarray(domain, float))
        di = _arr(np.asarray(domain_ignore, float))
        h = self.o.lib.or_grid_create(self.h, _ptr(d), nx, ny, nl, t0, dt, tan_a, tan_b,
                                      water_depth, swl, _ptr(di), int(ignore_at_init))
        return OracleGrid(self, h, nx, ny, nl)

    def lsgrid(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
               domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, init_only=0):
        """Legacy two-slot grid (lsGrid): initialize_surface_elevation(t0) +
        initialize_kinematics[_with_ignore]."""
        d = _arr(np.asarray(domain, float))
        di = _arr(np.asarray(domain_ignore, float))
        h = self.o.lib.or_lsgrid_create(self.h, _ptr(d), nx, ny, nl, t0, dt, tan_a, tan_b,
                                        water_depth, swl, _ptr(di), int(ignore_at_init), int(init_only))
        return OracleLsGrid(self, h)

    def column_fill(self, x, y, t, nl, tan_a, tan_b, water_depth, swl=0.0):
        x = _arr(x)
        y = _arr(y)
        n = x.size
        eta = np.zeros(n)
        ux, uy, uz = (np.zeros(n * nl) for _ in range(3))
        sec = self.o.lib.or_column_fill(self.h, n, _ptr(x), _ptr(y), t, nl, tan_a, tan_b,
                                        water_depth, swl, _ptr(eta), _ptr(ux), _ptr(uy), _ptr(uz))
        return sec, eta, ux.reshape(n, nl), uy.reshape(n, nl), uz.reshape(n, nl)


class OracleGrid:
    def __init__(self, sea: OracleSea, handle, nx, ny, nl):
        self.sea, self.h, self.nx, self.ny, self.nl = sea, handle, nx, ny, nl
        self.lib = sea.o.lib

    def __del__(self):
        try:
            self.lib.or_grid_destroy(self.h)
        except Exception:
            pass

    def force_update(self, t):
        return self.lib.or_grid_force_update(self.h, t)

    def state(self):
        ts = C.c_int()
        t0, dx, dy, ds = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self.lib.or_grid_state(self.h, C.byref(ts), C.byref(t0), C.byref(dx), C.byref(dy), C.byref(ds))
        return dict(tstep=ts.value, t0=t0.value, dx=dx.value, dy=dy.value, ds=ds.value)

    def export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else which
        n = self.lib.or_grid_export(self.h, w, None)
        out = np.zeros(n)
        self.lib.or_grid_export(self.h, w, _ptr(out))
        return out

    def query(self, t, x, y, z):
        t, x, y, z = (_arr(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        self.lib.or_grid_query(self.h, t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(z), _ptr(out))
        return out


class OracleLsGrid:
    """Legacy two-slot grid of the oracle; export(which) gives slot 0 then slot 1
    (which: 0 ETA, 4..6 UX/UY/UZ), query gives rows {eta, u, v, w}."""

    def __init__(self, sea, handle):
        self.sea, self.h = sea, handle
        self.lib = sea.lib if isinstance(sea, Oracle) else sea.o.lib

    def __del__(self):
        try:
            self.lib.or_lsgrid_destroy(self.h)
        except Exception:
            pass

    def check_time(self, t):
        return bool(self.lib.or_lsgrid_check_time(self.h, t))

    def update(self, t_target):
        return self.lib.or_lsgrid_update(self.h, t_target)

    @property
    def t0(self):
        return self.lib.or_lsgrid_t0(self.h)

    def export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else which
        n = self.lib.or_lsgrid_export(self.h, w, None)
        out = np.zeros(n)
        self.lib.or_lsgrid_export(self.h, w, _ptr(out))
        return out

    def query(self, t, x, y, z):
        t, x, y, z = (_arr(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        self.lib.or_lsgrid_query(self.h, t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(z), _ptr(out))
        return out


class RefHarness:
    """The compiled reference classes (oracle/_ref/libcfdwm_refh.so).

    Process-global state: one sea and one grid at a time, as in the reference.
    """

    def __init__(self, path=None):
        path = path or os.path.join(REF_DIR, "libcfdwm_refh.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference tree exists")
        L = self.lib = C.CDLL(path)
        L.refh_setup.restype = C.c_int
        L.refh_setup.argtypes = _SEA_ARGTYPES
        L.refh_get_components.argtypes = [_dp, _dp, _dp]
        L.refh_set_ramp.argtypes = [C.c_int, C.c_int, _dp]
        L.refh_ramp.restype = C.c_double
        L.refh_ramp.argtypes = [C.c_double] * 3
        L.refh_freq_ramp.restype = C.c_double
        L.refh_freq_ramp.argtypes = [C.c_double] * 4
        L.refh_eval.argtypes = [C.c_int, C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.refh_pair_count.restype = C.c_long
        L.refh_pair_export.argtypes = [_ip, _bp, _dp]
        L.refh_kmax.restype = C.c_double
        L.refh_set_vertical_theory.argtypes = [C.c_int]
        L.refh_pair_export_h.argtypes = [_dp]
        L.refh_grid_init.restype = C.c_double
        L.refh_grid_init.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                     C.c_int, C.c_int]
        L.refh_grid_force_update.restype = C.c_int
        L.refh_grid_force_update.argtypes = [C.c_double]
        L.refh_grid_state.argtypes = [_ip, _dp, _dp, _dp, _dp]
        L.refh_grid_export.restype = C.c_long
        L.refh_grid_export.argtypes = [C.c_int, _dp]
        L.refh_grid_query.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.refh_lsgrid_init.restype = C.c_double
        L.refh_lsgrid_init.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                       C.c_int, C.c_int]
        L.refh_lsgrid_update.argtypes = [C.c_double]
        L.refh_lsgrid_check_time.restype = C.c_int
        L.refh_lsgrid_check_time.argtypes = [C.c_double]
        L.refh_lsgrid_t0.restype = C.c_double
        L.refh_lsgrid_export.restype = C.c_long
        L.refh_lsgrid_export.argtypes = [C.c_int, _dp]
        L.refh_lsgrid_query.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.refh_lsgrid_write_vtu.restype = C.c_int
        L.refh_lsgrid_write_vtu.argtypes = [C.c_char_p, C.c_int]
        L.refh_sigmaS2cart.restype = C.c_double
        L.refh_sigmaS2cart.argtypes = [C.c_double] * 3
        L.refh_cart2sigmaS.restype = C.c_double
        L.refh_cart2sigmaS.argtypes = [C.c_double] * 3
        self.n = 0

    def setup(self, sea: dict):
        args, keep = _sea_args(sea)
        self._keep = keep
        self.n = len(sea["omega"])
        self.lib.refh_setup(*args)

    def components(self):
        A, ph = np.zeros(self.n), np.zeros(self.n)
        bw = C.c_double()
        self.lib.refh_get_components(_ptr(A), _ptr(ph), C.byref(bw))
        return A, ph, bw.value

    def set_ramp(self, enabled, flags, se12):
        self._se = _arr(np.asarray(se12, float))
        self.lib.refh_set_ramp(int(enabled), int(flags), _ptr(self._se))

    def eval(self, kind, t, x, y, z=None):
        t, x, y = np.broadcast_arrays(np.asarray(t, float), np.asarray(x, float), np.asarray(y, float))
        shape = t.shape
        t, x, y = (_arr(a.ravel()) for a in (t, x, y))
        zz = None if z is None else _arr(np.broadcast_to(np.asarray(z, float), shape).ravel())
        out = np.zeros(t.size)
        k = KIND[kind] if isinstance(kind, str) else int(kind)
        self.lib.refh_eval(k, t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(zz), _ptr(out))
        return out.reshape(shape)

    def pairs(self):
        n = self.lib.refh_pair_count()
        im = np.zeros((n, 2), dtype=np.int32)
        fl = np.zeros(n, dtype=np.uint8)
        v = np.zeros((n, 7))
        self.lib.refh_pair_export(im.ctypes.data_as(_ip), fl.ctypes.data_as(_bp), _ptr(v))
        return im, fl, v

    def kmax(self):
        return self.lib.refh_kmax()

    def set_vertical_theory(self, vt):
        self.lib.refh_set_vertical_theory(int(vt))

    def pairs_h(self):
        n = self.lib.refh_pair_count()
        h = np.zeros((n, 2))
        self.lib.refh_pair_export_h(_ptr(h))
        return h

    def grid_init(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
                  domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_domain=0, ignore_at_init=0):
        self._d = _arr(np.asarray(domain, float))
        self._di = _arr(np.asarray(domain_ignore, float))
        self.grid_shape = (nx, ny, nl)
        return self.lib.refh_grid_init(_ptr(self._d), nx, ny, nl, t0, dt, tan_a, tan_b,
                                       water_depth, swl, _ptr(self._di), int(ignore_domain),
                                       int(ignore_at_init))

    def grid_force_update(self, t):
        return self.lib.refh_grid_force_update(t)

    # -- legacy two-slot grid (lsGrid) ---------------------------------------
    def lsgrid_init(self, domain, nx, ny, nl, t0, dt, tan_a, tan_b, water_depth, swl=0.0,
                    domain_ignore=(1e12, -1e12, 1e12, -1e12), ignore_at_init=0, init_only=0):
        self._ld = _arr(np.asarray(domain, float))
        self._ldi = _arr(np.asarray(domain_ignore, float))
        return self.lib.refh_lsgrid_init(_ptr(self._ld), nx, ny, nl, t0, dt, tan_a, tan_b,
                                         water_depth, swl, _ptr(self._ldi), int(ignore_at_init), int(init_only))

    def lsgrid_update(self, t_target):
        self.lib.refh_lsgrid_update(t_target)

    def lsgrid_check_time(self, t):
        return bool(self.lib.refh_lsgrid_check_time(t))

    def lsgrid_t0(self):
        return self.lib.refh_lsgrid_t0()

    def lsgrid_export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else which
        n = self.lib.refh_lsgrid_export(w, None)
        out = np.zeros(n)
        self.lib.refh_lsgrid_export(w, _ptr(out))
        return out

    def lsgrid_query(self, t, x, y, z):
        t, x, y, z = (_arr(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        self.lib.refh_lsgrid_query(t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(z), _ptr(out))
        return out

    def lsgrid_write_vtu(self, path, last=False):
        return self.lib.refh_lsgrid_write_vtu(str(path).encode(), int(last))

    def grid_state(self):
        ts = C.c_int()
        t0, dx, dy, ds = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self.lib.refh_grid_state(C.byref(ts), C.byref(t0), C.byref(dx), C.byref(dy), C.byref(ds))
        return dict(tstep=ts.value, t0=t0.value, dx=dx.value, dy=dy.value, ds=ds.value)

    def grid_export(self, which):
        w = GRID_FIELDS.index(which) if isinstance(which, str) else which
        n = self.lib.refh_grid_export(w, None)
        out = np.zeros(n)
        self.lib.refh_grid_export(w, _ptr(out))
        return out

    def grid_query(self, t, x, y, z):
        t, x, y, z = (_arr(np.ravel(a)) for a in np.broadcast_arrays(t, x, y, z))
        out = np.zeros((t.size, 4))
        self.lib.refh_grid_query(t.size, _ptr(t), _ptr(x), _ptr(y), _ptr(z), _ptr(out))
        return out
