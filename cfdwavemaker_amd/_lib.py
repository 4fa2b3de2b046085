"""Loader and ctypes bindings for the in-tree libcfdwm.so.

The shared object is built by ``make -C cfdwavemaker_amd/csrc`` (or
``__graft_entry__.build()``).  There is no CPU fallback: if the library is
missing, importing the bindings raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcfdwm.so")

_d = C.c_double
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64 = C.c_int64

# public symbols of include/cfdwm.h (the export test checks this list against the header)
REFERENCE_API = [
    "CFDwavemaker_is_initialized", "wave_Initialize", "wave_Cleanup", "wave_VeloX", "wave_VeloY",
    "wave_VeloZ", "wave_DynPres", "wave_nhPres", "wave_SurfElev", "wave_VFrac", "wave_Seabed",
    "wave_Kinematics", "wave_piston_position", "wave_force_update", "wave_set_mpi_comm",
    "update_probes", "wave_phase_speed", "wave_mean_period", "wave_mean_length",
    "wave_water_depth", "VelocityX", "VelocityY", "VelocityZ", "DynamicPressure",
    "SurfaceElevation", "VolumeFraction", "Init", "Prepare", "Cleanup",
]
BATCH_API = [
    "wave_EvalPoints", "wave_EvalPointsDevice", "cfdwm_sea_create", "cfdwm_sea_destroy",
    "cfdwm_sea_terms", "cfdwm_eval", "cfdwm_grid_create", "cfdwm_grid_destroy",
    "cfdwm_grid_refill", "cfdwm_grid_force_update", "cfdwm_grid_state", "cfdwm_grid_export",
    "cfdwm_grid_query", "cfdwm_grid_query_device_f32", "cfdwm_sync", "cfdwm_last_error",
    "cfdwm_prof_enable", "cfdwm_prof_read", "cfdwm_grid_qform_info", "cfdwm_depth_basis", "cfdwm_profile_basis",
    "cfdwm_mlp_create", "cfdwm_mlp_destroy", "cfdwm_mlp_nparams", "cfdwm_mlp_get_params",
    "cfdwm_mlp_set_params", "cfdwm_mlp_set_norm", "cfdwm_mlp_get_norm", "cfdwm_mlp_forward_device",
    "cfdwm_mlp_loss_grad_device", "cfdwm_mlp_train_device", "cfdwm_surrogate_linear", "cfdwm_surrogate_features", "cfdwm_surrogate_add",
    "cfdwm_grids_refill_batched", "cfdwm_grid_release_workspace", "wave_PrefetchPoints", "wave_PrefetchClear",
    "wave_PrefetchStats", "cfdwm_eval_fields", "cfdwm_grid_write_vtu", "cfdwm_sea_set_vertical_theory",
    "cfdwm_grid_qform_clock", "cfdwm_grid_qform_mode", "cfdwm_sea_clone", "cfdwm_grid_kinematics_device_f32",
    "cfdwm_lsgrid_create", "cfdwm_lsgrid_destroy", "cfdwm_lsgrid_check_time", "cfdwm_lsgrid_update",
    "cfdwm_lsgrid_state", "cfdwm_lsgrid_export", "cfdwm_lsgrid_query", "cfdwm_lsgrid_query_device",
    "cfdwm_lsgrid_write_vtu", "cfdwm_lsgrid_set_vtk",
]


class BatchStatsC(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("groups", "batched", "single", "tiles", "blocks", "order", "tuned")] + \
               [("forms_ms", C.c_double), ("mfma_flop", C.c_double)] + \
               [("group_forms", C.c_int), ("group_ptiles", C.c_int), ("tables_ms", C.c_double),
                ("setup_ms", C.c_double)]


class SeaParams(C.Structure):
    _fields_ = [(n, _d) for n in ("depth", "g", "rho", "swl", "mtheta", "tofmax", "x0", "y0",
                                  "bandwidth", "cutoff", "ampl")] + \
               [(n, C.c_int) for n in ("order", "normalize", "bw_auto", "precision", "with_direct")]


class GridDescC(C.Structure):
    _fields_ = [("domain", _d * 4), ("nx", C.c_int), ("ny", C.c_int), ("nl", C.c_int),
                ("t0", _d), ("dt", _d), ("tan_a", _d), ("tan_b", _d), ("water_depth", _d),
                ("swl", _d), ("domain_ignore", _d * 4), ("ignore_at_init", C.c_int), ("defer_fill", C.c_int)]


_LIB = None


def load(path: str | None = None) -> C.CDLL:
    """Load libcfdwm.so once (RTLD_GLOBAL not needed) and declare signatures."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIBPATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `make -C cfdwavemaker_amd/csrc` "
                           "(there is no CPU fallback)")
    # torch bundles a HIP runtime whose soname is libamdhip64.so.7 but which
    # torch's own libraries request as "libamdhip64.so".  Loaded after
    # libcfdwm (which requests libamdhip64.so.7 from /opt/rocm), torch would
    # bring in a second runtime that sees no device.  Loading torch first makes
    # libcfdwm bind to the runtime already in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(p)
    xyzt = [_d, _d, _d, _d]
    L.CFDwavemaker_is_initialized.restype = C.c_int
    L.wave_Initialize.restype = C.c_int
    L.wave_Cleanup.restype = C.c_int
    for f in ("wave_VeloX", "wave_VeloY", "wave_VeloZ", "wave_DynPres", "wave_nhPres"):
        getattr(L, f).restype = _d
        getattr(L, f).argtypes = xyzt
    L.wave_SurfElev.restype = _d
    L.wave_SurfElev.argtypes = [_d, _d, _d]
    L.wave_VFrac.restype = _d
    L.wave_VFrac.argtypes = [_d] * 5
    L.wave_Seabed.restype = _d
    L.wave_Seabed.argtypes = [_d, _d]
    L.wave_Kinematics.restype = _dp
    L.wave_Kinematics.argtypes = xyzt
    L.wave_piston_position.restype = _d
    L.wave_piston_position.argtypes = [_d]
    L.wave_force_update.restype = None
    L.wave_force_update.argtypes = [_d]
    L.update_probes.restype = None
    L.update_probes.argtypes = [_d]
    L.wave_set_mpi_comm.restype = None
    L.wave_set_mpi_comm.argtypes = [C.c_int]
    for f in ("wave_phase_speed", "wave_mean_period", "wave_mean_length"):
        getattr(L, f).restype = _d
        getattr(L, f).argtypes = [C.c_int]
    L.wave_water_depth.restype = _d
    L.wave_EvalPoints.restype = C.c_int
    L.wave_EvalPoints.argtypes = [_i64] + [_dp] * 9 + [C.c_uint32]
    L.wave_EvalPointsDevice.restype = C.c_int
    L.wave_EvalPointsDevice.argtypes = [_i64] + [C.c_void_p] * 9 + [C.c_uint32]
    L.cfdwm_sea_create.restype = C.c_int
    L.cfdwm_sea_create.argtypes = [C.POINTER(C.c_void_p), C.c_int] + [_dp] * 8 + [C.POINTER(SeaParams)]
    L.cfdwm_sea_destroy.argtypes = [C.c_void_p]
    L.cfdwm_sea_clone.restype = C.c_int
    L.cfdwm_sea_clone.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_int]
    L.cfdwm_sea_set_vertical_theory.restype = C.c_int
    L.cfdwm_sea_set_vertical_theory.argtypes = [C.c_void_p, C.c_int]
    L.cfdwm_sea_terms.argtypes = [C.c_void_p] + [C.POINTER(_i64)] * 3
    L.cfdwm_eval.restype = C.c_int
    L.cfdwm_eval.argtypes = [C.c_void_p, C.c_int, _i64] + [_dp] * 5
    L.cfdwm_grid_create.restype = C.c_int
    L.cfdwm_grid_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(GridDescC)]
    L.cfdwm_grid_destroy.argtypes = [C.c_void_p]
    L.cfdwm_grid_refill.restype = C.c_int
    L.cfdwm_grid_refill.argtypes = [C.c_void_p]
    L.cfdwm_grid_force_update.restype = C.c_int
    L.cfdwm_grid_force_update.argtypes = [C.c_void_p, _d, C.POINTER(C.c_int)]
    L.cfdwm_grid_state.restype = C.c_int
    L.cfdwm_grid_state.argtypes = [C.c_void_p, C.POINTER(C.c_int)] + [_dp] * 4
    L.cfdwm_grid_export.restype = _i64
    L.cfdwm_grid_export.argtypes = [C.c_void_p, C.c_int, _dp]
    L.cfdwm_grid_query.restype = C.c_int
    L.cfdwm_grid_query.argtypes = [C.c_void_p, _i64] + [_dp] * 5
    L.cfdwm_grid_query_device_f32.restype = C.c_int
    L.cfdwm_grid_query_device_f32.argtypes = [C.c_void_p, _i64, _d] + [C.c_void_p] * 7
    L.cfdwm_grid_query_device_f32_noalloc.restype = C.c_int
    L.cfdwm_grid_query_device_f32_noalloc.argtypes = [C.c_void_p, _i64, _d] + [C.c_void_p] * 7
    L.cfdwm_grid_kinematics_device_f32.restype = C.c_int
    L.cfdwm_grid_kinematics_device_f32.argtypes = [C.c_void_p, _i64, _d] + [C.c_void_p] * 4
    L.cfdwm_sync.restype = C.c_int
    L.cfdwm_last_error.restype = C.c_int
    L.cfdwm_last_error.argtypes = [C.c_char_p, C.c_int]
    L.cfdwm_prof_enable.restype = C.c_int
    L.cfdwm_prof_enable.argtypes = [C.c_int]
    L.cfdwm_prof_read.restype = C.c_int
    L.cfdwm_prof_read.argtypes = [C.c_char_p, _dp, C.POINTER(_i64), C.c_int, C.c_int]
    L.cfdwm_grid_qform_info.restype = C.c_int
    L.cfdwm_grid_qform_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp]
    L.cfdwm_grid_qform_clock.restype = C.c_int
    L.cfdwm_grid_qform_clock.argtypes = [C.c_void_p, C.c_int, _dp]
    L.cfdwm_grid_qform_mode.restype = C.c_int
    L.cfdwm_grid_qform_mode.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp, C.POINTER(C.c_int)]
    L.cfdwm_depth_basis.restype = C.c_int
    L.cfdwm_depth_basis.argtypes = [_d, _d, _d, C.c_int, _dp, _dp, _dp]
    L.cfdwm_profile_basis.restype = C.c_int
    L.cfdwm_profile_basis.argtypes = [_d, _d, C.c_int, _d, C.c_int, _dp, _dp, _dp]
    _fp = C.POINTER(C.c_float)
    L.cfdwm_mlp_create.restype = C.c_int
    L.cfdwm_mlp_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]
    L.cfdwm_mlp_destroy.argtypes = [C.c_void_p]
    L.cfdwm_mlp_nparams.restype = _i64
    L.cfdwm_mlp_nparams.argtypes = [C.c_void_p]
    L.cfdwm_mlp_get_params.restype = C.c_int
    L.cfdwm_mlp_get_params.argtypes = [C.c_void_p, _fp]
    L.cfdwm_mlp_set_params.restype = C.c_int
    L.cfdwm_mlp_set_params.argtypes = [C.c_void_p, _fp]
    L.cfdwm_mlp_set_norm.restype = C.c_int
    L.cfdwm_mlp_set_norm.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp]
    L.cfdwm_mlp_get_norm.restype = C.c_int
    L.cfdwm_mlp_get_norm.argtypes = [C.c_void_p, _fp]
    L.cfdwm_mlp_forward_device.restype = C.c_int
    L.cfdwm_mlp_forward_device.argtypes = [C.c_void_p, _i64, C.c_void_p, C.c_void_p]
    L.cfdwm_mlp_loss_grad_device.restype = C.c_int
    L.cfdwm_mlp_loss_grad_device.argtypes = [C.c_void_p, _i64, C.c_void_p, C.c_void_p, _fp, _fp]
    L.cfdwm_mlp_train_device.restype = C.c_int
    L.cfdwm_mlp_train_device.argtypes = [C.c_void_p, _i64, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                                         C.c_uint64, C.c_int, _fp]
    L.cfdwm_surrogate_linear.restype = C.c_int
    L.cfdwm_surrogate_linear.argtypes = [C.c_void_p, _i64, C.c_void_p, C.c_void_p, _d, _d, C.c_int, _dp, C.c_int,
                                         _d, C.c_void_p, C.c_void_p]
    L.cfdwm_surrogate_features.restype = C.c_int
    L.cfdwm_surrogate_features.argtypes = [C.c_void_p, _i64, C.c_void_p, C.c_void_p, _d, _d, C.c_int, _dp,
                                           C.c_int, _d, C.c_int, C.c_void_p, C.c_void_p]
    L.cfdwm_surrogate_set_column.restype = C.c_int
    L.cfdwm_surrogate_set_column.argtypes = [_i64, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.cfdwm_surrogate_add.restype = C.c_int
    L.cfdwm_surrogate_add.argtypes = [_i64, C.c_void_p, C.c_void_p, C.c_void_p]
    L.cfdwm_grid_write_vtu.restype = C.c_int
    L.cfdwm_grid_write_vtu.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p]
    L.cfdwm_lsgrid_create.restype = C.c_int
    L.cfdwm_lsgrid_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(GridDescC), C.c_int]
    L.cfdwm_lsgrid_destroy.restype = C.c_int
    L.cfdwm_lsgrid_destroy.argtypes = [C.c_void_p]
    L.cfdwm_lsgrid_check_time.restype = C.c_int
    L.cfdwm_lsgrid_check_time.argtypes = [C.c_void_p, _d]
    L.cfdwm_lsgrid_update.restype = C.c_int
    L.cfdwm_lsgrid_update.argtypes = [C.c_void_p, _d, C.POINTER(C.c_int)]
    L.cfdwm_lsgrid_state.restype = C.c_int
    L.cfdwm_lsgrid_state.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int)]
    L.cfdwm_lsgrid_export.restype = _i64
    L.cfdwm_lsgrid_export.argtypes = [C.c_void_p, C.c_int, _dp]
    L.cfdwm_lsgrid_query.restype = C.c_int
    L.cfdwm_lsgrid_query.argtypes = [C.c_void_p, _i64] + [_dp] * 5
    L.cfdwm_lsgrid_query_device.restype = C.c_int
    L.cfdwm_lsgrid_query_device.argtypes = [C.c_void_p, _i64] + [C.c_void_p] * 5
    L.cfdwm_lsgrid_write_vtu.restype = C.c_int
    L.cfdwm_lsgrid_write_vtu.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p]
    L.cfdwm_lsgrid_set_vtk.restype = C.c_int
    L.cfdwm_lsgrid_set_vtk.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]
    L.cfdwm_eval_fields.restype = C.c_int
    L.cfdwm_eval_fields.argtypes = [C.c_void_p, _i64] + [C.c_void_p] * 10 + [C.c_int]
    L.cfdwm_grid_release_workspace.restype = C.c_int
    L.cfdwm_grid_release_workspace.argtypes = [C.c_void_p]
    L.cfdwm_grids_refill_batched.restype = C.c_int
    L.cfdwm_grids_refill_batched.argtypes = [C.POINTER(C.c_void_p), C.c_int, _d, C.c_int, C.POINTER(BatchStatsC)]
    if path is None:
        _LIB = L
    return L


def last_error(L=None) -> str:
    L = L or load()
    buf = C.create_string_buffer(1024)
    L.cfdwm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
