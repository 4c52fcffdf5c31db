"""ctypes binding of ``libaerolume_amd.so`` (the C ABI declared in ``include/aerolume_amd.h``).

The library is built in-tree (``aerolume_amd/build.py``). Loading fails loudly when it is
missing: there is no fallback implementation of the step.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_NAME = "libaerolume_amd.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

AERO_OK = 0
AERO_EINVAL = -1
AERO_EUNSUPPORTED = -2
AERO_EHIP = -3

BC = {"slip": 0, "noslip": 1, "moving": 2}
TH_MODE = {"adiabatic": 0, "temperature": 1, "heat_flux": 2, "inlet": 3, "outlet": 4, "periodic": 5}
TURB = {"none": 0, "smagorinsky": 1, "wale": 2, "vreman": 3}
MAX_GROUPS = 8

i32 = ctypes.c_int32
f64 = ctypes.c_double
vp = ctypes.c_void_p


class Desc2D(ctypes.Structure):
    _fields_ = [("nx", i32), ("ny", i32), ("dtype", i32), ("periodic", i32), ("closed", i32),
                ("advection", i32), ("turbulence", i32), ("bc_bottom", i32), ("bc_top", i32),
                ("bc_left", i32), ("bc_right", i32), ("fixed_dt", i32), ("win_j0", i32), ("win_j1", i32),
                ("win_i0", i32), ("win_i1", i32), ("nbody", i32), ("thermal", i32), ("buoyant", i32),
                ("wall_modes", i32 * 4), ("n_temp", i32), ("has_heat", i32), ("insulated", i32),
                ("dye_mode", i32), ("n_dye", i32), ("n_groups", i32), ("group_channel", i32 * MAX_GROUPS),
                ("h", f64), ("ground_speed", f64), ("cfl", f64), ("force_ref", f64), ("les_c", f64),
                ("gravity_x", f64), ("gravity_y", f64), ("prt", f64), ("diff_coef", f64),
                ("tile_rows", i32), ("pid_dt", i32), ("diagnostics", i32), ("wall_model", i32),
                ("n_flux", i32), ("reserved", i32)]


class Fields2D(ctypes.Structure):
    _fields_ = [("u", vp), ("v", vp), ("p", vp), ("theta", vp), ("dye", vp), ("t", vp), ("dt", vp), ("ctrl", vp)]


class Geom2D(ctypes.Structure):
    _fields_ = [(k, vp) for k in ("chi_u", "chi_v", "chi_c", "a_u", "a_v", "k_u", "k_v", "chi_t", "a_t", "w_t",
                                  "vol_u", "vol_v", "ins_u", "ins_v", "q_skin", "w_f")]


class Params2D(ctypes.Structure):
    _fields_ = [(k, vp) for k in ("nu", "eta", "u_inlet", "dt_fixed", "kappa", "ri", "eta_t", "theta_in",
                                  "wall_values", "heat", "heat_max", "dye_inflow", "group_weight", "group_target",
                                  "group_timing")]


class Outputs2D(ctypes.Structure):
    _fields_ = [("coeffs", vp), ("heat", vp), ("wall_heat", vp), ("heat_storage", vp), ("flow_heat", vp),
                ("hist", vp), ("hist_count", vp), ("hist_cap", ctypes.c_int64), ("aux", vp), ("aux_count", vp),
                ("aux_cap", ctypes.c_int64), ("diag", vp), ("diag_count", vp), ("diag_cap", ctypes.c_int64),
                ("theta_wall", vp)]


#: columns of a 2D diagnostics row (AERO2D_DIAG_COLS; float64)
DIAG_COLS = 22

class Trace2D(ctypes.Structure):
    _fields_ = [(k, vp * 3) for k in ("pen_u", "pen_v", "div", "psi", "proj_u", "proj_v", "theta")]


class ParticlesGrid2D(ctypes.Structure):
    _fields_ = [("dtype", i32), ("nx", i32), ("ny", i32), ("reserved", i32), ("x0", f64), ("y0", f64), ("h", f64)]


class Particles2D(ctypes.Structure):
    _fields_ = [(k, vp) for k in ("pos", "birth", "alive", "head", "acc", "last", "noz_pos", "noz_owner", "rate",
                                  "timing", "pulsed")] + [("n_nozzles", i32), ("n_emitters", i32), ("ring", i32),
                                                          ("reserved", i32)]


_LIB = None


def load() -> ctypes.CDLL:
    """The loaded library (raises when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(os.environ.get("AEROLUME_AMD_LIB", str(LIB_PATH)))
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
                           "or python -m aerolume_amd.build)")
    lib = ctypes.CDLL(str(path))
    P = ctypes.c_void_p
    lib.aero_last_error.restype = ctypes.c_char_p
    lib.aero_version.restype = ctypes.c_int
    lib.aero2d_plan_create.argtypes = [ctypes.POINTER(Desc2D), ctypes.POINTER(P)]
    lib.aero2d_plan_destroy.argtypes = [P]
    lib.aero2d_plan_workspace_bytes.argtypes = [P]
    lib.aero2d_plan_workspace_bytes.restype = ctypes.c_size_t
    lib.aero2d_step.argtypes = [P, P, ctypes.POINTER(Fields2D), ctypes.POINTER(Fields2D), ctypes.POINTER(Geom2D),
                                ctypes.POINTER(Params2D), ctypes.POINTER(Outputs2D)]
    lib.aero2d_step_chained.argtypes = lib.aero2d_step.argtypes
    lib.aero2d_step_chained.restype = ctypes.c_int
    lib.aero2d_project.argtypes = [P, P, P, P, P]
    lib.aero2d_poisson_solve.argtypes = [P, P, P, P]
    lib.aero2d_poisson_solve_static.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                                ctypes.c_int32, P, P]
    lib.aero2d_poisson_solve_static.restype = ctypes.c_int
    lib.aero2d_time_stage.argtypes = [P, P, ctypes.POINTER(Fields2D), ctypes.POINTER(Geom2D),
                                      ctypes.POINTER(Params2D), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    lib.aero2d_time_stage.restype = ctypes.c_int
    for name in ("aero2d_plan_create", "aero2d_plan_destroy", "aero2d_step", "aero2d_project",
                 "aero2d_poisson_solve"):
        getattr(lib, name).restype = ctypes.c_int
    lib.aero2d_step_trace.argtypes = lib.aero2d_step.argtypes + [ctypes.POINTER(Trace2D)]
    lib.aero2d_step_trace.restype = ctypes.c_int
    lib.aero_polygon_sdf.argtypes = [P, ctypes.c_int32, P, ctypes.c_int32, P, P, ctypes.c_int64, P]
    lib.aero_polygon_sdf.restype = ctypes.c_int
    lib.aero2d_particles_step.argtypes = [P, ctypes.POINTER(ParticlesGrid2D), P, P, P, P, P,
                                          ctypes.POINTER(Particles2D)]
    lib.aero2d_particles_step.restype = ctypes.c_int
    from . import _lib3d
    _lib3d.bind(lib)
    _bind_ensemble(lib)
    from . import _libcnn
    _libcnn.bind(lib)
    _LIB = lib
    return lib


def check(rc: int, what: str) -> None:
    """Raise with the library's message on a non-zero return code."""
    if rc != AERO_OK:
        msg = load().aero_last_error().decode(errors="replace")
        kind = {AERO_EINVAL: ValueError, AERO_EUNSUPPORTED: NotImplementedError}.get(rc, RuntimeError)
        raise kind(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device address of a tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_handle(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def _bind_ensemble(lib) -> None:
    P = ctypes.c_void_p
    lib.aero2d_ensemble_create.argtypes = [ctypes.c_int32, ctypes.POINTER(P), ctypes.POINTER(Fields2D),
                                           ctypes.POINTER(Geom2D), ctypes.POINTER(Params2D),
                                           ctypes.POINTER(Outputs2D), ctypes.POINTER(P)]
    lib.aero2d_ensemble_step.argtypes = [P, P, ctypes.c_int32]
    lib.aero2d_ensemble_tune.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    lib.aero2d_ensemble_destroy.argtypes = [P]
    for name in ("aero2d_ensemble_create", "aero2d_ensemble_step", "aero2d_ensemble_tune",
                 "aero2d_ensemble_destroy"):
        getattr(lib, name).restype = ctypes.c_int
