"""ctypes structures and bindings of the 3D entry points of ``libaerolume_amd.so``
(``aero3d_*`` in ``include/aerolume_amd.h``)."""
from __future__ import annotations

import ctypes

i32 = ctypes.c_int32
f64 = ctypes.c_double
vp = ctypes.c_void_p

MAX_BODIES = 8


class Desc3D(ctypes.Structure):
    _fields_ = [("nx", i32), ("ny", i32), ("nz", i32), ("dtype", i32), ("periodic_x", i32), ("periodic_y", i32),
                ("periodic_z", i32), ("closed", i32), ("advection", i32), ("turbulence", i32), ("fixed_dt", i32),
                ("win_k0", i32), ("win_k1", i32), ("win_j0", i32), ("win_j1", i32), ("win_i0", i32),
                ("win_i1", i32), ("nbody", i32), ("symmetry", i32), ("ghost_kind", i32 * 18),
                ("ghost_value", f64 * 18), ("h", f64), ("y0", f64), ("cfl", f64), ("les_c", f64),
                ("ref_area", f64), ("length_ref", f64), ("span_ref", f64), ("moment_ref", f64 * 3),
                ("nut_model", i32), ("reserved", i32), ("nut_trust", f64), ("dye", i32), ("dye_k0", i32),
                ("dye_k1", i32), ("dye_j0", i32), ("dye_j1", i32), ("dye_i0", i32), ("dye_i1", i32),
                ("reserved2", i32), ("thermal", i32), ("theta_side", i32 * 6), ("theta_q_k0", i32),
                ("theta_q_k1", i32), ("theta_q_j0", i32), ("theta_q_j1", i32), ("theta_q_i0", i32),
                ("theta_q_i1", i32), ("theta_nt", i32), ("theta_tbody", i32 * 8), ("theta_adia", i32),
                ("theta_nskin", i32), ("richardson", f64), ("prandtl_turbulent", f64),
                ("diagnostics", i32), ("pid_dt", i32), ("pid_target", f64)]


class Fields3D(ctypes.Structure):
    _fields_ = [(k, vp) for k in ("u", "v", "w", "p", "t", "dt", "dye", "theta", "ctrl")]


#: momentum window arrays (the Geometry3D tensors of the same names)
GEOM_MOMENTUM = ("chi_u", "chi_v", "chi_w", "chi_c", "a_u", "a_v", "a_w", "k_u", "k_v", "k_w")
GEOM_THERMAL = ("chi_t", "a_t", "w_t", "adia_u", "adia_v", "adia_w", "vol_w", "theta_q", "skin_idx", "skin_w",
                "skin_off", "body_theta", "flux_rate")


class Geom3D(ctypes.Structure):
    _fields_ = [(k, vp) for k in GEOM_MOMENTUM + GEOM_THERMAL]


class Params3D(ctypes.Structure):
    _fields_ = [(k, vp) for k in ("nu", "eta", "u_inlet", "dt_fixed", "nut_field", "dye_weight", "dye_pulse",
                                  "theta_scalars", "theta_side_values")]


class Outputs3D(ctypes.Structure):
    _fields_ = [("coeffs", vp), ("hist", vp), ("hist_count", vp), ("hist_cap", ctypes.c_int64), ("heat_rate", vp),
                ("theta_wall", vp), ("phist", vp), ("phist_walls", i32), ("reserved", i32), ("diag", vp),
                ("diag_count", vp), ("diag_cap", ctypes.c_int64)]


#: per-stage arrays of a traced step (aero3d_trace), three stage pointers each
TRACE_FIELDS = ("tend_u", "tend_v", "tend_w", "star_u", "star_v", "star_w", "pen_u", "pen_v", "pen_w", "div", "psi",
                "proj_u", "proj_v", "proj_w", "theta")


class Trace3D(ctypes.Structure):
    _fields_ = [(k, vp * 3) for k in TRACE_FIELDS]


class Shape3D(ctypes.Structure):
    """aero3d_shape: an analytic body (kind 0 rounded box, 1 cylinder along y) and its seven
    parameters."""
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("p", ctypes.c_double * 7)]


#: shape kinds (AERO3D_SHAPE_*)
SHAPE_KINDS = {"rounded_box": 0, "cylinder_y": 1}

#: columns of a diagnostics row (AERO3D_DIAG_COLS)
DIAG_COLS = 28


def bind(lib) -> None:
    P = ctypes.c_void_p
    lib.aero3d_plan_create.argtypes = [ctypes.POINTER(Desc3D), ctypes.POINTER(P)]
    lib.aero3d_plan_destroy.argtypes = [P]
    lib.aero3d_plan_workspace_bytes.argtypes = [P]
    lib.aero3d_plan_workspace_bytes.restype = ctypes.c_size_t
    lib.aero3d_step.argtypes = [P, P, ctypes.POINTER(Fields3D), ctypes.POINTER(Fields3D), ctypes.POINTER(Geom3D),
                                ctypes.POINTER(Params3D), ctypes.POINTER(Outputs3D)]
    lib.aero3d_step_chained.argtypes = lib.aero3d_step.argtypes
    lib.aero3d_step_trace.argtypes = lib.aero3d_step.argtypes + [ctypes.POINTER(Trace3D)]
    lib.aero3d_project.argtypes = [P, P, P, P, P, P]
    lib.aero3d_poisson_solve.argtypes = [P, P, P, P]
    lib.aero3d_cell_gradient.argtypes = [P, P, P, P, P, P, ctypes.c_double]
    lib.aero3d_time_stage.argtypes = [P, P, ctypes.POINTER(Fields3D), ctypes.POINTER(Geom3D),
                                      ctypes.POINTER(Params3D), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    lib.aero3d_analytic_fractions.argtypes = [P, ctypes.POINTER(Shape3D), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                              ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_double, ctypes.c_double, P]
    lib.aero3d_analytic_fractions.restype = ctypes.c_int
    for name in ("aero3d_plan_create", "aero3d_plan_destroy", "aero3d_step", "aero3d_step_chained", "aero3d_step_trace",
                 "aero3d_project", "aero3d_poisson_solve", "aero3d_time_stage", "aero3d_cell_gradient"):
        getattr(lib, name).restype = ctypes.c_int
