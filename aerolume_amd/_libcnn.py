"""ctypes bindings of the learned-closure network entry points (``aero_cnn_*`` in
``include/aerolume_amd.h``)."""
from __future__ import annotations

import ctypes

i32 = ctypes.c_int32


class CnnDesc(ctypes.Structure):
    _fields_ = [("nz", i32), ("ny", i32), ("nx", i32), ("n_layers", i32), ("channels", i32 * 9), ("act", i32 * 8)]


def bind(lib) -> None:
    P = ctypes.c_void_p
    f = ctypes.c_float
    lib.aero_cnn_create.argtypes = [ctypes.POINTER(CnnDesc), ctypes.POINTER(P)]
    lib.aero_cnn_destroy.argtypes = [P]
    lib.aero_cnn_param_count.argtypes = [P]
    lib.aero_cnn_workspace_bytes.argtypes = [P]
    lib.aero_cnn_workspace_bytes.restype = ctypes.c_size_t
    for name in ("aero_cnn_set_params", "aero_cnn_get_params", "aero_cnn_get_grads"):
        getattr(lib, name).argtypes = [P, P, P]
    lib.aero_cnn_forward.argtypes = [P, P, P, P]
    lib.aero_cnn_infer.argtypes = [P, P, P, P]
    lib.aero_cnn_backward.argtypes = [P, P, P, P]
    lib.aero_cnn_adam_step.argtypes = [P, P, f, f, f, f, ctypes.c_int]
    for name in ("aero_cnn_create", "aero_cnn_destroy", "aero_cnn_param_count", "aero_cnn_set_params",
                 "aero_cnn_get_params", "aero_cnn_get_grads", "aero_cnn_forward", "aero_cnn_infer",
                 "aero_cnn_backward", "aero_cnn_adam_step"):
        getattr(lib, name).restype = ctypes.c_int
