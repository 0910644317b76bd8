"""ctypes binding of libqp3d.so (the C ABI in include/qp3d.h).

The library is built in-tree by ``make -C gpu-c_3dqmodeling_amd`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there is no
fallback compute path.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libqp3d.so")

QP3D_OK = 0
QP3D_ERR_ARG = -1
QP3D_ERR_IO = -2
QP3D_ERR_ALLOC = -3
QP3D_ERR_HIP = -4
QP3D_ERR_NONFINITE = -5
QP3D_ERR_STATE = -6
QP3D_ERR_TIMEOUT = -7

STORAGE = {"fp32": 0, "bf16": 1, "fp16": 2}
NTILES = 7  # launchable tile shapes (QP3D_NTILES in include/qp3d.h)

_F = C.POINTER(C.c_float)


class Qp3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qp3d error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("pml", C.c_int), ("nt", C.c_int),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
        ("sz", C.c_int), ("storage", C.c_int), ("use_graph", C.c_int), ("snapflag", C.c_int),
        ("ntsnap", C.c_int), ("zchunk", C.c_int), ("device", C.c_int),
        ("tile_update", C.c_int), ("tile_grad", C.c_int), ("tile_pml", C.c_int), ("streams", C.c_int),
        ("tile_pml_y", C.c_int), ("snap_s", C.c_int), ("gmax_mode", C.c_int),
    ]


# (name, restype, argtypes)
_SIGS = [
    ("qp3d_last_error", C.c_char_p, []),
    ("qp3d_set_error", None, [C.c_char_p]),
    ("modeling3d", None, [C.c_int] * 5 + [C.c_float] * 4 + [C.c_int] * 5 + [_F] * 14 + [C.c_char_p]),
    ("qp3d_modeling3d", C.c_int, [C.c_int] * 5 + [C.c_float] * 4 + [C.c_int] * 5 + [_F] * 14 + [C.c_char_p]),
    ("qp3d_modeling3d_noalloc", C.c_int,
     [C.c_int] * 5 + [C.c_float] * 4 + [C.c_int] * 5 + [_F] * 14 + [C.c_char_p, C.c_void_p]),
    ("qp3d_fdcoff1", None, [_F]),
    ("qp3d_fdcoff2", None, [_F]),
    ("qp3d_pmlcoff", None, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float] + [_F] * 6),
    ("qp3d_source", None, [C.c_int] * 3 + [C.c_float] * 8 + [C.c_int] * 4 + [_F, _F]),
    ("qp3d_extend_model", None, [_F, _F, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("qp3d_absmax", C.c_float, [_F, C.c_long]),
    ("qp3d_readmodel3d", C.c_int, [C.c_char_p, _F] + [C.c_int] * 9),
    ("qp3d_genmodel", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_float, C.c_float, _F, _F, _F]),
    ("qp3d_plan_create", C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("qp3d_plan_destroy", C.c_int, [C.c_void_p]),
    ("qp3d_plan_upload", C.c_int, [C.c_void_p] + [_F] * 13),
    ("qp3d_plan_reset", C.c_int, [C.c_void_p]),
    ("qp3d_plan_run", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("qp3d_plan_sync", C.c_int, [C.c_void_p]),
    ("qp3d_plan_run_on", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("qp3d_plan_record", C.c_int, [C.c_void_p, _F]),
    ("qp3d_plan_field", C.c_int, [C.c_void_p, C.c_int, _F]),
    ("qp3d_plan_snapshot_count", C.c_int, [C.c_void_p]),
    ("qp3d_plan_snapshot", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _F]),
    ("qp3d_plan_snapshot_s", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _F]),
    ("qp3d_plan_step_maxima", C.c_int, [C.c_void_p, _F]),
    ("qp3d_plan_step_predictions", C.c_int, [C.c_void_p, _F]),
    ("qp3d_plan_set_amp_bound", C.c_int, [C.c_void_p, C.c_int]),
    ("qp3d_plan_gradient_tiles", C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong)]),
    ("qp3d_plan_update_cells", C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong)]),
    ("qp3d_plan_step_times", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("qp3d_plan_set_instrument", C.c_int, [C.c_void_p, C.c_int]),
    ("qp3d_plan_gmax_mode", C.c_int, [C.c_void_p]),
    ("qp3d_work_items", C.c_int, [C.POINTER(Config), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]),
    ("qp3d_plan_dims", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("qp3d_plan_kernel_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("qp3d_plan_step_bytes", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("qp3d_fwi_create", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("qp3d_fwi_destroy", C.c_int, [C.c_void_p]),
    ("qp3d_fwi_forward", C.c_int, [C.c_void_p]),
    ("qp3d_fwi_gradient", C.c_int, [C.c_void_p, _F] + [C.POINTER(C.c_double)] * 3),
    ("qp3d_fwi_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("qp3d_fwi_set_rest_map", C.c_int, [C.c_void_p, C.c_int]),
    ("qp3d_fwi_set_resident_segments", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("qp3d_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("qp3d_batch_destroy", C.c_int, [C.c_void_p]),
    ("qp3d_batch_reset", C.c_int, [C.c_void_p]),
    ("qp3d_batch_run", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("qp3d_batch_sync", C.c_int, [C.c_void_p]),
    ("qp3d_batch_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("qp3d_batch_step_bytes", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("qp3d_batch_per_shot", C.c_int, [C.c_void_p]),
    ("qp3d_batch_share", C.c_int, [C.c_void_p]),
    ("qp3d_batch_set_per_shot", C.c_int, [C.c_void_p, C.c_int]),
    ("qp3d_batch_persistent", C.c_int, [C.c_void_p]),
    ("qp3d_batch_set_persistent", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("qp3d_batch_set_wait_limit", C.c_int, [C.c_void_p, C.c_double]),
    ("qp3d_batch_persistent_choice", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def load():
    """Load libqp3d.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {PKG_ROOT}` or __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != QP3D_OK:
        raise Qp3dError(rc, load().qp3d_last_error().decode(errors="replace"))
    return rc


def fptr(a):
    """float* of a C-contiguous float32 numpy array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_F)
