"""ctypes binding of libfrhip.so (the C ABI declared in include/fr.h)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfrhip.so")

FR_LAYOUT_REF = 0
FR_LAYOUT_DEVICE = 1
FR_PREC_FP64 = 0
FR_PREC_FP32 = 1


class FrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libfrhip error {code}: {msg}")
        self.code = code


class FrConfig(ctypes.Structure):
    """Mirror of fr_config (include/fr.h)."""
    _fields_ = [
        ("equation", ctypes.c_int), ("nDims", ctypes.c_int), ("order", ctypes.c_int),
        ("riemannType", ctypes.c_int), ("viscous", ctypes.c_int), ("timeType", ctypes.c_int),
        ("sptsType", ctypes.c_int), ("vcjhScheme", ctypes.c_int), ("fixVis", ctypes.c_int),
        ("viscFix", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("gamma", ctypes.c_double), ("advectVx", ctypes.c_double), ("advectVy", ctypes.c_double),
        ("advectVz", ctypes.c_double), ("lambda_", ctypes.c_double), ("diffD", ctypes.c_double),
        ("rhoBound", ctypes.c_double), ("uBound", ctypes.c_double), ("vBound", ctypes.c_double),
        ("wBound", ctypes.c_double), ("pBound", ctypes.c_double),
        ("penFact", ctypes.c_double), ("tau", ctypes.c_double), ("mu_inf", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("rt_inf", ctypes.c_double), ("c_sth", ctypes.c_double),
        ("TWall", ctypes.c_double), ("RGas", ctypes.c_double), ("uWall", ctypes.c_double),
        ("vWall", ctypes.c_double), ("wWall", ctypes.c_double),
    ]


class FrMesh(ctypes.Structure):
    """Mirror of fr_mesh (include/fr.h)."""
    _fields_ = [
        ("nVerts", ctypes.c_int), ("xv", ctypes.POINTER(ctypes.c_double)),
        ("nEles", ctypes.c_int), ("nNodes", ctypes.c_int), ("c2v", ctypes.POINTER(ctypes.c_int)),
        ("vmap", ctypes.POINTER(ctypes.c_int)), ("nBounds", ctypes.c_int),
        ("bcType", ctypes.POINTER(ctypes.c_int)), ("bndOff", ctypes.POINTER(ctypes.c_int)),
        ("bndPts", ctypes.POINTER(ctypes.c_int)),
    ]


class FrErrorParams(ctypes.Structure):
    """Mirror of fr_error_params (include/fr.h)."""
    _fields_ = [
        ("testCase", ctypes.c_int), ("errorNorm", ctypes.c_int), ("quadOrder", ctypes.c_int),
        ("icType", ctypes.c_int),
        ("time", ctypes.c_double), ("rhoIC", ctypes.c_double), ("pIC", ctypes.c_double),
        ("vortexAngle", ctypes.c_double), ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
    ]


# every symbol include/fr.h declares, with (restype, argtypes)
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "fr_create": (ctypes.c_int, [ctypes.POINTER(FrConfig), ctypes.POINTER(FrMesh), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "fr_destroy": (ctypes.c_int, [_P]),
    "fr_last_error": (ctypes.c_char_p, []),
    "fr_get_sizes": (ctypes.c_int, [_P, _IP]),
    "fr_get_operator": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_get_geometry": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_get_faces": (ctypes.c_int, [_P, _IP]),
    "fr_get_element_order": (ctypes.c_int, [_P, _IP]),
    "fr_set_dt": (ctypes.c_int, [_P, ctypes.c_double]),
    "fr_calc_dt": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, _DP]),
    "fr_get_dt": (ctypes.c_int, [_P, _DP]),
    "fr_vjp_residual": (ctypes.c_int, [_P, _DP, _DP]),
    "fr_vjp_step": (ctypes.c_int, [_P, _DP, _DP]),
    "fr_vjp_steps": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _DP, _DP]),
    "fr_reduce_gradient": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
    "fr_batch_update": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_batch_tuning": (ctypes.c_int, [_P, _IP, _IP, _DP]),
    "fr_batch_group_tuning": (ctypes.c_int, [_P, ctypes.c_int, _IP, _IP, _IP, _IP, _DP]),
    "fr_batch_force_tile": (ctypes.c_int, [ctypes.c_int]),
    "fr_batch_destroy": (ctypes.c_int, [_P]),
    "fr_set_state": (ctypes.c_int, [_P, _DP, ctypes.c_int]),
    "fr_get_state": (ctypes.c_int, [_P, _DP, ctypes.c_int]),
    "fr_calc_residual": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_get_residual": (ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.c_int]),
    "fr_get_traces": (ctypes.c_int, [_P, _DP]),
    "fr_get_array": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_keep_final_residual": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_update": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_time_step_a": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
    "fr_time_step_b": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_copy_u_u0": (ctypes.c_int, [_P]),
    "fr_copy_u0_u": (ctypes.c_int, [_P]),
    "fr_reduce": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_set_stream": (ctypes.c_int, [_P, _P]),
    "fr_synchronize": (ctypes.c_int, [_P]),
    "fr_geometry_mode": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "fr_integrate_error": (ctypes.c_int, [_P, ctypes.POINTER(FrErrorParams), _DP]),
    "fr_reduce_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "fr_write_data": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "fr_read_restart": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, _DP]),
    "fr_write_checkpoint": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_double]),
    "fr_read_checkpoint": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), _DP]),
    "fr_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "fr_kernel_times": (ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.POINTER(ctypes.c_long)]),
    "fr_residual_bytes": (ctypes.c_int, [_P, _DP]),
}

_lib = None


def load():
    """Load libfrhip.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                                    "python fr_learn_amd/build.py")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        raise FrError(rc, load().fr_last_error().decode())
    return rc


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_IP)
