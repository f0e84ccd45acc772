"""ctypes binding of libgaussnufft.so (the C ABI in include/gaussnufft.h and
include/gaussnufft_solver.h).

torch is imported before the library is opened so that the HIP runtime torch
bundles is the one the library binds to.  There is no fallback: if the library
is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads torch's HIP runtime before ours)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgaussnufft.so")

GN_OK, GN_EINVAL, GN_ENOMEM, GN_EHIP, GN_EROCFFT, GN_ESTATE = 0, 1, 2, 3, 4, 5
GN_COIL_LAST, GN_COIL_MAJOR = 0, 1
GN_TYPE1, GN_TYPE2, GN_ADJOINT, GN_TYPE21 = 1, 2, 3, 21
GN_SENSE_TYPE2, GN_SENSE_TYPE1, GN_SENSE_ADJOINT, GN_SENSE_TYPE21 = 32, 31, 33, 321
GN_GRID_SPREAD, GN_GRID_GATHER = 41, 42
GN_FFT_AUTO, GN_FFT_ROCFFT = 0, 1
GN_FFT_FORWARD, GN_FFT_INVERSE, GN_FFT_SHIFT_IN, GN_FFT_SHIFT_OUT = 1, -1, 1, 2
STAGES = ("pad", "fft2", "gather", "spread", "fft1", "crop")
GN_F32, GN_C64, GN_C128 = 1, 2, 3
GN_TV_AMP_SQUARE, GN_TV_AMP_MODULUS = 0, 1

# every symbol declared in include/*.h, with its ctypes signature
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
SIGNATURES = {
    "gn_version": ([], ctypes.c_int),
    "gn_last_error": ([], ctypes.c_char_p),
    "gn_grid_params": ([ctypes.c_int, _i64p, ctypes.c_double, ctypes.c_int, _ip, _ip, _i64p,
                        ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "gn_plan_create": ([ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                        ctypes.c_int, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "gn_plan_set_stream": ([_vp, _vp], ctypes.c_int),
    "gn_set_points": ([_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, _vp], ctypes.c_int),
    "gn_execute": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "gn_execute_type21_noalloc": ([_vp, _vp, _vp], ctypes.c_int),
    "gn_plan_set_sense": ([_vp, _vp], ctypes.c_int),
    "gn_fft_c2c": ([ctypes.c_int, _i64p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp,
                    _vp], ctypes.c_int),
    "gn_plan_info": ([_vp, _i64p, _ip, ctypes.POINTER(ctypes.c_double), _i64p, _i64p, _ip], ctypes.c_int),
    "gn_plan_set_tiles": ([_vp, _ip], ctypes.c_int),
    "gn_plan_set_timing": ([_vp, ctypes.c_int], ctypes.c_int),
    "gn_plan_stage_times": ([_vp, ctypes.POINTER(ctypes.c_double), _i64p, ctypes.c_int], ctypes.c_int),
    "gn_plan_sample_order": ([_vp, _ip], ctypes.c_int),
    "gn_traj_grad": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "gn_traj_grad_ex": ([_vp, _vp, _vp, _vp, ctypes.c_int], ctypes.c_int),
    "gn_plan_set_fft_engine": ([_vp, ctypes.c_int], ctypes.c_int),
    "gn_plan_destroy": ([_vp], ctypes.c_int),
    # include/gaussnufft_solver.h
    "gn_reduce_work_bytes": ([], ctypes.c_int64),
    "gn_dot_f64": ([ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gn_sqdist_f64": ([ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gn_tv_work_bytes": ([ctypes.c_int, _i64p, ctypes.c_int64], ctypes.c_int64),
    "gn_tv_prox": ([ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int64, _vp, ctypes.c_double, ctypes.c_double,
                    ctypes.c_int, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
}



class GNInstance(ctypes.Structure):
    """gn_instance of include/gaussnufft_batch.h"""
    _fields_ = [("modes", ctypes.c_int64 * 3), ("ncoils", ctypes.c_int64), ("M", ctypes.c_int64),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p), ("dcf", ctypes.c_void_p)]


_vpp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES.update({
    # include/gaussnufft_batch.h
    "gn_batch_create": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(GNInstance), ctypes.c_int, ctypes.c_int,
                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, ctypes.POINTER(_vp)],
                        ctypes.c_int),
    "gn_batch_execute": ([_vp, ctypes.c_int, _vpp, _vpp], ctypes.c_int),
    "gn_batch_set_graph": ([_vp, ctypes.c_int], ctypes.c_int),
    "gn_batch_set_tiles": ([_vp, _ip], ctypes.c_int),
    "gn_batch_autotune": ([_vp, ctypes.c_int, _vpp, _vpp, ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_double)],
                          ctypes.c_int),
    "gn_batch_info": ([_vp, _ip, _ip, _ip], ctypes.c_int),
    "gn_batch_destroy": ([_vp], ctypes.c_int),
    "gn_plan_autotune": ([_vp, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _vp,
                          _vp, ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
})

_lib = None


class GNError(RuntimeError):
    pass


def load():
    """Open the library (once) and attach the signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("mripy_amd: %s is not built; run `python -m mripy_amd._build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    """Raise the Python exception matching a GN_* status."""
    if rc == GN_OK:
        return
    msg = load().gn_last_error().decode(errors="replace")
    if rc == GN_EINVAL:
        raise ValueError(msg)
    if rc == GN_ENOMEM:
        raise MemoryError(msg)
    raise GNError("gaussnufft status %d: %s" % (rc, msg))
