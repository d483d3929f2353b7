"""ctypes binding of the C ABI in include/itk_mattes_hip.h.

The HIP library is the only implementation of the metric: if it is missing,
or no HIP device is visible, every entry point raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes
import os

from .geometry import CGeometry

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libitk_mattes_hip.so")

MMI_OK = 0
MMI_ERR_ARGUMENT = -1
MMI_ERR_DEVICE = -2
MMI_ERR_DEGENERATE = -3
MMI_ERR_MEMORY = -4

MMI_F32, MMI_F64, MMI_BF16, MMI_F16 = 0, 1, 2, 3
MMI_U8, MMI_S8, MMI_U16, MMI_S16, MMI_U32, MMI_S32 = 4, 5, 6, 7, 8, 9
MMI_TF_IDENTITY, MMI_TF_AFFINE, MMI_TF_BSPLINE = 0, 1, 2
MMI_GRADIENT_CENTRAL_DIFFERENCE, MMI_GRADIENT_IMAGE = 0, 1
MMI_FLAG_RECORD_SAMPLE_BINS = 1
MMI_FLAG_EXACT_PATH_ONLY = 2
MMI_FLAG_TIME_PASSES = 4
MMI_FLAG_PASS2_RECLASSIFY = 8
MMI_SAMPLING_NONE, MMI_SAMPLING_REGULAR, MMI_SAMPLING_RANDOM = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_gp = ctypes.POINTER(CGeometry)

# name -> (restype, argtypes); the exported C ABI, declaration for declaration.
SIGNATURES = {
    "mmi_last_error": (ctypes.c_char_p, []),
    "mmi_context_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "mmi_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "mmi_context_destroy": (ctypes.c_int, [_vp]),
    "mmi_set_fixed_image": (ctypes.c_int, [_vp, _gp, ctypes.c_int32, _vp]),
    "mmi_set_moving_image": (ctypes.c_int, [_vp, _gp, ctypes.c_int32, _vp]),
    "mmi_set_moving_gradient": (ctypes.c_int, [_vp, _dp]),
    "mmi_compute_moving_gradient": (ctypes.c_int, [_vp]),
    "mmi_set_gradient_source": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmi_set_storage": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmi_set_smoothing_sigmas": (ctypes.c_int, [_vp, _dp, _dp]),
    "mmi_set_virtual_domain": (ctypes.c_int, [_vp, _gp]),
    "mmi_set_virtual_points": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.c_int64]),
    "mmi_metric_sample_points": (ctypes.c_int, [_gp, ctypes.c_int32, ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_uint32), _gp, _vp, _dp, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_int64)]),
    "mmi_set_fixed_mask": (ctypes.c_int, [_vp, _gp, _vp]),
    "mmi_set_moving_mask": (ctypes.c_int, [_vp, _gp, _vp]),
    "mmi_set_number_of_histogram_bins": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmi_set_launch_config": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "mmi_initialize": (ctypes.c_int, [_vp]),
    "mmi_set_fixed_transform": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.c_int32, _dp, ctypes.c_int64]),
    "mmi_set_moving_transform": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.c_int32, _dp, ctypes.c_int64]),
    # per-evaluation calls take plain addresses (ints): building typed ctypes
    # pointers costs microseconds per call
    "mmi_set_moving_parameters": (ctypes.c_int, [_vp, _vp, ctypes.c_int64]),
    "mmi_set_moving_initial_transforms": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
    "mmi_evaluate": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp]),
    "mmi_evaluate_static": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp, _vp]),
    "mmi_get_sample_bins": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int16), ctypes.POINTER(ctypes.c_int16)]),
    "mmi_get_moving_texels": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "mmi_get_histogram_parameters": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mmi_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "mmi_batch_destroy": (ctypes.c_int, [_vp]),
    "mmi_batch_set_graphs": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmi_batch_set_step_loop": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "mmi_set_histogram_copies": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmi_get_launch_config": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "mmi_load_tuning_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32)]),
    "mmi_batch_get_step_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "mmi_batch_evaluate": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_int32, ctypes.c_int32, _dp,
                                          ctypes.POINTER(_dp), ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int32)]),
    "mmi_batch_get_last_timing": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_int32)]),
    "mmi_regular_step_start": (ctypes.c_int, [_vp, _dp, ctypes.c_int32]),
    "mmi_batch_regular_step": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_int32, _vp, _vp, ctypes.c_int32]),
    "mmi_get_moving_parameters": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    "mmi_get_last_path": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "mmi_get_exception_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "mmi_measure_copy_bandwidth": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                                  ctypes.POINTER(ctypes.c_double)]),
    "mmi_get_last_timing": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


class RegularStepOptions(ctypes.Structure):
    _fields_ = [("learning_rate", ctypes.c_double), ("minimum_step_length", ctypes.c_double),
                ("relaxation_factor", ctypes.c_double), ("gradient_magnitude_tolerance", ctypes.c_double),
                ("number_of_iterations", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class RegularStepState(ctypes.Structure):
    _fields_ = [("gradient", ctypes.c_double * 12), ("previous_gradient", ctypes.c_double * 12),
                ("scales", ctypes.c_double * 12), ("relaxation", ctypes.c_double), ("value", ctypes.c_double),
                ("gradient_magnitude", ctypes.c_double), ("n_scales", ctypes.c_int32),
                ("scales_identity", ctypes.c_int32), ("iteration", ctypes.c_int32), ("stop", ctypes.c_int32),
                ("status", ctypes.c_int32), ("pad_", ctypes.c_int32)]


MMI_STOP_RUNNING = 0
MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS = 1
MMI_STOP_GRADIENT_MAGNITUDE_TOLERANCE = 2
MMI_STOP_STEP_TOO_SMALL = 3
MMI_STOP_METRIC_ERROR = 4


class MetricError(RuntimeError):
    """Raised for every non-zero status; the message is the C ABI's (the
    reference's exception text for degenerate inputs)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library (no device work).  Raises if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("the HIP metric library %s is missing: run `python -m itk_amd.build`" % path)
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int):
    if rc != MMI_OK:
        raise MetricError(rc, load().mmi_last_error().decode())
