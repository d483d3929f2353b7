"""Metric sample points of ``ImageRegistrationMethodv4`` (REGULAR / RANDOM
strategies, ``itkImageRegistrationMethodv4.hxx:867-1024``), generated on the
host by the native library (``mmi_metric_sample_points``)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .geometry import ImageGeometry

NONE, REGULAR, RANDOM = N.MMI_SAMPLING_NONE, N.MMI_SAMPLING_REGULAR, N.MMI_SAMPLING_RANDOM


def metric_sample_points(virtual_geom: ImageGeometry, strategy: int, percentage: float, seed: int,
                         fixed_mask=None) -> tuple:
    """(points (n, dim) in virtual space, next seed).  ``seed`` is the
    method's current random seed; the perturbation stream uses it and RANDOM's
    voxel stream the next one.  ``fixed_mask``: an ImageMaskSpatialObject or
    None."""
    L = N.load()
    g = virtual_geom.to_c()
    s = ctypes.c_uint32(int(seed) & 0xFFFFFFFF)
    bound = ctypes.c_int64()
    mg = mpx = None
    if fixed_mask is not None:
        mg = fixed_mask.geometry.to_c()
        mpx = np.ascontiguousarray(fixed_mask.inside_flags())
    mgp = ctypes.byref(mg) if mg is not None else None
    mpp = mpx.ctypes.data if mpx is not None else None
    N.check(L.mmi_metric_sample_points(ctypes.byref(g), strategy, float(percentage), ctypes.byref(s), mgp, mpp,
                                       None, 0, ctypes.byref(bound)))
    out = np.zeros((max(bound.value, 1), virtual_geom.dim))
    n = ctypes.c_int64()
    N.check(L.mmi_metric_sample_points(ctypes.byref(g), strategy, float(percentage), ctypes.byref(s), mgp, mpp,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.shape[0],
                                       ctypes.byref(n)))
    return np.ascontiguousarray(out[: n.value]), int(s.value)
