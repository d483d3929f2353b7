"""Python mirror of ``itk::MattesMutualInformationImageToImageMetricv4`` on the
MI355X C ABI.

Method names, argument meaning and error behaviour follow the reference class
(``Modules/Registration/Metricsv4/include/itkMattesMutualInformationImageToImageMetricv4.h``
and its base ``itkImageToImageMetricv4.h``): set images and transforms,
``Initialize()``, then ``GetValue()`` / ``GetValueAndDerivative()``.  The value
is -MI, the derivative is -d(value)/d(parameters), exactly as in ITK.
Degenerate inputs raise ``MetricError`` carrying the reference's exception text.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .geometry import ImageGeometry
from .transforms import TF_AFFINE, TF_BSPLINE, TF_COMPOSITE

_dp = ctypes.POINTER(ctypes.c_double)


def _dptr(a):
    return a.ctypes.data_as(_dp)


# numpy pixel dtypes and the C ABI's pixel type codes
_PIXEL_CODES = {np.float32: N.MMI_F32, np.float64: N.MMI_F64, np.uint8: N.MMI_U8, np.int8: N.MMI_S8,
                np.uint16: N.MMI_U16, np.int16: N.MMI_S16, np.uint32: N.MMI_U32, np.int32: N.MMI_S32}


def transform_fixed_parameters(tf) -> np.ndarray:
    if tf.kind == TF_AFFINE:
        return np.asarray(tf.center, dtype=np.float64)
    if tf.kind == TF_BSPLINE:
        g = tf.grid
        d = g.dim
        return np.concatenate([np.asarray(g.size, dtype=np.float64), np.asarray(g.origin), np.asarray(g.spacing),
                               g.direction.ravel(), g.physical_to_index.ravel()])
    raise ValueError("unsupported transform")


class MattesMutualInformationImageToImageMetricv4:
    """Mattes MI metric evaluated on one HIP device: dense over the virtual
    domain or sparse over a point set, with optional image masks."""

    def __init__(self, device: int = 0, record_sample_bins: bool = False, exact_path_only: bool = False,
                 time_passes: bool = False, pass2_reclassify: bool = False):
        self._L = N.load()
        self._ctx = ctypes.c_void_p()
        flags = (N.MMI_FLAG_RECORD_SAMPLE_BINS if record_sample_bins else 0) | (
            N.MMI_FLAG_EXACT_PATH_ONLY if exact_path_only else 0) | (N.MMI_FLAG_TIME_PASSES if time_passes else 0) | (
            N.MMI_FLAG_PASS2_RECLASSIFY if pass2_reclassify else 0)
        N.check(self._L.mmi_context_create(device, flags, ctypes.byref(self._ctx)))
        self._bins = 50
        self._launch = (-1, -1)  # library's launch configuration
        self._hist_copies = 0
        self._out = np.zeros(2)  # [value (double), valid count (int64 bits)]
        self._out_addr = self._out.ctypes.data
        self._fixed = self._moving = None
        self._virtual = None
        self._moving_tf = None
        self._fixed_tf = None
        self._use_gradient_filter = True
        self._gradient = None
        self._value = float(np.finfo(np.float64).max)
        self._n_valid = 0
        self._last_pdf = None
        self._storage = N.MMI_F32
        self._fixed_sigma = self._moving_sigma = None
        self._masks = [None, None]  # ImageMaskSpatialObject for fixed, moving
        self._use_sampled_points = False
        self._use_virtual_points = False
        self._virtual_points = None
        self._fixed_points = None
        self._sample_points = None  # the virtual points evaluated (sparse sampling)
        self._skipped_fixed_points = 0
        self._images_dirty = True    # host images not yet on the device
        self._gradient_dirty = True  # user gradient not yet on the device
        self._initialized = False

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self._ctx:
            self._L.mmi_context_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs -----------------------------------------------------------
    def SetFixedImage(self, pixels: np.ndarray, geometry: ImageGeometry):
        self._fixed = (self._as_pixels(pixels, geometry), geometry)
        self._images_dirty = True
        self._initialized = False

    def SetMovingImage(self, pixels: np.ndarray, geometry: ImageGeometry):
        self._moving = (self._as_pixels(pixels, geometry), geometry)
        self._gradient = None
        self._images_dirty = True
        self._initialized = False

    def SetMovingImageGradient(self, gradient: np.ndarray):
        """Use a precomputed moving-image gradient (dim doubles per voxel)."""
        self._gradient = np.ascontiguousarray(gradient, dtype=np.float64)
        self._gradient_dirty = True
        self._initialized = False

    def SetSmoothingSigmas(self, fixed_sigma=None, moving_sigma=None):
        """Smooth the fixed and moving images on the device at Initialize
        (SmoothingRecursiveGaussianImageFilter, physical sigma per axis or one
        value; None or 0 = no smoothing), as the registration method prepares
        each level.  The unsmoothed uploads stay resident."""
        def norm(s):
            if s is None:
                return None
            a = np.asarray(s, dtype=np.float64).reshape(-1)
            return None if np.all(a == 0) else a
        self._fixed_sigma = norm(fixed_sigma)
        self._moving_sigma = norm(moving_sigma)
        self._initialized = False

    def SetUseMovingImageGradientFilter(self, on: bool):
        self._use_gradient_filter = bool(on)
        self._initialized = False

    def GetUseMovingImageGradientFilter(self) -> bool:
        return self._use_gradient_filter

    _STORAGE = {"f32": N.MMI_F32, "bf16": N.MMI_BF16, "f16": N.MMI_F16}

    def SetImageStorage(self, storage: str):
        """Device storage of the fixed, moving and gradient images: "f32"
        (default), "bf16" or "f16".  16-bit storage halves the gather traffic;
        the metric is then that of the rounded images (no counterpart in the
        reference)."""
        if storage not in self._STORAGE:
            raise ValueError("storage must be one of %s" % sorted(self._STORAGE))
        self._storage = self._STORAGE[storage]
        self._initialized = False

    def GetImageStorage(self) -> str:
        return {v: k for k, v in self._STORAGE.items()}[self._storage]

    def SetNumberOfHistogramBins(self, bins: int):
        self._bins = max(5, int(bins))
        self._initialized = False

    def GetNumberOfHistogramBins(self) -> int:
        return self._bins

    def SetLaunchConfiguration(self, blocks_per_cu: int = -1, rows_walk: int = -1, hist_copies: int = 0):
        """Kernel launch configuration (no counterpart in ITK): workgroups per
        compute unit, rows per walk of the affine row-walk kernels (0: the
        sample-strided kernels) and pass 1's LDS histogram copies; -1 (0 for
        the copies) keeps the library's choice (the tuning table's row for the
        workload, else its defaults)."""
        self._launch = (int(blocks_per_cu), int(rows_walk))
        self._hist_copies = int(hist_copies)
        self._initialized = False

    def GetLaunchConfiguration(self) -> dict:
        """The launch configuration in effect after Initialize, and whether
        the tuning table chose it."""
        v = [ctypes.c_int32() for _ in range(4)]
        N.check(self._L.mmi_get_launch_config(self._ctx, *[ctypes.byref(x) for x in v]))
        return {"blocks_per_cu": v[0].value, "rows_walk": v[1].value, "hist_copies": v[2].value,
                "from_table": bool(v[3].value)}

    def SetVirtualDomain(self, geometry: ImageGeometry):
        self._virtual = geometry
        self._initialized = False

    # -- masks and sparse sampling (itkImageToImageMetricv4.h:87-101,420-484) --
    def SetFixedImageMask(self, mask):
        """An ImageMaskSpatialObject (itk_amd.masks), or None for no mask."""
        self._masks[0] = mask
        self._initialized = False

    def GetFixedImageMask(self):
        return self._masks[0]

    def SetMovingImageMask(self, mask):
        self._masks[1] = mask
        self._initialized = False

    def GetMovingImageMask(self):
        return self._masks[1]

    def SetFixedSampledPointSet(self, points):
        """Fixed-space points (n x dim); used when UseSampledPointSet is on and
        UseVirtualSampledPointSet is off."""
        self._fixed_points = None if points is None else np.array(points, dtype=np.float64)
        self._initialized = False

    def SetVirtualSampledPointSet(self, points):
        self._virtual_points = None if points is None else np.array(points, dtype=np.float64)
        self._initialized = False

    def SetUseSampledPointSet(self, on: bool):
        self._use_sampled_points = bool(on)
        self._initialized = False

    def GetUseSampledPointSet(self) -> bool:
        return self._use_sampled_points

    def SetUseVirtualSampledPointSet(self, on: bool):
        self._use_virtual_points = bool(on)
        self._initialized = False

    def GetUseVirtualSampledPointSet(self) -> bool:
        return self._use_virtual_points

    def UseSampledPointSetOn(self):
        self.SetUseSampledPointSet(True)

    def UseSampledPointSetOff(self):
        self.SetUseSampledPointSet(False)

    def UseVirtualSampledPointSetOn(self):
        self.SetUseVirtualSampledPointSet(True)

    def UseVirtualSampledPointSetOff(self):
        self.SetUseVirtualSampledPointSet(False)

    # -- settings of the reference class with no effect on the device -------
    def SetMaximumNumberOfWorkUnits(self, n: int):
        """Kept for interface compatibility: the evaluation's partition is the
        device's (the results do not depend on it)."""
        self._work_units = int(n)

    def GetMaximumNumberOfWorkUnits(self) -> int:
        return getattr(self, "_work_units", 0)

    def SetUseFixedImageGradientFilter(self, on: bool):
        """Mattes MI uses no fixed-image gradient; kept for compatibility."""
        self._use_fixed_gradient_filter = bool(on)

    def GetUseFixedImageGradientFilter(self) -> bool:
        return getattr(self, "_use_fixed_gradient_filter", True)

    def GetMetricCategory(self) -> str:
        """ObjectToObjectMetricBaseTemplateEnums::MetricCategory (itkImageToImageMetricv4.h:662-666)."""
        return "IMAGE_METRIC"

    def GetFixedTransform(self):
        return self._fixed_tf

    def GetVirtualSampledPointSet(self):
        return self._sample_points

    def GetNumberOfSkippedFixedSampledPoints(self) -> int:
        return self._skipped_fixed_points

    def GetNumberOfDomainPoints(self) -> int:
        """Samples per evaluation: the point set's size or the virtual voxels."""
        if self._sample_points is not None:
            return int(self._sample_points.shape[0])
        return (self._virtual or self._fixed[1]).num_pixels

    def _sampled_virtual_points(self):
        """The virtual point set of ImageToImageMetricv4::Initialize
        (itkImageToImageMetricv4.hxx:124-129): the user's virtual points, or
        the fixed points mapped by the inverse fixed transform and kept when
        their rounded virtual index is inside the virtual domain
        (MapFixedSampledPointSetToVirtual, itkImageToImageMetricv4.hxx:567-611).
        Fixed points are taken with the identity fixed transform only."""
        d = self._fixed[1].dim
        self._skipped_fixed_points = 0
        if self._use_virtual_points:
            if self._virtual_points is None:
                raise N.MetricError(N.MMI_ERR_ARGUMENT, "VirtualSampledPointSet is not present")
            return np.ascontiguousarray(self._virtual_points.reshape(-1, d))
        if self._fixed_points is None:
            raise N.MetricError(N.MMI_ERR_ARGUMENT, "Fixed Sample point set is empty.")
        ftf = self._fixed_tf
        if ftf is not None and not (ftf.kind == TF_AFFINE and np.array_equal(ftf.matrix, np.eye(d))
                                    and not np.any(ftf.offset)):
            raise NotImplementedError("fixed sampled point sets need the identity fixed transform")
        pts = self._fixed_points.reshape(-1, d)
        virt = self._virtual or self._fixed[1]
        _, keep = virt.physical_points_to_index(pts)
        # the reference takes the fixed range over every fixed point, the
        # device over the kept ones: they agree unless a dropped point lies
        # inside the fixed image
        _, in_fixed = self._fixed[1].physical_points_to_index(pts[~keep])
        if in_fixed.any():
            raise NotImplementedError("fixed sampled points outside the virtual domain but inside the fixed image")
        self._skipped_fixed_points = int((~keep).sum())
        if not keep.any():
            raise N.MetricError(N.MMI_ERR_ARGUMENT,
                                "The virtual sampled point set has zero points because no fixed sampled points were "
                                "within the virtual domain after mapping. There are no points to evaluate.")
        return np.ascontiguousarray(pts[keep])

    def SetFixedTransform(self, tf):
        self._fixed_tf = tf
        self._initialized = False

    def SetMovingTransform(self, tf):
        self._moving_tf = tf
        self._initialized = False

    def GetMovingTransform(self):
        return self._moving_tf

    @staticmethod
    def _as_pixels(pixels, geometry):
        """The pixels in a dtype of the C ABI: float32, float64 or an integer
        type (integer images keep the reference's truncation of interpolated
        values); other dtypes are taken as float32 images."""
        a = np.ascontiguousarray(pixels)
        if a.shape != geometry.shape:
            raise ValueError("pixel array shape %s does not match geometry %s" % (a.shape, geometry.shape))
        if a.dtype.type not in _PIXEL_CODES:
            a = a.astype(np.float32)
        return a

    # -- set-up -----------------------------------------------------------
    def Initialize(self):
        if self._fixed is None:
            raise N.MetricError(N.MMI_ERR_ARGUMENT, "FixedImage is not present")
        if self._moving is None:
            raise N.MetricError(N.MMI_ERR_ARGUMENT, "MovingImage is not present")
        if self._moving_tf is None:
            raise N.MetricError(N.MMI_ERR_ARGUMENT, "MovingTransform is not present")
        L, ctx = self._L, self._ctx
        N.check(L.mmi_set_storage(ctx, self._storage))
        if self._images_dirty:
            for setter, (px, g) in ((L.mmi_set_fixed_image, self._fixed), (L.mmi_set_moving_image, self._moving)):
                cg = g.to_c()
                N.check(setter(ctx, ctypes.byref(cg), _PIXEL_CODES[px.dtype.type], px.ctypes.data))
            self._images_dirty = False
            self._gradient_dirty = True
        N.check(L.mmi_set_gradient_source(
            ctx, N.MMI_GRADIENT_IMAGE if self._use_gradient_filter else N.MMI_GRADIENT_CENTRAL_DIFFERENCE))
        if self._use_gradient_filter and self._gradient is not None and self._gradient_dirty:
            N.check(L.mmi_set_moving_gradient(ctx, _dptr(self._gradient)))
            self._gradient_dirty = False
        N.check(L.mmi_set_number_of_histogram_bins(ctx, self._bins))
        N.check(L.mmi_set_launch_config(ctx, *self._launch))
        N.check(L.mmi_set_histogram_copies(ctx, self._hist_copies))
        d = self._fixed[1].dim
        fs = None if self._fixed_sigma is None else np.ascontiguousarray(np.broadcast_to(self._fixed_sigma, (d,)))
        ms = None if self._moving_sigma is None else np.ascontiguousarray(np.broadcast_to(self._moving_sigma, (d,)))
        N.check(L.mmi_set_smoothing_sigmas(ctx, _dptr(fs) if fs is not None else None,
                                           _dptr(ms) if ms is not None else None))
        cv = (self._virtual or self._fixed[1]).to_c()
        N.check(L.mmi_set_virtual_domain(ctx, ctypes.byref(cv)))
        for setter, mask in ((L.mmi_set_fixed_mask, self._masks[0]), (L.mmi_set_moving_mask, self._masks[1])):
            if mask is None:
                N.check(setter(ctx, None, None))
            else:
                flags = np.ascontiguousarray(mask.inside_flags())
                mg = mask.geometry.to_c()
                N.check(setter(ctx, ctypes.byref(mg), flags.ctypes.data))
        if self._use_sampled_points:
            self._sample_points = self._sampled_virtual_points()
            if self._sample_points.shape[0] == 0:
                raise N.MetricError(N.MMI_ERR_DEGENERATE, "the sampled point set has no points")
            N.check(L.mmi_set_virtual_points(ctx, d, _dptr(self._sample_points), self._sample_points.shape[0]))
        else:
            self._sample_points = None
            N.check(L.mmi_set_virtual_points(ctx, d, None, 0))
        if self._fixed_tf is not None:
            self._push_transform(L.mmi_set_fixed_transform, self._fixed_tf)
        self._push_moving_transform(self._moving_tf)
        N.check(L.mmi_initialize(ctx))
        self._initialized = True

    def _push_moving_transform(self, tf):
        """The moving transform; a composite as its optimized entry plus the
        earlier entries as matrix-offset stages."""
        if tf.kind == TF_COMPOSITE:
            m, o = tf.initial_stages()
            m = np.ascontiguousarray(m, dtype=np.float64)
            o = np.ascontiguousarray(o, dtype=np.float64)
            N.check(self._L.mmi_set_moving_initial_transforms(self._ctx, m.shape[0], _dptr(m), _dptr(o)))
            tf = tf.optimized
        else:
            N.check(self._L.mmi_set_moving_initial_transforms(self._ctx, 0, None, None))
        self._push_transform(self._L.mmi_set_moving_transform, tf)

    def _push_transform(self, setter, tf):
        if tf.kind == TF_COMPOSITE:
            raise ValueError("the fixed transform cannot be a composite")
        fp = transform_fixed_parameters(tf)
        p = np.ascontiguousarray(tf.get_parameters(), dtype=np.float64)
        N.check(setter(self._ctx, tf.kind, _dptr(fp), fp.size, _dptr(p), p.size))

    # -- ObjectToObjectMetricBaseTemplate interface -------------------------
    def GetNumberOfParameters(self) -> int:
        return self._moving_tf.number_of_parameters

    def GetNumberOfLocalParameters(self) -> int:
        return self._moving_tf.number_of_parameters

    def HasLocalSupport(self) -> bool:
        return False

    def GetParameters(self) -> np.ndarray:
        return self._moving_tf.get_parameters()

    def SetParameters(self, p):
        self._moving_tf.set_parameters(p)

    def UpdateTransformParameters(self, derivative, factor: float = 1.0):
        self._moving_tf.update_transform_parameters(derivative, factor)

    def _push_parameters(self) -> np.ndarray:
        p = np.ascontiguousarray(self._moving_tf.get_parameters(), dtype=np.float64)
        N.check(self._L.mmi_set_moving_parameters(self._ctx, p.ctypes.data, p.size))
        return p

    def _evaluate(self, want_derivative: bool, want_pdf: bool = False):
        if not self._initialized:
            raise N.MetricError(N.MMI_ERR_ARGUMENT, "Initialize() must be called first")
        p = self._push_parameters()
        out = self._out  # value and valid count, reused across calls
        out[:] = 0.0
        deriv = np.zeros(p.size) if want_derivative else None
        pdf = np.zeros(self._bins * self._bins) if want_pdf else None
        rc = self._L.mmi_evaluate(self._ctx, 1 if want_derivative else 0, self._out_addr,
                                  deriv.ctypes.data if deriv is not None else None, self._out_addr + 8,
                                  pdf.ctypes.data if pdf is not None else None, None)
        value = float(out[0])
        self._n_valid = int(out[1:2].view(np.int64)[0])
        N.check(rc)
        self._value = value
        if pdf is not None:
            self._last_pdf = pdf.reshape(self._bins, self._bins)
        return value, deriv

    def GetValue(self) -> float:
        return self._evaluate(False)[0]

    def GetValueAndDerivative(self):
        return self._evaluate(True)

    def GetDerivative(self) -> np.ndarray:
        return self._evaluate(True)[1]

    def GetCurrentValue(self) -> float:
        return self._value

    def GetNumberOfValidPoints(self) -> int:
        return self._n_valid

    def GetJointPDF(self) -> np.ndarray:
        """Normalised joint PDF of a fresh evaluation (row = fixed bin)."""
        self._evaluate(False, want_pdf=True)
        return self._last_pdf

    # -- diagnostics --------------------------------------------------------
    def GetSampleBins(self):
        n = self.GetNumberOfDomainPoints()
        fb = np.zeros(n, dtype=np.int16)
        mb = np.zeros(n, dtype=np.int16)
        N.check(self._L.mmi_get_sample_bins(self._ctx, fb.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                            mb.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))))
        return fb, mb

    def GetMovingTexels(self) -> np.ndarray:
        """The device's FP32 moving texels, shape moving.shape + (4,): value
        channel (the level's moving image) and physical gradient."""
        shape = self._moving[0].shape
        out = np.zeros(tuple(shape) + (4,), dtype=np.float32)
        N.check(self._L.mmi_get_moving_texels(self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def GetHistogramParameters(self) -> dict:
        vals = [ctypes.c_double() for _ in range(6)]
        N.check(self._L.mmi_get_histogram_parameters(self._ctx, *[ctypes.byref(v) for v in vals]))
        keys = ("fixed_min", "fixed_max", "moving_min", "moving_max", "fixed_bin_size", "moving_bin_size")
        return {k: v.value for k, v in zip(keys, vals)}

    def GetExceptionCount(self) -> tuple:
        """(samples the last fast-path evaluation sent to the exact path, or -1
        off the fast path; samples per evaluation)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        N.check(self._L.mmi_get_exception_count(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def GetLastEvaluationPath(self) -> str:
        """"exact", "fast" or "fast_overflow_exact" for the last evaluation."""
        v = ctypes.c_int32()
        N.check(self._L.mmi_get_last_path(self._ctx, ctypes.byref(v)))
        return ("exact", "fast", "fast_overflow_exact")[v.value]

    def GetLastTiming(self) -> dict:
        t = [ctypes.c_float() for _ in range(3)]
        N.check(self._L.mmi_get_last_timing(self._ctx, *[ctypes.byref(v) for v in t]))
        # pass1_ms / pass2_ms are -1 unless the metric was made with time_passes
        return {"total_ms": t[0].value, "pass1_ms": t[1].value, "pass2_ms": t[2].value}
