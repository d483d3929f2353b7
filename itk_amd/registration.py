"""ImageRegistrationMethodv4 for one image metric on the device.

Per level (Modules/Registration/RegistrationMethodsv4/include/itkImageRegistrationMethodv4.hxx:436-666):
the virtual domain is the fixed image's geometry shrunk by that level's
factors (ShrinkImageFilter output information, itkShrinkImageFilter.hxx:236-298),
the fixed and moving images are smoothed with that level's sigma
(SmoothingRecursiveGaussianImageFilter; sigmas in voxels are multiplied by
each image's spacing when SmoothingSigmasAreSpecifiedInPhysicalUnits is off),
the metric is re-initialised and the optimizer runs from the transform's
current parameters.  Images stay resident on the device across levels: only
the smoothing, gradient image and intensity ranges are recomputed there.

With a REGULAR or RANDOM metric sampling strategy each level draws its sample
points in the level's virtual domain (SetMetricSamplePoints,
itkImageRegistrationMethodv4.hxx:699-702,867-1024) and hands them to the
metric as its virtual sampled point set; the seed restarts from the method's
random seed at every Update (:791-794).
"""
from __future__ import annotations

import time

import numpy as np

from . import sampling
from .geometry import ImageGeometry


class ImageRegistrationMethodv4:
    def __init__(self):
        self.fixed = None
        self.moving = None
        self.metric = None
        self.optimizer = None
        self.transform = None
        self.virtual = None
        self.shrink_factors = [2, 1, 1]   # reference defaults (itkImageRegistrationMethodv4.hxx:93-104)
        self.smoothing_sigmas = [2.0, 1.0, 0.0]
        self.sigmas_in_physical_units = True
        self.level_log = []
        self.observers = []
        self.fixed_mask = None
        self.moving_mask = None
        self.sampling_strategy = sampling.NONE
        self.sampling_percentages = [1.0] * len(self.shrink_factors)
        # the reference seeds from the clock by default (GetNextSeed); so does
        # this method, with its own clock hash
        self.reseed = False
        self.random_seed = int(time.time_ns() & 0x7FFFFFFF)
        self.current_random_seed = self.random_seed

    # -- inputs ---------------------------------------------------------------
    def SetFixedImage(self, pixels: np.ndarray, geometry: ImageGeometry):
        self.fixed = (pixels, geometry)

    def SetMovingImage(self, pixels: np.ndarray, geometry: ImageGeometry):
        self.moving = (pixels, geometry)

    def SetMetric(self, metric):
        self.metric = metric

    def SetOptimizer(self, optimizer):
        self.optimizer = optimizer

    def SetInitialTransform(self, tf):
        """The transform optimised in place (InPlace on, the reference default)."""
        self.transform = tf

    def SetVirtualDomain(self, geometry: ImageGeometry):
        self.virtual = geometry

    def SetNumberOfLevels(self, n: int):
        n = int(n)
        self.shrink_factors = (list(self.shrink_factors) + [1] * n)[:n]
        self.smoothing_sigmas = (list(self.smoothing_sigmas) + [0.0] * n)[:n]
        # the reference resets the sampling percentages to 1 when the level
        # count changes (itkImageRegistrationMethodv4.hxx:851-863)
        if len(self.sampling_percentages) != n:
            self.sampling_percentages = [1.0] * n

    def SetFixedImageMask(self, mask):
        """ImageMaskSpatialObject (itk_amd.masks) given to the metric at every
        level, and used to drop sample points."""
        self.fixed_mask = mask

    def SetMovingImageMask(self, mask):
        self.moving_mask = mask

    def SetMetricSamplingStrategy(self, strategy: int):
        """sampling.NONE, sampling.REGULAR or sampling.RANDOM."""
        if strategy not in (sampling.NONE, sampling.REGULAR, sampling.RANDOM):
            raise ValueError("Invalid sampling strategy requested.")
        self.sampling_strategy = strategy

    def GetMetricSamplingStrategy(self) -> int:
        return self.sampling_strategy

    def SetMetricSamplingPercentage(self, p: float):
        self.SetMetricSamplingPercentagePerLevel([p] * self.GetNumberOfLevels())

    def SetMetricSamplingPercentagePerLevel(self, ps):
        ps = [float(p) for p in ps]
        for p in ps:
            if p <= 0.0 or p > 1.0:
                raise ValueError("sampling percentage outside of (0,1]")
        self.sampling_percentages = ps

    def GetMetricSamplingPercentagePerLevel(self):
        return list(self.sampling_percentages)

    def MetricSamplingReinitializeSeed(self, seed: int = None):
        """No argument: a fresh clock seed for every level's streams; with a
        seed: reproducible streams from that seed (itkImageRegistrationMethodv4.hxx:1225-1250)."""
        if seed is None:
            self.reseed = True
        else:
            self.reseed = False
            self.random_seed = int(seed)
            self.current_random_seed = int(seed)

    def _sample_points(self, level: int, virtual: ImageGeometry):
        if self.reseed:
            seed = int(time.time_ns() & 0xFFFFFFFF)
        else:
            seed = self.current_random_seed
        if level >= len(self.sampling_percentages):
            raise ValueError("one sampling percentage per level is required")
        pts, nxt = sampling.metric_sample_points(virtual, self.sampling_strategy, self.sampling_percentages[level],
                                                 seed, fixed_mask=self.fixed_mask)
        if not self.reseed:
            self.current_random_seed += nxt - (seed & 0xFFFFFFFF)
        return pts

    def GetNumberOfLevels(self) -> int:
        return len(self.shrink_factors)

    def SetShrinkFactorsPerLevel(self, factors):
        """One factor per level (all axes) or one list of per-axis factors per level."""
        self.shrink_factors = list(factors)

    def SetSmoothingSigmasPerLevel(self, sigmas):
        self.smoothing_sigmas = [float(s) for s in sigmas]

    def SetSmoothingSigmasAreSpecifiedInPhysicalUnits(self, on: bool):
        self.sigmas_in_physical_units = bool(on)

    def AddObserver(self, fn):
        """fn(method, level) is called at the start of each level."""
        self.observers.append(fn)

    def GetTransform(self):
        return self.transform

    # -- run ------------------------------------------------------------------
    def _level_sigmas(self, sigma: float, geom: ImageGeometry):
        if sigma <= 0.0:
            return None
        s = np.full(geom.dim, float(sigma))
        if not self.sigmas_in_physical_units:
            s = s * np.asarray(geom.spacing)
        return s

    def Update(self):
        if self.fixed is None or self.moving is None:
            raise ValueError("fixed and moving images are required")
        if self.metric is None or self.optimizer is None or self.transform is None:
            raise ValueError("metric, optimizer and initial transform are required")
        if len(self.shrink_factors) != len(self.smoothing_sigmas):
            raise ValueError("shrink factors and smoothing sigmas must have one entry per level")
        fixed_px, fixed_geom = self.fixed
        moving_px, moving_geom = self.moving
        base = self.virtual or fixed_geom
        m = self.metric
        m.SetFixedImage(fixed_px, fixed_geom)
        m.SetMovingImage(moving_px, moving_geom)
        m.SetMovingTransform(self.transform)
        self.level_log = []
        self.current_random_seed = self.random_seed  # the same streams on every Update
        if self.fixed_mask is not None or self.moving_mask is not None or hasattr(m, "SetFixedImageMask"):
            m.SetFixedImageMask(self.fixed_mask)
            m.SetMovingImageMask(self.moving_mask)
        for level, (factor, sigma) in enumerate(zip(self.shrink_factors, self.smoothing_sigmas)):
            for fn in self.observers:
                fn(self, level)
            t0 = time.perf_counter()
            factors = [int(factor)] * base.dim if np.isscalar(factor) else [int(f) for f in factor]
            virtual = base.shrink(factors)
            m.SetVirtualDomain(virtual)
            m.SetSmoothingSigmas(self._level_sigmas(sigma, fixed_geom), self._level_sigmas(sigma, moving_geom))
            if self.sampling_strategy != sampling.NONE:
                pts = self._sample_points(level, virtual)
                m.SetVirtualSampledPointSet(pts)
                m.SetUseSampledPointSet(True)
                m.SetUseVirtualSampledPointSet(True)
            m.Initialize()
            est = getattr(self.optimizer, "scales_estimator", None)
            if est is not None:
                est.SetTransform(self.transform)
                est.SetVirtualDomain(virtual)
            self.optimizer.SetMetric(m)
            self.optimizer.StartOptimization()
            self.level_log.append({
                "level": level, "shrink": factors, "sigma": sigma, "virtual_size": list(virtual.size),
                "samples": m.GetNumberOfDomainPoints() if hasattr(m, "GetNumberOfDomainPoints") else None,
                "iterations": self.optimizer.GetCurrentIteration(),
                "value": self.optimizer.GetCurrentMetricValue(),
                "stop": self.optimizer.GetStopConditionDescription(),
                "seconds": time.perf_counter() - t0,
            })
        return self.transform
