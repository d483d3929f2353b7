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
"""
from __future__ import annotations

import time

import numpy as np

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
        for level, (factor, sigma) in enumerate(zip(self.shrink_factors, self.smoothing_sigmas)):
            for fn in self.observers:
                fn(self, level)
            t0 = time.perf_counter()
            factors = [int(factor)] * base.dim if np.isscalar(factor) else [int(f) for f in factor]
            virtual = base.shrink(factors)
            m.SetVirtualDomain(virtual)
            m.SetSmoothingSigmas(self._level_sigmas(sigma, fixed_geom), self._level_sigmas(sigma, moving_geom))
            m.Initialize()
            est = getattr(self.optimizer, "scales_estimator", None)
            if est is not None:
                est.SetTransform(self.transform)
                est.SetVirtualDomain(virtual)
            self.optimizer.SetMetric(m)
            self.optimizer.StartOptimization()
            self.level_log.append({
                "level": level, "shrink": factors, "sigma": sigma, "virtual_size": list(virtual.size),
                "iterations": self.optimizer.GetCurrentIteration(),
                "value": self.optimizer.GetCurrentMetricValue(),
                "stop": self.optimizer.GetStopConditionDescription(),
                "seconds": time.perf_counter() - t0,
            })
        return self.transform
