"""The oracle behind the metric interface the registration method drives.

Test infrastructure: runs the CPU restatement of the reference (smoothing,
gradient filter and dense Mattes MI in FP64, float32 images like the
reference's) so a whole registration can be run once on the device metric and
once on the oracle, and the final transforms compared.
"""
from __future__ import annotations

import numpy as np


class OracleMattesMetric:
    def __init__(self, oracle_mod, bins: int = 50, threads: int = 8):
        self.o = oracle_mod
        self.bins = bins
        self.threads = threads
        self.fixed = self.moving = None
        self.virtual = None
        self.tf = None
        self.fixed_sigma = self.moving_sigma = None
        self.value = float(np.finfo(np.float64).max)
        self.points = None
        self.use_points = False
        self.masks = [None, None]

    def SetFixedImage(self, px, geom):
        self.fixed = (np.asarray(px, dtype=np.float32), geom)

    def SetMovingImage(self, px, geom):
        self.moving = (np.asarray(px, dtype=np.float32), geom)

    def SetMovingTransform(self, tf):
        self.tf = tf

    def SetVirtualDomain(self, geom):
        self.virtual = geom

    def SetNumberOfHistogramBins(self, b):
        self.bins = max(5, int(b))

    def SetVirtualSampledPointSet(self, pts):
        self.points = None if pts is None else np.asarray(pts, dtype=np.float64)

    def SetUseSampledPointSet(self, on):
        self.use_points = bool(on)

    def SetUseVirtualSampledPointSet(self, on):
        if not on:
            raise NotImplementedError("fixed-space point sets")

    def SetFixedImageMask(self, mask):
        self.masks[0] = mask

    def SetMovingImageMask(self, mask):
        self.masks[1] = mask

    @staticmethod
    def _mask(so):
        return None if so is None else (so.inside_flags(), so.geometry)

    def SetSmoothingSigmas(self, fixed_sigma=None, moving_sigma=None):
        self.fixed_sigma, self.moving_sigma = fixed_sigma, moving_sigma

    def Initialize(self):
        (f, fg), (m, mg) = self.fixed, self.moving
        if self.fixed_sigma is not None and np.any(np.asarray(self.fixed_sigma) > 0):
            f = self.o.smooth(f, fg, self.fixed_sigma, threads=self.threads)
        if self.moving_sigma is not None and np.any(np.asarray(self.moving_sigma) > 0):
            m = self.o.smooth(m, mg, self.moving_sigma, threads=self.threads)
        self.level_fixed, self.level_moving = f, m
        self.gradient = self.o.gradient_image(m, mg, threads=self.threads)

    def GetValueAndDerivative(self):
        (_, fg), (_, mg) = self.fixed, self.moving
        r = self.o.mattes(self.level_fixed, fg, self.level_moving, mg, self.tf, bins=self.bins,
                          virtual_geom=self.virtual, moving_gradient=self.gradient, threads=self.threads,
                          points=self.points if self.use_points else None,
                          fixed_mask=self._mask(self.masks[0]), moving_mask=self._mask(self.masks[1]))
        self.value = r["value"]
        return r["value"], r["derivative"]

    def GetNumberOfParameters(self):
        return self.tf.number_of_parameters

    def GetNumberOfLocalParameters(self):
        return self.tf.number_of_parameters

    def HasLocalSupport(self):
        return False

    def GetParameters(self):
        return self.tf.get_parameters()

    def SetParameters(self, p):
        self.tf.set_parameters(p)

    def UpdateTransformParameters(self, d, factor=1.0):
        self.tf.update_transform_parameters(d, factor)


def mean_displacement(tf_a, tf_b, geom, stride: int = 1) -> float:
    """Mean distance, in voxels of geom, between the two transforms' images of
    the voxel centres of geom (every stride-th voxel per axis)."""
    import itertools

    ranges = [range(0, s, stride) for s in geom.size]
    total = 0.0
    count = 0
    sp = np.asarray(geom.spacing)
    for idx in itertools.product(*ranges):
        p = geom.index_to_point(idx)
        a = np.asarray(tf_a.transform_point(p))
        b = np.asarray(tf_b.transform_point(p))
        total += float(np.sqrt(np.sum(((a - b) / sp) ** 2)))
        count += 1
    return total / count
