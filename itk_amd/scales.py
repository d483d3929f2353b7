"""RegistrationParameterScalesFromPhysicalShift for global linear transforms.

Follows Modules/Numerics/Optimizersv4/include/itkRegistrationParameterScalesFromShiftBase.hxx:31-107
(EstimateScales), :116-143 (EstimateStepScale) and
itkRegistrationParameterScalesFromPhysicalShift.hxx:40-90 (sample shifts), with
the corner sampling the reference chooses for general affine transforms
(itkRegistrationParameterScalesEstimator.hxx:348-366, 512-535).
"""
from __future__ import annotations

import numpy as np

from .geometry import ImageGeometry
from .transforms import AffineTransform


class RegistrationParameterScalesFromPhysicalShift:
    def __init__(self, small_parameter_variation: float = 0.01):
        self.small_parameter_variation = float(small_parameter_variation)
        self.transform = None
        self.virtual = None

    def SetTransform(self, tf):
        self.transform = tf

    def SetVirtualDomain(self, geometry: ImageGeometry):
        self.virtual = geometry

    def SetSmallParameterVariation(self, v: float):
        self.small_parameter_variation = float(v)

    def _check(self):
        if not isinstance(self.transform, AffineTransform):
            raise ValueError("physical-shift scales are implemented for affine transforms (corner sampling)")
        if self.virtual is None:
            raise ValueError("virtual domain is not set")

    def _corners(self):
        g = self.virtual
        d = g.dim
        pts = []
        for i in range(1 << d):
            idx = [g.start[k] + (((i >> k) & 1) * (g.size[k] - 1)) for k in range(d)]
            pts.append(g.index_to_point(idx))
        return pts

    def _max_shift(self, delta) -> float:
        tf = self.transform
        old = tf.get_parameters()
        pts = self._corners()
        before = [tf.transform_point(p) for p in pts]
        tf.update_transform_parameters(delta, 1.0)
        after = [tf.transform_point(p) for p in pts]
        tf.set_parameters(old)
        m = 0.0
        for a, b in zip(after, before):
            m = max(m, float(np.sqrt(np.sum((np.asarray(a) - np.asarray(b)) ** 2))))
        return m

    def EstimateScales(self) -> np.ndarray:
        self._check()
        P = self.transform.number_of_parameters
        eps = np.finfo(np.float64).eps
        scales = np.zeros(P)
        min_nonzero = np.finfo(np.float64).max
        for i in range(P):
            delta = np.zeros(P)
            delta[i] = self.small_parameter_variation
            s = self._max_shift(delta)
            scales[i] = s
            if eps < s < min_nonzero:
                min_nonzero = s
        if min_nonzero == np.finfo(np.float64).max:
            return np.ones(P)
        for i in range(P):
            if scales[i] <= eps:
                scales[i] = min_nonzero * min_nonzero
            else:
                scales[i] *= scales[i]
            scales[i] *= 1.0 / (self.small_parameter_variation * self.small_parameter_variation)
        return scales

    def EstimateStepScale(self, step) -> float:
        self._check()
        step = np.asarray(step, dtype=np.float64)
        max_step = float(np.max(np.abs(step))) if step.size else 0.0
        if max_step <= np.finfo(np.float64).eps:
            return 0.0
        factor = self.small_parameter_variation / max_step
        return self._max_shift(step * factor) / factor
