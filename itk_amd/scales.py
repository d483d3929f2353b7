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

    def _corners(self) -> np.ndarray:
        """The virtual domain's 2^D corner points, one per row."""
        g = self.virtual
        d = g.dim
        pts = []
        for i in range(1 << d):
            idx = [g.start[k] + (((i >> k) & 1) * (g.size[k] - 1)) for k in range(d)]
            pts.append(g.index_to_point(idx))
        return np.array(pts)

    @staticmethod
    def _map(tf, pts: np.ndarray) -> np.ndarray:
        """TransformPoint of every row: y[r] = (sum_c M[r][c] x[c]) + offset[r],
        summed in c order (element for element the transform's own arithmetic)."""
        m = tf.matrix
        off = tf.offset
        d = pts.shape[1]
        y = np.empty_like(pts)
        for r in range(d):
            s = np.zeros(pts.shape[0])
            for c in range(d):
                s = s + m[r, c] * pts[:, c]
            y[:, r] = s + off[r]
        return y

    def _max_shift(self, delta, pts=None, before=None) -> float:
        tf = self.transform
        if pts is None:
            pts = self._corners()
        if before is None:
            before = self._map(tf, pts)
        old = tf.get_parameters()
        tf.update_transform_parameters(delta, 1.0)
        after = self._map(tf, pts)
        tf.set_parameters(old)
        e = after - before
        sq = e[:, 0] * e[:, 0]
        for k in range(1, e.shape[1]):
            sq = sq + e[:, k] * e[:, k]
        return float(np.max(np.sqrt(sq))) if sq.size else 0.0

    def EstimateScales(self) -> np.ndarray:
        self._check()
        P = self.transform.number_of_parameters
        eps = np.finfo(np.float64).eps
        scales = np.zeros(P)
        min_nonzero = np.finfo(np.float64).max
        pts = self._corners()
        before = self._map(self.transform, pts)
        for i in range(P):
            delta = np.zeros(P)
            delta[i] = self.small_parameter_variation
            s = self._max_shift(delta, pts, before)
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
