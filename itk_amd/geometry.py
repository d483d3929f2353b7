"""Image geometry with ITK's conventions.

An image is a buffer with x fastest (ITK's ``Image`` memory order), a
buffered-region start index, an origin, a spacing and a direction matrix.

* ``index_to_physical = direction @ diag(spacing)``
  (``Modules/Core/Common/include/itkImageBase.hxx:165-177``)
* ``physical_to_index`` is its inverse.  ITK inverts with an SVD
  pseudo-inverse (``itkMatrix.h:330-337``); for a diagonal matrix that is
  exactly ``diag(1/spacing)`` up to the axis permutation, which is what this
  module returns for axis-aligned directions.  For rotated directions it uses
  ``numpy.linalg.inv``; a caller that needs bit-identical continuous indices
  to a running ITK process passes ITK's own matrix instead.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np


class CGeometry(ctypes.Structure):
    """C layout shared by the C-ABI (``include/itk_mattes_hip.h``) and the oracle."""

    _fields_ = [
        ("dim", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("size", ctypes.c_int64 * 3),
        ("start", ctypes.c_int64 * 3),
        ("origin", ctypes.c_double * 3),
        ("spacing", ctypes.c_double * 3),
        ("direction", ctypes.c_double * 9),
        ("index_to_physical", ctypes.c_double * 9),
        ("physical_to_index", ctypes.c_double * 9),
    ]


def _index_to_physical(direction: np.ndarray, spacing: np.ndarray) -> np.ndarray:
    d = direction.shape[0]
    out = np.zeros((d, d))
    for i in range(d):
        for j in range(d):
            s = 0.0
            for k in range(d):
                s += direction[i, k] * (spacing[j] if k == j else 0.0)
            out[i, j] = s
    return out


def _invert(m: np.ndarray) -> np.ndarray:
    d = m.shape[0]
    nz = np.abs(m) > 0
    if (nz.sum(axis=0) == 1).all() and (nz.sum(axis=1) == 1).all():
        inv = np.zeros_like(m)
        for i in range(d):
            j = int(np.nonzero(nz[i])[0][0])
            inv[j, i] = 1.0 / m[i, j]
        return inv
    return np.linalg.inv(m)


@dataclass
class ImageGeometry:
    size: tuple
    origin: tuple = None
    spacing: tuple = None
    direction: np.ndarray = None
    start: tuple = None
    physical_to_index: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        d = len(self.size)
        if d not in (2, 3):
            raise ValueError("only 2-D and 3-D images are supported")
        self.size = tuple(int(s) for s in self.size)
        self.origin = tuple(float(o) for o in (self.origin if self.origin is not None else (0.0,) * d))
        self.spacing = tuple(float(s) for s in (self.spacing if self.spacing is not None else (1.0,) * d))
        self.start = tuple(int(s) for s in (self.start if self.start is not None else (0,) * d))
        self.direction = np.eye(d) if self.direction is None else np.asarray(self.direction, dtype=np.float64)
        self.index_to_physical = _index_to_physical(self.direction, np.asarray(self.spacing))
        if self.physical_to_index is None:
            self.physical_to_index = _invert(self.index_to_physical)
        else:
            self.physical_to_index = np.asarray(self.physical_to_index, dtype=np.float64)

    @property
    def dim(self) -> int:
        return len(self.size)

    @property
    def num_pixels(self) -> int:
        return int(np.prod(self.size))

    @property
    def shape(self) -> tuple:
        """numpy shape of the buffer (slowest axis first)."""
        return tuple(reversed(self.size))

    def to_c(self) -> CGeometry:
        g = CGeometry()
        d = self.dim
        g.dim = d
        for i in range(d):
            g.size[i] = self.size[i]
            g.start[i] = self.start[i]
            g.origin[i] = self.origin[i]
            g.spacing[i] = self.spacing[i]
        for i in range(d):
            for j in range(d):
                g.direction[i * d + j] = self.direction[i, j]
                g.index_to_physical[i * d + j] = self.index_to_physical[i, j]
                g.physical_to_index[i * d + j] = self.physical_to_index[i, j]
        for i in range(d, 3):
            g.size[i] = 1
        return g

    def index_to_point(self, idx) -> np.ndarray:
        """ImageBase::TransformIndexToPhysicalPoint summation order."""
        d = self.dim
        p = np.zeros(d)
        for i in range(d):
            s = self.origin[i]
            for j in range(d):
                s += self.index_to_physical[i, j] * float(idx[j])
            p[i] = s
        return p

    def physical_points_to_index(self, points) -> tuple:
        """ImageBase::TransformPhysicalPointToIndex (itkImageBase.h:465-479)
        for an (n, dim) array: per row, sum += M[i][j] * (p[j] - origin[j])
        from zero, then Math::RoundHalfIntegerUp as built on x86-64
        (round-to-even of 2x + 0.5, halved; itkMathDetail.h:382-385).
        Returns (index (n, dim) int64, inside (n,) bool); an index too large
        for int64 counts as outside."""
        pts = np.asarray(points, dtype=np.float64).reshape(-1, self.dim)
        d = self.dim
        idx = np.zeros((pts.shape[0], d), dtype=np.int64)
        inside = np.ones(pts.shape[0], dtype=bool)
        with np.errstate(invalid="ignore"):
            for i in range(d):
                s = np.zeros(pts.shape[0])
                for j in range(d):
                    s = s + self.physical_to_index[i, j] * (pts[:, j] - self.origin[j])
                ok = np.abs(s) < 1.0e18
                r = np.where(ok, np.rint(2.0 * np.where(ok, s, 0.0) + 0.5), 0.0).astype(np.int64) >> 1
                idx[:, i] = r
                inside &= ok & (r >= self.start[i]) & (r < self.start[i] + self.size[i])
        return idx, inside

    def center_point(self) -> np.ndarray:
        """Physical centre of the buffer (the usual affine centre)."""
        return self.index_to_point([(s - 1) / 2.0 for s in self.size])

    def continuous_index_to_point(self, ci) -> np.ndarray:
        """ImageBase::TransformContinuousIndexToPhysicalPoint (itkImageBase.h:560-572):
        the matrix product first, then the origin."""
        d = self.dim
        p = np.zeros(d)
        for r in range(d):
            s = 0.0
            for c in range(d):
                s += self.index_to_physical[r, c] * float(ci[c])
            p[r] = s + self.origin[r]
        return p

    def shrink(self, factors) -> "ImageGeometry":
        """Output geometry of ShrinkImageFilter (itkShrinkImageFilter.hxx:236-298):
        spacing times the factor, size rounded down (at least 1), start rounded
        up, and the origin moved so the physical centres coincide."""
        d = self.dim
        f = [int(x) for x in (factors if not np.isscalar(factors) else [factors] * d)]
        if any(x < 1 for x in f):
            raise ValueError("shrink factors must be >= 1")
        spacing = [self.spacing[i] * float(f[i]) for i in range(d)]
        size = [max(1, int(math.floor(float(self.size[i]) / float(f[i])))) for i in range(d)]
        start = [int(math.ceil(float(self.start[i]) / float(f[i]))) for i in range(d)]
        in_center = [self.start[i] + (self.size[i] - 1) / 2.0 for i in range(d)]
        out_center = [start[i] + (size[i] - 1) / 2.0 for i in range(d)]
        provisional = ImageGeometry(size=size, origin=self.origin, spacing=spacing, direction=self.direction,
                                    start=start)
        a = self.continuous_index_to_point(in_center)
        b = provisional.continuous_index_to_point(out_center)
        origin = [self.origin[i] + (a[i] - b[i]) for i in range(d)]
        return ImageGeometry(size=size, origin=origin, spacing=spacing, direction=self.direction, start=start)
