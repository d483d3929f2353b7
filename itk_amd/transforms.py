"""Moving transforms of the registration, with ITK's parameter layouts.

* ``AffineTransform`` follows ``itk::AffineTransform`` /
  ``MatrixOffsetTransformBase``: parameters are the matrix (row-major)
  followed by the translation; the fixed parameters are the centre.  The
  offset is derived as in ``MatrixOffsetTransformBase::ComputeOffset``
  (``Modules/Core/Transform/include/itkMatrixOffsetTransformBase.hxx:615-631``).
* ``CompositeTransform`` follows ``itk::CompositeTransform`` as
  ``ImageRegistrationMethodv4`` sets it up (a moving initial transform, then
  the optimized one; only the most recent entry is optimized).
* ``BSplineTransform`` follows ``itk::BSplineTransform<double, D, 3>``:
  parameters are the coefficient images one dimension after another, each
  x-fastest over the grid of ``mesh + 3`` nodes; the grid geometry comes from
  ``SetFixedParametersFromTransformDomainInformation``
  (``Modules/Core/Transform/include/itkBSplineTransform.hxx:286-334``).
"""
from __future__ import annotations

import numpy as np

from .geometry import ImageGeometry

TF_IDENTITY = 0
TF_AFFINE = 1
TF_BSPLINE = 2
TF_COMPOSITE = 3  # Python-side only: the C ABI takes its entries separately


class AffineTransform:
    kind = TF_AFFINE

    def __init__(self, dim: int, center=None):
        self.dim = dim
        self.center = np.zeros(dim) if center is None else np.asarray(center, dtype=np.float64).copy()
        self.parameters = np.concatenate([np.eye(dim).ravel(), np.zeros(dim)])

    @property
    def number_of_parameters(self) -> int:
        return self.dim * self.dim + self.dim

    def set_parameters(self, p):
        p = np.asarray(p, dtype=np.float64)
        if p.size != self.number_of_parameters:
            raise ValueError("affine parameters must have %d entries" % self.number_of_parameters)
        self.parameters = p.copy()

    def get_parameters(self) -> np.ndarray:
        return self.parameters.copy()

    @property
    def matrix(self) -> np.ndarray:
        d = self.dim
        return self.parameters[: d * d].reshape(d, d).copy()

    @property
    def translation(self) -> np.ndarray:
        d = self.dim
        return self.parameters[d * d:].copy()

    @property
    def offset(self) -> np.ndarray:
        """ComputeOffset: offset[i] = t[i] + c[i] - sum_j M[i][j] c[j], in that order."""
        d = self.dim
        m = self.matrix
        t = self.translation
        off = np.zeros(d)
        for i in range(d):
            o = t[i] + self.center[i]
            for j in range(d):
                o -= m[i, j] * self.center[j]
            off[i] = o
        return off

    def update_transform_parameters(self, update, factor: float = 1.0):
        """Transform::UpdateTransformParameters (itkTransform.hxx:71-115): p += factor * update."""
        update = np.asarray(update, dtype=np.float64)
        if factor == 1.0:
            self.parameters = self.parameters + update
        else:
            self.parameters = self.parameters + update * factor

    def transform_point(self, x) -> np.ndarray:
        d = self.dim
        m = self.matrix
        off = self.offset
        y = np.zeros(d)
        for r in range(d):
            s = 0.0
            for c in range(d):
                s += m[r, c] * x[c]
            y[r] = s + off[r]
        return y


class BSplineTransform:
    kind = TF_BSPLINE
    order = 3

    def __init__(self, dim: int, grid: ImageGeometry):
        self.dim = dim
        self.grid = grid
        self.parameters = np.zeros(dim * grid.num_pixels)

    @classmethod
    def from_domain(cls, origin, physical_dimensions, mesh_size, direction=None):
        """SetFixedParametersFromTransformDomainInformation."""
        d = len(mesh_size)
        direction = np.eye(d) if direction is None else np.asarray(direction, dtype=np.float64)
        size = [int(m) + cls.order for m in mesh_size]
        spacing = [float(physical_dimensions[i]) / float(mesh_size[i]) for i in range(d)]
        o = np.array([-0.5 * spacing[i] * (cls.order - 1) for i in range(d)])
        ro = np.zeros(d)
        for i in range(d):
            s = 0.0
            for j in range(d):
                s += direction[i, j] * o[j]
            ro[i] = s
        grid_origin = [ro[i] + float(origin[i]) for i in range(d)]
        return cls(d, ImageGeometry(size=size, origin=grid_origin, spacing=spacing, direction=direction))

    @classmethod
    def for_image(cls, image_geometry: ImageGeometry, mesh_size):
        """Transform domain = the image's physical extent (the usual ITK set-up)."""
        d = image_geometry.dim
        phys = [image_geometry.spacing[i] * (image_geometry.size[i] - 1) for i in range(d)]
        origin = image_geometry.index_to_point(image_geometry.start)
        return cls.from_domain(origin, phys, mesh_size, image_geometry.direction)

    @property
    def number_of_parameters(self) -> int:
        return self.parameters.size

    @property
    def number_of_parameters_per_dimension(self) -> int:
        return self.grid.num_pixels

    def set_parameters(self, p):
        p = np.asarray(p, dtype=np.float64)
        if p.size != self.number_of_parameters:
            raise ValueError("B-spline parameters must have %d entries" % self.number_of_parameters)
        self.parameters = p.copy()

    def get_parameters(self) -> np.ndarray:
        return self.parameters.copy()

    def update_transform_parameters(self, update, factor: float = 1.0):
        update = np.asarray(update, dtype=np.float64)
        if factor == 1.0:
            self.parameters = self.parameters + update
        else:
            self.parameters = self.parameters + update * factor


class CompositeTransform:
    """itk::CompositeTransform with only its most recent transform optimized
    (``SetOnlyMostRecentTransformToOptimizeOn``, as
    ImageRegistrationMethodv4 builds it: itkImageRegistrationMethodv4.hxx:344-355).
    A point goes through the queue's last transform first, then the earlier
    ones (itkCompositeTransform.hxx:60-71); the parameters are the last
    transform's.  The entries before the last are affine (the metric takes
    them as matrix-offset stages)."""
    kind = TF_COMPOSITE

    def __init__(self, dim: int, transforms=()):
        self.dim = dim
        self.transforms = []
        for t in transforms:
            self.add_transform(t)

    def add_transform(self, tf):
        if tf.dim != self.dim:
            raise ValueError("transform dimension differs from the composite's")
        if tf.kind == TF_COMPOSITE:  # FlattenTransformQueue
            for t in tf.transforms:
                self.add_transform(t)
            return
        self.transforms.append(tf)

    @property
    def optimized(self):
        if not self.transforms:
            raise ValueError("the composite transform is empty")
        return self.transforms[-1]

    def initial_stages(self):
        """(matrices [n, D, D], offsets [n, D]) of the entries before the
        optimized one, in queue order."""
        stages = self.transforms[:-1]
        for t in stages:
            if t.kind != TF_AFFINE:
                raise ValueError("only affine transforms may precede the optimized transform")
        m = np.array([t.matrix for t in stages]).reshape(len(stages), self.dim, self.dim)
        o = np.array([t.offset for t in stages]).reshape(len(stages), self.dim)
        return m, o

    @property
    def number_of_parameters(self) -> int:
        return self.optimized.number_of_parameters

    def set_parameters(self, p):
        self.optimized.set_parameters(p)

    def get_parameters(self) -> np.ndarray:
        return self.optimized.get_parameters()

    def update_transform_parameters(self, update, factor: float = 1.0):
        self.optimized.update_transform_parameters(update, factor)

    def transform_point(self, x) -> np.ndarray:
        y = np.asarray(x, dtype=np.float64)
        for t in reversed(self.transforms):
            y = t.transform_point(y)
        return y
