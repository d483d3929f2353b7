"""``itk::ImageMaskSpatialObject`` as the v4 metrics use it.

A mask is a uint8-or-wider image with its own geometry.  A physical point is
inside when the image's rounded index of the point
(``ImageBase::TransformPhysicalPointToIndex``) lies in the buffer and the mask
pixel there is non-zero, or equals the mask value when ``UseMaskValue`` is on
(``Modules/Core/SpatialObjects/include/itkImageMaskSpatialObject.hxx:36-48``).
The object-to-world transform is the identity, so world and object points are
the same point (``itkSpatialObject.hxx:186-190``).  The device sees the mask
as one byte per voxel: 1 where the pixel counts as inside.
"""
from __future__ import annotations

import numpy as np

from .geometry import ImageGeometry


class ImageMaskSpatialObject:
    def __init__(self, pixels=None, geometry: ImageGeometry = None):
        self._px = None
        self._geom = None
        self._use_mask_value = False
        self._mask_value = 1
        if pixels is not None:
            self.SetImage(pixels, geometry)

    def SetImage(self, pixels, geometry: ImageGeometry):
        a = np.ascontiguousarray(pixels)
        if geometry is None or a.shape != geometry.shape:
            raise ValueError("mask pixel array does not match its geometry")
        self._px = a
        self._geom = geometry

    def GetImage(self):
        return self._px, self._geom

    def SetUseMaskValue(self, on: bool):
        self._use_mask_value = bool(on)

    def GetUseMaskValue(self) -> bool:
        return self._use_mask_value

    def SetMaskValue(self, v):
        self._mask_value = v

    def GetMaskValue(self):
        return self._mask_value

    def inside_flags(self) -> np.ndarray:
        """uint8 per voxel: 1 where the pixel counts as inside."""
        if self._px is None:
            raise ValueError("the mask has no image")
        if self._use_mask_value:
            return (self._px == self._mask_value).astype(np.uint8)
        return (self._px != 0).astype(np.uint8)

    @property
    def geometry(self) -> ImageGeometry:
        return self._geom

    def IsInsideInWorldSpace(self, point) -> bool:
        idx, ok = self._geom.physical_points_to_index(np.asarray(point, dtype=np.float64).reshape(1, -1))
        if not ok[0]:
            return False
        off = tuple(int(idx[0, k] - self._geom.start[k]) for k in reversed(range(self._geom.dim)))
        return bool(self.inside_flags()[off])
