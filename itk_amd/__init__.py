"""MI355X-native Mattes mutual-information v4 metric (value + derivative) for ITK registration."""
from .geometry import ImageGeometry
from .transforms import AffineTransform, BSplineTransform

__all__ = ["ImageGeometry", "AffineTransform", "BSplineTransform"]
