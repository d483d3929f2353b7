"""Inputs shared by the oracle and GPU parity tests."""
from __future__ import annotations

import numpy as np

from itk_amd import synth
from itk_amd.geometry import ImageGeometry
from itk_amd.transforms import AffineTransform, BSplineTransform

# Known metric values recorded in the reference's unit test for the perturbed
# parameter sets {p+d, p-d, p+2d, p-2d} with d = 1e-5 on the last parameter
# (Modules/Registration/Metricsv4/test/itkMattesMutualInformationImageToImageMetricv4Test.cxx:387-390).
REFERENCE_KNOWN_VALUES = (-1.297994750907097, -1.297995377157083, -1.297994448222851, -1.297995700722833)


def reference_test_images(dtype=np.float64):
    """The reference unit test's inputs (test.cxx:59-136): 100x100 images of
    unsigned-char-truncated Gaussians (sigma^2 = 2500/2 in index units), the
    moving one shifted by 5 pixels, spacing (3, 2), origin (0.001, -0.002)."""
    n = 100
    geom = ImageGeometry(size=(n, n), origin=(0.001, -0.002), spacing=(3.0, 2.0))
    yy, xx = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    c = n / 2.0
    s = n / 2.0

    def gauss(shift):
        x = xx - c + shift
        y = yy - c + shift
        return np.floor(200.0 * np.exp(-(x * x + y * y) / (s * s))).astype(np.uint8).astype(dtype)

    return gauss(0.0), gauss(5.0), geom


def affine_params(dim: int, rng: np.random.Generator, scale: float = 1.0) -> np.ndarray:
    """A random near-identity affine parameter vector."""
    m = np.eye(dim) + scale * rng.uniform(-0.04, 0.04, (dim, dim))
    t = scale * rng.uniform(-2.0, 2.0, dim)
    return np.concatenate([m.ravel(), t])


def blob_problem(shape, seed: int):
    p = synth.make_affine_problem(shape, seed, n_blobs=12 if len(shape) == 3 else 20)
    return p.fixed, p.moving, p.geometry


def affine_for(geom: ImageGeometry, params=None) -> AffineTransform:
    tf = AffineTransform(geom.dim, center=geom.index_to_point([(s - 1) / 2.0 for s in geom.size]))
    if params is not None:
        tf.set_parameters(params)
    return tf


def bspline_for(geom: ImageGeometry, mesh, seed: int, sigma: float = 0.8) -> BSplineTransform:
    tf = BSplineTransform.for_image(geom, mesh)
    rng = np.random.default_rng(seed)
    tf.set_parameters(rng.normal(0.0, sigma, tf.number_of_parameters))
    return tf
