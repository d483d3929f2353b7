"""Pixel types of the reference's images on the HIP path, against the oracle.

The reference assigns each sample's interpolated double to the image's pixel
type (itkImageToImageMetricv4.hxx:318,369): an integer type truncates toward
zero, a double keeps the value, a float rounds it.  Integer images (uint8,
int16) give the oracle's per-sample bins, valid count, value and derivative
when the oracle truncates the same way; double images that FP32 holds exactly
give the oracle's results with unrounded values; double pixels FP32 cannot
hold, and smoothing of non-float images, are refused with a clear error
instead of a silent rounding.
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem, bspline_for
from _gpu import _compare
from itk_amd._native import MetricError
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

pytestmark = pytest.mark.gpu


def _integer_images(kind):
    fixed, moving, geom = blob_problem((44, 40, 36), 900)
    if kind == "uint8":
        f = np.trunc(np.clip(fixed, 0, None) * 180.0).astype(np.uint8)
        m = np.trunc(np.clip(moving, 0, None) * 180.0).astype(np.uint8)
    else:  # CT-like int16 in Hounsfield-unit range
        f = np.trunc(fixed * 3000.0 - 1000.0).astype(np.int16)
        m = np.trunc(moving * 3000.0 - 1000.0).astype(np.int16)
    return f, m, geom


def _metric(fixed, moving, geom, tf, use_filter=True, gradient=None):
    m = Metric(record_sample_bins=True)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetUseMovingImageGradientFilter(use_filter)
    if gradient is not None:
        m.SetMovingImageGradient(gradient)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


@pytest.mark.parametrize("kind,transform,use_filter", [("uint8", "affine", True), ("int16", "affine", False),
                                                       ("int16", "bspline", True)])
def test_integer_pixels_truncate_like_the_reference(oracle_mod, kind, transform, use_filter):
    f, mv, geom = _integer_images(kind)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(901), scale=1.5)) if transform == "affine" else \
        bspline_for(geom, (4, 4, 3), seed=902, sigma=0.6)
    f64, m64 = f.astype(np.float64), mv.astype(np.float64)
    grad = oracle_mod.gradient_image(m64, geom) if use_filter else None
    m = _metric(f, mv, geom, tf, use_filter, gradient=grad)
    v, d = m.GetValueAndDerivative()
    assert m.GetLastEvaluationPath() == "exact"
    r = oracle_mod.mattes(f64, geom, m64, geom, tf, moving_gradient=grad, want_samples=True, threads=8,
                          integer_pixels=True)
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    fb, mb = m.GetSampleBins()
    assert np.array_equal(fb, r["fixed_bin"]) and np.array_equal(mb, r["moving_bin"])
    assert abs(v - r["value"]) <= 1e-5 * abs(r["value"])
    assert np.linalg.norm(d - r["derivative"]) <= 1e-4 * np.linalg.norm(r["derivative"])
    # the truncation matters: float semantics on the same values give other bins
    rf = oracle_mod.mattes(f64, geom, m64, geom, tf, moving_gradient=grad, want_samples=True, threads=8)
    assert not np.array_equal(mb, rf["moving_bin"])
    m.close()


def test_double_pixels_exact_in_float_keep_unrounded_values(oracle_mod):
    fixed, moving, geom = blob_problem((48, 44, 40), 903)
    f64, m64 = fixed.astype(np.float64), moving.astype(np.float64)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(904), scale=1.5))
    grad = oracle_mod.gradient_image(m64, geom)
    m = _metric(f64, m64, geom, tf, gradient=grad)
    _compare(oracle_mod, m, f64, m64, geom, tf, gradient=grad)
    assert m.GetLastEvaluationPath() == "fast"
    m.close()


def test_double_pixels_float_cannot_hold_are_refused():
    fixed, moving, geom = blob_problem((24, 20, 16), 905)
    m = Metric()
    with pytest.raises(MetricError, match="not exact in the device's FP32 storage"):
        m.SetFixedImage(fixed.astype(np.float64) + 1e-9, geom)
        m.SetMovingImage(moving, geom)
        m.SetMovingTransform(affine_for(geom))
        m.Initialize()
    m.close()


def test_smoothing_of_integer_images_is_refused():
    f, mv, geom = _integer_images("uint8")
    m = Metric()
    m.SetFixedImage(f, geom)
    m.SetMovingImage(mv, geom)
    m.SetSmoothingSigmas(1.0, 1.0)
    m.SetMovingTransform(affine_for(geom))
    with pytest.raises(MetricError, match="smoothing takes float images only"):
        m.Initialize()
    m.close()
