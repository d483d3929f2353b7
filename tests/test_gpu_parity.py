"""Parity of the HIP metric (through the C ABI) with the CPU oracle.

Bar: Parzen bin indices and valid-sample counts bit-exact; value within 1e-5
relative; derivative within 1e-4 relative (L2).
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import REFERENCE_KNOWN_VALUES, affine_for, affine_params, blob_problem, bspline_for, reference_test_images
from _gpu import _compare, _gpu_metric
from itk_amd import synth
from itk_amd._native import MetricError
from itk_amd.geometry import ImageGeometry
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric
from itk_amd.transforms import AffineTransform, BSplineTransform

pytestmark = pytest.mark.gpu

def test_reference_unit_test_case(oracle_mod):
    """The reference test's images and settings (central differences, 50 bins,
    spacing (3,2), origin (0.001,-0.002)) on float32 storage."""
    fixed, moving, geom = reference_test_images(np.float32)
    tf = AffineTransform(2)
    m = _gpu_metric(fixed, moving, geom, tf, use_filter=False)
    for k, dd in [(4, -3.5), (5, 1e-5), (0, 0.02), (1, -0.03)]:
        p = np.array([1.0, 0, 0, 1, 0, 0])
        p[k] += dd
        tf.set_parameters(p)
        _compare(oracle_mod, m, fixed, moving, geom, tf)
    p = np.array([1.0, 0, 0, 1, 0, 1e-5])
    tf.set_parameters(p)
    assert m.GetValue() == pytest.approx(REFERENCE_KNOWN_VALUES[0], rel=1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_affine_3d_gradient_image(oracle_mod, seed):
    fixed, moving, geom = blob_problem((40, 36, 32), seed)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(seed)))
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)
    tf.set_parameters(affine_params(3, np.random.default_rng(100 + seed), scale=2.0))
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)


def test_affine_2d_config0(oracle_mod):
    """2-D 256x256 blob registration problem (30 blobs), dense, 50 bins."""
    p = synth.make_affine_problem((256, 256), 42)
    grad = oracle_mod.gradient_image(p.moving, p.geometry)
    tf = affine_for(p.geometry, affine_params(2, np.random.default_rng(5)))
    m = _gpu_metric(p.fixed, p.moving, p.geometry, tf, gradient=grad)
    _compare(oracle_mod, m, p.fixed, p.moving, p.geometry, tf, gradient=grad)


def test_device_gradient_filter_matches_oracle(oracle_mod):
    """GradientRecursiveGaussianImageFilter on the device vs the oracle's, seen
    through the metric derivative (the value does not use the gradient)."""
    fixed, moving, geom = blob_problem((36, 30, 28), 4)
    geom = ImageGeometry(size=geom.size, spacing=(1.0, 1.25, 0.8), origin=(3.0, -2.0, 1.0))
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(8)))
    m = _gpu_metric(fixed, moving, geom, tf)  # no gradient given: computed on the device
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)


def test_central_difference_3d(oracle_mod):
    fixed, moving, geom = blob_problem((30, 28, 26), 6)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(9)))
    m = _gpu_metric(fixed, moving, geom, tf, use_filter=False)
    _compare(oracle_mod, m, fixed, moving, geom, tf)


def test_anisotropic_geometry_and_virtual_domain(oracle_mod):
    fixed, moving, _ = blob_problem((34, 30, 26), 10)
    geom = ImageGeometry(size=(34, 30, 26), spacing=(0.9, 1.1, 1.7), origin=(-5.0, 2.5, 11.0))
    virt = ImageGeometry(size=(20, 18, 12), start=(5, 4, 6), spacing=(1.3, 1.2, 2.1), origin=(-4.0, 3.0, 12.5))
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(11)))
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad, virtual=virt)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad, virtual=virt)


# 5 and 32: two LDS histogram copies in pass 1; 80: one copy; 100: the
# histogram exceeds LDS, exact path
@pytest.mark.parametrize("bins", [5, 32, 80, 100])
def test_histogram_sizes(oracle_mod, bins):
    fixed, moving, geom = blob_problem((24, 22, 20), 12)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(12)))
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad, bins=bins)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad, bins=bins)


def test_bspline_3d(oracle_mod):
    fixed, moving, geom = blob_problem((32, 30, 28), 13)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = bspline_for(geom, (4, 4, 4), seed=3, sigma=0.8)
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)


def test_bspline_2d(oracle_mod):
    p = synth.make_affine_problem((96, 80), 14)
    grad = oracle_mod.gradient_image(p.moving, p.geometry)
    tf = bspline_for(p.geometry, (6, 5), seed=4, sigma=1.0)
    m = _gpu_metric(p.fixed, p.moving, p.geometry, tf, gradient=grad)
    _compare(oracle_mod, m, p.fixed, p.moving, p.geometry, tf, gradient=grad)


def test_bspline_partial_domain_anisotropic(oracle_mod):
    """Separable B-spline derivative where the transform domain covers part of
    an anisotropic, offset virtual domain: samples outside the valid grid region
    keep zero displacement and no Jacobian, per axis of the support tables."""
    fixed, moving, _ = blob_problem((36, 30, 26), 21)
    geom = ImageGeometry(size=(36, 30, 26), spacing=(0.9, 1.1, 1.7), origin=(-5.0, 2.5, 11.0))
    virt = ImageGeometry(size=(28, 22, 14), start=(3, 4, 5), spacing=(1.0, 1.2, 1.9), origin=(-4.0, 3.0, 12.5))
    grad = oracle_mod.gradient_image(moving, geom)
    tf = BSplineTransform.from_domain(origin=(-1.0, 6.0, 16.0), physical_dimensions=(18.0, 17.0, 16.0),
                                      mesh_size=(3, 4, 2))
    tf.set_parameters(np.random.default_rng(22).normal(0.0, 0.7, tf.number_of_parameters))
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad, virtual=virt)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad, virtual=virt)


def test_bspline_oblique_direction(oracle_mod):
    """Rotated image and grid directions: the per-sample (non-separable)
    B-spline derivative path."""
    fixed, moving, _ = blob_problem((30, 28, 24), 23)
    rot = synth.rotation_matrix([8.0, -5.0, 12.0])
    geom = ImageGeometry(size=(30, 28, 24), spacing=(1.0, 1.1, 1.3), origin=(2.0, -3.0, 1.0), direction=rot)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = bspline_for(geom, (3, 3, 3), seed=24, sigma=0.7)
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)


def test_value_only_matches_value_and_derivative(oracle_mod):
    fixed, moving, geom = blob_problem((24, 24, 24), 15)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(13)))
    m = _gpu_metric(fixed, moving, geom, tf)
    v1 = m.GetValue()
    v2, _ = m.GetValueAndDerivative()
    assert v1 == v2


def test_repeatable_bitwise(oracle_mod):
    """Integer histogram accumulation and ordered partial sums: two
    evaluations of the affine metric agree to the last bit."""
    fixed, moving, geom = blob_problem((40, 40, 40), 16)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(14)))
    m = _gpu_metric(fixed, moving, geom, tf)
    a = m.GetValueAndDerivative()
    b = m.GetValueAndDerivative()
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])


def test_degenerate_inputs_raise():
    fixed, moving, geom = blob_problem((16, 16, 16), 17)
    tf = affine_for(geom)
    m = Metric()
    m.SetFixedImage(np.full_like(fixed, 2.0), geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    with pytest.raises(MetricError, match="All fixed-image samples in the analysis region have the same intensity"):
        m.Initialize()
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(np.zeros_like(moving), geom)
    m.SetMovingTransform(tf)
    with pytest.raises(MetricError, match="All moving-image samples in the analysis region have the same intensity"):
        m.Initialize()
    m = _gpu_metric(fixed, moving, geom, tf)
    p = tf.get_parameters()
    p[9] = 1e6
    tf.set_parameters(p)
    with pytest.raises(MetricError, match="do not sufficiently overlap"):
        m.GetValueAndDerivative()
    assert m.GetNumberOfValidPoints() == 0


def test_joint_pdf(oracle_mod):
    fixed, moving, geom = blob_problem((28, 26, 24), 18)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(15)))
    m = _gpu_metric(fixed, moving, geom, tf)
    pdf = m.GetJointPDF()
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, derivative=False)
    assert pdf.sum() == pytest.approx(1.0, abs=1e-12)
    # FP32 Parzen weights on the fast path: ~1e-7 relative per bin
    assert np.abs(pdf - r["joint_pdf"]).max() <= 1e-6 * r["joint_pdf"].max()


@pytest.mark.parametrize("shape,seed", [((64, 60, 56), 21), ((200, 180), 22)])
def test_fast_path_matches_exact_path(oracle_mod, shape, seed):
    """The certified FP32 fast path and the all-FP64 path give identical
    per-sample bins and valid counts, and values/derivatives to rounding."""
    fixed, moving, geom = blob_problem(shape, seed)
    grad = oracle_mod.gradient_image(moving, geom)
    d = len(shape)
    tf = affine_for(geom, affine_params(d, np.random.default_rng(seed), scale=3.0))
    out = []
    for exact in (False, True):
        m = Metric(record_sample_bins=True, exact_path_only=exact)
        m.SetFixedImage(fixed, geom)
        m.SetMovingImage(moving, geom)
        m.SetMovingImageGradient(grad)
        m.SetMovingTransform(tf)
        m.Initialize()
        v, dv = m.GetValueAndDerivative()
        assert m.GetLastEvaluationPath() == ("exact" if exact else "fast")
        out.append((v, dv, m.GetNumberOfValidPoints()) + m.GetSampleBins())
        m.close()
    (v0, d0, n0, f0, m0), (v1, d1, n1, f1, m1) = out
    assert n0 == n1
    assert np.array_equal(f0, f1) and np.array_equal(m0, m1)
    assert v0 == pytest.approx(v1, rel=1e-7)
    assert np.linalg.norm(d0 - d1) <= 1e-5 * np.linalg.norm(d1)


@pytest.mark.parametrize("shape,seed", [((48, 44, 40), 330), ((160, 150), 331)])
def test_plateaus_at_the_range_ends_stay_on_the_fast_path(oracle_mod, shape, seed):
    """Moving and fixed images clipped at their 10th and 90th percentiles:
    large plateaus whose interpolated values equal the range ends exactly.
    The reference counts those samples (a multilinear value of corners in
    [min, max] stays in [min, max]) and their Parzen terms sit on the clamped
    edges 2 and B - 2, so the row walks certify them without a range test and
    the exception lists do not overflow; bins, count, value and derivative
    equal the oracle's."""
    fixed, moving, geom = blob_problem(shape, seed)
    fixed = np.clip(fixed, *np.quantile(fixed, [0.1, 0.9])).astype(np.float32)
    moving = np.clip(moving, *np.quantile(moving, [0.1, 0.9])).astype(np.float32)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(geom.dim, np.random.default_rng(seed)))
    m = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)
    assert m.GetLastEvaluationPath() == "fast"


def test_full_size_config1_affine(oracle_mod):
    """256^3 float32, 80 blobs, affine: bins, count, value and derivative
    against the multithreaded oracle at full size."""
    p = synth.make_affine_problem((256, 256, 256), 2024)
    tf = affine_for(p.geometry, affine_params(3, np.random.default_rng(16)))
    m = _gpu_metric(p.fixed, p.moving, p.geometry, tf, record=True)
    grad = oracle_mod.gradient_image(p.moving, p.geometry, threads=16)
    _compare(oracle_mod, m, p.fixed, p.moving, p.geometry, tf, gradient=grad)


def test_device_smoothing_matches_oracle(oracle_mod):
    """Per-level smoothing on the device (SmoothingRecursiveGaussianImageFilter,
    last axis first) against the oracle's: the metric on the device-smoothed
    images has the oracle's bins on the oracle-smoothed images, bit for bit."""
    fixed, moving, _ = blob_problem((40, 36, 30), 41)
    geom = ImageGeometry(size=(40, 36, 30), spacing=(0.9, 1.2, 1.6), origin=(2.0, -1.0, 0.5))
    sig_f, sig_m = [1.2, 0.8, 1.5], [2.0, 2.0, 2.0]
    tf = affine_for(geom, affine_params(3, np.random.default_rng(42)))
    m = Metric(record_sample_bins=True)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetSmoothingSigmas(sig_f, sig_m)
    m.SetMovingTransform(tf)
    m.Initialize()
    fs = oracle_mod.smooth(fixed, geom, sig_f)
    ms = oracle_mod.smooth(moving, geom, sig_m)
    grad = oracle_mod.gradient_image(ms, geom)
    _compare(oracle_mod, m, fs, ms, geom, tf, gradient=grad)
    # back to unsmoothed images on the same context
    m.SetSmoothingSigmas(None, None)
    m.Initialize()
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=oracle_mod.gradient_image(moving, geom))


@pytest.mark.parametrize("shape,spacing,sigma", [
    ((40, 36, 30), (0.9, 1.2, 1.6), (1.2, 0.8, 1.5)),
    ((70, 33, 21), (1.0, 1.0, 1.0), (2.0, 2.0, 2.0)),  # lines of 70 along x: tiles 32 + 32 + 6
    ((37, 50), (1.1, 0.7), (1.5, 0.9)),  # 2-D: the last pass of both filters runs along x
])
def test_device_filters_bitwise(oracle_mod, shape, spacing, sigma):
    """Device per-level smoothing and gradient filter against the oracle's, bit
    for bit, through the FP32 texels: the value channel is the smoothed moving
    image, the gradient channels the filter's FP64 output rounded to FP32."""
    fixed, moving, _ = blob_problem(shape, 47)
    geom = ImageGeometry(size=shape, spacing=spacing, origin=(1.0,) * len(shape))
    tf = affine_for(geom, affine_params(len(shape), np.random.default_rng(48)))
    for sig in (None, sigma):
        m = Metric()
        m.SetFixedImage(fixed, geom)
        m.SetMovingImage(moving, geom)
        if sig is not None:
            m.SetSmoothingSigmas(sig, sig)
        m.SetMovingTransform(tf)
        m.Initialize()
        t = m.GetMovingTexels()
        ms = moving if sig is None else oracle_mod.smooth(moving, geom, sig)
        assert ms.dtype == np.float32
        grad = oracle_mod.gradient_image(ms, geom)
        assert np.array_equal(t[..., 0].view(np.uint32), ms.view(np.uint32))
        g32 = grad.astype(np.float32)
        for k in range(len(shape)):
            assert np.array_equal(t[..., 1 + k].view(np.uint32), g32[..., k].view(np.uint32)), k
        m.close()


def test_user_gradient_with_smoothing(oracle_mod):
    """A gradient image given by the caller stays as given while the moving
    image is smoothed for a level; the fast derivative pass reads the moving
    value from the texels, whose value channel follows the smoothed image:
    bins, count, value and derivative as the oracle's, on both levels."""
    fixed, moving, geom = blob_problem((48, 44, 40), 43)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(44)))
    m = Metric(record_sample_bins=True)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingImageGradient(grad)
    m.SetSmoothingSigmas([1.0] * 3, [1.5] * 3)
    m.SetMovingTransform(tf)
    m.Initialize()
    fs = oracle_mod.smooth(fixed, geom, [1.0] * 3)
    ms = oracle_mod.smooth(moving, geom, [1.5] * 3)
    _compare(oracle_mod, m, fs, ms, geom, tf, gradient=grad)
    assert m.GetLastEvaluationPath() == "fast"
    m.SetSmoothingSigmas(None, None)
    m.Initialize()
    _compare(oracle_mod, m, fixed, moving, geom, tf, gradient=grad)
    assert m.GetLastEvaluationPath() == "fast"


@pytest.mark.parametrize("factors", [2, 4, [2, 3, 1]])
def test_fast_path_on_shrunk_virtual_domain(oracle_mod, factors):
    """Coarse pyramid levels: virtual samples fall between fixed pixels, so the
    fast path interpolates the fixed image from per-axis cell tables.  Bins and
    counts equal the exact path's and the oracle's."""
    fixed, moving, geom = blob_problem((64, 60, 56), 61)
    grad = oracle_mod.gradient_image(moving, geom)
    virt = geom.shrink(factors)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(62), scale=2.0))
    m_fast = _gpu_metric(fixed, moving, geom, tf, gradient=grad, virtual=virt)
    m_exact = _gpu_metric(fixed, moving, geom, tf, gradient=grad, virtual=virt, exact=True)
    v0, d0 = m_fast.GetValueAndDerivative()
    assert m_fast.GetLastEvaluationPath() == "fast"
    v1, d1 = m_exact.GetValueAndDerivative()
    assert m_exact.GetLastEvaluationPath() == "exact"
    assert m_fast.GetNumberOfValidPoints() == m_exact.GetNumberOfValidPoints()
    for a, b in zip(m_fast.GetSampleBins(), m_exact.GetSampleBins()):
        assert np.array_equal(a, b)
    assert v0 == pytest.approx(v1, rel=1e-7)
    _compare(oracle_mod, m_fast, fixed, moving, geom, tf, gradient=grad, virtual=virt)


@pytest.mark.parametrize("shape,mesh", [((48, 44, 40), (4, 3, 5)), ((120, 100), (6, 5))])
def test_bspline_fast_path_matches_exact_path(oracle_mod, shape, mesh):
    """Certified B-spline fast path (displacement from per-row coefficient
    lines, FP32 interpolation, exception lists) against the all-FP64 path:
    identical bins and counts; and against the oracle."""
    fixed, moving, geom = blob_problem(shape, 63)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = bspline_for(geom, mesh, seed=64, sigma=1.2)
    m_fast = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    m_exact = _gpu_metric(fixed, moving, geom, tf, gradient=grad, exact=True)
    v0, d0 = m_fast.GetValueAndDerivative()
    assert m_fast.GetLastEvaluationPath() == "fast"
    v1, d1 = m_exact.GetValueAndDerivative()
    assert m_fast.GetNumberOfValidPoints() == m_exact.GetNumberOfValidPoints()
    for a, b in zip(m_fast.GetSampleBins(), m_exact.GetSampleBins()):
        assert np.array_equal(a, b)
    assert v0 == pytest.approx(v1, rel=1e-7)
    assert np.linalg.norm(d0 - d1) <= 1e-5 * np.linalg.norm(d1)
    _compare(oracle_mod, m_fast, fixed, moving, geom, tf, gradient=grad)


@pytest.mark.parametrize("case", ["3d_f32", "3d_bf16", "2d_f32", "3d_smoothed_user_gradient", "3d_80_bins",
                                  "3d_bspline"])
def test_pass2_from_pass1_classes_matches_reclassifying(oracle_mod, case):
    """Pass 2 reading pass 1's byte per sample and the moving values from the
    texels gives the bits of the pass 2 that classifies every sample again
    (with user texels under smoothing both run the reclassifying pass 2), both
    after the sample-strided pass 1.  The affine row-walk kernels (their own
    pass 1 and pass 2, rounding the weights from fixed-point coordinates) give
    the same bins and valid counts, value and derivative within 1e-9 and
    1e-6 of those bits (the bar against the oracle is 1e-5 and 1e-4)."""
    shape = (60, 52, 44) if case.startswith("3d") else (200, 180)
    fixed, moving, geom = blob_problem(shape, 140)
    d = geom.dim
    bs = case == "3d_bspline"
    tf = bspline_for(geom, (4, 4, 3), seed=141, sigma=0.7) if bs else \
        affine_for(geom, affine_params(d, np.random.default_rng(141), scale=2.0))
    p0 = tf.get_parameters()
    out = []
    for again, walk in ((False, 0), (True, 0), (False, -1)):
        m = Metric(pass2_reclassify=again)
        m.SetLaunchConfiguration(rows_walk=walk)
        if case == "3d_bf16":
            m.SetImageStorage("bf16")
        m.SetFixedImage(fixed, geom)
        m.SetMovingImage(moving, geom)
        if case == "3d_smoothed_user_gradient":
            m.SetMovingImageGradient(oracle_mod.gradient_image(moving, geom))
            m.SetSmoothingSigmas(1.0, 1.0)
        if case == "3d_80_bins":
            m.SetNumberOfHistogramBins(80)
        m.SetMovingTransform(tf)
        m.Initialize()
        vals = []
        for k in range(3):
            if bs:
                tf.set_parameters(p0 + np.random.default_rng(142 + k).normal(0, 0.05, p0.size))
            else:
                tf.set_parameters(affine_params(d, np.random.default_rng(142 + k), scale=2.0))
            v, dv = m.GetValueAndDerivative()
            assert m.GetLastEvaluationPath() == "fast"
            vals.append((v, dv.copy(), m.GetNumberOfValidPoints()))
        out.append(vals)
    for (v0, d0, n0), (v1, d1, n1), (v2, d2, n2) in zip(*out):
        assert n0 == n1 and v0 == v1
        assert np.array_equal(d0, d1)
        assert n2 == n0
        assert abs(v2 - v0) <= 1e-9 * abs(v0)
        assert np.linalg.norm(d2 - d0) <= 1e-6 * np.linalg.norm(d0)
