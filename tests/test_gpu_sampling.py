"""Sparse point-set sampling and image masks on the device, against the
oracle: per-sample bins and valid counts bit-exact, value within 1e-5 and
derivative within 1e-4 relative (L2), intensity ranges equal to the oracle's
(ImageToImageMetricv4 masks and sampled point sets,
itkImageToImageMetricv4.hxx:298-309,351-362,567-611;
itkMattesMutualInformationImageToImageMetricv4.hxx:97-220; the sparse
threader, itkImageToImageMetricv4GetValueAndDerivativeThreader.hxx:44-63)."""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem, bspline_for
from _gpu import DERIV_RTOL, VALUE_RTOL
from itk_amd._native import MMI_ERR_DEGENERATE, MetricError
from itk_amd.batch import BatchEvaluator
from itk_amd.geometry import ImageGeometry
from itk_amd.masks import ImageMaskSpatialObject
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

pytestmark = pytest.mark.gpu


def _voxel_points(geom):
    d = geom.dim
    grids = np.meshgrid(*[np.arange(geom.start[k], geom.start[k] + geom.size[k]) for k in range(d)], indexing="ij")
    idx = np.stack([g.transpose(tuple(reversed(range(d)))).ravel() for g in grids], axis=1)
    pts = np.zeros((idx.shape[0], d))
    for i in range(d):
        s = np.full(idx.shape[0], geom.origin[i])
        for j in range(d):
            s = s + geom.index_to_physical[i, j] * idx[:, j].astype(np.float64)
        pts[:, i] = s
    return pts


def _random_points(geom, n, seed):
    rng = np.random.default_rng(seed)
    pts = _voxel_points(geom)[rng.choice(geom.num_pixels, n, replace=False)]
    pts = pts + rng.uniform(-0.45, 0.45, pts.shape) * np.asarray(geom.spacing)
    far = np.full((2, geom.dim), -1e3)
    return np.concatenate([pts, far])  # two points outside every image


def _metric(fixed, moving, geom, tf, *, grad=None, points=None, fmask=None, mmask=None, bins=50):
    m = Metric(record_sample_bins=True)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    if grad is not None:
        m.SetMovingImageGradient(grad)
    m.SetNumberOfHistogramBins(bins)
    m.SetMovingTransform(tf)
    if points is not None:
        m.SetVirtualSampledPointSet(points)
        m.SetUseSampledPointSet(True)
        m.SetUseVirtualSampledPointSet(True)
    if fmask is not None:
        m.SetFixedImageMask(ImageMaskSpatialObject(*fmask))
    if mmask is not None:
        m.SetMovingImageMask(ImageMaskSpatialObject(*mmask))
    m.Initialize()
    return m


def _check(oracle_mod, m, fixed, moving, geom, tf, *, grad=None, points=None, fmask=None, mmask=None, bins=50):
    v, d = m.GetValueAndDerivative()
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, bins=bins, moving_gradient=grad, want_samples=True,
                          threads=8, points=points, fixed_mask=fmask, moving_mask=mmask)
    hp = m.GetHistogramParameters()
    fr = oracle_mod.intensity_range(fixed, geom, mask=None if points is not None else fmask, points=points)
    mr = oracle_mod.intensity_range(moving, geom, mask=mmask)
    print("value gpu %.17g oracle %.17g  nvalid %d/%d path %s" % (
        v, r["value"], m.GetNumberOfValidPoints(), r["n_valid"], m.GetLastEvaluationPath()))
    assert (hp["fixed_min"], hp["fixed_max"]) == fr[:2]
    assert (hp["moving_min"], hp["moving_max"]) == mr[:2]
    # point sets and B-spline masks run on the exact path, affine masks on the
    # certified fast path
    want = "fast" if (points is None and tf.kind == 1 and (fmask is not None or mmask is not None)) else "exact"
    assert m.GetLastEvaluationPath() == want
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    fb, mb = m.GetSampleBins()
    assert np.array_equal(fb, r["fixed_bin"])
    assert np.array_equal(mb, r["moving_bin"])
    assert abs(v - r["value"]) <= VALUE_RTOL * abs(r["value"]) + 1e-11
    err = np.linalg.norm(d - r["derivative"]) / max(np.linalg.norm(r["derivative"]), 1e-9)
    print("derivative rel L2 error %.3e" % err)
    assert err <= DERIV_RTOL
    return v, d


def _coarse_mask(geom, seed, factor=2, fill=0.7):
    rng = np.random.default_rng(seed)
    d = geom.dim
    mg = ImageGeometry(size=tuple(s // factor + 2 for s in geom.size),
                       origin=tuple(o - 0.37 * s for o, s in zip(geom.origin, geom.spacing)),
                       spacing=tuple(factor * s for s in geom.spacing))
    px = (rng.uniform(size=mg.shape) < fill).astype(np.uint8) * rng.integers(1, 255, mg.shape, dtype=np.uint8)
    assert mg.dim == d
    return px, mg


@pytest.mark.parametrize("kind", ["affine", "bspline"])
def test_virtual_point_set(oracle_mod, kind):
    fixed, moving, geom = blob_problem((40, 36, 30), 51)
    grad = oracle_mod.gradient_image(moving, geom)
    if kind == "affine":
        tf = affine_for(geom, affine_params(3, np.random.default_rng(52), scale=2.0))
    else:
        tf = bspline_for(geom, (4, 4, 3), seed=53, sigma=0.7)
    pts = _random_points(geom, 9000, 54)
    m = _metric(fixed, moving, geom, tf, grad=grad, points=pts)
    assert m.GetNumberOfDomainPoints() == pts.shape[0]
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, points=pts)
    # a new parameter vector on the same point set
    p = tf.get_parameters()
    tf.set_parameters(p + np.random.default_rng(55).normal(0, 0.05, p.size))
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, points=pts)


def test_point_set_of_every_voxel_equals_dense(oracle_mod):
    """Every virtual voxel centre, in buffer order, as a point set gives the
    dense evaluation bit for bit (same exact path, same sample order)."""
    fixed, moving, geom = blob_problem((32, 28, 24), 56)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(57)))
    dense = Metric(exact_path_only=True)
    for mm in (dense,):
        mm.SetFixedImage(fixed, geom)
        mm.SetMovingImage(moving, geom)
        mm.SetMovingTransform(tf)
        mm.Initialize()
    sparse = _metric(fixed, moving, geom, tf, points=_voxel_points(geom))
    v0, d0 = dense.GetValueAndDerivative()
    v1, d1 = sparse.GetValueAndDerivative()
    assert dense.GetNumberOfValidPoints() == sparse.GetNumberOfValidPoints()
    assert v0 == v1
    assert np.allclose(d0, d1, rtol=1e-12, atol=1e-15)


def test_point_set_2d(oracle_mod):
    fixed, moving, geom = blob_problem((120, 100), 58)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(2, np.random.default_rng(59), scale=2.0))
    pts = _random_points(geom, 3000, 60)
    m = _metric(fixed, moving, geom, tf, grad=grad, points=pts)
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, points=pts)


@pytest.mark.parametrize("which", ["fixed", "moving", "both"])
def test_image_masks(oracle_mod, which):
    fixed, moving, geom = blob_problem((44, 40, 34), 61)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(62), scale=2.0))
    fmask = _coarse_mask(geom, 63) if which in ("fixed", "both") else None
    mmask = _coarse_mask(geom, 64, factor=3) if which in ("moving", "both") else None
    m = _metric(fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)


def test_masks_bspline_and_points(oracle_mod):
    """B-spline transform with both masks (separable derivative path), then
    with a point set as well (the fixed mask no longer sets the range)."""
    fixed, moving, geom = blob_problem((40, 38, 30), 65)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = bspline_for(geom, (3, 4, 3), seed=66, sigma=0.6)
    fmask, mmask = _coarse_mask(geom, 67), _coarse_mask(geom, 68, fill=0.8)
    m = _metric(fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)
    pts = _random_points(geom, 7000, 69)
    m = _metric(fixed, moving, geom, tf, grad=grad, points=pts, fmask=fmask, mmask=mmask)
    _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, points=pts, fmask=fmask, mmask=mmask)


def test_mask_value_and_fixed_point_set(oracle_mod):
    """UseMaskValue selects the pixels equal to the mask value; a fixed-space
    point set (identity fixed transform) is the same virtual point set, with
    the points outside the virtual domain skipped."""
    fixed, moving, geom = blob_problem((36, 32, 28), 70)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(71)))
    rng = np.random.default_rng(72)
    labels = rng.integers(0, 4, geom.shape).astype(np.int16)
    so = ImageMaskSpatialObject(labels, geom)
    so.SetUseMaskValue(True)
    so.SetMaskValue(2)
    m = Metric(record_sample_bins=True)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    m.SetMovingImageMask(so)
    m.Initialize()
    mm = ((labels == 2).astype(np.uint8), geom)
    v, _ = m.GetValueAndDerivative()
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_mask=mm, threads=8)
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    assert abs(v - r["value"]) <= VALUE_RTOL * abs(r["value"])

    pts = _random_points(geom, 4000, 73)  # the last two lie outside the virtual domain
    a = _metric(fixed, moving, geom, tf, points=pts[:-2])
    b = Metric(record_sample_bins=True)
    b.SetFixedImage(fixed, geom)
    b.SetMovingImage(moving, geom)
    b.SetMovingTransform(tf)
    b.SetFixedSampledPointSet(pts)
    b.SetUseSampledPointSet(True)
    b.Initialize()
    assert b.GetNumberOfSkippedFixedSampledPoints() == 2
    va, da = a.GetValueAndDerivative()
    vb, db = b.GetValueAndDerivative()
    assert va == vb and np.allclose(da, db, rtol=1e-12, atol=1e-15)


def test_empty_selection_is_degenerate():
    fixed, moving, geom = blob_problem((24, 20, 16), 74)
    tf = affine_for(geom)
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    m.SetFixedImageMask(ImageMaskSpatialObject(np.zeros(geom.shape, dtype=np.uint8), geom))
    with pytest.raises(MetricError) as e:
        m.Initialize()
    assert e.value.code == MMI_ERR_DEGENERATE
    # removing the mask restores the dense metric
    m.SetFixedImageMask(None)
    m.Initialize()
    m.GetValue()


def test_masked_member_in_batch():
    """A masked context in a batch runs its own (masked fast-path) kernels and
    gives the same results as alone; the fused members are unaffected."""
    fixed, moving, geom = blob_problem((40, 36, 30), 75)
    tf1 = affine_for(geom, affine_params(3, np.random.default_rng(76)))
    tf2 = affine_for(geom, affine_params(3, np.random.default_rng(77)))
    plain = _metric(fixed, moving, geom, tf1)
    masked = _metric(fixed, moving, geom, tf2, fmask=_coarse_mask(geom, 78))
    be = BatchEvaluator()
    got = be.evaluate([plain, masked])
    assert be.last_timing()["fused"] == 1
    for m, (v, d, nv, err) in zip([plain, masked], got):
        assert err is None
        v1, d1 = m.GetValueAndDerivative()
        assert (v, nv) == (v1, m.GetNumberOfValidPoints())
        assert np.array_equal(d, d1)
    be.close()


def test_mask_on_half_integer_indices(oracle_mod):
    """A mask on a grid of twice the spacing along x with the image's origin:
    every odd x index maps exactly onto a half-integer mask index, where ITK
    rounds up.  The fast path cannot certify those samples (they go to the
    exact path; so many that the lists overflow and the evaluation is redone
    exactly), and the result still equals the oracle's."""
    fixed, moving, geom = blob_problem((40, 32, 28), 79)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(80)))
    mg = ImageGeometry(size=(21, 32, 28), origin=geom.origin,
                       spacing=(2.0 * geom.spacing[0], geom.spacing[1], geom.spacing[2]))
    rng = np.random.default_rng(81)
    mpx = (rng.uniform(size=mg.shape) < 0.6).astype(np.uint8)
    m = _metric(fixed, moving, geom, tf, grad=grad, fmask=(mpx, mg))
    v, d = m.GetValueAndDerivative()
    assert m.GetLastEvaluationPath() in ("fast", "fast_overflow_exact")
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_gradient=grad, want_samples=True, threads=8,
                          fixed_mask=(mpx, mg))
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    fb, mb = m.GetSampleBins()
    assert np.array_equal(fb, r["fixed_bin"]) and np.array_equal(mb, r["moving_bin"])
    assert abs(v - r["value"]) <= VALUE_RTOL * abs(r["value"])
    assert np.linalg.norm(d - r["derivative"]) <= DERIV_RTOL * np.linalg.norm(r["derivative"])


def test_interface_switches():
    """The reference's On/Off switches and the settings with no device effect."""
    fixed, moving, geom = blob_problem((24, 20, 16), 90)
    tf = affine_for(geom)
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    m.SetMaximumNumberOfWorkUnits(3)
    m.SetUseFixedImageGradientFilter(False)
    assert m.GetMaximumNumberOfWorkUnits() == 3 and not m.GetUseFixedImageGradientFilter()
    assert m.GetMetricCategory() == "IMAGE_METRIC" and m.GetFixedTransform() is None
    m.SetVirtualSampledPointSet(_random_points(geom, 500, 91)[:-2])
    m.UseSampledPointSetOn()
    m.UseVirtualSampledPointSetOn()
    m.Initialize()
    assert m.GetNumberOfDomainPoints() == 500
    m.UseSampledPointSetOff()
    m.Initialize()
    assert m.GetNumberOfDomainPoints() == geom.num_pixels
    m.GetValueAndDerivative()


@pytest.mark.parametrize("case", ["2d_f32", "3d_bf16"])
def test_masks_fast_path_2d_and_16bit(oracle_mod, case):
    """Both masks on the fast path in 2-D, and with BF16 storage (against the
    oracle on the host-rounded images)."""
    shape = (150, 130) if case == "2d_f32" else (40, 36, 30)
    fixed, moving, geom = blob_problem(shape, 92)
    d = geom.dim
    tf = affine_for(geom, affine_params(d, np.random.default_rng(93), scale=2.0))
    fmask, mmask = _coarse_mask(geom, 94), _coarse_mask(geom, 95, factor=3)
    if case == "3d_bf16":
        rnd = lambda a: (np.asarray(a, np.float32).view(np.uint32) + 0x7FFF +
                         ((np.asarray(a, np.float32).view(np.uint32) >> 16) & 1)).astype(np.uint32) \
            & np.uint32(0xFFFF0000)
        r32 = lambda a: rnd(a).view(np.float32)
        grad = oracle_mod.gradient_image(moving, geom)
        grad = r32(grad.astype(np.float32)).astype(np.float64)
        m = Metric(record_sample_bins=True)
        m.SetImageStorage("bf16")
        m.SetFixedImage(fixed, geom)
        m.SetMovingImage(moving, geom)
        m.SetMovingImageGradient(grad)
        m.SetMovingTransform(tf)
        m.SetFixedImageMask(ImageMaskSpatialObject(*fmask))
        m.SetMovingImageMask(ImageMaskSpatialObject(*mmask))
        m.Initialize()
        _check(oracle_mod, m, r32(fixed), r32(moving), geom, tf, grad=grad, fmask=fmask, mmask=mmask)
    else:
        grad = oracle_mod.gradient_image(moving, geom)
        m = _metric(fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)
        _check(oracle_mod, m, fixed, moving, geom, tf, grad=grad, fmask=fmask, mmask=mmask)
