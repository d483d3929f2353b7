"""Pins the CPU oracle (oracle/mattes_oracle.cpp) to the reference's own known
answers and to what the mathematics fixes.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import ndimage

from _cases import (REFERENCE_KNOWN_VALUES, affine_for, affine_params, blob_problem, bspline_for,
                    reference_test_images)
from itk_amd.geometry import ImageGeometry
from itk_amd.transforms import AffineTransform, BSplineTransform


def _perturbed_values(oracle, fixed, moving, geom, tf, k, delta=1e-5):
    base = tf.get_parameters()
    vals = []
    for dd in (delta, -delta, 2 * delta, -2 * delta):
        p = base.copy()
        p[k] += dd
        tf.set_parameters(p)
        vals.append(oracle.mattes(fixed, geom, moving, geom, tf, bins=50, derivative=False)["value"])
    tf.set_parameters(base)
    return vals


def _fd4(vals, delta=1e-5):
    m1p, m1m, m2p, m2m = vals
    # the reference's 4th-order central difference of -value (test.cxx:405-409)
    return -1.0 * ((-1.0 / 12.0) * m2p + (1.0 / 12.0) * m2m + (2.0 / 3.0) * m1p + (-2.0 / 3.0) * m1m) / delta


def test_reference_known_values(oracle_mod):
    """Reference test.cxx:387-401: perturbed-parameter values.  The recorded
    numbers are those of the last parameter; the oracle reproduces them to
    rounding (the reference only demands 5 %)."""
    fixed, moving, geom = reference_test_images()
    tf = AffineTransform(2)
    P = tf.number_of_parameters
    for k in range(P):
        vals = _perturbed_values(oracle_mod, fixed, moving, geom, tf, k)
        for v, kv in zip(vals, REFERENCE_KNOWN_VALUES):
            assert abs(abs(v / kv) - 1.0) <= 0.05
        if k == P - 1:
            for v, kv in zip(vals, REFERENCE_KNOWN_VALUES):
                assert abs(v - kv) <= 4 * abs(kv) * np.finfo(float).eps, (v, kv)


def test_reference_finite_difference_derivative(oracle_mod):
    """Reference test.cxx:349-427: analytic derivative vs 4th-order FD, dense
    tolerance 0.014, central-difference gradients."""
    fixed, moving, geom = reference_test_images()
    tf = AffineTransform(2)
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, bins=50)
    assert r["n_valid"] == 100 * 100
    for k in range(tf.number_of_parameters):
        fd = _fd4(_perturbed_values(oracle_mod, fixed, moving, geom, tf, k))
        assert abs(r["derivative"][k] / fd - 1.0) <= 0.014, (k, r["derivative"][k], fd)


def test_value_only_equals_value_with_derivative(oracle_mod):
    """Reference test.cxx:304-329 (8 ULP)."""
    fixed, moving, geom = reference_test_images()
    tf = AffineTransform(2)
    p = tf.get_parameters()
    p[4] = -3.5
    tf.set_parameters(p)
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf)["value"]
    b = oracle_mod.mattes(fixed, geom, moving, geom, tf, derivative=False)["value"]
    assert abs(a - b) <= 8 * abs(a) * np.finfo(float).eps


def test_constant_images_raise(oracle_mod):
    """Reference test.cxx:477-544: degenerate intensity ranges."""
    with pytest.raises(oracle_mod.OracleError, match="have the same intensity value"):
        oracle_mod.bin_parameters(3.0, 3.0, 50)
    fixed, moving, geom = reference_test_images()
    tf = AffineTransform(2)
    tf.set_parameters([1, 0, 0, 1, 1e6, 0])  # everything maps outside the moving buffer
    with pytest.raises(oracle_mod.OracleError, match="do not sufficiently overlap"):
        oracle_mod.mattes(fixed, geom, moving, geom, tf)


def test_kernel_closed_forms(oracle_mod):
    L = oracle_mod.lib()
    assert L.orc_bspline3(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert L.orc_bspline3(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert L.orc_bspline3(2.0) == 0.0
    assert L.orc_bspline3_derivative(0.0) == 0.0
    assert L.orc_bspline3_derivative(1.0) == pytest.approx(-0.5, abs=1e-15)
    rng = np.random.default_rng(0)
    for x in rng.uniform(0, 1, 50):
        # partition of unity and its derivative
        s = sum(L.orc_bspline3(x - k) for k in range(-2, 3))
        ds = sum(L.orc_bspline3_derivative(x - k) for k in range(-2, 3))
        assert s == pytest.approx(1.0, abs=1e-14)
        assert ds == pytest.approx(0.0, abs=1e-14)
    for u in rng.uniform(-2.2, 2.2, 50):
        h = 1e-6
        num = (L.orc_bspline3(u + h) - L.orc_bspline3(u - h)) / (2 * h)
        assert L.orc_bspline3_derivative(u) == pytest.approx(num, abs=1e-8)
        # the quadratic/cubic closed form of the cubic B-spline
        a = abs(u)
        ref = (2 / 3 - a * a + a ** 3 / 2) if a < 1 else ((2 - a) ** 3 / 6 if a < 2 else 0.0)
        assert L.orc_bspline3(u) == pytest.approx(ref, abs=1e-14)


def test_itk_floor_is_floor(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(1)
    xs = list(rng.uniform(-300, 300, 2000)) + [float(i) for i in range(-5, 6)] + [
        np.nextafter(float(i), -1e9) for i in range(-5, 6) if i != 0] + [-0.5, 0.5, 1.5, -1.5, 255.5]
    for x in xs:
        assert L.orc_itk_floor(float(x)) == math.floor(x), x
    # RoundHalfIntegerToEven(2x - 0.5) >> 1 is not floor for |x| below half an
    # ulp of 0.5: the reference maps tiny negatives to 0.
    assert L.orc_itk_floor(-5e-324) == 0
    assert L.orc_itk_floor(-1e-17) == 0
    assert L.orc_itk_floor(-1e-16) == -1


def test_bin_special_cases(oracle_mod):
    L = oracle_mod.lib()
    B = 50
    tmin, tmax = -0.37, 1.91
    bs, nm = oracle_mod.bin_parameters(tmin, tmax, B)
    assert bs == pytest.approx((tmax - tmin) / (B - 4))
    assert L.orc_parzen_index(tmin, bs, nm, B) == 2
    assert L.orc_parzen_index(tmax, bs, nm, B) == B - 3
    vals = np.linspace(tmin, tmax, 997)
    idx = [L.orc_parzen_index(v, bs, nm, B) for v in vals]
    assert all(b >= a for a, b in zip(idx, idx[1:]))
    assert min(idx) == 2 and max(idx) == B - 3


def test_joint_pdf_properties(oracle_mod):
    fixed, moving, geom = blob_problem((20, 18, 16), 5)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(2)))
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, want_samples=True)
    pdf = r["joint_pdf"]
    assert pdf.sum() == pytest.approx(1.0, abs=1e-12)
    # partition of unity: each sample adds exactly 1 to its fixed row
    assert np.allclose(pdf.sum(axis=1), r["fixed_marginal"], atol=1e-12)
    assert np.allclose(pdf.sum(axis=0), r["moving_marginal"], atol=1e-15)
    fb, mb = r["fixed_bin"], r["moving_bin"]
    assert ((fb >= 2) & (fb <= 47)).sum() == r["n_valid"] == ((mb >= 2) & (mb <= 47)).sum()
    counts = np.bincount(fb[fb >= 0], minlength=50) / r["joint_pdf_sum"]
    assert np.allclose(counts, r["fixed_marginal"], atol=1e-15)


def test_intensity_remap_invariance(oracle_mod):
    """Bins follow the intensity range, so a positive affine remap of the moving
    image leaves the metric and its derivative unchanged."""
    fixed, moving, geom = blob_problem((18, 16, 14), 7)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(3)))
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf)
    moving2 = (moving.astype(np.float64) * 2.5 + 7.0)
    b = oracle_mod.mattes(fixed.astype(np.float64), geom, moving2, geom, tf)
    assert b["value"] == pytest.approx(a["value"], rel=1e-5)
    assert np.linalg.norm(b["derivative"] - a["derivative"]) <= 1e-4 * np.linalg.norm(a["derivative"])


def test_gradient_image_matches_central_difference_on_ramp(oracle_mod):
    """For a moving image linear in space, linear interpolation is exact, so the
    gradient-image source (a constant field) and central differences agree."""
    n = 24
    geom = ImageGeometry(size=(n, n, n), spacing=(1.0, 1.5, 0.75), origin=(2.0, -1.0, 0.5))
    fixed, _, _ = blob_problem((n, n, n), 9)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64)] * 3, indexing="ij")
    slope = np.array([0.3, -0.2, 0.45])  # per physical unit
    px = geom.origin[0] + geom.spacing[0] * x
    py = geom.origin[1] + geom.spacing[1] * y
    pz = geom.origin[2] + geom.spacing[2] * z
    moving = (slope[0] * px + slope[1] * py + slope[2] * pz).astype(np.float64)
    grad = np.broadcast_to(slope, moving.shape + (3,)).copy()
    virt = ImageGeometry(size=(n - 8, n - 8, n - 8), start=(4, 4, 4), spacing=geom.spacing, origin=geom.origin)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(4), scale=0.3))
    a = oracle_mod.mattes(fixed.astype(np.float64), geom, moving, geom, tf, virtual_geom=virt)
    b = oracle_mod.mattes(fixed.astype(np.float64), geom, moving, geom, tf, virtual_geom=virt,
                          moving_gradient=grad)
    assert a["n_valid"] == b["n_valid"] == virt.num_pixels
    assert b["value"] == a["value"]
    assert np.allclose(b["derivative"], a["derivative"], rtol=1e-9, atol=1e-12)


def _ramp_problem(shape, seed):
    """Smooth blob fixed image and a moving image linear in every axis.  Linear
    interpolation and central differences are then exact inside the buffer,
    so Mattes' analytic derivative must equal the derivative of the value:
    the finite difference pins the Jacobian, the dPDF accumulation and the
    contraction with log(p/pM) to ~1e-7 instead of the reference's 1.4 %."""
    from itk_amd import synth

    fixed = synth.blob_image(shape, 6, seed, noise=0.0).astype(np.float64)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in reversed(shape)], indexing="ij")
    moving = 0.3 * x - 0.2 * y + 0.45 * z
    geom = ImageGeometry(size=shape)
    virt = ImageGeometry(size=tuple(s - 12 for s in shape), start=(6, 6, 6))
    return fixed, moving, geom, virt


def _fd_check(oracle_mod, fixed, moving, geom, virt, tf, params, rel):
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, virtual_geom=virt)
    assert r["n_valid"] == virt.num_pixels
    base = tf.get_parameters()
    scale = np.abs(r["derivative"]).max()
    for k in params:
        vals = []
        for dd in (1e-5, -1e-5, 2e-5, -2e-5):
            p = base.copy()
            p[k] += dd
            tf.set_parameters(p)
            vals.append(oracle_mod.mattes(fixed, geom, moving, geom, tf, derivative=False, virtual_geom=virt)["value"])
        tf.set_parameters(base)
        assert r["derivative"][k] == pytest.approx(_fd4(vals), rel=rel, abs=rel * scale), k
    return r


def test_affine_3d_finite_difference(oracle_mod):
    fixed, moving, geom, virt = _ramp_problem((32, 30, 28), 11)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(5), scale=0.5))
    _fd_check(oracle_mod, fixed, moving, geom, virt, tf, range(12), 1e-6)


def test_bspline_finite_difference(oracle_mod):
    fixed, moving, geom, virt = _ramp_problem((28, 28, 28), 13)
    tf = bspline_for(geom, (2, 2, 2), seed=6, sigma=0.5)
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, virtual_geom=virt)
    d = r["derivative"]
    assert d.shape == (3 * 5 ** 3,)
    _fd_check(oracle_mod, fixed, moving, geom, virt, tf, np.argsort(-np.abs(d))[:10], 1e-5)


def test_thread_count_invariance(oracle_mod):
    fixed, moving, geom = blob_problem((16, 16, 24), 15)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(7)))
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf, threads=1)
    b = oracle_mod.mattes(fixed, geom, moving, geom, tf, threads=5)
    assert a["n_valid"] == b["n_valid"]
    assert b["value"] == pytest.approx(a["value"], rel=1e-13)
    assert np.allclose(a["derivative"], b["derivative"], rtol=1e-11, atol=1e-14)


def test_affine_transform_point(oracle_mod):
    rng = np.random.default_rng(8)
    tf = AffineTransform(3, center=[3.0, -2.0, 7.5])
    tf.set_parameters(affine_params(3, rng))
    for _ in range(20):
        x = rng.uniform(-50, 50, 3)
        ref = tf.matrix @ (x - tf.center) + tf.center + tf.translation
        assert np.allclose(oracle_mod.transform_point(tf, x), ref, rtol=1e-13, atol=1e-12)


def test_bspline_transform_partition_of_unity(oracle_mod):
    geom = ImageGeometry(size=(30, 30, 30), spacing=(1.0, 2.0, 1.5))
    tf = BSplineTransform.for_image(geom, (3, 4, 5))
    n = tf.number_of_parameters_per_dimension
    shift = np.array([0.7, -1.3, 2.1])
    tf.set_parameters(np.concatenate([np.full(n, s) for s in shift]))
    rng = np.random.default_rng(9)
    for _ in range(30):
        x = geom.index_to_point(rng.uniform(0, 28.9, 3))
        assert np.allclose(oracle_mod.transform_point(tf, x), x + shift, atol=1e-12)
    far = np.array([1e4, 0.0, 0.0])
    assert np.array_equal(oracle_mod.transform_point(tf, far), far)


def test_linear_interpolation_reproduces_linear_functions(oracle_mod):
    geom = ImageGeometry(size=(9, 8, 7), spacing=(1.0, 2.0, 0.5), origin=(1.0, 2.0, 3.0))
    z, y, x = np.meshgrid(np.arange(7.0), np.arange(8.0), np.arange(9.0), indexing="ij")
    img = 1.5 * x - 0.25 * y + 3.0 * z + 0.5
    rng = np.random.default_rng(10)
    for _ in range(50):
        ci = rng.uniform(-0.49, [8.49, 7.49, 6.49])
        p = geom.index_to_point(ci)
        v = oracle_mod.interpolate(img, geom, p)
        cc = np.clip(ci, 0, [8, 7, 6])  # values beyond the last pixel are clamped
        assert v == pytest.approx(1.5 * cc[0] - 0.25 * cc[1] + 3.0 * cc[2] + 0.5, abs=1e-12)
    assert oracle_mod.interpolate(img, geom, geom.index_to_point([8.5, 0, 0])) is None
    assert oracle_mod.interpolate(img, geom, geom.index_to_point([-0.5, 0, 0])) is not None


def test_recursive_gaussian_closed_forms(oracle_mod):
    n = 64
    geom = ImageGeometry(size=(n, n, 8))
    const = np.full((8, n, n), 3.25)
    sm = oracle_mod.smooth(const, geom, 2.0)
    assert np.allclose(sm, 3.25, rtol=1e-12)
    # gradient of a ramp: exact slope away from the edge-extended borders
    z, y, x = np.meshgrid(np.arange(8.0), np.arange(n * 1.0), np.arange(n * 1.0), indexing="ij")
    ramp = 0.7 * x - 0.2 * y
    g = oracle_mod.gradient_image(ramp, geom, sigma=1.0)
    inner = g[:, 20:44, 20:44]
    assert np.allclose(inner[..., 0], 0.7, rtol=1e-9)
    assert np.allclose(inner[..., 1], -0.2, rtol=1e-9)
    assert np.allclose(inner[..., 2], 0.0, atol=1e-12)
    # a linear function survives smoothing in the interior (symmetric kernel)
    sr = oracle_mod.smooth(ramp, geom, 1.5)
    assert np.allclose(sr[:, 20:44, 20:44], ramp[:, 20:44, 20:44], rtol=1e-9, atol=1e-9)


def test_recursive_gaussian_approximates_gaussian(oracle_mod):
    """Farneback-Westin coefficients approximate a true Gaussian to ~1e-3."""
    n = 96
    geom = ImageGeometry(size=(n, n))
    rng = np.random.default_rng(11)
    img = ndimage.gaussian_filter(rng.normal(size=(n, n)), 3.0)
    sig = 2.5
    ours = oracle_mod.smooth(img, geom, sig)
    ref = ndimage.gaussian_filter(img, sig, mode="nearest", truncate=8.0)
    inner = (slice(20, -20), slice(20, -20))

    def rel(a, b):
        return np.linalg.norm(a[inner] - b[inner]) / np.linalg.norm(b[inner])

    assert rel(ours, ref) <= 1.5e-2
    # first order, normalised across scale: sigma * dG/dx
    gx = oracle_mod.gradient_image(img, geom, sigma=sig)[..., 0]
    refg = ndimage.gaussian_filter(img, sig, order=(0, 1), mode="nearest", truncate=8.0) * sig
    assert rel(gx, refg) <= 2e-2


def test_bspline_node_layout_single_coefficient(oracle_mod):
    """One coefficient (dimension 1, node (2, 3, 1)) set, all others zero: the
    displacement is c * B3(u_x - 2) B3(u_y - 3) B3(u_z - 1) in dimension 1 only,
    u the point's continuous index in the grid (itkBSplineTransform.hxx:508-590,
    parameters dimension-major, nodes x fastest).  A permuted node layout or
    dimension order moves the bump elsewhere."""
    geom = ImageGeometry(size=(20, 18, 16), spacing=(1.0, 1.5, 2.0), origin=(-3.0, 1.0, 2.0))
    tf = BSplineTransform.for_image(geom, [4, 3, 2])
    G = tf.grid
    node, dim, c = (2, 3, 1), 1, 0.37
    p = np.zeros(tf.number_of_parameters)
    p[dim * G.num_pixels + node[0] + G.size[0] * (node[1] + G.size[1] * node[2])] = c

    def b3(u):
        u = abs(u)
        return (4 - 6 * u * u + 3 * u ** 3) / 6 if u < 1 else ((2 - u) ** 3 / 6 if u < 2 else 0.0)

    tf.set_parameters(p)
    rng = np.random.default_rng(12)
    inside_support = 0
    for _ in range(300):
        x = geom.index_to_point(rng.uniform(0, np.array(geom.size) - 1))
        u = (np.array(x) - np.array(G.origin)) / np.array(G.spacing)
        w = b3(u[0] - node[0]) * b3(u[1] - node[1]) * b3(u[2] - node[2])
        inside_support += w > 0
        expect = np.zeros(3)
        expect[dim] = c * w
        assert np.allclose(np.array(oracle_mod.transform_point(tf, x)) - np.array(x), expect, atol=1e-13)
    assert inside_support > 50


def test_recursive_gaussian_borders_extend_the_edge(oracle_mod):
    """Near the line ends the recursive filters (zero and first order, with the
    border coefficients BN / BM of itkRecursiveGaussianImageFilter.hxx:56-300)
    follow a Gaussian of the edge-extended line: within 1 % of scipy's
    mode='nearest' in a 6-pixel band at both ends, where zero, reflected or
    mirrored extension differ by 5-90 % (the first order from reflected or
    zero extension, the zero order from zero extension)."""
    n = 96
    geom = ImageGeometry(size=(n, n))
    rng = np.random.default_rng(13)
    img = ndimage.gaussian_filter(rng.normal(size=(n, n)), 3.0)
    sig = 2.5
    band = (slice(20, -20), np.r_[0:6, n - 6:n])

    def rel(a, b):
        return np.linalg.norm(a[band] - b[band]) / np.linalg.norm(b[band])

    gx = oracle_mod.gradient_image(img, geom, sigma=sig)[..., 0]
    sm = oracle_mod.smooth(img, geom, sig)
    for mode in ("nearest", "reflect", "constant"):
        refg = ndimage.gaussian_filter(img, sig, order=(0, 1), mode=mode, truncate=8.0) * sig
        ref0 = ndimage.gaussian_filter(img, sig, mode=mode, truncate=8.0)
        if mode == "nearest":
            assert rel(gx, refg) <= 1e-2 and rel(sm, ref0) <= 1e-2
        else:
            assert rel(gx, refg) >= 0.05
            if mode == "constant":
                assert rel(sm, ref0) >= 0.05


def _rotation_stage(geom, angle, scale, shift):
    """An affine transform rotating about the image centre (x-y plane),
    scaled slightly, then shifted: a moving initial transform."""
    c = geom.index_to_point([(s - 1) / 2.0 for s in geom.size])
    ca, sa = np.cos(angle), np.sin(angle)
    m = np.eye(3)
    m[:2, :2] = [[ca, -sa], [sa, ca]]
    m = m * scale
    tf = affine_for(geom)
    tf.center = np.asarray(c, dtype=np.float64)
    tf.set_parameters(np.concatenate([m.ravel(), shift]))
    return tf


def test_composite_with_identity_stage_is_the_moving_transform(oracle_mod):
    """A composite whose earlier entry is the identity affine maps every point
    and forms every Jacobian column with exact products by 1 and 0: value and
    derivative equal the moving transform's alone, bit for bit."""
    from itk_amd.transforms import CompositeTransform

    fixed, moving, geom = blob_problem((30, 28, 26), 41)
    for tf in (affine_for(geom, affine_params(3, np.random.default_rng(42), scale=1.0)),
               bspline_for(geom, (3, 3, 2), seed=43, sigma=0.6)):
        ident = affine_for(geom)
        comp = CompositeTransform(3, [ident, tf])
        a = oracle_mod.mattes(fixed, geom, moving, geom, tf, want_samples=True)
        b = oracle_mod.mattes(fixed, geom, moving, geom, comp, want_samples=True)
        assert a["value"] == b["value"] and a["n_valid"] == b["n_valid"]
        assert np.array_equal(a["derivative"], b["derivative"])
        assert np.array_equal(a["moving_bin"], b["moving_bin"])


def test_composite_maps_through_the_earlier_entry(oracle_mod):
    """The composite's samples are the moving transform's points mapped by
    the earlier entry: with a pure translation stage of a whole number of
    voxels on a moving image shifted the same way, every sample's value, bin
    and the metric value are those of the moving transform alone."""
    from itk_amd.transforms import CompositeTransform

    fixed, moving, geom = blob_problem((30, 28, 26), 44)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(45), scale=0.5))
    shift = np.array([3.0, -2.0, 4.0]) * np.asarray(geom.spacing)
    stage = affine_for(geom)
    stage.set_parameters(np.concatenate([np.eye(3).ravel(), shift]))
    moved = ImageGeometry(size=geom.size, origin=tuple(np.asarray(geom.origin) + shift), spacing=geom.spacing,
                          direction=geom.direction)
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf, want_samples=True, derivative=False)
    b = oracle_mod.mattes(fixed, geom, moving, moved, CompositeTransform(3, [stage, tf]), want_samples=True,
                          derivative=False)
    assert a["n_valid"] == b["n_valid"]
    assert np.array_equal(a["moving_bin"], b["moving_bin"]) and np.array_equal(a["fixed_bin"], b["fixed_bin"])
    assert b["value"] == pytest.approx(a["value"], rel=1e-12)


def test_composite_affine_finite_difference(oracle_mod):
    """The composite's derivative (stage matrix times the affine Jacobian) is
    the derivative of its value."""
    from itk_amd.transforms import CompositeTransform

    fixed, moving, geom, virt = _ramp_problem((32, 30, 28), 46)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(47), scale=0.5))
    comp = CompositeTransform(3, [_rotation_stage(geom, 0.08, 1.03, np.array([0.4, -0.3, 0.2])), tf])
    _fd_check(oracle_mod, fixed, moving, geom, virt, comp, range(12), 1e-6)


def test_composite_bspline_finite_difference(oracle_mod):
    from itk_amd.transforms import CompositeTransform

    fixed, moving, geom, virt = _ramp_problem((28, 28, 28), 48)
    tf = bspline_for(geom, (2, 2, 2), seed=49, sigma=0.5)
    comp = CompositeTransform(3, [_rotation_stage(geom, -0.06, 0.98, np.array([-0.3, 0.2, 0.1])), tf])
    r = oracle_mod.mattes(fixed, geom, moving, geom, comp, virtual_geom=virt)
    d = r["derivative"]
    _fd_check(oracle_mod, fixed, moving, geom, virt, comp, np.argsort(-np.abs(d))[:10], 1e-5)
