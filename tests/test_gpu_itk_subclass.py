"""The ITK metric subclass (include/itkMattesMutualInformationImageToImageMetricv4HIP.h,
compiled against the ITK interface mock) on the GPU: its Initialize and
GetValueAndDerivative through the C++ layer and the C ABI give the bits of the
Python binding on the same inputs (both drive the same library), and the
oracle's valid count, value and derivative to the parity bar.  Cases: the
reference unit test's Image<double, 2> inputs with central differences,
Image<float, 3> affine with the gradient filter, Image<float, 3> cubic
B-spline; each transform inside a one-entry CompositeTransform, as the
registration method hands it over, and two parameter sets per case."""
from __future__ import annotations

import os

import numpy as np
import pytest

import _itk_subclass
from _cases import affine_for, affine_params, blob_problem, bspline_for, reference_test_images
from _gpu import DERIV_RTOL, VALUE_RTOL
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric
from itk_amd.metric import transform_fixed_parameters

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "2d_f64_reference_test":
        fixed, moving, geom = reference_test_images(np.float64)
        tf = affine_for(geom)
        rng = np.random.default_rng(800)
        sets = [affine_params(2, rng, 0.5), affine_params(2, rng, 0.5)]
        return fixed, moving, geom, tf, sets, False
    if name == "3d_f32_affine":
        fixed, moving, geom = blob_problem((48, 44, 40), 801)
        tf = affine_for(geom)
        rng = np.random.default_rng(802)
        sets = [affine_params(3, rng, 1.5), affine_params(3, rng, 1.5)]
        return fixed, moving, geom, tf, sets, True
    fixed, moving, geom = blob_problem((40, 36, 32), 803)
    tf = bspline_for(geom, (4, 4, 3), seed=804, sigma=0.6)
    rng = np.random.default_rng(805)
    sets = [tf.get_parameters(), tf.get_parameters() + rng.normal(0, 0.1, tf.number_of_parameters)]
    return fixed, moving, geom, tf, sets, True


@pytest.mark.parametrize("name", ["2d_f64_reference_test", "3d_f32_affine", "3d_f32_bspline"])
def test_itk_subclass_matches_binding_and_oracle(oracle_mod, tmp_path, name):
    if not os.path.exists(_itk_subclass.BIN):
        pytest.fail("tests/cpp/itk_subclass_check is not built: run the build step (make -C tests/cpp)")
    fixed, moving, geom, tf, sets, use_filter = _case(name)
    fp = transform_fixed_parameters(tf)
    if tf.kind == 2:
        fp = fp[: 3 * geom.dim + geom.dim * geom.dim]  # ITK's BSplineTransform fixed parameters
    print("case", name, flush=True)
    rc, out, err = _itk_subclass.run(tmp_path, fixed, moving, geom, "affine" if tf.kind == 1 else "bspline", fp, sets,
                                     use_filter=use_filter)
    assert rc == 0, (out, err)
    assert len(out) == len(sets)
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetUseMovingImageGradientFilter(use_filter)
    m.SetMovingTransform(tf)
    m.Initialize()
    grad = oracle_mod.gradient_image(moving, geom) if use_filter else None
    for p, got in zip(sets, out):
        tf.set_parameters(p)
        v, d = m.GetValueAndDerivative()
        assert got["n_valid"] == m.GetNumberOfValidPoints()
        assert got["value"] == v
        assert np.array_equal(np.array(got["derivative"]), d)
        r = oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_gradient=grad, threads=8)
        assert got["n_valid"] == r["n_valid"]
        assert abs(got["value"] - r["value"]) <= VALUE_RTOL * abs(r["value"])
        assert np.linalg.norm(np.array(got["derivative"]) - r["derivative"]) <= DERIV_RTOL * np.linalg.norm(
            r["derivative"])
    m.close()


def test_itk_subclass_composite_with_initial_transform(oracle_mod, tmp_path):
    """A CompositeTransform of an affine moving initial transform and the
    optimized B-spline (only the most recent entry optimized, as
    ImageRegistrationMethodv4 sets it up): the subclass hands the initial
    entry over as a matrix-offset stage and matches the binding bit for bit
    and the oracle to the bar."""
    from itk_amd.transforms import CompositeTransform

    fixed, moving, geom = blob_problem((40, 36, 32), 806)
    tf = bspline_for(geom, (4, 4, 3), seed=807, sigma=0.5)
    init = affine_for(geom, affine_params(3, np.random.default_rng(808), scale=0.6))
    comp = CompositeTransform(3, [init, tf])
    rng = np.random.default_rng(809)
    sets = [tf.get_parameters(), tf.get_parameters() + rng.normal(0, 0.1, tf.number_of_parameters)]
    fp = transform_fixed_parameters(tf)[: 3 * 3 + 9]
    rc, out, err = _itk_subclass.run(tmp_path, fixed, moving, geom, "bspline", fp, sets, initial=[init])
    assert rc == 0, (out, err)
    assert len(out) == len(sets)
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(comp)
    m.Initialize()
    grad = oracle_mod.gradient_image(moving, geom)
    for p, got in zip(sets, out):
        tf.set_parameters(p)
        v, d = m.GetValueAndDerivative()
        assert got["n_valid"] == m.GetNumberOfValidPoints()
        assert got["value"] == v
        assert np.array_equal(np.array(got["derivative"]), d)
        r = oracle_mod.mattes(fixed, geom, moving, geom, comp, moving_gradient=grad, threads=8)
        assert got["n_valid"] == r["n_valid"]
        assert abs(got["value"] - r["value"]) <= VALUE_RTOL * abs(r["value"])
        assert np.linalg.norm(np.array(got["derivative"]) - r["derivative"]) <= DERIV_RTOL * np.linalg.norm(
            r["derivative"])
    m.close()
