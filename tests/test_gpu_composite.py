"""Composite moving transforms on the HIP path against the oracle: a moving
initial transform (affine, or a chain of two) before the optimized affine or
B-spline transform, as ImageRegistrationMethodv4 composes them
(itkImageRegistrationMethodv4.hxx:344-355; itkCompositeTransform.hxx:60-71,
467-560).  Bar: bins and valid count bit-exact, value 1e-5, derivative 1e-4
(L2); the derivative takes the composite's Jacobian, so it differs from the
optimized transform's alone."""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem, bspline_for
from _gpu import _compare, _gpu_metric
from itk_amd.transforms import CompositeTransform

pytestmark = pytest.mark.gpu


def _stage(geom, angle, scale, shift):
    c = geom.index_to_point([(s - 1) / 2.0 for s in geom.size])
    ca, sa = np.cos(angle), np.sin(angle)
    m = np.eye(3)
    m[1:, 1:] = [[ca, -sa], [sa, ca]]
    tf = affine_for(geom)
    tf.center = np.asarray(c, dtype=np.float64)
    tf.set_parameters(np.concatenate([(m * scale).ravel(), shift]))
    return tf


@pytest.mark.parametrize("use_filter", [True, False])
def test_affine_initial_then_affine(oracle_mod, use_filter):
    fixed, moving, geom = blob_problem((48, 44, 40), 510)
    grad = oracle_mod.gradient_image(moving, geom) if use_filter else None
    tf = affine_for(geom, affine_params(3, np.random.default_rng(511), scale=1.0))
    comp = CompositeTransform(3, [_stage(geom, 0.07, 1.02, np.array([0.8, -0.5, 0.3])), tf])
    m = _gpu_metric(fixed, moving, geom, comp, gradient=grad, use_filter=use_filter)
    _, d, _ = _compare(oracle_mod, m, fixed, moving, geom, comp, gradient=grad)
    assert m.GetLastEvaluationPath() == "exact"
    # the stage's Jacobian matters: the optimized transform alone differs
    r1 = oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_gradient=grad, threads=8)
    assert np.linalg.norm(d - r1["derivative"]) > 1e-3 * np.linalg.norm(r1["derivative"])
    # a new parameter vector after Initialize (the optimizer's step)
    tf.set_parameters(affine_params(3, np.random.default_rng(512), scale=1.0))
    _compare(oracle_mod, m, fixed, moving, geom, comp, gradient=grad)
    m.close()


def test_two_initial_stages_then_bspline(oracle_mod):
    fixed, moving, geom = blob_problem((44, 40, 36), 513)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = bspline_for(geom, (4, 4, 3), seed=514, sigma=0.6)
    comp = CompositeTransform(3, [_stage(geom, -0.05, 1.0, np.array([0.2, 0.4, -0.6])),
                                  _stage(geom, 0.04, 0.99, np.array([-0.3, 0.1, 0.2])), tf])
    m = _gpu_metric(fixed, moving, geom, comp, gradient=grad)
    _compare(oracle_mod, m, fixed, moving, geom, comp, gradient=grad)
    m.close()


def test_one_entry_composite_is_the_transform_itself(oracle_mod):
    """A composite holding only the optimized transform takes the fast path
    and gives the transform's own results bit for bit."""
    fixed, moving, geom = blob_problem((48, 44, 40), 515)
    grad = oracle_mod.gradient_image(moving, geom)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(516), scale=1.0))
    a = _gpu_metric(fixed, moving, geom, tf, gradient=grad)
    b = _gpu_metric(fixed, moving, geom, CompositeTransform(3, [tf]), gradient=grad)
    va, da = a.GetValueAndDerivative()
    vb, db = b.GetValueAndDerivative()
    assert b.GetLastEvaluationPath() == "fast"
    assert va == vb and np.array_equal(da, db)
    a.close()
    b.close()
