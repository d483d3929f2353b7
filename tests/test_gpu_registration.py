"""Whole registrations on the device metric against the same registrations on
the oracle: ImageRegistrationMethodv4 with per-level shrink and smoothing, the
RegularStep gradient-descent and L-BFGS-B optimizers.

Bar: final transforms within 0.1 voxel mean displacement of the oracle run's
(the bins match exactly and derivatives to ~1e-6, so trajectories stay
together).  The recovered transform's distance to the generating affine is
printed and loosely bounded.
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for
from _oracle_metric import OracleMattesMetric, mean_displacement
from itk_amd import synth
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric
from itk_amd.optimizers import LBFGSBOptimizerv4, RegularStepGradientDescentOptimizerv4
from itk_amd.registration import ImageRegistrationMethodv4
from itk_amd.scales import RegistrationParameterScalesFromPhysicalShift
from itk_amd.transforms import AffineTransform, BSplineTransform

pytestmark = pytest.mark.gpu


def truth_transform(problem) -> AffineTransform:
    """The transform registration should find: T(x) = G^-1 (x - c - t) + c for
    moving(y) = fixed(G (y - c) + c + t)."""
    tr = problem.truth
    ginv = np.linalg.inv(tr.matrix)
    d = ginv.shape[0]
    tf = AffineTransform(d, center=tr.center)
    tf.set_parameters(np.concatenate([ginv.reshape(-1), -ginv @ tr.translation]))
    return tf


def run_affine(problem, metric, iterations, shrink, sigmas, lr=2.0, sampling=None):
    geom = problem.geometry
    tf = affine_for(geom)
    opt = RegularStepGradientDescentOptimizerv4(learning_rate=lr, minimum_step_length=1e-3,
                                                relaxation_factor=0.5, number_of_iterations=iterations,
                                                scales_estimator=RegistrationParameterScalesFromPhysicalShift())
    reg = ImageRegistrationMethodv4()
    reg.SetFixedImage(problem.fixed, geom)
    reg.SetMovingImage(problem.moving, geom)
    reg.SetMetric(metric)
    reg.SetOptimizer(opt)
    reg.SetInitialTransform(tf)
    reg.SetShrinkFactorsPerLevel(shrink)
    reg.SetSmoothingSigmasPerLevel(sigmas)
    reg.SetSmoothingSigmasAreSpecifiedInPhysicalUnits(False)
    if sampling is not None:
        strategy, pct, seed, mask = sampling
        reg.SetMetricSamplingStrategy(strategy)
        reg.SetMetricSamplingPercentage(pct)
        reg.MetricSamplingReinitializeSeed(seed)
        reg.SetFixedImageMask(mask)
    reg.Update()
    return tf, reg.level_log


@pytest.mark.parametrize("shape,iterations", [((128, 112), 60), ((48, 44, 40), 30)])
def test_affine_registration_matches_oracle(oracle_mod, shape, iterations):
    p = synth.make_affine_problem(shape, 51)
    shrink, sigmas = [2, 1], [1.0, 0.0]
    tf_gpu, log_gpu = run_affine(p, Metric(), iterations, shrink, sigmas)
    tf_cpu, log_cpu = run_affine(p, OracleMattesMetric(oracle_mod), iterations, shrink, sigmas)
    for a, b in zip(log_gpu, log_cpu):
        print("level %d: gpu %d its value %.6f | oracle %d its value %.6f"
              % (a["level"], a["iterations"], a["value"], b["iterations"], b["value"]))
    stride = 1 if len(shape) == 2 else 4
    d_oracle = mean_displacement(tf_gpu, tf_cpu, p.geometry, stride)
    d_truth = mean_displacement(tf_gpu, truth_transform(p), p.geometry, stride)
    print("mean displacement vs oracle run %.3e voxel, vs generating transform %.3f voxel" % (d_oracle, d_truth))
    assert d_oracle <= 0.1
    assert d_truth <= 1.5


@pytest.mark.parametrize("shape,strategy,pct,masked", [((128, 112), "random", 0.25, True),
                                                        ((48, 44, 40), "regular", 0.2, False)])
def test_sampled_registration_matches_oracle(oracle_mod, shape, strategy, pct, masked):
    """REGULAR / RANDOM metric sampling (with a fixed mask in 2-D): the same
    point sets per level on both metrics, final transforms within 0.1 voxel."""
    from itk_amd import sampling
    from itk_amd.masks import ImageMaskSpatialObject

    p = synth.make_affine_problem(shape, 53)
    mask = None
    if masked:
        mpx = np.zeros(p.geometry.shape, dtype=np.uint8)
        sl = tuple(slice(s // 8, s - s // 8) for s in p.geometry.shape)
        mpx[sl] = 1
        mask = ImageMaskSpatialObject(mpx, p.geometry)
    code = sampling.RANDOM if strategy == "random" else sampling.REGULAR
    shrink, sigmas = [2, 1], [1.0, 0.0]
    logs, tfs = [], []
    for metric in (Metric(), OracleMattesMetric(oracle_mod)):
        tf, log = run_affine(p, metric, 40, shrink, sigmas, sampling=(code, pct, 20240, mask))
        tfs.append(tf)
        logs.append(log)
    for a, b in zip(*logs):
        print("level %d: %s samples, gpu %d its value %.6f | oracle %d its value %.6f"
              % (a["level"], a["samples"], a["iterations"], a["value"], b["iterations"], b["value"]))
        assert a["samples"] > 0
    stride = 1 if len(shape) == 2 else 4
    d_oracle = mean_displacement(tfs[0], tfs[1], p.geometry, stride)
    d_truth = mean_displacement(tfs[0], truth_transform(p), p.geometry, stride)
    print("mean displacement vs oracle run %.3e voxel, vs generating transform %.3f voxel" % (d_oracle, d_truth))
    assert d_oracle <= 0.1
    assert d_truth <= 2.0


def test_bspline_registration_lbfgsb_matches_oracle(oracle_mod):
    p = synth.make_bspline_problem((36, 34, 32), 52, mesh=4, sigma_vox=1.5, n_blobs=20)
    geom = p.geometry
    finals = []
    for metric in (Metric(), OracleMattesMetric(oracle_mod)):
        tf = BSplineTransform.for_image(geom, (4, 4, 4))
        opt = LBFGSBOptimizerv4(number_of_iterations=12)
        reg = ImageRegistrationMethodv4()
        reg.SetFixedImage(p.fixed, geom)
        reg.SetMovingImage(p.moving, geom)
        reg.SetMetric(metric)
        reg.SetOptimizer(opt)
        reg.SetInitialTransform(tf)
        reg.SetShrinkFactorsPerLevel([1])
        reg.SetSmoothingSigmasPerLevel([0.0])
        reg.Update()
        print("L-BFGS-B: %d iterations, %d evaluations, value %.6f"
              % (opt.GetCurrentIteration(), opt.evaluations, opt.GetCurrentMetricValue()))
        finals.append(tf)
    total = 0.0
    count = 0
    for idx in np.ndindex(*[s // 4 for s in geom.size]):
        q = geom.index_to_point([4 * i for i in idx])
        a = oracle_mod.transform_point(finals[0], q)
        b = oracle_mod.transform_point(finals[1], q)
        total += float(np.linalg.norm(a - b))
        count += 1
    d = total / count
    print("B-spline mean displacement vs oracle run %.3e voxel" % d)
    assert d <= 0.1
