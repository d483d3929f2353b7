"""Batched evaluation (mmi_batch): every context's results equal its own
evaluation bit for bit — fused fast-path contexts keep their own workgroup
partition, the others run the single-context code — across graph replays with
eager work in between, mixed with contexts off the fast path and with a
degenerate member."""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem, bspline_for
from itk_amd._native import MMI_ERR_DEGENERATE, MetricError
from itk_amd.batch import BatchEvaluator
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

pytestmark = pytest.mark.gpu


def _metric(fixed, moving, geom, tf, **kw):
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    if kw.get("virtual") is not None:
        m.SetVirtualDomain(kw["virtual"])
    if kw.get("cd"):
        m.SetUseMovingImageGradientFilter(False)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


def _problems():
    out = []
    for k, shape in enumerate([(48, 44, 40), (64, 60, 56), (40, 36, 34), (56, 50, 44)]):
        fixed, moving, geom = blob_problem(shape, 70 + k)
        tf = affine_for(geom, affine_params(3, np.random.default_rng(80 + k), scale=2.0))
        out.append((fixed, moving, geom, tf))
    return out


def _alone(m):
    v, d = m.GetValueAndDerivative()
    return v, d, m.GetNumberOfValidPoints(), m.GetLastEvaluationPath()


def test_fused_batch_equals_single_evaluations():
    probs = _problems()
    ms = [_metric(*p) for p in probs]
    ms.append(_metric(*probs[1][:3], probs[1][3], virtual=probs[1][2].shrink(2)))  # coarse level, fast path
    be = BatchEvaluator()
    rng = np.random.default_rng(90)
    for rep in range(3):  # replays of one captured graph, with eager work between them
        for (_, _, geom, tf), m in zip(probs + [probs[1]], ms):
            tf.set_parameters(affine_params(3, rng, scale=2.0))
        got = be.evaluate(ms)
        t = be.last_timing()
        assert t["fused"] == len(ms)
        for m, (v, d, nv, err) in zip(ms, got):
            assert err is None
            v1, d1, nv1, path = _alone(m)
            assert path == "fast"
            assert nv == nv1
            assert v == v1
            assert np.array_equal(d, d1)
    be.close()


def test_mixed_batch_and_degenerate_member():
    probs = _problems()
    fixed, moving, geom, tf = probs[0]
    members = [
        _metric(*probs[0]),                                    # fused
        _metric(*probs[2][:3], probs[2][3], cd=True),          # central differences: single
        _metric(fixed, moving, geom, bspline_for(geom, (3, 3, 3), seed=91, sigma=0.6)),  # B-spline: single
        _metric(*probs[3]),                                    # fused
    ]
    far = affine_for(geom, affine_params(3, np.random.default_rng(92)))
    p = far.get_parameters()
    p[-3:] += 1e4
    far.set_parameters(p)
    members.append(_metric(fixed, moving, geom, far))          # every sample outside: degenerate (fused)
    members.append(_metric(fixed, moving, geom, far, cd=True))  # the same off the fused launch (worker thread)
    be = BatchEvaluator()
    got = be.evaluate(members)
    assert be.last_timing()["fused"] == 3
    for m, (v, d, nv, err) in zip(members[:4], got[:4]):
        assert err is None
        v1, d1, nv1, _ = _alone(m)
        assert (v, nv) == (v1, nv1)
        assert np.array_equal(d, d1)
    for k in (4, 5):
        err = got[k][3]
        assert err is not None and err.code == MMI_ERR_DEGENERATE
        # the batch reports the reference's exception text, as the member's own evaluation raises it
        with pytest.raises(MetricError) as alone:
            members[k].GetValueAndDerivative()
        assert str(err) == str(alone.value)
        assert "All samples map outside moving image buffer" in str(err)
    be.close()


def test_batch_mixed_histogram_sizes():
    """Members with 32, 50 and 80 bins in one fused launch: pass 1's LDS holds
    two histogram copies for the smaller ones and one for 80 bins, and the
    launch is sized for the largest; every member equals its own evaluation."""
    probs = _problems()
    ms = []
    for (fixed, moving, geom, tf), bins in zip(probs[:3], (32, 80, 50)):
        m = _metric(fixed, moving, geom, tf)
        m.SetNumberOfHistogramBins(bins)
        m.Initialize()
        ms.append(m)
    be = BatchEvaluator()
    for rep in range(2):
        got = be.evaluate(ms)
        assert be.last_timing()["fused"] == 3
        for m, (v, d, nv, err) in zip(ms, got):
            assert err is None
            v1, d1, nv1, path = _alone(m)
            assert path == "fast"
            assert (v, nv) == (v1, nv1)
            assert np.array_equal(d, d1)
    be.close()


def test_batch_value_only_and_without_graphs():
    probs = _problems()
    ms = [_metric(*p) for p in probs[:3]]
    be = BatchEvaluator(graphs=False)
    got = be.evaluate(ms, want_derivative=False)
    for m, (v, d, nv, err) in zip(ms, got):
        assert err is None and d is None
        assert v == m.GetValue()
    be.close()


def _rsgd(native: bool, metric, iterations=40):
    from itk_amd.optimizers import RegularStepGradientDescentOptimizerv4
    from itk_amd.scales import RegistrationParameterScalesFromPhysicalShift

    est = RegistrationParameterScalesFromPhysicalShift()
    est.SetTransform(metric.GetMovingTransform())
    est.SetVirtualDomain(metric._virtual or metric._fixed[1])
    opt = RegularStepGradientDescentOptimizerv4(learning_rate=2.0, minimum_step_length=1e-3,
                                                number_of_iterations=iterations, scales=est.EstimateScales())
    opt.use_native_loop = native
    opt.SetMetric(metric)
    return opt


def test_native_regular_step_matches_python_loop():
    """The C ABI's batched RegularStep loop gives the Python optimizer's
    iterates bit for bit, one registration at a time and several at once."""
    from itk_amd import _native as N
    from itk_amd import synth

    probs = []
    for k, shape in enumerate([(48, 44, 40), (40, 36, 34), (56, 50, 44)]):
        p = synth.make_affine_problem(shape, 120 + k, n_blobs=12)
        probs.append(p)
    ref = []
    for p in probs:
        tf = affine_for(p.geometry)
        m = _metric(p.fixed, p.moving, p.geometry, tf)
        opt = _rsgd(False, m)
        opt.StartOptimization()
        ref.append((tf.get_parameters(), opt.GetCurrentIteration(), opt.GetCurrentMetricValue(),
                    opt.GetStopCondition()))
    # one at a time, native
    for p, (params, its, val, stop) in zip(probs, ref):
        tf = affine_for(p.geometry)
        m = _metric(p.fixed, p.moving, p.geometry, tf)
        opt = _rsgd(True, m)
        opt.StartOptimization()
        assert opt.GetCurrentIteration() == its and opt.GetStopCondition() == stop
        assert opt.GetCurrentMetricValue() == val
        assert np.array_equal(tf.get_parameters(), params)
    # all together, in chunks of 7 iterations
    be = BatchEvaluator()
    ms, states = [], []
    for p in probs:
        tf = affine_for(p.geometry)
        m = _metric(p.fixed, p.moving, p.geometry, tf)
        opt = _rsgd(True, m)
        opt._prepare_scales()
        st = N.RegularStepState()
        sc = np.ascontiguousarray(opt.scales)
        N.check(N.load().mmi_regular_step_start(__import__("ctypes").byref(st), sc.ctypes.data_as(N._dp), sc.size))
        ms.append(m)
        states.append(st)
    opts = N.RegularStepOptions(2.0, 1e-3, 0.5, 1e-4, 40, 0)
    while any(s.stop == N.MMI_STOP_RUNNING for s in states):
        be.regular_step(ms, opts, states, 7)
    for m, st, (params, its, val, _) in zip(ms, states, ref):
        assert st.iteration == its and st.value == val
        assert np.array_equal(m.GetMovingTransform().get_parameters(), params)
    be.close()
