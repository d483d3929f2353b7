"""The RegularStep loop on the device (mmi_batch_regular_step with the step
kernel between captured evaluations) against the host loop: the same states
(iteration, value, gradient, previous gradient, relaxation, stop condition)
and parameters bit for bit, for each stop condition, for chunk sizes that do
and do not divide the call's step budget, for batches that mix fused and
non-fused members, and for members the device hands back to the host (an
exception-list overflow before a step, a step after which the fast path's
bounds no longer hold)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem
from itk_amd import _native as N
from itk_amd.batch import BatchEvaluator
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric
from itk_amd.scales import RegistrationParameterScalesFromPhysicalShift

pytestmark = pytest.mark.gpu


def _metric(fixed, moving, fgeom, mgeom, params, cd=False):
    tf = affine_for(fgeom, params)
    m = Metric()
    m.SetFixedImage(fixed, fgeom)
    m.SetMovingImage(moving, mgeom)
    if cd:
        m.SetUseMovingImageGradientFilter(False)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


def _problems():
    out = []
    for k, shape in enumerate([(48, 44, 40), (40, 36, 34), (56, 50, 44)]):
        fixed, moving, geom = blob_problem(shape, 310 + k)
        out.append((fixed, moving, geom, geom, affine_params(3, np.random.default_rng(320 + k), scale=1.5)))
    return out


def _start(m):
    est = RegistrationParameterScalesFromPhysicalShift()
    est.SetTransform(m.GetMovingTransform())
    est.SetVirtualDomain(m._fixed[1])
    sc = np.ascontiguousarray(est.EstimateScales(), dtype=np.float64)
    st = N.RegularStepState()
    N.check(N.load().mmi_regular_step_start(ctypes.byref(st), sc.ctypes.data_as(N._dp), sc.size))
    return st


def _run(cases, opts, on_device, chunk=8, max_steps=7, cd=()):
    be = BatchEvaluator()
    be.set_step_loop(on_device, chunk)
    ms = [_metric(*c, cd=(i in cd)) for i, c in enumerate(cases)]
    states = [_start(m) for m in ms]
    device_steps = 0
    calls = 0
    while any(s.stop == N.MMI_STOP_RUNNING for s in states) and calls < 50:
        be.regular_step(ms, opts, states, max_steps)
        device_steps += be.device_steps()
        calls += 1
    out = [(_state_tuple(s), m.GetMovingTransform().get_parameters().copy()) for s, m in zip(states, ms)]
    for m in ms:
        m.close()
    be.close()
    return out, device_steps


def _state_tuple(s):
    return (s.iteration, s.stop, s.status, s.value, s.relaxation, s.gradient_magnitude,
            tuple(s.gradient), tuple(s.previous_gradient))


def _same(a, b):
    assert len(a) == len(b)
    for (sa, pa), (sb, pb) in zip(a, b):
        assert sa == sb
        assert np.array_equal(pa, pb)


OPTS = {
    # a full run to the iteration limit
    "max_iterations": N.RegularStepOptions(2.0, 1e-4, 0.5, 1e-8, 23, 0),
    # a step too small after a few direction changes
    "step_too_small": N.RegularStepOptions(2.0, 0.3, 0.5, 1e-8, 200, 0),
    # the gradient tolerance
    "gradient_tolerance": N.RegularStepOptions(1.0, 1e-4, 0.5, 5e-4, 200, 0),
}


@pytest.mark.parametrize("which", sorted(OPTS))
def test_device_loop_equals_host_loop(which):
    cases = _problems()
    opts = OPTS[which]
    host, hsteps = _run(cases, opts, on_device=False)
    assert hsteps == 0
    dev, dsteps = _run(cases, opts, on_device=True)
    assert dsteps > 0
    _same(dev, host)
    stops = {s[1] for s, _ in host}
    want = {"max_iterations": N.MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS,
            "step_too_small": N.MMI_STOP_STEP_TOO_SMALL,
            "gradient_tolerance": N.MMI_STOP_GRADIENT_MAGNITUDE_TOLERANCE}[which]
    assert want in stops


@pytest.mark.parametrize("chunk,max_steps", [(1, 5), (3, 7), (8, 8), (16, 40)])
def test_chunking_does_not_change_the_iterates(chunk, max_steps):
    cases = _problems()[:2]
    opts = OPTS["max_iterations"]
    host, _ = _run(cases, opts, on_device=False, max_steps=max_steps)
    dev, dsteps = _run(cases, opts, on_device=True, chunk=chunk, max_steps=max_steps)
    assert dsteps == sum(s[0] for s, _ in host)
    _same(dev, host)


def test_mixed_batch_with_a_member_off_the_fused_launch():
    cases = _problems()
    opts = OPTS["max_iterations"]
    # member 1 takes central differences: the host loop iterates it
    host, _ = _run(cases, opts, on_device=False, cd=(1,))
    dev, dsteps = _run(cases, opts, on_device=True, cd=(1,))
    _same(dev, host)
    assert 0 < dsteps < sum(s[0] for s, _ in host)


def test_list_overflow_hands_the_member_back():
    """A moving image of integer plateaus 1..45 with one voxel at 0 and one at
    46 and 50 bins (moving bin size exactly 1): inside a plateau the
    interpolated value is an integer, its Parzen term sits on an interior bin
    edge, every such sample is uncertain and the exception lists overflow; the
    host redoes those evaluations exactly."""
    fixed, moving, geom = blob_problem((48, 44, 40), 330)
    q = (1.0 + np.round(np.clip(moving, 0.0, None) / moving.max() * 44.0)).astype(np.float32)
    q.flat[0] = 0.0
    q.flat[-1] = 46.0
    params = affine_params(3, np.random.default_rng(331), scale=1.0)
    m = _metric(fixed, q, geom, geom, params)
    m.GetValueAndDerivative()
    assert m.GetLastEvaluationPath() == "fast_overflow_exact"
    m.close()
    cases = [(fixed, q, geom, geom, params)] + _problems()[:1]
    opts = OPTS["max_iterations"]
    host, _ = _run(cases, opts, on_device=False)
    dev, dsteps = _run(cases, opts, on_device=True)
    _same(dev, host)
    assert dsteps > 0


def test_step_off_the_certified_path_hands_the_member_back():
    """A learning rate so large that the first step scales the matrix by
    millions: the map's error bound leaves the fast path, the device hands the
    member to the host loop, whose exact evaluation then fails (no sample maps
    inside), as in the host-only loop."""
    cases = _problems()[:2]
    opts = N.RegularStepOptions(1e7, 1e-4, 0.5, 1e-8, 20, 0)
    host, _ = _run(cases, opts, on_device=False)
    dev, dsteps = _run(cases, opts, on_device=True)
    _same(dev, host)
    assert dsteps == len(cases)  # one step each on the device, the rest on the host
    for s, _ in host:
        assert s[0] == 1 and s[1] == N.MMI_STOP_METRIC_ERROR and s[2] == N.MMI_ERR_DEGENERATE
