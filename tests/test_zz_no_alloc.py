"""mmi_evaluate_static, the evaluation that allocates no device memory (its
scratch is one static device array per GPU), against mmi_evaluate on the same
contexts, through the C ABI: two calls in a row, the second on a smaller
problem, each with value and derivative and with the value alone.  The same
kernels run on the same inputs, so the valid counts, values and derivatives
are equal bit for bit (the bar of integer work, stronger than the 1e-5 /
1e-4 bar of the floating-point results)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem
from itk_amd import _native as N
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

pytestmark = pytest.mark.gpu


def _static(m, want_derivative):
    out = np.zeros(2)
    deriv = np.zeros(m.GetNumberOfParameters()) if want_derivative else None
    rc = m._L.mmi_evaluate_static(m._ctx, 1 if want_derivative else 0, out.ctypes.data,
                                  deriv.ctypes.data if deriv is not None else None, out.ctypes.data + 8)
    N.check(rc)
    return float(out[0]), deriv, int(out[1:2].view(np.int64)[0])


def _metric(shape, seed):
    fixed, moving, geom = blob_problem(shape, seed)
    tf = affine_for(geom, affine_params(geom.dim, np.random.default_rng(seed + 1), scale=1.5))
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


@pytest.mark.parametrize("shapes", [((96, 88, 80), (40, 36, 32)), ((256, 240), (64, 60))])
def test_static_scratch_matches_mmi_evaluate(shapes):
    for shape in shapes:  # the second call on a smaller problem
        print("mmi_evaluate_static vs mmi_evaluate, shape", shape, flush=True)
        m = _metric(shape, 300 + len(shape))
        v, d = m.GetValueAndDerivative()
        n = m.GetNumberOfValidPoints()
        vs, ds, ns = _static(m, True)
        assert ns == n and vs == v
        assert np.array_equal(ds, d)
        v0 = m.GetValue()
        vs0, _, ns0 = _static(m, False)
        assert ns0 == n and vs0 == v0 == v
        m.close()


def test_static_scratch_rejects_what_it_cannot_hold():
    m = _metric((40, 36, 32), 310)
    m.SetNumberOfHistogramBins(200)
    m.Initialize()
    with pytest.raises(N.MetricError, match="128 histogram bins"):
        _static(m, True)
    m.close()
