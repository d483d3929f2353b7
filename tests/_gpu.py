"""Helpers of the GPU parity tests: build the HIP metric, compare with the oracle.

Bar: Parzen bin indices and valid-sample counts bit-exact; value within 1e-5
relative; derivative within 1e-4 relative (L2).
"""
from __future__ import annotations

import numpy as np

from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

VALUE_RTOL = 1e-5
DERIV_RTOL = 1e-4


def _gpu_metric(fixed, moving, geom, tf, *, gradient=None, use_filter=True, virtual=None, bins=50, record=True,
                storage="f32", exact=False):
    m = Metric(record_sample_bins=record, exact_path_only=exact)
    m.SetImageStorage(storage)
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetUseMovingImageGradientFilter(use_filter)
    if gradient is not None:
        m.SetMovingImageGradient(gradient)
    if virtual is not None:
        m.SetVirtualDomain(virtual)
    m.SetNumberOfHistogramBins(bins)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


def _compare(oracle_mod, m, fixed, moving, geom, tf, *, gradient=None, virtual=None, bins=50, check_bins=True):
    v, d = m.GetValueAndDerivative()
    r = oracle_mod.mattes(fixed, geom, moving, geom, tf, bins=bins, virtual_geom=virtual,
                          moving_gradient=gradient, want_samples=check_bins, threads=8)
    print("value gpu %.17g oracle %.17g  nvalid %d/%d" % (v, r["value"], m.GetNumberOfValidPoints(), r["n_valid"]))
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    if check_bins:
        fb, mb = m.GetSampleBins()
        assert np.array_equal(fb, r["fixed_bin"])
        assert np.array_equal(mb, r["moving_bin"])
    # relative bar, with an absolute floor for metrics that are ~0 (e.g. 5 bins)
    assert abs(v - r["value"]) <= VALUE_RTOL * abs(r["value"]) + 1e-11
    err = np.linalg.norm(d - r["derivative"]) / max(np.linalg.norm(r["derivative"]), 1e-9)
    print("derivative rel L2 error %.3e" % err)
    assert err <= DERIV_RTOL
    return v, d, r
