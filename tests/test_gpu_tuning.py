"""The tuning table on the device: a row for the workload sets the launch
configuration that mmi_initialize uses (GetLaunchConfiguration reports it
and that it came from the table), explicit settings override it, and every
configuration gives the same value (to 1e-9; bit for bit for the same
launch settings)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem
from itk_amd import _native as N
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric

pytestmark = pytest.mark.gpu


def _load(path):
    n = ctypes.c_int32()
    N.check(N.load().mmi_load_tuning_table(path.encode() if path else None, ctypes.byref(n)))
    return n.value


def test_table_row_sets_the_launch_configuration(tmp_path):
    fixed, moving, geom = blob_problem((64, 60, 56), 701)  # 215040 samples: log2 17
    tf = affine_for(geom, affine_params(3, np.random.default_rng(702), scale=1.0))
    m = Metric()
    m.SetFixedImage(fixed, geom)
    m.SetMovingImage(moving, geom)
    m.SetMovingTransform(tf)
    try:
        p = tmp_path / "t.tsv"
        p.write_text("affine f32 3 image 64 17 3 8 1\naffine f32 3 image 64 24 6 32 2\n"
                     "affine f32 3 cd 64 17 5 4 4\n")
        assert _load(str(p)) == 3
        m.Initialize()
        cfg = m.GetLaunchConfiguration()
        assert cfg == {"blocks_per_cu": 3, "rows_walk": 8, "hist_copies": 1, "from_table": True}
        v1, d1 = m.GetValueAndDerivative()
        # the caller's choice wins for what it sets
        m.SetLaunchConfiguration(2, -1, 0)
        m.Initialize()
        cfg = m.GetLaunchConfiguration()
        assert cfg["blocks_per_cu"] == 2 and cfg["rows_walk"] == 8 and cfg["hist_copies"] == 1
        v2, d2 = m.GetValueAndDerivative()
        # 40 bins take the 64-bin row as well
        m.SetLaunchConfiguration(-1, -1, 0)
        m.SetNumberOfHistogramBins(40)
        m.Initialize()
        assert m.GetLaunchConfiguration()["from_table"]
        m.SetNumberOfHistogramBins(50)
        m.Initialize()
        v3, d3 = m.GetValueAndDerivative()
        assert v1 == v3 and abs(v2 - v1) <= 1e-9 * abs(v1)
        assert np.allclose(d1, d2, rtol=1e-7, atol=0) and np.array_equal(d1, d3)
        # central differences take their own row
        m.SetUseMovingImageGradientFilter(False)
        m.Initialize()
        assert m.GetLaunchConfiguration() == {"blocks_per_cu": 5, "rows_walk": 4, "hist_copies": 4,
                                              "from_table": True}
        m.SetUseMovingImageGradientFilter(True)
        p.write_text("bspline f32 3 image 64 17 3 8 1\n")
        _load(str(p))
        m.Initialize()
        cfg = m.GetLaunchConfiguration()
        assert not cfg["from_table"] and cfg["blocks_per_cu"] == 4 and cfg["rows_walk"] == 16
        assert abs(m.GetValueAndDerivative()[0] - v1) <= 1e-9 * abs(v1)
    finally:
        _load(None)
        m.close()


def test_copy_probe_reports_a_rate():
    """The copy-rate diagnostic the benchmark reports beside the roofline:
    positive and below the device's published 8 TB/s."""
    import ctypes

    from itk_amd import _native as N

    g = ctypes.c_double(0.0)
    N.check(N.load().mmi_measure_copy_bandwidth(0, 64 << 20, 3, ctypes.byref(g)))
    assert 100.0 < g.value < 8000.0
    with pytest.raises(N.MetricError):
        N.check(N.load().mmi_measure_copy_bandwidth(0, 8, 3, ctypes.byref(g)))
