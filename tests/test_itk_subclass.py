"""CPU checks of the ITK metric subclass (include/itkMattesMutualInformationImageToImageMetricv4HIP.h):
it compiles for Image<float, 3>, Image<double, 2> and the other pixel and
dimension pairs against the ITK interface mock, links to the HIP library, and
without a HIP device its Initialize raises an itk::ExceptionObject carrying
the library's message (no device calls)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import _itk_subclass
from _cases import affine_for, reference_test_images
from itk_amd.metric import transform_fixed_parameters


def _device_visible() -> bool:
    try:
        lib = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return lib.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_subclass_compiles_and_links():
    assert _itk_subclass.build()


@pytest.mark.skipif(_device_visible(), reason="checks the no-device error path")
def test_initialize_without_device_raises_exception_object(tmp_path):
    _itk_subclass.build()
    fixed, moving, geom = reference_test_images(np.float64)
    tf = affine_for(geom)
    rc, out, _ = _itk_subclass.run(tmp_path, fixed, moving, geom, "affine", transform_fixed_parameters(tf),
                                   [tf.get_parameters()])
    assert rc == 3 and len(out) == 1 and "error" in out[0]
    assert out[0]["error"]
