"""ITK text transform files (itk_amd.transformio): the number format against
the reference's own NumberToString tests
(Modules/Core/Common/test/itkNumberToStringGTest.cxx:53-110,158-200) and the
ECMAScript Number-to-string rule it follows; the reference's example file
(Examples/Data/IdentityTransform.tfm, committed as a fixture); the reference
reader test's illegal file (itkIOTransformTxtTest.cxx:195-240); bit-exact
round trips of affine and B-spline transforms.  CPU only."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from _cases import affine_for, affine_params, bspline_for
from itk_amd import transformio as T
from itk_amd.geometry import ImageGeometry

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_number_to_string_reference_cases():
    for k in range(1, 21):  # +-10^k for k = 1..20 in decimal notation
        v = float(10 ** k)
        assert T.number_to_string(v) == "1" + "0" * k
        assert T.number_to_string(-v) == "-1" + "0" * k
    for k in range(1, 7):
        v = 1.0 / float(10 ** k)
        assert T.number_to_string(v) == "0." + "0" * (k - 1) + "1"
        assert T.number_to_string(-v) == "-0." + "0" * (k - 1) + "1"
    assert T.number_to_string(float("nan")) == "NaN"
    assert T.number_to_string(float("inf")) == "Infinity"
    assert T.number_to_string(float("-inf")) == "-Infinity"
    assert T.number_to_string(0.0) == "0" and T.number_to_string(-0.0) == "0"
    for d in range(10):
        assert T.number_to_string(float(d)) == str(d)
    fi = sys.float_info
    for v in (-fi.max, fi.epsilon, 0.5, 5e-324, fi.min, fi.max):
        assert float(T.number_to_string(v)) == v


def test_number_to_string_ecmascript_switch_points():
    # decimal while the decimal point lies in (-6, 21], else exponential with a sign
    assert T.number_to_string(1e21) == "1e+21"
    assert T.number_to_string(1.5e21) == "1.5e+21"
    assert T.number_to_string(1e-7) == "1e-7"
    assert T.number_to_string(1.25e-7) == "1.25e-7"
    assert T.number_to_string(123456789012345680000.0) == "123456789012345680000"
    assert T.number_to_string(0.000123) == "0.000123"
    assert T.number_to_string(-2.5e-10) == "-2.5e-10"
    assert T.number_to_string(12345.678) == "12345.678"
    rng = np.random.default_rng(1)
    for v in rng.normal(0, 1, 2000) * 10.0 ** rng.integers(-30, 30, 2000):
        s = T.number_to_string(v)
        assert float(s) == v
        assert len(s.replace("-", "").replace(".", "").split("e")[0].strip("0")) <= 17


def test_reference_identity_example():
    tfs = T.ReadTransform(os.path.join(GOLDEN, "IdentityTransform.tfm"))
    assert len(tfs) == 1
    tf = tfs[0]
    assert tf.dim == 3
    assert np.array_equal(tf.get_parameters(), np.concatenate([np.eye(3).ravel(), np.zeros(3)]))
    assert np.array_equal(tf.center, np.zeros(3))


def test_reference_illegal_file_raises(tmp_path):
    p = tmp_path / "IllegalTransform.txt"
    p.write_text("#Insight Transform File V1.0\n#Transform 0\nTransform: AffineTransform_double_10_10\n"
                 "Parameters:   0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 20 21 22\n"
                 "FixedParameters: 0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18")
    with pytest.raises(ValueError):
        T.ReadTransform(str(p))
    p.write_text("#Insight Transform File V1.0\nTransform AffineTransform_double_3_3\n")
    with pytest.raises(ValueError, match="Tags must be delimited by :"):
        T.ReadTransform(str(p))


def test_float_typed_file_reads_as_double(tmp_path):
    # a reader of double transforms takes a _float_ class name as its double
    # counterpart and keeps every digit of the text
    # (Modules/IO/TransformBase/include/itkTransformIOBase.h:203, CorrectTransformPrecisionType)
    p = tmp_path / "f.tfm"
    p.write_text("#Insight Transform File V1.0\n#Transform 0\nTransform: AffineTransform_float_2_2\n"
                 "Parameters: 1.1000000000000001 0.2 -0.3 0.9 4.25 -1.0000000000000002\n"
                 "FixedParameters: 0.5 -0.5\n")
    (tf,) = T.ReadTransform(str(p))
    assert tf.kind == T.TF_AFFINE and tf.dim == 2
    assert np.array_equal(tf.get_parameters(), [1.1000000000000001, 0.2, -0.3, 0.9, 4.25, -1.0000000000000002])
    assert np.array_equal(tf.center, [0.5, -0.5])
    assert T.type_string(tf) == "AffineTransform_double_2_2"


def test_round_trips(tmp_path):
    geom = ImageGeometry(size=(30, 26, 20), spacing=(0.9, 1.2, 1.7), origin=(-3.0, 4.5, 0.25))
    aff = affine_for(geom, affine_params(3, np.random.default_rng(2), scale=3.0))
    aff2 = affine_for(ImageGeometry(size=(40, 30)), affine_params(2, np.random.default_rng(3)))
    bsp2 = bspline_for(ImageGeometry(size=(33, 21), spacing=(0.7, 1.3), origin=(2.0, -1.5)), (5, 4), seed=7)
    bsp = bspline_for(geom, (4, 3, 5), seed=4, sigma=1.3)
    for tf in (aff, aff2, bsp2, bsp):
        p = str(tmp_path / "t.tfm")
        T.WriteTransform(p, tf)
        (back,) = T.ReadTransform(p)
        assert back.kind == tf.kind and back.dim == tf.dim
        assert np.array_equal(back.get_parameters(), tf.get_parameters())
        assert np.array_equal(T.fixed_parameters(back), T.fixed_parameters(tf))
        x = np.array([1.3, -2.2, 5.1][: tf.dim])
        if tf.kind == T.TF_AFFINE:
            assert np.array_equal(back.transform_point(x), tf.transform_point(x))
    text = open(p).read().splitlines()
    assert text[0] == "#Insight Transform File V1.0"
    assert text[2] == "Transform: BSplineTransform_double_3_3"
    assert text[4].startswith("FixedParameters: 7 6 8 ")


def test_composite_file(tmp_path):
    geom = ImageGeometry(size=(20, 18, 16))
    a = affine_for(geom, affine_params(3, np.random.default_rng(5)))
    b = bspline_for(geom, (3, 3, 3), seed=6)
    p = str(tmp_path / "c.tfm")
    T.WriteTransform(p, [a, b], composite=True)
    lines = open(p).read().splitlines()
    assert lines[1:3] == ["#Transform 0", "Transform: CompositeTransform_double_3_3"]
    assert lines[3] == "#Transform 1" and lines[4] == "Transform: AffineTransform_double_3_3"
    back = T.ReadTransform(p)
    assert [t.kind for t in back] == [T.TF_AFFINE, T.TF_BSPLINE]
    assert np.array_equal(back[1].get_parameters(), b.get_parameters())
