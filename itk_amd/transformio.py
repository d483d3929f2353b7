"""ITK's text transform files (``.tfm`` / ``.txt``), so registration results
load back into stock ITK and ITK's files load here.

Format of ``TxtTransformIOTemplate``
(``Modules/IO/TransformInsightLegacy/src/itkTxtTransformIO.cxx:110-307``): a
``#Insight Transform File V1.0`` line, then per transform ``#Transform k``,
``Transform: <Class>_<double|float>_<in>_<out>``, ``Parameters: ...`` and
``FixedParameters: ...``.  A composite transform is written as its name
followed by its components.  Numbers are written as ``ConvertNumberToString``
does (``Modules/Core/Common/src/itkNumberToString.cxx``): the shortest
round-trip digits in ECMAScript notation, decimal while the decimal point
lies in (-6, 21], else ``d.ddde±x``; ``NaN``, ``Infinity``, a single ``0``
for both zeros.

Transforms: ``AffineTransform`` (parameters: matrix row-major then
translation; fixed parameters: centre) and ``BSplineTransform`` of order 3
(parameters: coefficients dimension-major; fixed parameters: grid size,
origin, spacing, direction row-major,
``itkBSplineTransform.hxx:95-218``).
"""
from __future__ import annotations

import math
import re
from decimal import Decimal

import numpy as np

from .geometry import ImageGeometry
from .transforms import TF_AFFINE, TF_BSPLINE, AffineTransform, BSplineTransform

HEADER = "#Insight Transform File V1.0"


def number_to_string(x: float) -> str:
    """ConvertNumberToString for a double."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "0"
    # repr gives the shortest digits that read back to x (the nearest on ties)
    t = Decimal(repr(abs(x))).as_tuple()
    digits = "".join(str(k) for k in t.digits).lstrip("0")
    exp = t.exponent
    stripped = digits.rstrip("0")
    exp += len(digits) - len(stripped)
    digits = stripped
    point = len(digits) + exp  # value = 0.<digits> * 10^point
    if -6 < point <= 21:
        if point <= 0:
            body = "0." + "0" * (-point) + digits
        elif point >= len(digits):
            body = digits + "0" * (point - len(digits))
        else:
            body = digits[:point] + "." + digits[point:]
    else:
        xe = point - 1
        body = digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e" + ("+" if xe >= 0 else "-") + str(abs(xe))
    return ("-" if x < 0 else "") + body


def _vector(v) -> str:
    return " ".join(number_to_string(x) for x in v)


def type_string(tf) -> str:
    """GetTransformTypeAsString (itkTransform.hxx:40-47, itkBSplineTransform.hxx:65-76)."""
    name = {TF_AFFINE: "AffineTransform", TF_BSPLINE: "BSplineTransform"}.get(tf.kind)
    if name is None:
        raise ValueError("unsupported transform")
    return "%s_double_%d_%d" % (name, tf.dim, tf.dim)


def fixed_parameters(tf) -> np.ndarray:
    if tf.kind == TF_AFFINE:
        return np.asarray(tf.center, dtype=np.float64)
    g = tf.grid
    return np.concatenate([np.asarray(g.size, dtype=np.float64), np.asarray(g.origin, dtype=np.float64),
                           np.asarray(g.spacing, dtype=np.float64), np.asarray(g.direction, dtype=np.float64).ravel()])


def format_transforms(transforms, composite: bool = False) -> str:
    """The file text for a list of transforms; composite=True writes them as
    the components of one CompositeTransform (the registration method's
    output form)."""
    transforms = list(transforms)
    if not transforms:
        raise ValueError("no transform to write")
    lines = [HEADER]
    k = 0
    if composite:
        d = transforms[0].dim
        lines += ["#Transform 0", "Transform: CompositeTransform_double_%d_%d" % (d, d)]
        k = 1
    for tf in transforms:
        lines += ["#Transform %d" % k, "Transform: " + type_string(tf),
                  "Parameters: " + _vector(tf.get_parameters()),
                  "FixedParameters: " + _vector(fixed_parameters(tf))]
        k += 1
    return "\n".join(lines) + "\n"


def WriteTransform(path: str, transforms, composite: bool = False):
    """TransformFileWriter::Update for one transform or a list."""
    if not isinstance(transforms, (list, tuple)):
        transforms = [transforms]
    with open(path, "w") as f:
        f.write(format_transforms(transforms, composite))


_TYPE = re.compile(r"^(\w+?)_(double|float)_(\d+)_(\d+)$")


def _parse_numbers(text: str) -> np.ndarray:
    vals = []
    for tok in text.split():
        t = tok.strip()
        if t in ("NaN", "nan"):
            vals.append(float("nan"))
        elif t in ("Infinity", "inf"):
            vals.append(float("inf"))
        elif t in ("-Infinity", "-inf"):
            vals.append(float("-inf"))
        else:
            vals.append(float(t))
    return np.asarray(vals, dtype=np.float64)


def _make(type_name: str, params, fixed):
    m = _TYPE.match(type_name)
    if not m:
        raise ValueError("unknown transform type " + type_name)
    cls, prec, din, dout = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
    if din != dout or din not in (2, 3):
        raise ValueError("unsupported transform dimensions in " + type_name)
    d = din  # a float transform is read as double (CorrectTransformPrecisionType)
    if cls in ("AffineTransform", "MatrixOffsetTransformBase"):
        if params.size != d * d + d or fixed.size != d:
            raise ValueError("affine transform needs %d parameters and %d fixed parameters" % (d * d + d, d))
        tf = AffineTransform(d, center=fixed)
        tf.set_parameters(params)
        return tf
    if cls == "BSplineTransform":
        if fixed.size != 3 * d + d * d:
            raise ValueError("B-spline transform needs %d fixed parameters" % (3 * d + d * d))
        size = [int(round(v)) for v in fixed[:d]]
        grid = ImageGeometry(size=size, origin=fixed[d:2 * d], spacing=fixed[2 * d:3 * d],
                             direction=fixed[3 * d:].reshape(d, d))
        tf = BSplineTransform(d, grid)
        tf.set_parameters(params)
        return tf
    raise ValueError("unsupported transform type " + type_name)


def parse_transforms(text: str) -> list:
    """TxtTransformIOTemplate::Read: '#' lines and blank lines skipped, every
    other line 'Name: value'; Parameters and FixedParameters in either order
    complete the transform named last.  Composite entries only group what
    follows them, so the components are returned in file order."""
    out = []
    pending_type = None
    params = fixed = None
    for raw in text.splitlines():
        line = raw.strip(" \t\r\n")
        if not line or line[0] == "#":
            continue
        if ":" not in line:
            raise ValueError("Tags must be delimited by :")
        name, value = line.split(":", 1)
        name, value = name.strip(), value.strip()
        if name == "Transform":
            if value.startswith("CompositeTransform"):
                pending_type = None
                continue
            pending_type = value
            params = fixed = None
        elif name in ("Parameters", "FixedParameters"):
            if pending_type is None:
                raise ValueError("parameters before a transform type")
            if name == "Parameters":
                params = _parse_numbers(value)
            else:
                fixed = _parse_numbers(value)
            if params is not None and fixed is not None:
                out.append(_make(pending_type, params, fixed))
                pending_type = None
                params = fixed = None
        elif name == "ComponentTransformFile":
            raise ValueError("component transform files are not supported")
    return out


def ReadTransform(path: str) -> list:
    """TransformFileReader::Update: the transforms of a text transform file."""
    with open(path) as f:
        return parse_transforms(f.read())
