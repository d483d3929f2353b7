"""Runs tests/cpp/itk_subclass_check (the ITK metric subclass over the ITK
interface mock, linked to the HIP library) on a problem written to a
directory; returns its JSON lines."""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "itk_subclass_check")


def build() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "itk_subclass_check"], check=True)
    return BIN


def run(tmpdir, fixed, moving, geom, kind, fixed_params, param_sets, bins=50, use_filter=True, timeout=120,
        initial=()):
    """initial: AffineTransforms before the moving transform in the
    composite's queue."""
    d = geom.dim
    pixel = "f64" if fixed.dtype == np.float64 else "f32"
    np.ascontiguousarray(fixed).tofile(os.path.join(tmpdir, "fixed.raw"))
    np.ascontiguousarray(moving).tofile(os.path.join(tmpdir, "moving.raw"))
    fmt = lambda v: " ".join("%.17g" % x for x in v)  # noqa: E731
    lines = ["%d %s" % (d, pixel), fmt(geom.size), fmt(geom.spacing), fmt(geom.origin),
             "%s %d %s" % (kind, len(fixed_params), fmt(fixed_params)),
             "%d %s" % (len(param_sets[0]), fmt(param_sets[0])),
             "%d %d %d" % (bins, 1 if use_filter else 0, len(param_sets))]
    lines += [fmt(p) for p in param_sets[1:]]
    if initial:
        lines.append("%d" % len(initial))
        lines += ["%s %s" % (fmt(t.center), fmt(t.get_parameters())) for t in initial]
    with open(os.path.join(tmpdir, "spec.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(tmpdir)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=timeout)
    out = [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")]
    return r.returncode, out, r.stderr
