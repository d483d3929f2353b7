"""Launch-configuration tuner for the evaluation kernels.

Times one GetValueAndDerivative (the device time between the HIP events
around its kernels) over a grid of workgroups per compute unit, rows per walk
of the affine row walks and pass-1 LDS histogram copies, for each workload
class: transform kind, image storage, gradient source, virtual sample count
and histogram bins.  Writes the fastest setting of every class as the text
table the library reads at mmi_initialize (mmi_load_tuning_table):

    kind storage dim gradient bins log2_samples blocks_per_cu rows_walk hist_copies

    python -m itk_amd.tune [--out itk_amd/tuning/mi355x.tsv] [--quick]

Every setting gives the same bins and valid count; the value agrees to
about 1e-12 (the walk length widens the map's certified error bound, which
moves a few samples between the FP32 and the exact FP64 Parzen weights), which
the tuner checks to 1e-9.
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys

import numpy as np

from . import synth
from .metric import MattesMutualInformationImageToImageMetricv4 as Metric
from .transforms import AffineTransform, BSplineTransform

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning", "mi355x.tsv")


def _transform(kind, geom, rng):
    if kind == "affine":
        tf = AffineTransform(3, center=geom.index_to_point([(s - 1) / 2.0 for s in geom.size]))
        m = np.eye(3) + rng.uniform(-0.04, 0.04, (3, 3))
        tf.set_parameters(np.concatenate([m.ravel(), rng.uniform(-2, 2, 3)]))
        return tf
    tf = BSplineTransform.for_image(geom, (8, 8, 8))
    tf.set_parameters(rng.normal(0.0, 0.5, tf.number_of_parameters))
    return tf


def _time(m, reps):
    times, value = [], None
    for k in range(reps + 2):
        v, _ = m.GetValueAndDerivative()
        if value is None:
            value = v
        elif v != value:
            raise RuntimeError("the value changed between evaluations: %r != %r" % (v, value))
        if k >= 2:
            times.append(m.GetLastTiming()["total_ms"])
    return float(np.median(times)), value


def tune_class(kind, storage, gradient, problem, bins, quick=False, log=print):
    p = problem
    size = p.geometry.size
    rng = np.random.default_rng(8)
    m = Metric()
    m.SetImageStorage(storage)
    m.SetFixedImage(p.fixed, p.geometry)
    m.SetMovingImage(p.moving, p.geometry)
    m.SetUseMovingImageGradientFilter(gradient == "image")
    m.SetNumberOfHistogramBins(bins)
    m.SetMovingTransform(_transform(kind, p.geometry, rng))
    bpcs = (2, 4, 8) if quick else (1, 2, 3, 4, 6, 8)
    walks = ((8, 16, 32) if quick else (4, 8, 16, 32)) if kind == "affine" else (16,)
    # copies whose LDS a workgroup of the device holds (160 KiB on MI355X)
    copies = [c for c in (1, 2, 4) if c * bins * bins * 8 + 8 <= 160 * 1024]
    results = []
    ref = None
    for bpc in bpcs:
        for walk in walks:
            for c in copies:
                m.SetLaunchConfiguration(bpc, walk, c)
                m.Initialize()
                ms, v = _time(m, 3 if quick else 5)
                if ref is None:
                    ref = v
                elif abs(v - ref) > 1e-9 * abs(ref):
                    raise RuntimeError("value differs across launch settings: %r, %r" % (v, ref))
                results.append((ms, bpc, walk, c))
    results.sort()
    best = results[0]
    # the library's built-in choice (4 workgroups per CU, walks of 16 rows,
    # two copies when they fit in 64 KiB)
    m.SetLaunchConfiguration(4, 16, 2 if 2 * bins * bins * 8 + 8 <= 64 * 1024 else 1)
    m.Initialize()
    lib_ms, _ = _time(m, 3)
    m.close()
    n = int(np.prod(size))
    log("%-7s %-4s %-5s %-11s bins %3d: best %.4f ms (bpc %d walk %d copies %d), built-in choice %.4f ms"
        % (kind, storage, gradient, "x".join(map(str, size)), bins, best[0], best[1], best[2], best[3], lib_ms))
    return {"kind": kind, "storage": storage, "dim": 3, "gradient": gradient, "bins": bins,
            "log2n": int(np.floor(np.log2(n))), "bpc": best[1], "walk": best[2], "copies": best[3],
            "ms": best[0], "default_ms": lib_ms}


def run(quick=False, log=print):
    """Every class, each problem made once per size."""
    sizes = [(64, 64, 64), (128, 128, 128), (256, 256, 256)] if quick else \
        [(64, 64, 64), (96, 96, 96), (128, 128, 128), (192, 192, 192), (256, 256, 256)]
    rows = []
    for size in sizes:
        p = synth.make_affine_problem(size, 7)
        for storage in ("f32",) if quick else ("f32", "bf16", "f16"):
            for gradient in ("image", "cd"):
                for bins in (50,) if quick else (50, 80):
                    rows.append(tune_class("affine", storage, gradient, p, bins, quick, log))
        if size[0] <= 192:
            pb = synth.make_bspline_problem(size, 7, mesh=8)
            rows.append(tune_class("bspline", "f32", "image", pb, 50, quick, log))
    return rows


def bins_bucket(bins):
    """The largest bins a row serves: 64 (tuned at 50) and 90 (tuned at 80;
    the fast path's LDS histogram holds at most 90 bins)."""
    return 64 if bins <= 64 else (90 if bins <= 90 else 253)


def write_table(rows, path, command):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("# Launch settings of the evaluation kernels per workload class, timed on one MI355X by\n")
        f.write("#   %s\n" % command)
        f.write("# on %s.  Columns: kind storage dim gradient bins log2_samples blocks_per_cu rows_walk\n"
                % datetime.date.today().isoformat())
        f.write("# hist_copies; then the tuned time and the time of the built-in choice (4, 16, copies by\n")
        f.write("# size), device ms per evaluation.\n")
        for r in rows:
            f.write("%s %s %d %s %d %d %d %d %d  # %.4f %.4f\n" % (
                r["kind"], r["storage"], r["dim"], r["gradient"], bins_bucket(r["bins"]), r["log2n"], r["bpc"],
                r["walk"], r["copies"], r["ms"], r["default_ms"]))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true", help="fewer classes and settings")
    args = ap.parse_args(argv)
    rows = run(args.quick, log=lambda s: print(s, flush=True))
    write_table(rows, args.out, "python -m itk_amd.tune" + (" --quick" if args.quick else "") +
                (" --out %s" % args.out if args.out != DEFAULT_OUT else ""))
    print("wrote %d rows to %s" % (len(rows), args.out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
