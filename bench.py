"""Benchmark of the MI355X Mattes MI v4 hot path.

One step = one dense GetValueAndDerivative of Mattes mutual information on
the 256^3 float32 affine problem (80 Gaussian blobs, remapped moving image,
50 bins, moving-image gradient from the recursive Gaussian filter), i.e. the
finest pyramid level of the 3-D affine registration workload.  Inputs are
resident in HBM before timing starts; each step sets new transform
parameters (as an optimizer iteration does) and returns value and the
12-entry derivative to the host.

Prints one JSON line.  --gpus above 1 is rejected: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_SAMPLE_F32_3D = 20  # fixed 4 + moving 4 + gradient 3 x 4 (one compulsory read each)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline(problem, tf, seconds: float):
    """The oracle (CPU FP64 restatement of the reference path, multithreaded
    like the reference's slab threader) on the same 256^3 evaluation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = min(16, os.cpu_count() or 1)
    grad = oracle.gradient_image(problem.moving, problem.geometry, threads=threads)
    n = 0
    t0 = time.perf_counter()
    while True:
        oracle.mattes(problem.fixed, problem.geometry, problem.moving, problem.geometry, tf, moving_gradient=grad,
                      threads=threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 50:
            break
    return {"value": n / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "%d full 256^3 GetValueAndDerivative evaluation(s) of the same problem in %.1f s "
                      "(oracle, %d threads, slab partition)" % (n, el, threads)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2
    from itk_amd import synth
    from itk_amd.metric import MattesMutualInformationImageToImageMetricv4
    from itk_amd.transforms import AffineTransform

    s = args.size
    t0 = time.perf_counter()
    problem = synth.make_affine_problem((s, s, s), args.seed)
    geom = problem.geometry
    print("# synthesized %d^3 problem in %.1f s" % (s, time.perf_counter() - t0), file=sys.stderr, flush=True)
    tf = AffineTransform(3, center=geom.index_to_point([(x - 1) / 2.0 for x in geom.size]))
    base = tf.get_parameters()
    rng = np.random.default_rng(args.seed)
    params = [base + np.concatenate([rng.uniform(-0.02, 0.02, 9), rng.uniform(-1.5, 1.5, 3)])
              for _ in range(16)]

    m = MattesMutualInformationImageToImageMetricv4(device=0)
    m.SetFixedImage(problem.fixed, geom)
    m.SetMovingImage(problem.moving, geom)
    m.SetNumberOfHistogramBins(50)
    m.SetMovingTransform(tf)
    t0 = time.perf_counter()
    m.Initialize()  # uploads, device gradient filter, intensity ranges
    print("# initialize (upload + device gradient filter) %.2f s" % (time.perf_counter() - t0), file=sys.stderr,
          flush=True)

    for i in range(args.warmup):
        tf.set_parameters(params[i % len(params)])
        m.GetValueAndDerivative()
    t_pass1 = t_pass2 = t_total = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        tf.set_parameters(params[i % len(params)])
        v, d = m.GetValueAndDerivative()
        tm = m.GetLastTiming()
        t_total += tm["total_ms"]
        t_pass1 += tm["pass1_ms"]
        t_pass2 += tm["pass2_ms"]
    wall = time.perf_counter() - t0
    K = args.steps
    ms_step = wall * 1e3 / K
    n = geom.num_pixels
    alg_bytes = n * BYTES_PER_SAMPLE_F32_3D
    dev_ms = t_total / K
    achieved = alg_bytes / (dev_ms * 1e-3) / 1e9
    line = {
        "metric": "MI value+grad evals/s & HBM GB/s; batched registrations/s + p99 latency",
        "value": K / wall,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 images, f64 coordinates/PDF",
        "data": "synthetic (seeded blob volumes, see itk_amd/synth.py)",
        "config": {"workload": "mattes_mi_dense_value_and_derivative_affine_%dcube" % s, "bins": 50,
                   "samples_per_eval": n, "gradient_source": "recursive Gaussian gradient image (device)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole evaluation (pass1 + finalize + pass2), HIP events on the metric stream",
                     "algorithmic_bytes_per_eval": alg_bytes, "device_ms_per_eval": dev_ms,
                     "pass1_ms": t_pass1 / K, "pass2_ms": t_pass2 / K},
        "last_value": v,
    }
    m.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(problem, tf, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
