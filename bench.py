"""Benchmark of the MI355X Mattes MI v4 hot path.

One step = one dense GetValueAndDerivative of Mattes mutual information on
the 256^3 float32 affine problem (80 Gaussian blobs, remapped moving image,
50 bins, moving-image gradient from the recursive Gaussian filter), i.e. the
finest pyramid level of the 3-D affine registration workload.  Inputs are
resident in HBM before timing starts; each step sets new transform
parameters (as an optimizer iteration does) and returns value and the
12-entry derivative to the host.

--workload bspline measures the 192^3 cubic B-spline case instead (mesh 16^3,
19^3 x 3 = 20,577 parameters, moving image warped by an 8^3-mesh B-spline).

--serving measures whole registrations instead: a Poisson stream of requests
(sizes uniform over 128^3, 160x160x96, 192^3, 256x256x128, 256^3; 70% affine
over a 3-level pyramid, 30% cubic B-spline with an 8^3 or 16^3 mesh), served
by the continuous-batching server, against the same requests run one at a time
on the plain metric.

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
# One compulsory read each of the fixed, moving and gradient (3 components)
# images per sample: 4 + 4 + 12 bytes in FP32 storage, half in BF16/FP16.
BYTES_PER_SAMPLE_3D = {"f32": 20, "bf16": 10, "f16": 10}
# central differences: pass 2 reads the moving value again instead of a texel
BYTES_PER_SAMPLE_3D_CD = {"f32": 12, "bf16": 6, "f16": 6}
# HBM bytes per evaluation (FETCH_SIZE + WRITE_SIZE summed over the evaluation's
# kernels, scaled by the counters' reading for a kernel of known traffic), from the
# rocprofv3 --pmc runs summarised in profiles/mattes_affine_256cube_fast_path.md,
# profiles/mattes_affine_256cube_central_differences.md and
# profiles/mattes_bspline_192cube_mesh16.md.  None where not profiled.
TRAFFIC_BYTES_PER_EVAL = {
    ("mattes_mi_dense_value_and_derivative_affine_256cube", "f32", "image"): 432.3e6,
    ("mattes_mi_dense_value_and_derivative_affine_256cube", "f32", "cd"): 214.7e6,
    ("mattes_mi_dense_value_and_derivative_bspline_192cube_mesh16", "f32", "image"): 541.4e6,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("affine", "bspline"), default="affine")
    ap.add_argument("--size", type=int, default=0, help="cube edge (default 256 affine, 192 B-spline)")
    ap.add_argument("--mesh", type=int, default=16, help="B-spline mesh size per axis")
    ap.add_argument("--storage", choices=("f32", "bf16", "f16"), default="f32",
                    help="device storage of the fixed, moving and gradient images")
    ap.add_argument("--gradient", choices=("image", "cd"), default="image",
                    help="moving gradient: the device gradient image (the metric's default) or central "
                         "differences (the registration method's default; affine workload)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pass2-reclassify", action="store_true",
                    help="pass 2 classifies every sample again (comparison runs)")
    ap.add_argument("--serving", action="store_true", help="registration serving workload")
    ap.add_argument("--requests", type=int, default=32, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serving: mean arrivals per second")
    ap.add_argument("--inflight", type=int, default=3, help="serving: registrations in flight at most")
    ap.add_argument("--serving-requests", type=int, default=16,
                    help="requests of the serving summary added to the metric line (0: none)")
    return ap.parse_args()


SERVING_SIZES = [(128, 128, 128), (160, 160, 96), (192, 192, 192), (256, 256, 128), (256, 256, 256)]


def _serving_spec(rid: int, seed: int):
    """Kind, size, mesh and image seed of request rid (splitmix64 stream of the seed)."""
    from itk_amd import synth

    u = synth.uniform(seed + rid, 23, 4)
    size = SERVING_SIZES[min(4, int(u[0] * 5))]
    kind = "affine" if u[1] < 0.7 else "bspline"
    mesh = 8 if u[2] < 0.5 else 16
    return kind, size, mesh, seed * 1000 + rid


def _synthesize(args):
    rid, seed = args
    from itk_amd import synth

    kind, size, mesh, img_seed = _serving_spec(rid, seed)
    if kind == "affine":
        p = synth.make_affine_problem(size, img_seed)
    else:
        p = synth.make_bspline_problem(size, img_seed, mesh=8, sigma_vox=2.0)
    return rid, kind, size, mesh, p.fixed, p.moving


def synthesize_requests(args):
    """The serving requests' images, made by a pool of worker processes (run
    before this process starts the GPU runtime)."""
    import concurrent.futures as cf

    t0 = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        made = list(ex.map(_synthesize, [(i, args.seed) for i in range(args.requests)]))
    print("# synthesized %d requests in %.1f s" % (len(made), time.perf_counter() - t0), file=sys.stderr, flush=True)
    return made


def serving_summary_child(args, timeout_s: float = 300.0) -> dict:
    """The serving summary of the metric line, from `bench.py --serving` run as
    a child process with a time limit (started before this process touches the
    GPU), so a serving problem cannot take the metric line with it."""
    import subprocess

    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serving", "--requests", str(args.serving_requests),
           "--rate", str(args.rate), "--inflight", str(args.inflight), "--seed", str(args.seed)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=timeout_s, text=True)
    except subprocess.TimeoutExpired:
        return {"error": "serving run exceeded %.0f s" % timeout_s}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "serving run failed (exit %d)" % r.returncode}
    d = json.loads(lines[-1])
    c = d["config"]
    return dict(d["serving"], requests=c["requests"], kinds=c["kinds"], arrival_rate_per_s=c["arrival_rate_per_s"],
                max_in_flight=c["max_in_flight"])


def serving(args):
    """Registrations/s and latency of the continuous-batching server on a
    Poisson stream, against the same requests one at a time; prints the JSON
    line."""
    from itk_amd.geometry import ImageGeometry
    from itk_amd.serving import RegistrationRequest, RegistrationServer, latency_summary, run_sequential

    made = synthesize_requests(args)
    from itk_amd import synth

    gaps = -np.log(1.0 - synth.uniform(args.seed, 29, args.requests)) / args.rate
    arrivals = np.cumsum(gaps) - gaps[0]

    def requests(poisson: bool):
        out = []
        for (rid, kind, size, mesh, fixed, moving), t in zip(made, arrivals):
            out.append(RegistrationRequest(rid, kind, fixed, moving, ImageGeometry(size=size),
                                           arrival=float(t) if poisson else 0.0, mesh=mesh,
                                           iterations=100 if kind == "affine" else 50))
        return out

    seq, busy = run_sequential(requests(False))
    print("# sequential: %d registrations in %.2f s" % (len(seq), busy), file=sys.stderr, flush=True)
    _, busy_native = run_sequential(requests(False), native_loop=True)
    print("# sequential, C optimizer loop: %.2f s" % busy_native, file=sys.stderr, flush=True)
    server = RegistrationServer(max_inflight=args.inflight)
    sat = server.run(requests(False))
    sat_sum = latency_summary(sat)
    sat_stats = {"batches": server.batches, "evaluations": server.batched_evaluations,
                 "largest_batch": server.max_batch, "peak_in_flight": server.peak_inflight}
    server_max_batch = server.max_batch
    print("# saturated batched: %s %s" % (sat_sum, sat_stats), file=sys.stderr, flush=True)
    server = RegistrationServer(max_inflight=args.inflight)
    poi = server.run(requests(True))
    poi_sum = latency_summary(poi)
    print("# poisson batched: %s" % poi_sum, file=sys.stderr, flush=True)
    seq_poisson, _ = run_sequential(requests(True))
    seq_poi_sum = latency_summary(seq_poisson)
    same = all(a.error is None and b.error is None and np.array_equal(a.parameters, b.parameters)
               for a, b in zip(seq, sat))
    seq_rate = len(seq) / busy
    kinds = {"affine": sum(1 for m in made if m[1] == "affine"), "bspline": sum(1 for m in made if m[1] == "bspline")}
    line = {
        "metric": "MI value+grad evals/s & HBM GB/s; batched registrations/s + p99 latency",
        "value": sat_sum["registrations_per_s"],
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": 0,
        "ms_per_step": 1e3 / sat_sum["registrations_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 images, f64 coordinates/PDF",
        "data": "synthetic (seeded blob volumes per request id, see itk_amd/synth.py)",
        "config": {"workload": "registration_serving_mixed_sizes", "requests": args.requests, "kinds": kinds,
                   "arrival_rate_per_s": args.rate, "max_in_flight": args.inflight,
                   "affine": "3 levels (shrink 4,2,1; sigma 2,1,0 voxels), RegularStep GD, <=100 its/level",
                   "bspline": "mesh 8^3 or 16^3, L-BFGS-B, <=50 iterations", "parallelism": "single GPU"},
        "serving": {
            "sequential_registrations_per_s": seq_rate,
            "sequential_c_loop_registrations_per_s": len(seq) / busy_native,
            "batched_registrations_per_s": sat_sum["registrations_per_s"],
            # against the one-at-a-time service with the same C / device
            # optimizer loop, the fair baseline; the Python-loop ratio beside it
            "speedup_vs_sequential_c_loop": sat_sum["registrations_per_s"] / (len(seq) / busy_native),
            "speedup_vs_sequential": sat_sum["registrations_per_s"] / seq_rate,
            # "batched" only when registrations shared a fused launch
            "mode": "batched" if server_max_batch > 1 else "concurrent, no shared launch",
            "batched_saturated": dict(sat_sum, **sat_stats),
            "batched_poisson": poi_sum,
            "sequential_poisson": seq_poi_sum,
            "results_identical_alone_and_batched": bool(same),
        },
    }
    print(json.dumps(line), flush=True)
    return 0


class Workload:
    """Problem, transform and the parameter sets the timed steps cycle through."""

    def __init__(self, args):
        from itk_amd import synth
        from itk_amd.transforms import AffineTransform, BSplineTransform

        rng = np.random.default_rng(args.seed)
        self.kind = args.workload
        t0 = time.perf_counter()
        if args.workload == "affine":
            s = args.size or 256
            self.problem = synth.make_affine_problem((s, s, s), args.seed)
            geom = self.problem.geometry
            self.tf = AffineTransform(3, center=geom.index_to_point([(x - 1) / 2.0 for x in geom.size]))
            base = self.tf.get_parameters()
            self.params = [base + np.concatenate([rng.uniform(-0.02, 0.02, 9), rng.uniform(-1.5, 1.5, 3)])
                           for _ in range(16)]
            self.name = "mattes_mi_dense_value_and_derivative_affine_%dcube" % s
            self.extra = {}
        else:
            s = args.size or 192
            self.problem = synth.make_bspline_problem((s, s, s), args.seed, mesh=8, sigma_vox=2.0)
            geom = self.problem.geometry
            self.tf = BSplineTransform.for_image(geom, [args.mesh] * 3)
            P = self.tf.number_of_parameters
            self.params = [rng.normal(0.0, 0.5, P) for _ in range(4)]
            self.name = "mattes_mi_dense_value_and_derivative_bspline_%dcube_mesh%d" % (s, args.mesh)
            self.extra = {"parameters": P}
        self.size = s
        print("# synthesized %s in %.1f s" % (self.name, time.perf_counter() - t0), file=sys.stderr, flush=True)


def cpu_baseline(wl: Workload, seconds: float):
    """The oracle (CPU FP64 restatement of the reference path, multithreaded
    like the reference's slab threader) on the same evaluation.  The B-spline
    case is sampled on a slab of the virtual domain (its per-sample cost is the
    whole cost) and scaled to full evaluations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from itk_amd.geometry import ImageGeometry

    problem, tf = wl.problem, wl.tf
    geom = problem.geometry
    threads = min(16, os.cpu_count() or 1)
    grad = oracle.gradient_image(problem.moving, geom, threads=threads)
    virtual, scale, what = geom, 1.0, "the full %d^3 problem" % wl.size
    if wl.kind == "bspline":
        nz = max(2, geom.size[2] // 16)
        virtual = ImageGeometry(size=[geom.size[0], geom.size[1], nz], origin=geom.origin, spacing=geom.spacing,
                                direction=geom.direction, start=geom.start)
        scale = nz / float(geom.size[2])
        what = "a %d-slice slab (%.1f%% of the samples) of the %d^3 problem" % (nz, 100 * scale, wl.size)
    n = 0
    t0 = time.perf_counter()
    while True:
        oracle.mattes(problem.fixed, geom, problem.moving, geom, tf, moving_gradient=grad, threads=threads,
                      virtual_geom=virtual)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 50:
            break
    return {"value": n / el * scale, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "%d GetValueAndDerivative evaluation(s) over %s in %.1f s, scaled to full "
                      "evaluations (oracle, %d threads, slab partition)" % (n, what, el, threads)}


def pass_breakdown(args, wl, n):
    """Pass 1 and pass 2 timed apart (HIP events between them, a context with
    pass timing; outside the timed region): pass 1 reads a fixed and a moving
    value per sample, pass 2 one moving texel (value and gradient)."""
    from itk_amd.metric import MattesMutualInformationImageToImageMetricv4

    geom = wl.problem.geometry
    m = MattesMutualInformationImageToImageMetricv4(device=0, time_passes=True)
    m.SetImageStorage(args.storage)
    m.SetFixedImage(wl.problem.fixed, geom)
    m.SetMovingImage(wl.problem.moving, geom)
    m.SetNumberOfHistogramBins(50)
    m.SetUseMovingImageGradientFilter(args.gradient == "image")
    m.SetMovingTransform(wl.tf)
    m.Initialize()
    p1 = p2 = 0.0
    reps = 10
    for i in range(reps + 2):
        wl.tf.set_parameters(wl.params[i % len(wl.params)])
        m.GetValueAndDerivative()
        if i >= 2:
            t = m.GetLastTiming()
            p1 += t["pass1_ms"] / reps
            p2 += t["pass2_ms"] / reps
    m.close()
    px = 4 if args.storage == "f32" else 2
    b1, b2 = n * 2 * px, n * (4 if args.gradient == "image" else 1) * px
    out = {"pass1_ms": p1, "pass1_algorithmic_bytes": b1, "pass1_frac": b1 / (p1 * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "pass2_ms": p2, "pass2_algorithmic_bytes": b2, "pass2_frac": b2 / (p2 * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if wl.kind != "affine":
        out["note"] = "pass 1 includes the coefficient lines and the exception kernel, pass 2 the contraction"
    return out


def copy_bandwidth_gbs():
    """Device-to-device copy rate of a 1 GiB buffer (the library's plain copy
    probe, read + written bytes over its HIP-event time, best of 10): the
    practical HBM ceiling the roofline fraction can also be read against."""
    import ctypes

    from itk_amd import _native as N

    g = ctypes.c_double(0.0)
    N.check(N.load().mmi_measure_copy_bandwidth(0, 1 << 30, 10, ctypes.byref(g)))
    return g.value


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2
    if args.serving:
        return serving(args)
    served = serving_summary_child(args) if args.serving_requests > 0 else None
    from itk_amd.metric import MattesMutualInformationImageToImageMetricv4

    wl = Workload(args)
    problem, tf = wl.problem, wl.tf
    geom = problem.geometry
    m = MattesMutualInformationImageToImageMetricv4(device=0, pass2_reclassify=args.pass2_reclassify)
    m.SetImageStorage(args.storage)
    m.SetFixedImage(problem.fixed, geom)
    m.SetMovingImage(problem.moving, geom)
    m.SetNumberOfHistogramBins(50)
    m.SetUseMovingImageGradientFilter(args.gradient == "image")
    m.SetMovingTransform(tf)
    t0 = time.perf_counter()
    m.Initialize()  # uploads, device gradient filter, intensity ranges
    print("# initialize (upload + device gradient filter) %.2f s" % (time.perf_counter() - t0), file=sys.stderr,
          flush=True)

    for i in range(args.warmup):
        tf.set_parameters(wl.params[i % len(wl.params)])
        m.GetValueAndDerivative()
    t_total = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        tf.set_parameters(wl.params[i % len(wl.params)])
        v, d = m.GetValueAndDerivative()
        t_total += m.GetLastTiming()["total_ms"]
    wall = time.perf_counter() - t0
    K = args.steps
    ms_step = wall * 1e3 / K
    n = geom.num_pixels
    alg_bytes = n * (BYTES_PER_SAMPLE_3D if args.gradient == "image" else BYTES_PER_SAMPLE_3D_CD)[args.storage]
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
        "dtype": "%s images, f64 coordinates/PDF" % args.storage,
        "data": "synthetic (seeded blob volumes, see itk_amd/synth.py)",
        "config": dict({"workload": wl.name + ("" if args.storage == "f32" else "_" + args.storage) +
                        ("" if args.gradient == "image" else "_central_differences"), "bins": 50,
                        "samples_per_eval": n, "storage": args.storage,
                        "gradient_source": "recursive Gaussian gradient image (device)" if args.gradient == "image"
                        else "central differences of the moving image",
                        "parallelism": "single GPU"}, **wl.extra),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_EVAL.get((wl.name, args.storage, args.gradient)),
                     "kernel": "whole evaluation (clear, pass 1, finalize, pass 2, derivative), HIP events on "
                               "the metric stream around each evaluation",
                     "algorithmic_bytes_per_eval": alg_bytes, "device_ms_per_eval": dev_ms},
        "last_value": v,
    }
    m.close()
    line["roofline"]["passes"] = pass_breakdown(args, wl, n)
    copy = copy_bandwidth_gbs()
    line["roofline"]["copy_gbs"] = copy
    line["roofline"]["frac_of_copy"] = achieved / copy if copy else None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    if served is not None:
        # the metric's second half: registrations of mixed sizes served by the
        # continuous-batching server (a short run of the --serving workload)
        line["serving"] = served
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
