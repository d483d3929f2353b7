"""Continuous-batching registration server on one GPU.

Requests (a fixed and a moving image, a transform kind) arrive over time.
A request is admitted when its device-memory estimate fits the HBM budget and
a slot is free; it then runs the ordinary registration method
(``ImageRegistrationMethodv4`` with its optimizer) in a worker thread whose
metric forwards every evaluation to the scheduler.  The scheduler gathers the
evaluations every running registration is waiting for and runs them as one
batch (``BatchEvaluator``: one fused, graph-replayed launch for the fast-path
metrics, the rest while it runs), so registrations of different sizes, kinds
and pyramid levels share launches; requests that arrive meanwhile join the
next batch.  Images, pyramids (smoothed levels, gradient images), histograms
and optimizer state stay resident until the request completes.  Because a
batched evaluation is bit-identical to a single one, each registration's
result does not depend on what it was batched with.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import BatchEvaluator
from .geometry import ImageGeometry
from .metric import MattesMutualInformationImageToImageMetricv4
from .optimizers import LBFGSBOptimizerv4, RegularStepGradientDescentOptimizerv4
from .registration import ImageRegistrationMethodv4
from .scales import RegistrationParameterScalesFromPhysicalShift
from .transforms import AffineTransform, BSplineTransform


@dataclass
class RegistrationRequest:
    rid: int
    kind: str                 # "affine" or "bspline"
    fixed: np.ndarray
    moving: np.ndarray
    geometry: ImageGeometry
    arrival: float = 0.0      # seconds after the run starts
    mesh: int = 8             # B-spline mesh size per axis
    shrink: list = field(default_factory=lambda: [4, 2, 1])
    sigmas: list = field(default_factory=lambda: [2.0, 1.0, 0.0])
    iterations: int = 100     # per level (affine) / in total (B-spline)
    learning_rate: float = 2.0

    def device_bytes(self) -> int:
        """Resident estimate: FP32 uploads and smoothed levels (16 B), gradient
        texels (16 B), the gradient filter's FP64 temporaries (40 B) per voxel."""
        return int(self.geometry.num_pixels) * 72 + (64 << 20)


def build_registration(req: RegistrationRequest, metric):
    """The registration each request runs: affine with RegularStep gradient
    descent over the pyramid, or a cubic B-spline with L-BFGS-B on one level."""
    geom = req.geometry
    reg = ImageRegistrationMethodv4()
    reg.SetFixedImage(req.fixed, geom)
    reg.SetMovingImage(req.moving, geom)
    reg.SetMetric(metric)
    if req.kind == "affine":
        tf = AffineTransform(geom.dim, center=geom.center_point())
        opt = RegularStepGradientDescentOptimizerv4(
            learning_rate=req.learning_rate, minimum_step_length=1e-3, relaxation_factor=0.5,
            number_of_iterations=req.iterations, scales_estimator=RegistrationParameterScalesFromPhysicalShift())
        reg.SetShrinkFactorsPerLevel(req.shrink)
        reg.SetSmoothingSigmasPerLevel(req.sigmas)
        reg.SetSmoothingSigmasAreSpecifiedInPhysicalUnits(False)
    elif req.kind == "bspline":
        tf = BSplineTransform.for_image(geom, [req.mesh] * geom.dim)
        opt = LBFGSBOptimizerv4(number_of_iterations=req.iterations)
        reg.SetShrinkFactorsPerLevel([1])
        reg.SetSmoothingSigmasPerLevel([0.0])
    else:
        raise ValueError("unknown request kind %r" % req.kind)
    reg.SetOptimizer(opt)
    reg.SetInitialTransform(tf)
    return reg, tf, opt


@dataclass
class RegistrationResult:
    rid: int
    kind: str
    parameters: np.ndarray
    arrival: float
    admitted: float
    finished: float
    evaluations: int
    error: Exception = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _BatchingMetric:
    """The device metric with its evaluations routed through the scheduler."""

    def __init__(self, metric, server, job):
        self._m = metric
        self._server = server
        self._job = job

    def __getattr__(self, name):
        return getattr(self._m, name)

    def GetValueAndDerivative(self):
        return self._server._evaluate(self._job, True)

    def GetValue(self):
        return self._server._evaluate(self._job, False)[0]

    def GetDerivative(self):
        return self._server._evaluate(self._job, True)[1]


class _Job:
    def __init__(self, server, req: RegistrationRequest, device: int, storage: str):
        self.req = req
        self.metric = MattesMutualInformationImageToImageMetricv4(device=device)
        self.metric.SetImageStorage(storage)
        self.proxy = _BatchingMetric(self.metric, server, self)
        self.server = server
        self.admitted = 0.0
        self.evaluations = 0
        self.result = None
        self.thread = threading.Thread(target=self._run, daemon=True)

    def _run(self):
        err = None
        tf = None
        try:
            reg, tf, _ = build_registration(self.req, self.proxy)
            reg.Update()
        except Exception as e:  # reported in the result
            err = e
        params = tf.get_parameters() if tf is not None else None
        self.server._finish(self, params, err)


class RegistrationServer:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 64 << 30, max_inflight: int = 32,
                 max_wait_s: float = 0.002, graphs: bool = True, storage: str = "f32"):
        self.device = device
        self.budget = int(hbm_budget_bytes)
        self.max_inflight = int(max_inflight)
        self.max_wait = float(max_wait_s)
        self.storage = storage
        self.batch = BatchEvaluator(device, graphs=graphs)
        self.cv = threading.Condition()
        self.pending = []      # [job, want_derivative, slot]
        self.running = set()
        self.results = []
        self.batches = 0
        self.batched_evaluations = 0
        self.max_batch = 0
        self.peak_inflight = 0

    # -- called from worker threads ----------------------------------------------
    def _evaluate(self, job, want_derivative: bool):
        slot = {"done": False}
        with self.cv:
            self.pending.append((job, want_derivative, slot))
            self.cv.notify_all()
            while not slot["done"]:
                self.cv.wait()
        job.evaluations += 1
        if slot["err"] is not None:
            raise slot["err"]
        return slot["value"], slot["deriv"]

    def _finish(self, job, params, err):
        with self.cv:
            job.result = RegistrationResult(job.req.rid, job.req.kind, params, job.req.arrival, job.admitted,
                                            self._now(), job.evaluations, err)
            self.results.append(job.result)
            self.running.discard(job)
            self.used -= job.req.device_bytes()
            self.cv.notify_all()
        job.metric.close()

    # -- scheduler ---------------------------------------------------------------
    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _run_batch(self, items):
        for want in (True, False):
            group = [it for it in items if it[1] == want]
            if not group:
                continue
            out = self.batch.evaluate([job.metric for job, _, _ in group], want_derivative=want)
            self.batches += 1
            self.batched_evaluations += len(group)
            self.max_batch = max(self.max_batch, len(group))
            with self.cv:
                for (job, _, slot), (v, d, nv, err) in zip(group, out):
                    slot.update(value=v, deriv=d, err=err, done=True)
                self.cv.notify_all()

    def run(self, requests):
        """Serve the requests (arrival times relative to now); returns the results."""
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self.t0 = time.perf_counter()
        self.used = 0
        total = len(queue)
        while True:
            with self.cv:
                now = self._now()
                # admission: arrived, within budget and slots, in arrival order
                while queue and queue[0].arrival <= now and len(self.running) < self.max_inflight and (
                        self.used + queue[0].device_bytes() <= self.budget or not self.running):
                    req = queue.pop(0)
                    job = _Job(self, req, self.device, self.storage)
                    job.admitted = now
                    self.used += req.device_bytes()
                    self.running.add(job)
                    self.peak_inflight = max(self.peak_inflight, len(self.running))
                    job.thread.start()
                if len(self.results) == total:
                    break
                if not self.running:
                    # idle until the next arrival
                    self.cv.wait(max(1e-4, queue[0].arrival - self._now()))
                    continue
                # gather: every running registration waiting, or the first
                # waiter has waited max_wait, or a new request has arrived
                t_first = None
                while len(self.pending) < len(self.running):
                    if self.pending and t_first is None:
                        t_first = time.perf_counter()
                    if t_first is not None and time.perf_counter() - t_first >= self.max_wait:
                        break
                    if queue and queue[0].arrival <= self._now() and len(self.running) < self.max_inflight:
                        break
                    timeout = self.max_wait if self.pending else 0.01
                    if queue:
                        timeout = min(timeout, max(1e-4, queue[0].arrival - self._now()))
                    self.cv.wait(timeout)
                    if len(self.results) == total:
                        break
                items = self.pending
                self.pending = []
            if items:
                self._run_batch(items)
        self.batch.close()
        return sorted(self.results, key=lambda r: r.rid)


def run_sequential(requests, device: int = 0, storage: str = "f32"):
    """The same requests one at a time on the plain metric, in arrival order
    (each starts at max(arrival, previous finish))."""
    queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
    t0 = time.perf_counter()
    results = []
    busy = 0.0
    for req in queue:
        now = time.perf_counter() - t0
        if now < req.arrival:
            time.sleep(req.arrival - now)
        start = time.perf_counter() - t0
        m = MattesMutualInformationImageToImageMetricv4(device=device)
        m.SetImageStorage(storage)
        err = None
        tf = None
        calls = [0]
        orig = m.GetValueAndDerivative

        def counted():
            calls[0] += 1
            return orig()
        m.GetValueAndDerivative = counted
        try:
            reg, tf, _ = build_registration(req, m)
            reg.Update()
        except Exception as e:
            err = e
        m.close()
        end = time.perf_counter() - t0
        busy += end - start
        results.append(RegistrationResult(req.rid, req.kind, tf.get_parameters() if tf is not None else None,
                                          req.arrival, start, end, calls[0], err))
    return sorted(results, key=lambda r: r.rid), busy


def latency_summary(results) -> dict:
    lat = np.array([r.latency for r in results if r.error is None])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {
        "completed": int(lat.size),
        "failed": sum(1 for r in results if r.error is not None),
        "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
        "mean_s": float(lat.mean()) if lat.size else None,
        "makespan_s": float(span),
        "registrations_per_s": float(lat.size / span) if span > 0 else None,
    }
