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
        """Resident estimate of the registration's context: per voxel the FP32
        uploads and smoothed levels (16 B), the gradient texels (16 B), the
        filter chains' FP64 volumes (6 x 8 B) and FP64 gradient (24 B), the
        per-sample bytes (fixed bins, classes: 2 B), and for an affine
        registration the row walks' bordered texels and values (20 B), for a
        B-spline one its per-sample derivative factors (12 B); plus 64 MB."""
        per = 16 + 16 + 48 + 24 + 2 + (20 if self.kind == "affine" else 12)
        return int(self.geometry.num_pixels) * per + (64 << 20)


def build_registration(req: RegistrationRequest, metric, native_loop: bool = True):
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
        opt.use_native_loop = native_loop
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
    """The device metric with its evaluations routed through the scheduler
    (a job without a lane evaluates on its own context and stream)."""

    def __init__(self, metric, server, job):
        self._m = metric
        self._server = server
        self._job = job

    def __getattr__(self, name):
        return getattr(self._m, name)

    def _eval(self, want_derivative: bool):
        if self._job.lane is None:
            self._job.evaluations += 1
            return self._m.GetValueAndDerivative() if want_derivative else (self._m.GetValue(), None)
        return self._server._evaluate(self._job, want_derivative)

    def GetValueAndDerivative(self):
        return self._eval(True)

    def GetValue(self):
        return self._eval(False)[0]

    def GetDerivative(self):
        return self._eval(True)[1]


class _Job:
    def __init__(self, server, lane, req: RegistrationRequest, metric):
        self.req = req
        self.lane = lane
        self.metric = metric
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
            reg, tf, opt = build_registration(self.req, self.proxy)
            if isinstance(opt, RegularStepGradientDescentOptimizerv4):
                opt.native_runner = lambda m, o, st: self.server._optimize(self, o, st)
            reg.Update()
        except Exception as e:  # reported in the result
            err = e
        params = tf.get_parameters() if tf is not None else None
        self.server._finish(self, params, err)


class _Lane:
    """One batching loop with its own stream: it gathers the evaluations its
    registrations wait for and runs them as one batch.  With two lanes, one
    lane's registrations take their host-side optimizer steps while the
    other's batch runs on the device."""

    def __init__(self, server, device: int, graphs: bool):
        self.server = server
        self.batch = BatchEvaluator(device, graphs=graphs)
        self.cv = threading.Condition(server.lock)
        self.pending = []
        self.first = None
        self.optimizing = []  # [job, options, state, event]: RegularStep levels run in the C loop
        self.running = set()
        self.thread = threading.Thread(target=self._loop, daemon=True)

    def _loop(self):
        srv = self.server
        while True:
            with srv.lock:
                while True:
                    if srv.stopping and not self.running and not self.pending:
                        self.batch.close()
                        return
                    waiting = len(self.pending) + len(self.optimizing)
                    if waiting and (waiting >= len(self.running) or
                                    time.perf_counter() - self.first >= srv.max_wait):
                        break
                    self.cv.wait(srv.max_wait if waiting else 0.05)
                items, self.pending = self.pending, []
                self.first = time.perf_counter() if self.optimizing else None
                opt = list(self.optimizing)
            # a failure of a launch fails the registrations it served (their
            # threads raise it) instead of stopping the lane with them waiting
            if opt:
                try:
                    self._run_optimizing(opt)
                except Exception as e:
                    with srv.lock:  # the levels not finished before the failure
                        failed = [it for it in opt if it in self.optimizing]
                        for it in failed:
                            self.optimizing.remove(it)
                    for it in failed:
                        it.append(e)
                        it[3].set()
            if items:
                try:
                    self._run_batch(items)
                except Exception as e:
                    for job, _, slot in items:
                        if not slot["event"].is_set():
                            slot["value"], slot["deriv"], slot["err"] = None, None, e
                            slot["event"].set()

    def _run_optimizing(self, jobs):
        """A few RegularStep iterations of every registration optimizing in this
        lane, grouped by optimizer settings, in the C loop."""
        srv = self.server
        groups = {}
        for it in jobs:
            o = it[1]
            key = (o.learning_rate, o.minimum_step_length, o.relaxation_factor, o.gradient_magnitude_tolerance,
                   o.number_of_iterations)
            groups.setdefault(key, []).append(it)
        for group in groups.values():
            states = [it[2] for it in group]
            before = sum(int(st.iteration) for st in states)
            self.batch.regular_step([it[0].metric for it in group], group[0][1], states, srv.chunk_steps)
            done = []
            with srv.lock:
                srv.batches += srv.chunk_steps
                srv.batched_evaluations += sum(int(st.iteration) for st in states) - before
                srv.max_batch = max(srv.max_batch, len(group))
                for it, st in zip(group, states):
                    it[2] = st
                    if st.stop != 0:
                        self.optimizing.remove(it)
                        done.append(it)
            for it in done:
                it[3].set()

    def _run_batch(self, items):
        srv = self.server
        for want in (True, False):
            group = [it for it in items if it[1] == want]
            if not group:
                continue
            out = self.batch.evaluate([job.metric for job, _, _ in group], want_derivative=want)
            with srv.lock:
                srv.batches += 1
                srv.batched_evaluations += len(group)
                srv.max_batch = max(srv.max_batch, len(group))
            for (job, _, slot), (v, d, nv, err) in zip(group, out):
                slot["value"], slot["deriv"], slot["err"] = v, d, err
                slot["event"].set()


class RegistrationServer:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 64 << 30, max_inflight: int = 3,
                 max_wait_s: float = 0.0005, graphs: bool = True, storage: str = "f32", lanes: int = 1,
                 chunk_steps: int = 8, wait_limit_s: float = 600.0):
        self.device = device
        # a registration whose evaluation is not served within this time fails
        # (reported in its result) instead of waiting forever
        self.wait_limit = float(wait_limit_s)
        self.budget = int(hbm_budget_bytes)
        self.max_inflight = int(max_inflight)
        self.max_wait = float(max_wait_s)
        self.storage = storage
        self.graphs = graphs
        self.n_lanes = max(1, int(lanes))
        self.chunk_steps = max(1, int(chunk_steps))
        self.lock = threading.Lock()
        self.main_cv = threading.Condition(self.lock)
        self.results = []
        self.batches = 0
        self.batched_evaluations = 0
        self.max_batch = 0
        self.peak_inflight = 0
        self.stopping = False
        # idle metrics: contexts whose device buffers only grow; their resident
        # bytes (the largest registration each has held) stay in `used`
        self.pool = []
        self.peak_used = 0

    # -- called from worker threads ----------------------------------------------
    def _evaluate(self, job, want_derivative: bool):
        slot = {"event": threading.Event()}
        lane = job.lane
        with self.lock:
            if not lane.pending:
                lane.first = time.perf_counter()
            lane.pending.append((job, want_derivative, slot))
            if len(lane.pending) >= len(lane.running):
                lane.cv.notify()
        if not slot["event"].wait(self.wait_limit):
            raise RuntimeError("batched evaluation not served within %.0f s" % self.wait_limit)
        job.evaluations += 1
        if slot["err"] is not None:
            raise slot["err"]
        return slot["value"], slot["deriv"]

    def _optimize(self, job, options, state):
        """Hand one RegularStep level to the job's lane; returns the final state."""
        it = [job, options, state, threading.Event()]
        lane = job.lane
        with self.lock:
            if not lane.pending and not lane.optimizing:
                lane.first = time.perf_counter()
            lane.optimizing.append(it)
            lane.cv.notify()
        if not it[3].wait(self.wait_limit):
            raise RuntimeError("batched optimizer level not served within %.0f s" % self.wait_limit)
        if len(it) > 4:
            raise it[4]
        job.evaluations += int(it[2].iteration)  # one evaluation per iteration
        return it[2]

    def _finish(self, job, params, err):
        with self.lock:
            self.pool.append(job.metric)  # device buffers kept for the next registration
            job.result = RegistrationResult(job.req.rid, job.req.kind, params, job.req.arrival, job.admitted,
                                            self._now(), job.evaluations, err)
            self.results.append(job.result)
            if job.lane is not None:
                job.lane.running.discard(job)
                job.lane.cv.notify()
            self.running -= 1
            self.main_cv.notify()

    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _pooled_metric(self, need: int):
        """The idle metric a registration of `need` resident bytes takes (the
        smallest that holds it, else the largest, which grows; None: a new
        one) and the bytes `used` grows by.  Other idle metrics are released,
        largest first, while the registration would not fit the budget."""
        fits = [m for m in self.pool if m._resident_bytes >= need]
        if fits:
            metric = min(fits, key=lambda m: m._resident_bytes)
        elif self.pool:
            metric = max(self.pool, key=lambda m: m._resident_bytes)
        else:
            metric = None
        extra = need if metric is None else max(0, need - metric._resident_bytes)
        while self.used + extra > self.budget:
            others = [m for m in self.pool if m is not metric]
            if not others:
                break
            victim = max(others, key=lambda m: m._resident_bytes)
            self.pool.remove(victim)
            self.used -= victim._resident_bytes
            victim.close()
        return metric, extra

    def run(self, requests):
        """Serve the requests (arrival times relative to now); returns the results."""
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self.t0 = time.perf_counter()
        self.used = 0
        self.running = 0
        self.stopping = False
        lanes = [_Lane(self, self.device, self.graphs) for _ in range(self.n_lanes)]
        total = len(queue)
        for ln in lanes:
            ln.thread.start()
        with self.lock:
            while len(self.results) < total:
                now = self._now()
                # admission: arrived, within budget and slots, in arrival order;
                # each registration joins the lane with fewer running ones
                while queue and queue[0].arrival <= now and self.running < self.max_inflight:
                    req = queue[0]
                    need = req.device_bytes()
                    metric, extra = self._pooled_metric(need)
                    if self.used + extra > self.budget and self.running > 0:
                        break
                    queue.pop(0)
                    # affine levels share the lanes' fused RegularStep loops;
                    # a B-spline registration (L-BFGS-B steps on the host between
                    # evaluations) runs on its own stream, so no lane waits for it
                    lane = min(lanes, key=lambda ln: len(ln.running)) if req.kind == "affine" else None
                    if metric is not None:
                        self.pool.remove(metric)
                    else:
                        metric = MattesMutualInformationImageToImageMetricv4(device=self.device)
                        metric.SetImageStorage(self.storage)
                    metric._resident_bytes = max(getattr(metric, "_resident_bytes", 0), need)
                    job = _Job(self, lane, req, metric)
                    job.admitted = now
                    self.used += extra
                    self.peak_used = max(self.peak_used, self.used)
                    self.running += 1
                    if lane is not None:
                        lane.running.add(job)
                    self.peak_inflight = max(self.peak_inflight, self.running)
                    job.thread.start()
                timeout = 0.05
                if queue:
                    timeout = min(timeout, max(1e-4, queue[0].arrival - self._now()))
                self.main_cv.wait(timeout)
            self.stopping = True
            for ln in lanes:
                ln.cv.notify()
        for ln in lanes:
            ln.thread.join()
        for m in self.pool:
            m.close()
        self.pool = []
        self.used = 0
        return sorted(self.results, key=lambda r: r.rid)


def run_sequential(requests, device: int = 0, storage: str = "f32", native_loop: bool = False):
    """The same requests one at a time on the plain metric (one metric reused,
    as a one-at-a-time service would), in arrival order: each starts at
    max(arrival, previous finish).  By default the optimizer calls the metric
    once per iteration from Python, as with the reference's optimizers;
    native_loop runs affine levels in the C loop instead."""
    queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
    m = MattesMutualInformationImageToImageMetricv4(device=device)
    m.SetImageStorage(storage)
    calls = [0]
    orig = m.GetValueAndDerivative

    def counted():
        calls[0] += 1
        return orig()
    m.GetValueAndDerivative = counted
    t0 = time.perf_counter()
    results = []
    busy = 0.0
    for req in queue:
        now = time.perf_counter() - t0
        if now < req.arrival:
            time.sleep(req.arrival - now)
        start = time.perf_counter() - t0
        calls[0] = 0
        err = None
        tf = None
        try:
            reg, tf, _ = build_registration(req, m, native_loop=native_loop)
            reg.Update()
        except Exception as e:
            err = e
        end = time.perf_counter() - t0
        busy += end - start
        results.append(RegistrationResult(req.rid, req.kind, tf.get_parameters() if tf is not None else None,
                                          req.arrival, start, end, calls[0], err))
    m.close()
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
