"""Admission accounting of the registration server (host logic, no GPU): idle
pooled contexts keep their grow-only device buffers, so their resident bytes
stay in the budget until a registration that would not fit releases them."""
from __future__ import annotations

import numpy as np

from itk_amd.geometry import ImageGeometry
from itk_amd.serving import RegistrationRequest, RegistrationServer


class _Idle:
    def __init__(self, resident):
        self._resident_bytes = resident
        self.closed = False

    def close(self):
        self.closed = True


def _server(budget, pool):
    s = RegistrationServer(hbm_budget_bytes=budget)
    s.pool = list(pool)
    s.used = sum(m._resident_bytes for m in pool)
    return s


def test_request_estimate_counts_every_resident_buffer():
    g = ImageGeometry(size=(64, 64, 64))
    img = np.zeros((64, 64, 64), np.float32)
    n = 64 ** 3
    affine = RegistrationRequest(0, "affine", img, img, g)
    bspline = RegistrationRequest(1, "bspline", img, img, g)
    # uploads + smoothed (16), texels (16), filter volumes (48), FP64 gradient (24), per-sample bytes (2),
    # and the affine bordered copies (20) or the B-spline factors (12)
    assert affine.device_bytes() == n * 126 + (64 << 20)
    assert bspline.device_bytes() == n * 118 + (64 << 20)


def test_pooled_context_that_holds_the_request_adds_nothing():
    small, large = _Idle(100), _Idle(300)
    s = _server(1000, [small, large])
    metric, extra = s._pooled_metric(250)
    assert metric is large and extra == 0          # the smallest that holds it
    metric, extra = s._pooled_metric(50)
    assert metric is small and extra == 0


def test_growing_a_pooled_context_counts_the_growth():
    a = _Idle(100)
    s = _server(1000, [a])
    metric, extra = s._pooled_metric(400)
    assert metric is a and extra == 300


def test_idle_contexts_are_released_when_the_budget_needs_them():
    a, b, c = _Idle(400), _Idle(300), _Idle(200)
    s = _server(1000, [a, b, c])                    # 900 resident
    metric, extra = s._pooled_metric(700)           # grows the largest (a) by 300
    assert metric is a and extra == 300
    assert b.closed and not c.closed                # the largest other idle context goes first
    assert s.used + extra <= 1000
    assert s.pool == [a, c]


def test_new_context_when_nothing_is_pooled():
    s = _server(1000, [])
    metric, extra = s._pooled_metric(600)
    assert metric is None and extra == 600


class _Job:
    def __init__(self, lane):
        self.lane, self.evaluations, self.metric = lane, 0, object()


class _FailingBatch:
    def regular_step(self, *a, **k):
        raise RuntimeError("launch failed")

    def evaluate(self, *a, **k):
        raise RuntimeError("launch failed")

    def close(self):
        pass


def test_a_failed_launch_fails_its_registrations_instead_of_stopping_the_lane():
    """A launch that raises in a batching lane is delivered to the
    registrations it served; the lane keeps serving the next ones."""
    import threading
    import types

    import pytest

    from itk_amd.serving import _Lane

    s = RegistrationServer(hbm_budget_bytes=1 << 30, wait_limit_s=10.0)
    lane = object.__new__(_Lane)  # no device: the lane's own fields with a batch that fails
    lane.server, lane.batch = s, _FailingBatch()
    lane.cv = threading.Condition(s.lock)
    lane.pending, lane.first, lane.optimizing, lane.running = [], None, [], set()
    lane.thread = threading.Thread(target=lane._loop, daemon=True)
    lane.thread.start()
    job = _Job(lane)
    with s.lock:
        lane.running.add(job)
    opts = types.SimpleNamespace(learning_rate=1.0, minimum_step_length=1e-3, relaxation_factor=0.5,
                                 gradient_magnitude_tolerance=1e-4, number_of_iterations=10)
    for _ in range(2):
        with pytest.raises(RuntimeError, match="launch failed"):
            s._optimize(job, opts, types.SimpleNamespace(iteration=0, stop=0))
        with pytest.raises(RuntimeError, match="launch failed"):
            s._evaluate(job, True)
    with s.lock:
        s.stopping = True
        lane.running.clear()
        lane.cv.notify()
    lane.thread.join(5.0)
    assert not lane.thread.is_alive()
