"""Batched evaluation of several device metrics (the C ABI's mmi_batch).

Metrics on the certified fast path share one launch of each kernel, captured
as a HIP graph and replayed while the batch keeps its shape; the others are
evaluated one by one while the fused launch runs.  Results are identical to
evaluating each metric on its own.
"""
from __future__ import annotations

import atexit
import ctypes

import numpy as np

from . import _native as N


class BatchEvaluator:
    def __init__(self, device: int = 0, graphs: bool = True):
        self._L = N.load()
        self._b = ctypes.c_void_p()
        N.check(self._L.mmi_batch_create(device, ctypes.byref(self._b)))
        N.check(self._L.mmi_batch_set_graphs(self._b, 1 if graphs else 0))

    def close(self):
        if self._b:
            self._L.mmi_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, metrics, want_derivative: bool = True):
        """GetValueAndDerivative (or GetValue) of every metric.  Returns one
        (value, derivative, n_valid, error) tuple per metric; error is None or
        the MetricError that metric's own evaluation would raise."""
        n = len(metrics)
        if n == 0:
            return []
        ctxs = (ctypes.c_void_p * n)()
        derivs = []
        dptrs = (N._dp * n)()
        for i, m in enumerate(metrics):
            if not m._initialized:
                raise N.MetricError(N.MMI_ERR_ARGUMENT, "Initialize() must be called first")
            m._push_parameters()
            ctxs[i] = m._ctx
            d = np.zeros(m.GetNumberOfParameters()) if want_derivative else None
            derivs.append(d)
            if d is not None:
                dptrs[i] = d.ctypes.data_as(N._dp)
        values = np.zeros(n)
        nvalid = np.zeros(n, dtype=np.int64)
        status = np.zeros(n, dtype=np.int32)
        N.check(self._L.mmi_batch_evaluate(self._b, ctxs, n, 1 if want_derivative else 0,
                                           values.ctypes.data_as(N._dp), dptrs if want_derivative else None,
                                           nvalid.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        out = []
        for i, m in enumerate(metrics):
            err = None
            m._n_valid = int(nvalid[i])
            if status[i] != N.MMI_OK:
                # the message this metric's own evaluation would raise
                msg = self._L.mmi_context_last_error(m._ctx).decode()
                err = N.MetricError(int(status[i]), msg or "evaluation failed with status %d" % status[i])
            else:
                m._value = float(values[i])
            out.append((float(values[i]), derivs[i], int(nvalid[i]), err))
        return out

    def regular_step(self, metrics, options, states, max_steps: int):
        """Up to max_steps RegularStep iterations for every metric whose state
        is still running (mmi_batch_regular_step); the metrics' transforms
        receive the new parameters."""
        n = len(metrics)
        if n == 0:
            return
        ctxs = (ctypes.c_void_p * n)()
        for i, m in enumerate(metrics):
            if not m._initialized:
                raise N.MetricError(N.MMI_ERR_ARGUMENT, "Initialize() must be called first")
            m._push_parameters()
            ctxs[i] = m._ctx
        arr = (N.RegularStepState * n)(*states)
        N.check(self._L.mmi_batch_regular_step(self._b, ctxs, n, ctypes.byref(options), arr, int(max_steps)))
        for i, m in enumerate(metrics):
            states[i] = arr[i]
            p = np.zeros(m.GetNumberOfParameters())
            N.check(self._L.mmi_get_moving_parameters(m._ctx, p.ctypes.data_as(N._dp), p.size))
            m._moving_tf.set_parameters(p)

    def set_step_loop(self, on_device: bool = True, chunk: int = 8):
        """Where regular_step iterates fused metrics: on the device in
        captured chunks of `chunk` iterations (default), or in the host loop;
        the iterates are the same bits either way."""
        N.check(self._L.mmi_batch_set_step_loop(self._b, 1 if on_device else 0, int(chunk)))

    def device_steps(self) -> int:
        """Iterations the last regular_step call ran in the device loop."""
        n = ctypes.c_int32()
        N.check(self._L.mmi_batch_get_step_count(self._b, ctypes.byref(n)))
        return n.value

    def last_timing(self) -> dict:
        ms = ctypes.c_float()
        nf = ctypes.c_int32()
        N.check(self._L.mmi_batch_get_last_timing(self._b, ctypes.byref(ms), ctypes.byref(nf)))
        return {"device_ms": ms.value, "fused": nf.value}


_shared = {}


def shared_evaluator(device: int = 0) -> BatchEvaluator:
    """One evaluator per device for single registrations' native loops."""
    if device not in _shared:
        _shared[device] = BatchEvaluator(device)
    return _shared[device]


@atexit.register
def _close_shared():
    for ev in _shared.values():
        ev.close()
    _shared.clear()
