"""Continuous-batching server: registrations of mixed sizes and kinds that
arrive while others run give the same final transforms as when each runs
alone (batched evaluations are bit-identical to single ones), and admission
respects the HBM budget."""
from __future__ import annotations

import numpy as np
import pytest

from itk_amd import synth
from itk_amd.serving import RegistrationRequest, RegistrationServer, latency_summary, run_sequential

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    shapes = [(40, 36, 32), (48, 44, 40), (32, 32, 28), (44, 40, 36)]
    for i, shape in enumerate(shapes):
        p = synth.make_affine_problem(shape, 100 + i, n_blobs=12)
        # arrivals close together: each registration takes only milliseconds,
        # and the test needs them to overlap
        reqs.append(RegistrationRequest(i, "affine", p.fixed, p.moving, p.geometry, arrival=0.0005 * i,
                                        shrink=[2, 1], sigmas=[1.0, 0.0], iterations=25))
    for j, shape in enumerate([(36, 34, 30), (32, 30, 28)]):
        p = synth.make_bspline_problem(shape, 110 + j, mesh=4, sigma_vox=1.0, n_blobs=12)
        reqs.append(RegistrationRequest(10 + j, "bspline", p.fixed, p.moving, p.geometry, arrival=0.00025 + 0.0005 * j,
                                        mesh=3, iterations=8))
    return reqs


def test_results_same_alone_and_in_the_mix():
    reqs = _requests()
    alone, _ = run_sequential(reqs)
    # a batching window long enough that concurrent registrations meet in a
    # lane however fast a single evaluation is
    server = RegistrationServer(max_wait_s=0.05, max_inflight=8)
    mixed = server.run(reqs)
    print("batches %d, evaluations %d, largest batch %d, peak in flight %d"
          % (server.batches, server.batched_evaluations, server.max_batch, server.peak_inflight))
    print(latency_summary(mixed))
    assert server.max_batch > 1
    for a, b in zip(alone, mixed):
        assert a.rid == b.rid and a.error is None and b.error is None
        assert a.evaluations == b.evaluations
        assert np.array_equal(a.parameters, b.parameters), a.rid


def test_admission_respects_hbm_budget():
    reqs = _requests()[:3]
    budget = max(r.device_bytes() for r in reqs) + 1
    server = RegistrationServer(hbm_budget_bytes=budget, max_inflight=8)  # the budget, not the cap, limits
    out = server.run(reqs)
    assert server.peak_inflight == 1
    # pooled contexts keep their buffers: the resident total never exceeds the budget
    assert server.peak_used <= budget
    alone, _ = run_sequential(reqs)
    for a, b in zip(alone, out):
        assert np.array_equal(a.parameters, b.parameters)
