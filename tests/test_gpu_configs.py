"""The benchmark workloads at their own shapes on the HIP path, against the
oracle (the FP64 restatement of the reference's dense-sampling metric).

* 192^3 cubic B-spline, mesh 16^3 (19^3 x 3 = 20,577 parameters): at the
  full shape every sample's Parzen bins and the valid count are bit-exact and
  the value within 1e-5; the derivative (1e-4, L2) over a 24-slice slab of
  the virtual domain with the full images and the full grid (the oracle's
  dense dPDF is 411 MB per thread, a full-volume derivative takes it minutes).
* BF16 / FP16 storage: the error of the 16-bit metric against the FP64 oracle
  on the unrounded images, worst of 10 random parameter sets, for the 256^3
  affine problem (value and derivative at full size) and the 192^3 B-spline
  problem (value at full size, derivative over a 16-slice slab); FP32 storage
  at the same parameters stays within the parity bar.
* Registrations of the serving mix's sizes (128^3 ... 256x256x128, affine and
  B-spline with 8^3 and 16^3 meshes) run several at once through the
  continuous-batching server: batched evaluations happen, and every final
  transform equals the same registration run alone bit for bit.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from _gpu import DERIV_RTOL, VALUE_RTOL
from itk_amd import synth
from itk_amd.geometry import ImageGeometry
from itk_amd.metric import MattesMutualInformationImageToImageMetricv4 as Metric
from itk_amd.serving import RegistrationRequest, RegistrationServer, latency_summary, run_sequential
from itk_amd.transforms import AffineTransform, BSplineTransform

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
# 16-bit storage against the FP64 oracle: relative value error and relative L2
# derivative error.  A half-swapped texel, a channel mix-up or a format mix-up
# (BF16 read as FP16) gives errors of order 1.
STORAGE_BOUND = {"bf16": (5e-3, 5e-2), "f16": (2e-3, 2e-2)}


def _slab(geom: ImageGeometry, z0: int, nz: int) -> ImageGeometry:
    """Slices z0 .. z0 + nz - 1 of geom (the same physical points)."""
    return ImageGeometry(size=(geom.size[0], geom.size[1], nz), origin=geom.origin, spacing=geom.spacing,
                         direction=geom.direction, start=(0, 0, z0))


def _metric(p, geom, tf, storage="f32", record=False):
    m = Metric(record_sample_bins=record)
    m.SetImageStorage(storage)
    m.SetFixedImage(p.fixed, geom)
    m.SetMovingImage(p.moving, geom)
    m.SetMovingTransform(tf)
    m.Initialize()
    return m


def _rel(a, b):
    return abs(a - b) / abs(b)


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def bspline192(oracle_mod):
    print("synthesizing the 192^3 B-spline problem", flush=True)
    p = synth.make_bspline_problem((192, 192, 192), 2024, mesh=8, sigma_vox=2.0)
    grad = oracle_mod.gradient_image(p.moving, p.geometry, threads=THREADS)
    return p, p.geometry, grad


@pytest.fixture(scope="module")
def affine256(oracle_mod):
    print("synthesizing the 256^3 affine problem", flush=True)
    p = synth.make_affine_problem((256, 256, 256), 2024)
    grad = oracle_mod.gradient_image(p.moving, p.geometry, threads=THREADS)
    return p, p.geometry, grad


@pytest.mark.parametrize("k", [0, 1])
def test_config2_bspline_192cube_mesh16(oracle_mod, bspline192, k):
    p, geom, grad = bspline192
    tf = BSplineTransform.for_image(geom, [16, 16, 16])
    assert tf.number_of_parameters == 20577
    tf.set_parameters(np.random.default_rng(500 + k).normal(0.0, 0.5, tf.number_of_parameters))
    m = _metric(p, geom, tf, record=True)
    v = m.GetValue()
    assert m.GetLastEvaluationPath() == "fast"
    print("full shape: value", v, flush=True)
    r = oracle_mod.mattes(p.fixed, geom, p.moving, geom, tf, moving_gradient=grad, derivative=False,
                          want_samples=True, threads=THREADS)
    assert m.GetNumberOfValidPoints() == r["n_valid"]
    fb, mb = m.GetSampleBins()
    assert np.array_equal(fb, r["fixed_bin"])
    assert np.array_equal(mb, r["moving_bin"])
    assert _rel(v, r["value"]) <= VALUE_RTOL
    m.close()
    slab = _slab(geom, 84, 24)
    m = _metric(p, geom, tf)
    m.SetVirtualDomain(slab)
    m.Initialize()
    vs, ds = m.GetValueAndDerivative()
    assert m.GetLastEvaluationPath() == "fast"
    rs = oracle_mod.mattes(p.fixed, geom, p.moving, geom, tf, moving_gradient=grad, virtual_geom=slab,
                           threads=THREADS)
    print("slab: value %.15g oracle %.15g, derivative rel L2 %.3e, nonzero %d of %d"
          % (vs, rs["value"], _rel_l2(ds, rs["derivative"]), np.count_nonzero(rs["derivative"]), ds.size))
    assert m.GetNumberOfValidPoints() == rs["n_valid"]
    assert _rel(vs, rs["value"]) <= VALUE_RTOL
    assert _rel_l2(ds, rs["derivative"]) <= DERIV_RTOL
    m.close()


def _storage_errors(oracle_mod, p, geom, grad, tf, params, virtual=None, value_only_full=False):
    """Worst value and derivative errors of f32 / bf16 / f16 storage against
    the oracle over the parameter sets."""
    ms = {s: _metric(p, geom, tf, storage=s) for s in ("f32", "bf16", "f16")}
    if virtual is not None:
        mv = {s: _metric(p, geom, tf, storage=s) for s in ms}
        for m in mv.values():
            m.SetVirtualDomain(virtual)
            m.Initialize()
    worst = {s: [0.0, 0.0] for s in ms}
    for i, q in enumerate(params):
        tf.set_parameters(q)
        if value_only_full:
            r = oracle_mod.mattes(p.fixed, geom, p.moving, geom, tf, moving_gradient=grad, derivative=False,
                                  threads=THREADS)
            rd = oracle_mod.mattes(p.fixed, geom, p.moving, geom, tf, moving_gradient=grad, virtual_geom=virtual,
                                   threads=THREADS)
        else:
            r = rd = oracle_mod.mattes(p.fixed, geom, p.moving, geom, tf, moving_gradient=grad, threads=THREADS)
        for s in ms:
            if value_only_full:
                v = ms[s].GetValue()
                _, d = mv[s].GetValueAndDerivative()
            else:
                v, d = ms[s].GetValueAndDerivative()
            worst[s][0] = max(worst[s][0], _rel(v, r["value"]))
            worst[s][1] = max(worst[s][1], _rel_l2(d, rd["derivative"]))
        print("parameter set %d: %s" % (i, {s: "%.2e / %.2e" % tuple(w) for s, w in worst.items()}), flush=True)
    for m in list(ms.values()) + (list(mv.values()) if virtual is not None else []):
        m.close()
    return worst


def test_config3_affine_256cube_16bit_storage_error(oracle_mod, affine256):
    p, geom, grad = affine256
    tf = AffineTransform(3, center=geom.index_to_point([(x - 1) / 2.0 for x in geom.size]))
    base = tf.get_parameters()
    rng = np.random.default_rng(600)
    params = [base + np.concatenate([rng.uniform(-0.02, 0.02, 9), rng.uniform(-1.5, 1.5, 3)]) for _ in range(10)]
    worst = _storage_errors(oracle_mod, p, geom, grad, tf, params)
    print("256^3 affine, worst of 10 (value / derivative):", worst)
    assert worst["f32"][0] <= VALUE_RTOL and worst["f32"][1] <= DERIV_RTOL
    for s in ("bf16", "f16"):
        assert worst[s][0] <= STORAGE_BOUND[s][0] and worst[s][1] <= STORAGE_BOUND[s][1], s


def test_config3_bspline_192cube_16bit_storage_error(oracle_mod, bspline192):
    p, geom, grad = bspline192
    tf = BSplineTransform.for_image(geom, [16, 16, 16])
    rng = np.random.default_rng(700)
    params = [rng.normal(0.0, 0.5, tf.number_of_parameters) for _ in range(10)]
    worst = _storage_errors(oracle_mod, p, geom, grad, tf, params, virtual=_slab(geom, 88, 16), value_only_full=True)
    print("192^3 B-spline mesh 16, worst of 10 (value full / derivative slab):", worst)
    assert worst["f32"][0] <= VALUE_RTOL and worst["f32"][1] <= DERIV_RTOL
    for s in ("bf16", "f16"):
        assert worst[s][0] <= STORAGE_BOUND[s][0] and worst[s][1] <= STORAGE_BOUND[s][1], s


def _serving_mix():
    specs = [(0, "affine", (128, 128, 128), 8), (1, "affine", (160, 160, 96), 8), (2, "bspline", (192, 192, 192), 8),
             (3, "bspline", (128, 128, 128), 16), (4, "affine", (256, 256, 128), 8)]
    reqs = []
    for rid, kind, size, mesh in specs:
        print("synthesizing request %d: %s %s" % (rid, kind, size), flush=True)
        if kind == "affine":
            p = synth.make_affine_problem(size, 900 + rid)
            reqs.append(RegistrationRequest(rid, kind, p.fixed, p.moving, p.geometry, arrival=0.0))
        else:
            p = synth.make_bspline_problem(size, 900 + rid, mesh=8, sigma_vox=2.0)
            reqs.append(RegistrationRequest(rid, kind, p.fixed, p.moving, p.geometry, arrival=0.0, mesh=mesh,
                                            iterations=50))
    return reqs


def test_config4_serving_mix_same_alone_and_batched():
    reqs = _serving_mix()
    alone, _ = run_sequential(reqs)
    print("alone:", latency_summary(alone), flush=True)
    server = RegistrationServer(max_inflight=3)  # the server's defaults: three in flight, one fused lane
    mixed = server.run(reqs)
    print("batches %d, evaluations %d, largest batch %d, peak in flight %d"
          % (server.batches, server.batched_evaluations, server.max_batch, server.peak_inflight))
    print("mixed:", latency_summary(mixed), flush=True)
    assert server.peak_inflight > 1 and server.max_batch > 1
    for a, b in zip(alone, mixed):
        assert a.rid == b.rid and a.error is None and b.error is None, (a.error, b.error)
        assert a.evaluations == b.evaluations, a.rid
        assert np.array_equal(a.parameters, b.parameters), a.rid
