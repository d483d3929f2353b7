"""Pins the oracle's sparse point-set sampling, image masks and mask- or
point-aware intensity ranges (oracle/mattes_oracle.cpp, orc_mattes_evaluate_ex
and orc_intensity_range) to equivalences the reference's definitions fix:

* the point set of every virtual voxel centre, in buffer order, is the dense
  evaluation;
* a fixed mask that holds a box of voxels is the dense evaluation over that
  box as a virtual sub-region;
* a moving mask is the point set of the voxels whose mapped moving point
  rounds into a non-zero mask pixel (worked out here in numpy);
* ranges are numpy min / max over the selected pixels.

Also pins RoundHalfIntegerUp and TransformPhysicalPointToIndex (known answers,
and numpy's floor(x + 0.5) away from ties).  CPU only.
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import affine_for, affine_params, blob_problem, bspline_for
from itk_amd.geometry import ImageGeometry
from itk_amd.masks import ImageMaskSpatialObject


def _voxel_points(geom: ImageGeometry) -> np.ndarray:
    """Physical points of every voxel, buffer order (x fastest), as
    ImageBase::TransformIndexToPhysicalPoint sums them."""
    d = geom.dim
    grids = np.meshgrid(*[np.arange(geom.start[k], geom.start[k] + geom.size[k]) for k in range(d)], indexing="ij")
    idx = np.stack([g.transpose(tuple(reversed(range(d)))).ravel() for g in grids], axis=1)
    pts = np.zeros((idx.shape[0], d))
    for i in range(d):
        s = np.full(idx.shape[0], geom.origin[i])
        for j in range(d):
            s = s + geom.index_to_physical[i, j] * idx[:, j].astype(np.float64)
        pts[:, i] = s
    return pts


def _same(a, b):
    assert a["n_valid"] == b["n_valid"]
    assert a["value"] == b["value"]
    assert np.array_equal(a["derivative"], b["derivative"])
    assert np.array_equal(a["joint_pdf"], b["joint_pdf"])


def test_round_half_integer_up_known_answers(oracle_mod):
    # ties go up; the x86-64 build rounds 2x + 0.5 to even and halves, so the
    # odd integer 2^52 + 1 stays itself (x + 0.5 would round to 2^52 + 2)
    cases = {0.5: 1, -0.5: 0, 1.5: 2, -1.5: -1, 2.5: 3, -2.5: -2, 0.0: 0, -0.0: 0, 3.2: 3, -3.7: -4,
             7.0: 7, -7.0: -7, 1e9 + 0.5: 1000000001, 4503599627370497.0: 4503599627370497}
    for x, want in cases.items():
        assert oracle_mod.itk_round(x) == want, x


def test_physical_point_to_index_matches_rounding(oracle_mod):
    rng = np.random.default_rng(3)
    geom = ImageGeometry(size=(7, 9, 5), origin=(1.5, -2.0, 0.25), spacing=(0.8, 1.3, 2.1), start=(2, -1, 0))
    pts = rng.uniform(-6.0, 16.0, (400, 3))
    ci = (pts - np.asarray(geom.origin)) / np.asarray(geom.spacing)
    keep = np.all(np.abs(ci - np.round(ci) - 0.5) > 1e-6, axis=1) & np.all(np.abs(ci - np.round(ci) + 0.5) > 1e-6,
                                                                            axis=1)
    want_idx = np.floor(ci + 0.5).astype(np.int64)
    lo, hi = np.asarray(geom.start), np.asarray(geom.start) + np.asarray(geom.size)
    want_in = np.all((want_idx >= lo) & (want_idx < hi), axis=1)
    host_idx, host_in = geom.physical_points_to_index(pts)
    for k in np.nonzero(keep)[0]:
        ok, idx = oracle_mod.physical_point_to_index(geom, pts[k])
        assert ok == want_in[k]
        assert idx == tuple(want_idx[k])
        # the host-side restatement agrees with the oracle
        assert host_in[k] == ok and tuple(host_idx[k]) == idx
    assert want_in.sum() > 20 and (~want_in).sum() > 20


def test_point_set_of_every_voxel_is_dense(oracle_mod):
    fixed, moving, geom = blob_problem((14, 12, 10), 5)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(6), scale=2.0))
    dense = oracle_mod.mattes(fixed, geom, moving, geom, tf, want_samples=True)
    pts = _voxel_points(geom)
    sparse = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=pts, want_samples=True)
    _same(dense, sparse)
    assert np.array_equal(dense["fixed_bin"], sparse["fixed_bin"])
    assert np.array_equal(dense["moving_bin"], sparse["moving_bin"])
    # threads split the point list; sums change only by rounding
    four = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=pts, threads=4)
    assert four["n_valid"] == dense["n_valid"]
    assert four["value"] == pytest.approx(dense["value"], rel=1e-13)
    assert np.allclose(four["derivative"], dense["derivative"], rtol=1e-11, atol=1e-14)


def test_point_set_2d_bspline(oracle_mod):
    fixed, moving, geom = blob_problem((30, 26), 8)
    tf = bspline_for(geom, (3, 3), seed=9, sigma=0.6)
    dense = oracle_mod.mattes(fixed, geom, moving, geom, tf)
    sparse = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=_voxel_points(geom))
    _same(dense, sparse)


def test_point_subset_and_range(oracle_mod):
    """A random subset in random order: the fixed range is numpy's over the
    pixels the points round to, and reversing the list changes nothing but
    the sample order."""
    fixed, moving, geom = blob_problem((16, 14, 12), 11)
    rng = np.random.default_rng(12)
    pts = _voxel_points(geom)[rng.choice(geom.num_pixels, 700, replace=False)]
    pts = pts + rng.uniform(-0.3, 0.3, pts.shape) * np.asarray(geom.spacing)
    pts = np.concatenate([pts, [[-50.0, -50.0, -50.0]]])  # one point outside everything
    lo, hi, n = oracle_mod.intensity_range(fixed, geom, points=pts)
    idx, inside = geom.physical_points_to_index(pts)
    vals = np.array([fixed[tuple(reversed(tuple(i)))] for i in idx[inside]])
    assert n == inside.sum() == 700
    assert (lo, hi) == (float(vals.min()), float(vals.max()))
    tf = affine_for(geom, affine_params(3, rng))
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=pts, want_samples=True)
    b = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=pts[::-1].copy(), want_samples=True)
    assert a["n_valid"] == b["n_valid"] and 0 < a["n_valid"] <= 700
    assert np.allclose(a["joint_pdf"], b["joint_pdf"], rtol=1e-12, atol=1e-15)
    assert np.array_equal(a["fixed_bin"], b["fixed_bin"][::-1])
    assert np.array_equal(a["moving_bin"], b["moving_bin"][::-1])
    assert b["value"] == pytest.approx(a["value"], rel=1e-12)
    assert a["fixed_bin"][-1] == -1


def _box_mask(geom, box):
    m = np.zeros(geom.shape, dtype=np.uint8)
    (x0, x1), (y0, y1), (z0, z1) = box
    m[z0:z1, y0:y1, x0:x1] = 1
    return m


def test_fixed_box_mask_is_virtual_subregion(oracle_mod):
    fixed, moving, geom = blob_problem((18, 16, 14), 21)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(22), scale=2.0))
    box = ((3, 15), (2, 11), (4, 13))
    mask = (_box_mask(geom, box), geom)
    ranges = (oracle_mod.image_range(fixed), oracle_mod.image_range(moving))
    masked = oracle_mod.mattes(fixed, geom, moving, geom, tf, fixed_mask=mask, ranges=ranges)
    sub = ImageGeometry(size=tuple(b - a for a, b in box), origin=geom.origin, spacing=geom.spacing,
                        direction=geom.direction, start=tuple(a for a, _ in box),
                        physical_to_index=geom.physical_to_index)
    dense = oracle_mod.mattes(fixed, geom, moving, geom, tf, virtual_geom=sub, ranges=ranges)
    _same(masked, dense)
    # the masked range is numpy's over the box
    lo, hi, n = oracle_mod.intensity_range(fixed, geom, mask=mask)
    sel = fixed[4:13, 2:11, 3:15]
    assert n == sel.size and (lo, hi) == (float(sel.min()), float(sel.max()))


def test_moving_mask_is_point_selection(oracle_mod):
    """The moving mask is tested at the mapped moving point: the dense masked
    evaluation equals the point set of the voxels whose moving point rounds
    into a set mask pixel (mask on a coarser grid of its own)."""
    fixed, moving, geom = blob_problem((16, 15, 13), 31)
    rng = np.random.default_rng(32)
    tf = affine_for(geom, affine_params(3, rng, scale=2.0))
    mgeom = ImageGeometry(size=(9, 8, 8), origin=(geom.origin[0] + 0.3, geom.origin[1] - 0.2, geom.origin[2]),
                          spacing=tuple(2.0 * s for s in geom.spacing))
    mpx = (rng.uniform(size=mgeom.shape) < 0.6).astype(np.uint8) * rng.integers(1, 255, mgeom.shape).astype(np.uint8)
    ranges = (oracle_mod.image_range(fixed), oracle_mod.image_range(moving))
    masked = oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_mask=(mpx, mgeom), ranges=ranges)
    pts = _voxel_points(geom)
    mp = np.array([oracle_mod.transform_point(tf, p) for p in pts])
    idx, inside = mgeom.physical_points_to_index(mp)
    keep = inside.copy()
    for k in np.nonzero(inside)[0]:
        keep[k] = mpx[tuple(reversed(tuple(idx[k] - np.asarray(mgeom.start))))] != 0
    assert 0 < keep.sum() < keep.size
    sel = oracle_mod.mattes(fixed, geom, moving, geom, tf, points=pts[keep], ranges=ranges)
    _same(masked, sel)
    # masks that hold everything / nothing
    ones = (np.ones(geom.shape, dtype=np.uint8), geom)
    _same(oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_mask=ones, fixed_mask=ones),
          oracle_mod.mattes(fixed, geom, moving, geom, tf))
    zeros = (np.zeros(geom.shape, dtype=np.uint8), geom)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.mattes(fixed, geom, moving, geom, tf, moving_mask=zeros, ranges=ranges)


def test_empty_selection_range(oracle_mod):
    fixed, _, geom = blob_problem((8, 8, 8), 41)
    lo, hi, n = oracle_mod.intensity_range(fixed, geom, mask=(np.zeros(geom.shape, dtype=np.uint8), geom))
    f32max = float(np.finfo(np.float32).max)
    assert n == 0 and lo == f32max and hi == -f32max


def test_mask_spatial_object_mask_value():
    geom = ImageGeometry(size=(4, 3), spacing=(2.0, 1.0))
    px = np.array([[0, 1, 2, 3], [2, 2, 0, 7], [5, 2, 2, 0]], dtype=np.int16)
    m = ImageMaskSpatialObject(px, geom)
    assert np.array_equal(m.inside_flags(), (px != 0).astype(np.uint8))
    m.SetUseMaskValue(True)
    m.SetMaskValue(2)
    assert np.array_equal(m.inside_flags(), (px == 2).astype(np.uint8))
    assert m.IsInsideInWorldSpace([4.0, 0.0])       # index (2, 0): pixel 2
    assert not m.IsInsideInWorldSpace([2.0, 0.0])   # index (1, 0): pixel 1
    assert m.IsInsideInWorldSpace([0.9, 1.49])      # rounds to (0, 1): pixel 2
    assert m.IsInsideInWorldSpace([-1.0, 1.0])      # x index -0.5 rounds up to 0: (0, 1), pixel 2
    assert not m.IsInsideInWorldSpace([-1.2, 1.0])  # x index -0.6 rounds to -1: outside


def test_masks_thread_invariance_and_value_only(oracle_mod):
    """With masks, the slab partition changes only the derivative's rounding,
    and the value-only evaluation gives the same value and count."""
    fixed, moving, geom = blob_problem((20, 18, 16), 51)
    tf = affine_for(geom, affine_params(3, np.random.default_rng(52), scale=2.0))
    rng = np.random.default_rng(53)
    fm = ((rng.uniform(size=geom.shape) < 0.7).astype(np.uint8), geom)
    mm = ((rng.uniform(size=geom.shape) < 0.8).astype(np.uint8), geom)
    a = oracle_mod.mattes(fixed, geom, moving, geom, tf, fixed_mask=fm, moving_mask=mm, threads=1)
    b = oracle_mod.mattes(fixed, geom, moving, geom, tf, fixed_mask=fm, moving_mask=mm, threads=5)
    c = oracle_mod.mattes(fixed, geom, moving, geom, tf, fixed_mask=fm, moving_mask=mm, derivative=False)
    assert a["n_valid"] == b["n_valid"] == c["n_valid"]
    assert b["value"] == pytest.approx(a["value"], rel=1e-13)
    assert c["value"] == a["value"]
    assert np.allclose(a["derivative"], b["derivative"], rtol=1e-11, atol=1e-14)
