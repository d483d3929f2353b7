"""Deterministic synthetic registration problems.

Fixed image: a sum of Gaussian blobs (centres uniform over the image,
sigma U[4,20] voxels, amplitude U[0.2,1]) plus N(0, 0.02) noise.  Moving
image: the fixed image resampled with linear interpolation through a known
transform (an affine about the image centre, or a cubic B-spline
displacement), remapped by m = 1 - max(x,0)^1.5, plus N(0, 0.02) noise.

Random numbers come from a counter-based splitmix64 stream and an explicit
Box-Muller transform, so every platform produces the same volumes from the
same seed.  Spacing 1, origin 0, identity direction.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import ndimage

from .geometry import ImageGeometry

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 started at state seed ^ (stream * golden)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream & 0xFFFFFFFF) * _GOLDEN)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """N(0,1) by Box-Muller on two uniform streams."""
    u1 = 1.0 - uniform(seed, stream, n)  # (0, 1]
    u2 = uniform(seed, stream + 1, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def blob_image(shape_xyz, n_blobs: int, seed: int, noise: float = 0.02) -> np.ndarray:
    """Sum of Gaussian blobs plus noise; returns float32 in numpy order (z,y,x)."""
    d = len(shape_xyz)
    u = uniform(seed, 1, n_blobs * (d + 2)).reshape(n_blobs, d + 2)
    img = np.zeros(tuple(reversed(shape_xyz)), dtype=np.float64)
    axes = [np.arange(s, dtype=np.float64) for s in shape_xyz]
    for b in range(n_blobs):
        c = [u[b, i] * shape_xyz[i] for i in range(d)]
        sigma = 4.0 + 16.0 * u[b, d]
        amp = 0.2 + 0.8 * u[b, d + 1]
        lo = [max(0, int(c[i] - 4 * sigma)) for i in range(d)]
        hi = [min(shape_xyz[i], int(c[i] + 4 * sigma) + 1) for i in range(d)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        g = [np.exp(-((axes[i][lo[i]:hi[i]] - c[i]) ** 2) / (2.0 * sigma * sigma)) for i in range(d)]
        if d == 2:
            img[lo[1]:hi[1], lo[0]:hi[0]] += amp * np.outer(g[1], g[0])
        else:
            img[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += amp * (g[2][:, None, None] * g[1][None, :, None]
                                                                  * g[0][None, None, :])
    if noise > 0:
        img += noise * normal(seed, 3, img.size).reshape(img.shape)
    return img.astype(np.float32)


def rotation_matrix(angles_deg) -> np.ndarray:
    a = [math.radians(x) for x in angles_deg]
    if len(a) == 1:
        c, s = math.cos(a[0]), math.sin(a[0])
        return np.array([[c, -s], [s, c]])
    cx, sx = math.cos(a[0]), math.sin(a[0])
    cy, sy = math.cos(a[1]), math.sin(a[1])
    cz, sz = math.cos(a[2]), math.sin(a[2])
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


@dataclass
class AffineTruth:
    matrix: np.ndarray       # G: moving(y) = fixed(matrix @ (y - c) + c + translation)
    translation: np.ndarray
    center: np.ndarray


def random_affine(dim: int, seed: int, shape_xyz) -> AffineTruth:
    u = uniform(seed, 7, 8)
    if dim == 2:
        rot = rotation_matrix([-10 + 20 * u[0]])
    else:
        rot = rotation_matrix([-10 + 20 * u[i] for i in range(3)])
    scale = 0.95 + 0.1 * u[3]
    t = np.array([-8 + 16 * u[4 + i] for i in range(dim)])
    c = np.array([(s - 1) / 2.0 for s in shape_xyz])
    return AffineTruth(rot * scale, t, c)


def remap(img: np.ndarray, seed: int, noise: float = 0.02) -> np.ndarray:
    x = np.maximum(img.astype(np.float64), 0.0)
    m = 1.0 - x ** 1.5
    if noise > 0:
        m += noise * normal(seed, 11, m.size).reshape(m.shape)
    return m.astype(np.float32)


def resample_affine(img: np.ndarray, truth: AffineTruth) -> np.ndarray:
    """out(y) = img(G (y - c) + c + t), linear interpolation, edge clamp."""
    d = img.ndim
    # numpy axes are reversed (z,y,x); convert the xyz affine.
    perm = list(range(d))[::-1]
    m = truth.matrix[np.ix_(perm, perm)]
    c = truth.center[perm]
    t = truth.translation[perm]
    offset = c + t - m @ c
    return ndimage.affine_transform(img.astype(np.float64), m, offset=offset, order=1, mode="nearest").astype(
        np.float32)


def bspline_basis(n: int, mesh: int) -> np.ndarray:
    """(n, mesh+3) cubic B-spline weights of voxel i on a grid spanning [0, n-1]."""
    spacing = (n - 1) / float(mesh)
    w = np.zeros((n, mesh + 3))
    for i in range(n):
        ci = i / spacing + 1.0  # grid origin is one spacing before the domain
        start = int(math.floor(ci - 1.0))
        x = ci - start
        for k in range(4):
            u = abs(x - k)
            if u < 1:
                v = (4 - 6 * u * u + 3 * u ** 3) / 6
            elif u < 2:
                v = (2 - u) ** 3 / 6
            else:
                v = 0.0
            if 0 <= start + k < mesh + 3:
                w[i, start + k] = v
    return w


def resample_bspline(img: np.ndarray, mesh: int, sigma_vox: float, seed: int) -> np.ndarray:
    """Warp by a random cubic B-spline displacement (coefficients N(0, sigma) voxels)."""
    shape = img.shape  # (z, y, x)
    d = img.ndim
    g = mesh + 3
    coeffs = sigma_vox * normal(seed, 13, d * g ** d).reshape((d,) + (g,) * d)  # [comp, z, y, x]
    basis = [bspline_basis(shape[a], mesh) for a in range(d)]
    disp = []
    for comp in range(d):
        c = coeffs[comp]
        if d == 3:
            u = np.einsum("zyx,Zz->Zyx", c, basis[0])
            u = np.einsum("Zyx,Yy->ZYx", u, basis[1])
            u = np.einsum("ZYx,Xx->ZYX", u, basis[2])
        else:
            u = np.einsum("yx,Yy->Yx", c, basis[0])
            u = np.einsum("Yx,Xx->YX", u, basis[1])
        disp.append(u)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    coords = [grids[a] + disp[d - 1 - a] for a in range(d)]  # comp 0 is x = numpy axis d-1
    return ndimage.map_coordinates(img.astype(np.float64), coords, order=1, mode="nearest").astype(np.float32)


@dataclass
class Problem:
    fixed: np.ndarray
    moving: np.ndarray
    geometry: ImageGeometry
    truth: object


def make_affine_problem(shape_xyz, seed: int, n_blobs: int = None) -> Problem:
    d = len(shape_xyz)
    n_blobs = n_blobs or (30 if d == 2 else 80)
    fixed = blob_image(shape_xyz, n_blobs, seed)
    truth = random_affine(d, seed, shape_xyz)
    moving = remap(resample_affine(fixed, truth), seed)
    return Problem(fixed, moving, ImageGeometry(size=shape_xyz), truth)


def make_bspline_problem(shape_xyz, seed: int, mesh: int = 8, sigma_vox: float = 2.0, n_blobs: int = 80) -> Problem:
    fixed = blob_image(shape_xyz, n_blobs, seed)
    moving = remap(resample_bspline(fixed, mesh, sigma_vox, seed), seed)
    return Problem(fixed, moving, ImageGeometry(size=shape_xyz), ("bspline", mesh, sigma_vox))
