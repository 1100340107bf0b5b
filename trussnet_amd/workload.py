"""Synthetic relaxation workloads: label volume -> grid -> (device) seeds.

The grid builder here prepares the relaxation's input in v2mesh's Grid layout
(src/v2m_grid.h:45-81): per 8^3 brick the labels within R voxels (ascending),
one 512-value slot of the smoothed indicator phi_l per label of a mixed brick,
the sizing field h and its 256 log grades. It follows the structure of
build_grid_cpu (src/v2m_grid.cpp:108-246: brick labels, slots, a separable
normalised Gaussian of the label indicators with the exterior outside the
grid; grades as v2m_grade, src/opencl/v2m_grid_body.cl:498-506) but is not a
bit-for-bit restatement of it: the sizing is a given field (the user sizing
field hvox of src/v2m_grid.h:40-42) and there is no thin-layer blend. The
relaxation's parity is always checked on the grid both sides are given.

Workloads:
  * ``plain_small``  -- 64^3 two-sphere volume, uniform 3-voxel size;
  * ``plain_large``  -- 256^3 six-label nested ellipsoids with perturbed
    boundaries, sizes graded from 1.5 to 6 voxels by distance to the nearest
    interface (about 2 M seeds).
"""
from __future__ import annotations

import numpy as np

from . import volumes
from ._abi import BL, BS, NOLAB, SLOT, Grid


def _dilate_any(ind: np.ndarray, R: int, outside: bool) -> np.ndarray:
    """Per voxel x: any(ind) over [x - R, x + R + 1] on each axis -- so a brick's
    voxels [8b, 8b+7] see [8b - R, 8b + 8 + R], v2m_brick_labels' window
    (src/opencl/v2m_grid_body.cl:32-35)."""
    from scipy.ndimage import maximum_filter1d

    out = ind.astype(np.uint8)
    for ax in range(3):
        out = maximum_filter1d(out, size=2 * R + 2, origin=-1, axis=ax, mode="constant", cval=1 if outside else 0)
    return out.astype(bool)


def _pad_to_bricks(a: np.ndarray, fill) -> np.ndarray:
    nz, ny, nx = a.shape
    pz, py, px = (-nz) % BS, (-ny) % BS, (-nx) % BS
    if pz or py or px:
        a = np.pad(a, ((0, pz), (0, py), (0, px)), constant_values=fill)
    return a


def _bricks_of(a: np.ndarray) -> np.ndarray:
    """(nz, ny, nx) padded to bricks -> (nbz, nby, nbx, 8, 8, 8), slot order (k, j, i)."""
    nz, ny, nx = a.shape
    return a.reshape(nz // BS, BS, ny // BS, BS, nx // BS, BS).transpose(0, 2, 4, 1, 3, 5)


def grade_of(h: np.ndarray, hmin: float, hmax: float) -> np.ndarray:
    """v2m_grade (src/opencl/v2m_grid_body.cl:498-506): 256 log grades in [hmin, hmax]."""
    hmin, hmax = np.float32(hmin), np.float32(hmax)
    if hmax <= hmin * np.float32(1.0001):
        return np.zeros(h.shape, np.uint8)
    t = np.log(h.astype(np.float32) / hmin) / np.log(hmax / hmin)
    t = np.clip(t, 0.0, 1.0).astype(np.float32)
    return (t * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def build_grid(labels: np.ndarray, h, vs=(1.0, 1.0, 1.0), sigma: float = 1.0, R: int = 6,
               hmin: float | None = None, hmax: float | None = None, prob: np.ndarray | None = None) -> Grid:
    """A relaxation grid from a (nz, ny, nx) label volume and a sizing field (array or scalar, mm).

    R: the brick label radius (Grid::R: the reference's default is 6, from its
    wider curvature smoothing); phi itself is smoothed with radius ceil(3 sigma).
    prob: per-label probabilities (C, nz, ny, nx) of a tissue-probability input;
    label l's field is then its smoothed probability instead of its smoothed
    indicator (build_grid_cpu's TPM source, src/v2m_grid.cpp:154-168).
    """
    from scipy.ndimage import correlate1d

    L = np.ascontiguousarray(labels, np.uint16)
    nz, ny, nx = L.shape
    nbx, nby, nbz = (nx + BS - 1) // BS, (ny + BS - 1) // BS, (nz + BS - 1) // BS
    nb = nbx * nby * nbz
    present = [int(v) for v in np.unique(L)]
    if 0 not in present:
        present = [0] + present  # the exterior surrounds the grid
    # 1. labels within R voxels of each brick (outside the grid: the exterior)
    near = np.zeros((len(present), nb), bool)
    for k, l in enumerate(present):
        d = _pad_to_bricks(_dilate_any(L == l, R, outside=(l == 0)), l == 0)
        near[k] = _bricks_of(d).reshape(nb, -1).any(axis=1)
    cnt = near.sum(axis=0).astype(np.int32)
    if (cnt > BL).any():
        raise ValueError("more labels near a brick than a brick holds")
    bl_lab = np.full((nb, BL), NOLAB, np.uint16)
    order = np.argsort(np.array(present))
    pos = np.zeros(nb, np.int32)
    for k in order:  # ascending labels
        sel = near[k]
        bl_lab[sel, pos[sel]] = present[k]
        pos[sel] += 1
    # 2. slots of the mixed bricks
    mixed = cnt > 1
    bl_slot = np.full(nb, -1, np.int32)
    starts = np.cumsum(np.where(mixed, cnt, 0)) - np.where(mixed, cnt, 0)
    bl_slot[mixed] = starts[mixed]
    nslot = int(np.where(mixed, cnt, 0).sum())
    phi = np.zeros((nslot, SLOT), np.float32)
    # 3. the smoothed indicators (separable normalised Gaussian, exterior outside)
    Rs = min(6, max(1, int(np.ceil(3.0 * sigma))))
    taps = np.exp(-0.5 * np.arange(-Rs, Rs + 1, dtype=np.float32) ** 2 / np.float32(sigma * sigma))
    taps = (taps / taps.sum()).astype(np.float32)
    mixed_idx = np.flatnonzero(mixed)
    for k in order:
        l = present[k]
        rows = mixed_idx[near[k, mixed_idx]]
        if rows.size == 0:
            continue
        if prob is None:
            f = (L == l).astype(np.float32)
        else:
            f = np.array(prob[l], np.float32) if l < prob.shape[0] else np.zeros(L.shape, np.float32)
        for ax in range(3):
            f = correlate1d(f, taps, axis=ax, mode="constant", cval=1.0 if l == 0 else 0.0)
        blocks = _bricks_of(_pad_to_bricks(f, np.float32(1.0 if l == 0 else 0.0))).reshape(nb, SLOT)
        # the slot of label l in brick b: bl_slot[b] + its rank among the brick's labels
        rank = (bl_lab[rows] < l).sum(axis=1)
        phi[bl_slot[rows] + rank] = blocks[rows]
    # 4. sizing
    if np.isscalar(h):
        hv = np.full(L.shape, np.float32(h), np.float32)
    else:
        hv = np.ascontiguousarray(h, np.float32).reshape(L.shape)
    hmin = float(hv.min()) if hmin is None else hmin
    hmax = float(hv.max()) if hmax is None else hmax
    return Grid(nx, ny, nz, tuple(float(v) for v in vs), L.ravel(), cnt, bl_lab.ravel(), bl_slot, phi.ravel(),
                hv.ravel(), hmin, hmax, grade_of(hv, hmin, hmax).ravel())


def interface_distance(labels: np.ndarray) -> np.ndarray:
    """Distance (voxels) from each voxel centre to the nearest voxel with another label next to it."""
    from scipy.ndimage import distance_transform_edt

    edge = np.zeros(labels.shape, bool)
    for ax in range(3):
        d = np.diff(labels.astype(np.int32), axis=ax) != 0
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        edge[tuple(lo)] |= d
        edge[tuple(hi)] |= d
    if not edge.any():
        return np.full(labels.shape, np.float32(1e6), np.float32)
    return distance_transform_edt(~edge).astype(np.float32)


def graded_size(labels: np.ndarray, lo: float, hi: float, slope: float) -> np.ndarray:
    """h = lo at the interfaces, growing by `slope` per voxel of distance, capped at hi."""
    return np.minimum(np.float32(hi), np.float32(lo) + np.float32(slope) * interface_distance(labels)).astype(
        np.float32)


# --- named workloads -------------------------------------------------------------------

def plain_small() -> Grid:
    """64^3, label 1 = ball r 24, label 2 = ball r 12 (centred), uniform size 3 voxels."""
    return build_grid(volumes.two_spheres(64, 24.0, 12.0), 3.0)


LARGE_SLOPE = 0.03   # voxel growth of h per voxel of distance: ~2 M seeds on the large volume


def plain_large(seed: int = 0, n: int = 256) -> Grid:
    """256^3, six nested perturbed ellipsoid labels, h graded 1.5 -> 6 voxels off the interfaces."""
    lab = volumes.nested_ellipsoids(n, nlab=6, seed=seed, axes=(20.0, 120.0), grf_sigma=4.0, grf_amp=3.0)
    return build_grid(lab, graded_size(lab, 1.5, 6.0, LARGE_SLOPE), hmin=1.5, hmax=6.0)


def prob_large(seed: int = 1, shape=(182, 218, 182), h: float = 1.8) -> Grid:
    """The probability-input workload: a 182x218x182x6 softmax of six Gaussian random
    fields (sigma 6 voxels, channel 0 the exterior), labels by argmax, each label's
    interface field its smoothed probability; uniform size h voxels."""
    prob, lab = volumes.tpm_softmax(shape, 6, 6.0, seed)
    return build_grid(lab, np.float32(h), hmin=h, hmax=h, prob=prob)


def gray_large(n: int = 256, seed: int = 2, device: int = 0) -> Grid:
    """The gray-scale workload: an n^3 intensity 1 - r / (0.45 n) plus 0.06 x a
    Gaussian random field (sigma 3 voxels), cut by the thresholds 0.2 / 0.5 / 0.75
    into three nested perturbed shells; the grid built on the device with the
    reference's defaults and hbase 4.5, hmin 1.5, hmax 6 voxels."""
    from . import gridbuild
    from ._abi import GridParams

    gray = volumes.gray_ball(n, seed=seed)
    lab = volumes.gray_labels(gray, volumes.GRAY_THRESHOLDS)
    return gridbuild.build(lab, prm=GridParams(hbase=4.5, hmin=1.5, hmax=6.0), gray=gray,
                           thresholds=volumes.GRAY_THRESHOLDS, device=device)


def adjoint_volume(n: int = 128, seed: int = 5, h: float = 1.5) -> Grid:
    """The differentiation workload: 128^3, three nested perturbed ellipsoid labels
    (semi-axes U(20, 60) voxels), uniform size h voxels (about 200 k seeds)."""
    lab = volumes.nested_ellipsoids(n, nlab=3, seed=seed, axes=(20.0, 60.0), grf_sigma=3.0, grf_amp=2.0,
                                    centre_jitter=3.0)
    return build_grid(lab, np.float32(h), hmin=h, hmax=h)


def batch_volumes(count: int = 64, seed: int = 0) -> list:
    """Many independent volumes: edge U(32, 160) voxels, 2-4 nested perturbed ellipsoid
    labels (semi-axes U(0.25, 0.5) x edge), a uniform node size U(2, 5) voxels per
    volume."""
    rng = np.random.default_rng(seed)
    grids = []
    for k in range(count):
        n = int(rng.integers(32, 161))
        nlab = int(rng.integers(2, 5))
        h = float(rng.uniform(2.0, 5.0))
        lab = volumes.nested_ellipsoids(n, nlab=nlab, seed=int(rng.integers(1 << 30)),
                                        axes=(0.25 * n, 0.5 * n), grf_sigma=3.0, grf_amp=2.0,
                                        centre_jitter=0.03 * n)
        grids.append(build_grid(lab, np.float32(h), hmin=h, hmax=h))
    return grids


def estimate_seeds(grid: Grid, nseed: int = 8) -> float:
    """The expected lattice seed count: HCP density sqrt(2)/s^3 over the non-exterior voxels."""
    lev = np.floor(grid.grade.astype(np.float32) * np.float32(nseed - 1) / np.float32(255.0) + np.float32(0.5))
    s = grid.hmin * (grid.hmax / grid.hmin) ** (lev / max(nseed - 1, 1))
    inside = grid.L != 0
    vol = float(np.prod(grid.vs))
    return float((np.sqrt(2.0) / s[inside] ** 3).sum() * vol)
