"""Synthetic label volumes for the relaxation workloads (numpy, x fastest in memory).

Arrays are (nz, ny, nx) uint16 so that ``.ravel()`` is the x-fastest layout of
v2mesh's LabelVolume (src/v2m_volume.h:22-27). Label 0 is the exterior.
"""
from __future__ import annotations

import numpy as np


def _centres(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.float64)


def two_spheres(n: int = 64, r1: float = 24.0, r2: float = 12.0) -> np.ndarray:
    """Label 1 = ball of radius r1 about the volume centre, label 2 = concentric ball r2.

    The 64^3 case with r 24/12 and a uniform 3-voxel size is the small plain workload.
    """
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(_centres(n), _centres(n), _centres(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    lab = np.zeros((n, n, n), np.uint16)
    lab[r < r1] = 1
    lab[r < r2] = 2
    return lab


def sectors(n: int = 48, r1: float = 18.0, r2: float = 10.0) -> np.ndarray:
    """A ball (label 1) holding a smaller ball cut into three sectors (labels 2, 3, 4).

    Triple lines run where the sectors meet each other and label 1; points where
    four labels meet (0 excluded) sit where the sector axis leaves the inner ball,
    so the seeds hold interface, junction and corner nodes.
    """
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(_centres(n), _centres(n), _centres(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    lab = np.zeros((n, n, n), np.uint16)
    lab[r < r1] = 1
    inner = r < r2
    xs, ys = x - c - 0.3, y - c + 0.2
    lab[inner & (xs < 0)] = 2
    lab[inner & (xs >= 0) & (ys < 0)] = 3
    lab[inner & (xs >= 0) & (ys >= 0)] = 4
    return lab


def _rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def gaussian_field(shape, sigma: float, rng: np.random.Generator) -> np.ndarray:
    """White noise smoothed by a Gaussian of `sigma` voxels, scaled to unit std (float32)."""
    from scipy.ndimage import gaussian_filter

    f = gaussian_filter(rng.standard_normal(shape).astype(np.float32), sigma, mode="wrap")
    f /= max(float(f.std()), 1e-12)
    return f


def nested_ellipsoids(n: int = 256, nlab: int = 6, seed: int = 0, axes=(20.0, 120.0), grf_sigma: float = 4.0,
                      grf_amp: float = 3.0, centre_jitter: float = 6.0) -> np.ndarray:
    """Nested, randomly oriented ellipsoids (labels 1..nlab), boundaries perturbed.

    Semi-axes are drawn uniformly in `axes` voxels and sorted per axis, so
    ellipsoid k's axes are each at least ellipsoid k+1's; ellipsoid k is painted
    over k-1, largest first, so each later one nests inside the earlier ones as
    far as its orientation allows. Every boundary is displaced by a smoothed
    Gaussian random field (std 1 after smoothing with `grf_sigma`, times
    `grf_amp` voxels).
    """
    rng = np.random.default_rng(seed)
    ax = -np.sort(-rng.uniform(axes[0], axes[1], size=(nlab, 3)), axis=0)
    rots = [_rotation(rng) for _ in range(nlab)]
    cen = (n - 1) / 2.0 + rng.uniform(-centre_jitter, centre_jitter, size=(nlab, 3))
    grf = gaussian_field((n, n, n), grf_sigma, rng) * grf_amp
    lab = np.zeros((n, n, n), np.uint16)
    zz = np.arange(n, dtype=np.float32)
    for k in range(nlab):
        R = rots[k].astype(np.float32)
        a = ax[k].astype(np.float32)
        for z0 in range(0, n, 32):
            z1 = min(n, z0 + 32)
            z, y, x = np.meshgrid(zz[z0:z1], zz, zz, indexing="ij")
            d = np.stack([x - cen[k, 0], y - cen[k, 1], z - cen[k, 2]], axis=-1).astype(np.float32)
            u = d @ R  # body coordinates
            q = np.sqrt((u[..., 0] / a[0]) ** 2 + (u[..., 1] / a[1]) ** 2 + (u[..., 2] / a[2]) ** 2)
            # signed distance (approximately, in voxels) to the ellipsoid surface, perturbed
            s = (q - 1.0) * np.float32(a.mean()) + grf[z0:z1]
            sl = lab[z0:z1]
            sl[s < 0] = k + 1
    return lab


_M64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) of each 64-bit state: the
    generator's output for counter x (state = x * golden gamma)."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def jittered_points(n: int = 64, spacing: float = 3.0, jitter: float = 0.25, seed: int = 0,
                    vs=(1.0, 1.0, 1.0)) -> np.ndarray:
    """A cubic grid of points at `spacing` (mm) over the volume's voxel centres
    [0, (n-1) v], each coordinate moved by U(-jitter, jitter) * spacing from a
    counter-based generator: point i (x fastest), axis a draws counter
    seed * 2^32 + 3 i + a of SplitMix64, whose top 24 bits give u in [0, 1).
    Returns float32 (m, 3) positions in mm (the small plain workload's seeds before
    labelling and classification)."""
    ext = (n - 1) * np.asarray(vs, np.float64)
    axes = [np.arange(0.0, e + 1e-9, spacing) for e in ext]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    base = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    m = base.shape[0]
    ctr = (np.uint64(seed) << np.uint64(32)) + np.arange(3 * m, dtype=np.uint64)
    u = (splitmix64(ctr) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24
    off = (2.0 * u - 1.0).reshape(m, 3) * jitter * spacing
    return (base + off).astype(np.float32)


def tpm_softmax(shape=(182, 218, 182), channels: int = 6, sigma: float = 6.0, seed: int = 1,
                gain: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """A tissue-probability map: the softmax over `channels` Gaussian random fields
    (white noise smoothed by `sigma` voxels, unit std, times `gain`), channel 0 the
    exterior. `shape` is (nx, ny, nz). Returns (prob float32 (C, nz, ny, nx),
    labels uint16 (nz, ny, nx) = the argmax channel)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    f = np.stack([gaussian_field((nz, ny, nx), sigma, rng) * np.float32(gain) for _ in range(channels)])
    f -= f.max(axis=0, keepdims=True)
    np.exp(f, out=f)
    f /= f.sum(axis=0, keepdims=True)
    return f.astype(np.float32), np.argmax(f, axis=0).astype(np.uint16)


GRAY_THRESHOLDS = (0.2, 0.5, 0.75)


def gray_ball(n: int = 40, radius: float | None = None, seed: int = 2, amp: float = 0.06,
              sigma: float = 3.0) -> np.ndarray:
    """A gray-scale intensity (float32, (n, n, n)): 1 - r / radius about the volume
    centre plus `amp` times a Gaussian random field (sigma voxels), so that the
    thresholds GRAY_THRESHOLDS cut three nested, perturbed shells (labels 1..3, the
    number of thresholds <= the intensity)."""
    radius = 0.45 * n if radius is None else radius
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(_centres(n), _centres(n), _centres(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    g = gaussian_field((n, n, n), sigma, np.random.default_rng(seed))
    return (1.0 - r / radius + amp * g).astype(np.float32)


def gray_labels(gray: np.ndarray, thresholds) -> np.ndarray:
    """The label of each voxel: the number of thresholds <= its intensity
    (apply_thresholds without smoothing, src/v2m_volume.cpp:140-149)."""
    t = np.sort(np.asarray(thresholds, np.float32))
    return (np.asarray(gray, np.float32)[..., None] >= t).sum(-1).astype(np.uint16)
