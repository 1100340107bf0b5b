"""Small problems of the differentiable relaxation for the adjoint tests."""
import numpy as np

from trussnet_amd._abi import INTERFACE as TN_INTERFACE, INTERIOR as TN_INTERIOR, K


def small_problem(n_side: int = 7, h: float = 1.5, seed: int = 0):
    """A jittered cubic block of nodes (the outer shell fixed, typed as interface),
    the K nearest within 1.6 h as lists, a Delaunay tetrahedralisation, and a
    positive sizing grid around h. Returns (x0, nbr, nnb, typ, tets, theta,
    origin, spacing)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    g = np.arange(n_side, dtype=np.float64) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    x0 = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    x0 += rng.uniform(-0.2, 0.2, size=x0.shape) * h
    n = x0.shape[0]
    edge = (x.ravel() == 0) | (x.ravel() == g[-1]) | (y.ravel() == 0) | (y.ravel() == g[-1]) | \
           (z.ravel() == 0) | (z.ravel() == g[-1])
    typ = np.where(edge, TN_INTERFACE, TN_INTERIOR).astype(np.uint8)
    d = np.linalg.norm(x0[:, None, :] - x0[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    nbr = np.full((n, K), -1, np.int32)
    nnb = np.zeros(n, np.int32)
    for i in range(n):
        order = np.argsort(d[i], kind="stable")
        sel = [j for j in order[:K] if d[i, j] < 1.6 * h]
        nbr[i, :len(sel)] = sel
        nnb[i] = len(sel)
    tets = Delaunay(x0).simplices.astype(np.int32)
    gs = (5, 6, 7)   # gz, gy, gx
    origin = (-0.5 * h, -0.4 * h, -0.6 * h)
    ext = g[-1] + h
    spacing = (ext / (gs[2] - 1), ext / (gs[1] - 1), ext / (gs[0] - 1))
    theta = (h * (1.0 + 0.2 * rng.uniform(-1.0, 1.0, size=gs))).astype(np.float32)
    return x0.astype(np.float32), nbr, nnb, typ, tets, theta, origin, spacing
