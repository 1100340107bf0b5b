"""The FP64 adjoint of the differentiable relaxation (oracle/adjoint.c) pinned by
central finite differences of its own double-precision forward, on a small block
of nodes: directional derivatives along random directions in the positions and
in the sizing grid, and single components. The step is 1e-6: the smoothed quality
of near-flat tets is strongly curved, and the central difference's O(eps^2)
error is 1e-4 relative at eps = 1e-4 but 1e-8 at 1e-6."""
import numpy as np
import pytest

from adjoint_cases import small_problem
from trussnet_amd.adjoint import AdjParams, problem_struct


@pytest.fixture(scope="module")
def case(oracle_lib):
    x0, nbr, nnb, typ, tets, theta, origin, spacing = small_problem()
    prm = AdjParams(iters=12, ckpt_every=4)
    pc, keep = problem_struct(nbr, nnb, typ, tets, theta.shape, origin, spacing, prm)
    return oracle_lib, pc, keep, x0.astype(np.float64), theta.astype(np.float64)


def fd(orc, pc, x0, th, dx, dt, eps):
    lp = orc.adj_loss(pc, x0 + eps * dx, th + eps * dt)[0]
    lm = orc.adj_loss(pc, x0 - eps * dx, th - eps * dt)[0]
    return (lp - lm) / (2 * eps)


def test_loss_terms_are_sane(case):
    orc, pc, keep, x0, th = case
    L, xT, terms = orc.adj_loss(pc, x0, th)
    assert np.isfinite(L) and L == pytest.approx(terms[1] + terms[2])   # wq = 1
    assert 0.0 < terms[2] < 1.0   # mean (1 - q) of a jittered block
    moved = np.linalg.norm(xT - x0, axis=1)
    fixed = keep[2] != 0
    assert np.all(moved[fixed] == 0) and moved[~fixed].max() > 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gradient_matches_central_differences(case, seed):
    orc, pc, keep, x0, th = case
    L, gx, gt = orc.adj_grad(pc, x0, th)
    assert L == pytest.approx(orc.adj_loss(pc, x0, th)[0], rel=1e-14)
    rng = np.random.default_rng(seed)
    dx = rng.standard_normal(x0.shape)
    dt = rng.standard_normal(th.shape)
    for ux, ut in ((dx, 0 * dt), (0 * dx, dt), (dx, dt)):
        want = fd(orc, pc, x0, th, ux, ut, 1e-6)
        got = float((gx * ux).sum() + (gt * ut).sum())
        assert got == pytest.approx(want, rel=1e-6, abs=1e-10)


def test_single_components(case):
    orc, pc, keep, x0, th = case
    _, gx, gt = orc.adj_grad(pc, x0, th)
    interior = np.flatnonzero(keep[2] == 0)
    for i in interior[:3]:
        for a in range(3):
            e = np.zeros_like(x0)
            e[i, a] = 1.0
            assert gx[i, a] == pytest.approx(fd(orc, pc, x0, th, e, 0 * th, 1e-6), rel=1e-5, abs=1e-10)
    big = np.argsort(-np.abs(gt.ravel()))[:3]
    for c in big:
        e = np.zeros(th.size)
        e[c] = 1.0
        assert gt.ravel()[c] == pytest.approx(fd(orc, pc, x0, th, 0 * x0, e.reshape(th.shape), 1e-6), rel=1e-5)
