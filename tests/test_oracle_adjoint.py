"""The FP64 adjoint of the differentiable relaxation (oracle/adjoint.c) pinned by
central finite differences of its own double-precision forward, on a small block
of nodes: directional derivatives along random directions in the positions and
in the sizing grid (step 1e-6), and single components (step 1e-4: the smallest
components, ~3e-5, lose 1e-5 of their digits to cancellation at 1e-6, while the
central difference's O(eps^2) error is ~1e-8 relative at 1e-4). The quality term
is pinned by closed forms of the Joe-Liu quality."""
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
            assert gx[i, a] == pytest.approx(fd(orc, pc, x0, th, e, 0 * th, 1e-4), rel=1e-5, abs=1e-10)
    big = np.argsort(-np.abs(gt.ravel()))[:3]
    for c in big:
        e = np.zeros(th.size)
        e[c] = 1.0
        assert gt.ravel()[c] == pytest.approx(fd(orc, pc, x0, th, 0 * x0, e.reshape(th.shape), 1e-4), rel=1e-5)


def quality_loss(orc, pts, tets, eps_v=1e-9):
    """The loss of nodes pts with no bars and no steps: wq * mean (1 - q) over tets."""
    n = pts.shape[0]
    nbr = np.full((n, 24), -1, np.int32)
    nnb = np.zeros(n, np.int32)
    typ = np.ones(n, np.uint8)   # fixed
    theta = np.ones((2, 2, 2), np.float32)
    prm = AdjParams(iters=0, ckpt_every=1, wq=1.0, eps_v=eps_v)
    pc, keep = problem_struct(nbr, nnb, typ, np.asarray(tets, np.int32), theta.shape, (0.0, 0.0, 0.0),
                              (100.0, 100.0, 100.0), prm)
    L, _, terms = orc.adj_loss(pc, np.asarray(pts, np.float64), theta.astype(np.float64))
    return L, terms


def test_quality_term_closed_forms(oracle_lib):
    """The quality term is Joe-Liu's 12 (3V)^(2/3) / sum l^2 (src/v2m_tetra.cpp:100-120)
    as eps_v -> 0: 1 for a regular tet at any scale and orientation, 12 (1/2)^(2/3) / 9
    for the unit corner tet, 0 for a flat one, and each tet counted once in the mean."""
    orc = oracle_lib
    reg = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64)
    L, terms = quality_loss(orc, reg, [[0, 1, 2, 3]])
    assert terms[1] == 0.0 and abs(terms[2]) < 1e-12   # q = 1
    rng = np.random.default_rng(4)
    q_, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    L2, t2 = quality_loss(orc, 3.7 * reg @ q_.T + np.array([5.0, -2.0, 9.0]), [[0, 1, 2, 3]])
    assert abs(t2[2]) < 1e-12
    corner = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    _, t3 = quality_loss(orc, corner, [[0, 1, 2, 3]])
    assert t3[2] == pytest.approx(1.0 - 12.0 * 0.5 ** (2.0 / 3.0) / 9.0, rel=1e-12)
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float64)
    _, t4 = quality_loss(orc, flat, [[0, 1, 2, 3]])
    assert t4[2] == pytest.approx(1.0, abs=1e-5)   # q -> 0 (eps_v = 1e-9)
    both = np.concatenate([reg, corner + 10.0])
    _, t5 = quality_loss(orc, both, [[0, 1, 2, 3], [4, 5, 6, 7]])
    assert t5[2] == pytest.approx(0.5 * t3[2], rel=1e-12)


def test_quality_matches_the_reference_tet_quality(oracle_lib):
    """The quality term against the reference's own tet_quality (src/v2m_tetra.cpp:
    100-150), compiled from its sources: 1 - q per tet equals 1 - its Joe-Liu quality
    for random tets of every shape, as eps_v -> 0."""
    from oracle import ref

    if not ref.available():
        pytest.skip("the reference's sources are not built here (oracle/_ref)")
    rng = np.random.default_rng(11)
    for k in range(40):
        pts = rng.standard_normal((4, 3)) * rng.uniform(0.2, 5.0, size=3) + rng.uniform(-10, 10, size=3)
        _, terms = quality_loss(oracle_lib, pts, [[0, 1, 2, 3]], eps_v=1e-12)
        _, jl, vol = ref.tet_quality(pts)
        assert vol > 1e-6
        assert terms[2] == pytest.approx(1.0 - jl, rel=1e-9, abs=1e-12), (k, terms[2], 1.0 - jl)
