"""The device adjoint of the differentiable relaxation (tn_adj_*) against the FP64
CPU gradient (oracle/adjoint.c, itself pinned by finite differences), through the
C ABI and through the torch.autograd.Function."""
import numpy as np
import pytest

from adjoint_cases import small_problem
from trussnet_amd.adjoint import AdjParams, problem_struct

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible to torch")
    return torch


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


@pytest.mark.parametrize("iters,ckpt", [(0, 1), (1, 1), (12, 1), (12, 4), (13, 5)])
def test_device_gradient_equals_fp64(torch_dev, oracle_lib, iters, ckpt):
    from trussnet_amd.adjoint import Problem

    torch = torch_dev
    x0, nbr, nnb, typ, tets, theta, origin, spacing = small_problem()
    prm = AdjParams(iters=iters, ckpt_every=ckpt)
    print(f"[adjoint] small block: {x0.shape[0]} nodes, {tets.shape[0]} tets, {iters} steps, ckpt {ckpt}", flush=True)
    pb = Problem(nbr, nnb, typ, tets, theta.shape, origin, spacing, prm, device=0)
    xd = torch.from_numpy(x0).cuda()
    td = torch.from_numpy(theta).cuda()
    xT = torch.empty_like(xd)
    loss = pb.forward_ptr(xd.data_ptr(), td.data_ptr(), xT.data_ptr())
    gx = torch.empty_like(xd)
    gt = torch.empty_like(td)
    pb.backward_ptr(td.data_ptr(), gx.data_ptr(), gt.data_ptr())
    torch.cuda.synchronize()
    pc, keep = problem_struct(nbr, nnb, typ, tets, theta.shape, origin, spacing, prm)
    L, xT64, terms = oracle_lib.adj_loss(pc, x0, theta)
    L2, gx64, gt64 = oracle_lib.adj_grad(pc, x0, theta)
    assert loss == pytest.approx(L, rel=1e-5)
    assert rel(xT.cpu().numpy(), xT64) < 1e-5
    assert rel(gx.cpu().numpy(), gx64) < 1e-4, rel(gx.cpu().numpy(), gx64)
    assert rel(gt.cpu().numpy(), gt64) < 1e-4, rel(gt.cpu().numpy(), gt64)
    pb.close()


def test_autograd_function(torch_dev):
    from trussnet_amd.adjoint import Problem, RelaxLoss

    torch = torch_dev
    x0, nbr, nnb, typ, tets, theta, origin, spacing = small_problem(seed=4)
    prm = AdjParams(iters=10, ckpt_every=3)
    pb = Problem(nbr, nnb, typ, tets, theta.shape, origin, spacing, prm, device=0)
    xd = torch.from_numpy(x0).cuda().requires_grad_(True)
    td = torch.from_numpy(theta).cuda().requires_grad_(True)
    loss = RelaxLoss.apply(xd, td, pb)
    (2.0 * loss).backward()
    gx = torch.empty_like(xd)
    gt = torch.empty_like(td)
    L = pb.forward_ptr(xd.detach().data_ptr(), td.detach().data_ptr())
    pb.backward_ptr(td.detach().data_ptr(), gx.data_ptr(), gt.data_ptr())
    torch.cuda.synchronize()
    assert float(loss.detach()) == L
    assert torch.equal(xd.grad, 2.0 * gx) and torch.equal(td.grad, 2.0 * gt)
    pb.close()
