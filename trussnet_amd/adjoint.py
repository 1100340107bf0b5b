"""The differentiable relaxation (tn_adj_* of the C ABI) and a torch.autograd.Function over it.

No counterpart in the reference, whose relaxation is not differentiable (hinge
bars, nearest-voxel sizes, projections, FIRE resets). The surrogate keeps its bar
force and Jacobi step (src/opencl/v2m_particle_body.cl:161-184, 844-870) with a
softplus bar law, a smooth active-bar count and a sizing grid theta interpolated
trilinearly in smoothstep cell fractions (C1 across cell faces);
interior nodes move, the others stay, the lists and tets are fixed. After `iters`
steps the loss is the mean truss energy plus wq times the mean (1 - smoothed
Joe-Liu quality) of the tets (src/v2m_tetra.cpp:100-120). The gradient with
respect to the start positions and to theta comes from the hand-written adjoint
kernels, which recompute the steps from checkpoints every `ckpt_every` steps.

torch is used for device memory and streams only: the library receives raw
device pointers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._abi import K
from .relax import _check, lib


class AdjProblemC(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nbr", C.c_void_p), ("nnb", C.c_void_p), ("typ", C.c_void_p),
        ("ntet", C.c_int64), ("tet", C.c_void_p),
        ("gx", C.c_int32), ("gy", C.c_int32), ("gz", C.c_int32),
        ("origin", C.c_float * 3), ("spacing", C.c_float * 3),
        ("dt", C.c_float), ("fscale", C.c_float), ("fsurf", C.c_float), ("beta", C.c_float), ("wq", C.c_float),
        ("eps_v", C.c_float), ("iters", C.c_int32), ("ckpt_every", C.c_int32),
    ]


@dataclass
class AdjParams:
    iters: int = 64
    ckpt_every: int = 8
    dt: float = 0.5
    fscale: float = 1.2
    fsurf: float = 1.0
    beta: float = 4.0     # softplus sharpness, 1/mm
    wq: float = 1.0       # weight of the quality term
    eps_v: float = 1e-3   # volume smoothing, mm^3


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp = C.c_void_p
    L.tn_adj_create.argtypes = [C.POINTER(AdjProblemC), C.c_int32, C.POINTER(vp)]
    L.tn_adj_forward.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double), vp]
    L.tn_adj_backward.argtypes = [vp, vp, vp, vp, vp]
    L.tn_adj_loss_terms.argtypes = [vp, C.POINTER(C.c_double)]
    L.tn_adj_destroy.argtypes = [vp]
    for name in ("tn_adj_create", "tn_adj_forward", "tn_adj_backward", "tn_adj_loss_terms", "tn_adj_destroy"):
        getattr(L, name).restype = C.c_int
    _bound = True
    return L


def problem_struct(nbr, nnb, typ, tets, grid_shape, origin, spacing, prm: AdjParams):
    """The C struct of a problem over host arrays: (struct, the arrays it points at)."""
    keep = (np.ascontiguousarray(nbr, np.int32).reshape(-1, K), np.ascontiguousarray(nnb, np.int32).ravel(),
            np.ascontiguousarray(typ, np.uint8).ravel(), np.ascontiguousarray(tets, np.int32).reshape(-1, 4))
    c = AdjProblemC()
    c.n = int(keep[1].size)
    c.nbr, c.nnb, c.typ = keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data
    c.ntet = keep[3].shape[0]
    c.tet = keep[3].ctypes.data if c.ntet else None
    c.gz, c.gy, c.gx = (int(g) for g in grid_shape)
    c.origin[:] = tuple(float(o) for o in origin)
    c.spacing[:] = tuple(float(v) for v in spacing)
    c.dt, c.fscale, c.fsurf, c.beta, c.wq, c.eps_v = prm.dt, prm.fscale, prm.fsurf, prm.beta, prm.wq, prm.eps_v
    c.iters, c.ckpt_every = prm.iters, prm.ckpt_every
    return c, keep


class Problem:
    """Fixed lists, types and tets, a sizing-grid geometry and the parameters,
    resident on one device (tn_adj_create)."""

    def __init__(self, nbr: np.ndarray, nnb: np.ndarray, typ: np.ndarray, tets: np.ndarray, grid_shape,
                 origin, spacing, prm: AdjParams | None = None, device: int = -1):
        self.prm = prm or AdjParams()
        self.nbr = np.ascontiguousarray(nbr, np.int32).reshape(-1, K)
        self.nnb = np.ascontiguousarray(nnb, np.int32).ravel()
        self.typ = np.ascontiguousarray(typ, np.uint8).ravel()
        self.tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        self.n = int(self.nnb.size)
        self.grid_shape = tuple(int(g) for g in grid_shape)   # (gz, gy, gx)
        self.origin = tuple(float(o) for o in origin)
        self.spacing = tuple(float(s) for s in spacing)
        self.c = self.to_c()
        self._h = C.c_void_p()
        _check(_bind().tn_adj_create(C.byref(self.c), device, C.byref(self._h)))

    def to_c(self) -> AdjProblemC:
        c, self._keep = problem_struct(self.nbr, self.nnb, self.typ, self.tets, self.grid_shape, self.origin,
                                       self.spacing, self.prm)
        return c

    def close(self):
        if self._h:
            _bind().tn_adj_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # raw device pointers (ints); stream: a hipStream_t as an int, or 0
    def forward_ptr(self, x0: int, theta: int, xT: int = 0, stream: int = 0) -> float:
        loss = C.c_double()
        _check(_bind().tn_adj_forward(self._h, x0, theta, xT or None, C.byref(loss), stream or None))
        return float(loss.value)

    def backward_ptr(self, theta: int, grad_x0: int, grad_theta: int, stream: int = 0):
        _check(_bind().tn_adj_backward(self._h, theta, grad_x0 or None, grad_theta or None, stream or None))

    def terms(self) -> tuple[float, float, float]:
        t = (C.c_double * 3)()
        _check(_bind().tn_adj_loss_terms(self._h, t))
        return float(t[0]), float(t[1]), float(t[2])

    def problem_arrays(self) -> dict:
        """The problem as host arrays (for the CPU gradient)."""
        return {"nbr": self.nbr, "nnb": self.nnb, "typ": self.typ, "tets": self.tets,
                "grid_shape": self.grid_shape, "origin": self.origin, "spacing": self.spacing, "prm": self.prm}


def _torch():
    import torch

    return torch


class RelaxLoss:
    """loss = RelaxLoss.apply(x0, theta, problem): x0 (n, 3) and theta (gz, gy, gx)
    float32 CUDA tensors; the loss is a float64 scalar tensor. Gradients flow to both."""

    _fn = None

    @classmethod
    def apply(cls, x0, theta, problem: Problem):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, x0, theta, problem):
                    x0c = x0.detach().contiguous()
                    th = theta.detach().contiguous()
                    if x0c.dtype != torch.float32 or th.dtype != torch.float32 or not x0c.is_cuda:
                        raise TypeError("RelaxLoss: x0 and theta must be float32 device tensors")
                    if x0c.shape != (problem.n, 3) or tuple(th.shape) != problem.grid_shape:
                        raise ValueError("RelaxLoss: x0 must be (n, 3) and theta the problem's grid shape")
                    stream = torch.cuda.current_stream(x0c.device).cuda_stream
                    loss = problem.forward_ptr(x0c.data_ptr(), th.data_ptr(), 0, stream)
                    ctx.save_for_backward(x0c, th)
                    ctx.problem = problem
                    return torch.tensor(loss, dtype=torch.float64, device=x0c.device)

                @staticmethod
                def backward(ctx, g):
                    x0c, th = ctx.saved_tensors
                    problem = ctx.problem
                    stream = torch.cuda.current_stream(x0c.device).cuda_stream
                    # the checkpoints on the handle belong to the last forward: redo it for these inputs
                    problem.forward_ptr(x0c.data_ptr(), th.data_ptr(), 0, stream)
                    gx = torch.empty_like(x0c)
                    gt = torch.empty_like(th)
                    problem.backward_ptr(th.data_ptr(), gx.data_ptr(), gt.data_ptr(), stream)
                    s = g.to(torch.float32)
                    return gx * s, gt * s, None

            cls._fn = _Fn
        return cls._fn.apply(x0, theta, problem)



def build_problem(grid, seeds, prm: AdjParams | None = None, grid_pts: int = 32, device: int = -1,
                  seed: int = 0) -> tuple["Problem", np.ndarray, np.ndarray]:
    """A problem over relaxation seeds: the device's K-nearest lists at the seeds (a
    plan's first rebuild), a Delaunay tetrahedralisation of the seeds (scipy), and a
    grid_pts^3 sizing grid spanning the volume, started from the grid's sizes with a
    smooth +-10 % perturbation. Returns (problem, x0 (n, 3) float32, theta)."""
    from scipy.spatial import Delaunay

    from .relax import Plan

    with Plan(grid, seeds, device=device) as plan:
        rb = plan.rebuild()
    x0 = np.ascontiguousarray(seeds.P, np.float32)
    tets = Delaunay(x0.astype(np.float64)).simplices.astype(np.int32)
    ext = np.array([(grid.nx - 1) * grid.vs[0], (grid.ny - 1) * grid.vs[1], (grid.nz - 1) * grid.vs[2]])
    spacing = tuple(float(e / (grid_pts - 1)) for e in ext)
    # the grid's own sizes at the sizing grid's points (nearest voxel), perturbed
    g = np.arange(grid_pts, dtype=np.float64)
    iz, iy, ix = np.meshgrid(g, g, g, indexing="ij")
    vx = np.clip(np.rint(ix * spacing[0] / grid.vs[0]), 0, grid.nx - 1).astype(np.int64)
    vy = np.clip(np.rint(iy * spacing[1] / grid.vs[1]), 0, grid.ny - 1).astype(np.int64)
    vz = np.clip(np.rint(iz * spacing[2] / grid.vs[2]), 0, grid.nz - 1).astype(np.int64)
    h = np.asarray(grid.h, np.float32)[vx + grid.nx * (vy + grid.ny * vz)]
    rng = np.random.default_rng(seed)
    wave = np.sin(2 * np.pi * (ix / grid_pts + rng.uniform())) * np.cos(2 * np.pi * (iy / grid_pts + rng.uniform()))
    theta = (h * (1.0 + 0.1 * wave * np.cos(2 * np.pi * iz / grid_pts))).astype(np.float32)
    pb = Problem(rb["nbr"], rb["nnb"], seeds.typ, tets, theta.shape, (0.0, 0.0, 0.0), spacing, prm, device)
    return pb, x0, theta
