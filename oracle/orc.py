"""ctypes loader of oracle/lib/liboracle.so, the CPU restatement of the relaxation.

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared with. Built by
``make -C oracle`` from oracle/oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from trussnet_amd._abi import K, Grid, GridView, HashInfoC, Nodes, NodesView, RelaxParams, RelaxParamsC, RelaxStatsC

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")


class IterTrace(C.Structure):
    _fields_ = [("it", C.c_int32), ("rebuilt", C.c_int32), ("nhalf", C.c_int32), ("nrefresh", C.c_int32),
                ("mmax", C.c_float), ("p99", C.c_float), ("fdt", C.c_float), ("falpha", C.c_float),
                ("power", C.c_double), ("hist", C.c_int32 * 32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_hash_geometry.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(HashInfoC)]
        L.orc_rebuild.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView)] + [C.c_void_p] * 6
        L.orc_force.argtypes = [C.POINTER(RelaxParamsC), C.POINTER(NodesView)] + [C.c_void_p] * 4
        L.orc_move.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView), C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.orc_relax.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView),
                                C.POINTER(RelaxStatsC), C.POINTER(IterTrace), C.c_int32]
        L.orc_field_points.argtypes = [C.POINTER(GridView), C.c_int32, C.c_int32, C.c_int64] + [C.c_void_p] * 6
        L.orc_sdf_points.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_hist_bin.argtypes = [C.c_float]
        L.orc_hist_bin.restype = C.c_int32
        L.orc_thin_nodes.restype = C.c_int64
        L.orc_thin_nodes.argtypes = [C.POINTER(GridView), C.c_float, C.c_int64] + [C.c_void_p] * 6
        L.orc_protect_features.restype = C.c_int64
        L.orc_protect_features.argtypes = [C.POINTER(GridView), C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def hash_info(grid: Grid, prm: RelaxParams) -> HashInfoC:
    hi = HashInfoC()
    gv, pc = grid.view(), prm.to_c()
    lib().orc_hash_geometry(C.byref(gv), C.byref(pc), C.byref(hi))
    return hi


def rebuild(grid: Grid, nodes: Nodes, prm: RelaxParams) -> dict:
    n = nodes.n
    nk = hash_info(grid, prm).nkeys
    r = {"key": np.zeros(n, np.int32), "sorted": np.zeros(n, np.int32), "cstart": np.zeros(nk + 1, np.int32),
         "nbr": np.zeros(n * K, np.int32), "nnb": np.zeros(n, np.int32), "hn": np.zeros(n, np.float32)}
    gv, pc, nv = grid.view(), prm.to_c(), nodes.view()
    _check(lib().orc_rebuild(C.byref(gv), C.byref(pc), C.byref(nv), _p(r["key"]), _p(r["sorted"]), _p(r["cstart"]),
                             _p(r["nbr"]), _p(r["nnb"]), _p(r["hn"])))
    return r


def force(nodes: Nodes, prm: RelaxParams, hn, nbr, nnb) -> np.ndarray:
    F = np.zeros(nodes.n * 4, np.float32)
    pc, nv = prm.to_c(), nodes.view()
    lib().orc_force(C.byref(pc), C.byref(nv), _p(hn), _p(nbr), _p(nnb), _p(F))
    return F


def move(grid: Grid, nodes: Nodes, prm: RelaxParams, hn, F, V, fdt, falpha):
    out = nodes.copy()
    hn = np.array(hn, np.float32)
    V = np.array(V, np.float32)
    pw = np.zeros(nodes.n, np.float32)
    mv = np.zeros(nodes.n, np.float32)
    gv, pc, nv = grid.view(), prm.to_c(), out.view()
    _check(lib().orc_move(C.byref(gv), C.byref(pc), C.byref(nv), _p(hn), _p(F), _p(V), fdt, falpha, _p(pw), _p(mv)))
    return out, hn, V, pw, mv


def relax(grid: Grid, nodes: Nodes, prm: RelaxParams, *, trace: bool = False, nthreads: int = 0):
    """relax_cpu's restatement on a copy of the nodes: (nodes', stats dict[, trace list])."""
    out = nodes.copy()
    st = RelaxStatsC()
    tr = (IterTrace * max(prm.max_iters, 1))() if trace else None
    gv, pc, nv = grid.view(), prm.to_c(), out.view()
    _check(lib().orc_relax(C.byref(gv), C.byref(pc), C.byref(nv), C.byref(st),
                           tr if tr is not None else None, nthreads))
    if trace:
        recs = [{"it": r.it, "rebuilt": r.rebuilt, "nhalf": r.nhalf, "nrefresh": r.nrefresh, "mmax": r.mmax,
                 "p99": r.p99, "fdt": r.fdt, "falpha": r.falpha, "power": r.power, "hist": list(r.hist)}
                for r in list(tr)[: st.iters]]
        return out, st.as_dict(), recs
    return out, st.as_dict()


def field_points(grid: Grid, a: int, b: int, X: np.ndarray):
    X = np.ascontiguousarray(X, np.float32)
    m = X.shape[0]
    phi = np.zeros(m, np.float32)
    psi = np.zeros(m, np.float32)
    grad = np.zeros((m, 3), np.float32)
    lab = np.zeros(m, np.int32)
    sec = np.zeros(m, np.int32)
    gv = grid.view()
    lib().orc_field_points(C.byref(gv), a, b, m, _p(X), _p(phi), _p(psi), _p(grad), _p(lab), _p(sec))
    return phi, psi, grad, lab, sec


def thin_keep(grid: Grid, nodes: Nodes, thin: float) -> np.ndarray:
    """thin_nodes (src/v2m_particles.cpp:1223-1357): the keep flags of the nodes."""
    keep = np.zeros(nodes.n, np.uint8)
    gv = grid.view()
    p3 = None if nodes.part3 is None else np.ascontiguousarray(nodes.part3, np.uint16)
    lib().orc_thin_nodes(C.byref(gv), thin, nodes.n, _p(nodes.P), _p(nodes.lab), _p(nodes.typ), _p(nodes.part),
                         _p(p3) if p3 is not None else None, _p(keep))
    return keep.astype(bool)


def protect_keep(grid: Grid, feat: np.ndarray, nodes: Nodes) -> np.ndarray:
    """protect_features (src/v2m_particles.cpp:1002-1221) with feature curves `feat`: keep flags."""
    keep = np.zeros(nodes.n, np.uint8)
    gv = grid.view()
    feat = np.ascontiguousarray(feat, np.float32)
    lib().orc_protect_features(C.byref(gv), _p(feat), feat.size, nodes.n, _p(nodes.P), _p(nodes.typ), _p(keep))
    return keep.astype(bool)


def hist_bin(m: float) -> int:
    return int(lib().orc_hist_bin(m))


def _adj_bind():
    L = lib()
    if not getattr(L, "_adj_bound", False):
        from trussnet_amd.adjoint import AdjProblemC

        L.orc_adj_loss.restype = C.c_double
        L.orc_adj_loss.argtypes = [C.POINTER(AdjProblemC), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_adj_grad.restype = C.c_double
        L.orc_adj_grad.argtypes = [C.POINTER(AdjProblemC)] + [C.c_void_p] * 5
        L._adj_bound = True
    return L


def adj_loss(pc, x0: np.ndarray, theta: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
    """The differentiable relaxation's loss in double precision (oracle/adjoint.c):
    (loss, x_T, [loss, energy, quality term]). pc: an AdjProblemC."""
    x0 = np.ascontiguousarray(x0, np.float64)
    th = np.ascontiguousarray(theta, np.float64)
    xT = np.zeros_like(x0)
    terms = np.zeros(3)
    L = _adj_bind().orc_adj_loss(C.byref(pc), _p(x0), _p(th), _p(xT), _p(terms))
    return float(L), xT, terms


def adj_grad(pc, x0: np.ndarray, theta: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
    """(loss, dL/dx0, dL/dtheta) in double precision (reverse mode, every state kept)."""
    x0 = np.ascontiguousarray(x0, np.float64)
    th = np.ascontiguousarray(theta, np.float64)
    gx = np.zeros_like(x0)
    gt = np.zeros_like(th)
    terms = np.zeros(3)
    L = _adj_bind().orc_adj_grad(C.byref(pc), _p(x0), _p(th), _p(gx), _p(gt), _p(terms))
    return float(L), gx, gt


def sdf_points(prog: np.ndarray, l: int, X: np.ndarray, grad: bool = False):
    """Label l's analytic field of a shape program at points X (n, 3): (values, gradients or None)."""
    prog = np.ascontiguousarray(prog, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    m = X.shape[0]
    v = np.zeros(m, np.float32)
    g = np.zeros((m, 3), np.float32) if grad else None
    lib().orc_sdf_points(_p(prog), l, m, _p(X), int(grad), _p(v), _p(g) if grad else None)
    return v, g
