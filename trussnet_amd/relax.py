"""Python entry to the gfx950 relaxation: ctypes over lib/libtrussnet_amd.so.

``relax(grid, nodes, params, device)`` has relax_cl's contract
(src/v2m_gpu.h:16-18): it relaxes a copy of the nodes on the device and returns
it with the RelaxStats; any failure raises :class:`DeviceError` (relax_cl
throws std::runtime_error), and the caller decides whether to fall back, as
v2mesh's pipeline does (src/v2m_pipeline.cpp:497-526). There is no silent CPU
path: without the built library every call raises.

``Plan`` keeps the grid and the seeds resident on one device and exposes the
stages of a relaxation (rebuild, forces, single iterations) for checking.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import (K, TN_ERR_NODEVICE, TN_ERR_UNSUPPORTED, TN_OK, Grid, GridView, HashInfoC, KernelTimeC, Nodes,
                   NodesView, RelaxParams, RelaxParamsC, RelaxStatsC)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TN_LIBRARY") or os.path.join(HERE, "lib", "libtrussnet_amd.so")

# every entry point of include/trussnet_amd.h
SYMBOLS = ["tn_relax_params_default", "tn_last_error", "tn_device_count", "tn_relax", "tn_relax_noalloc",
           "tn_relax_batch",
           "tn_seed_count", "tn_seed",
           "tn_plan_create", "tn_plan_run", "tn_plan_download", "tn_plan_destroy", "tn_plan_hash_info",
           "tn_plan_reset", "tn_plan_rebuild", "tn_plan_force", "tn_plan_step", "tn_plan_set_timing",
           "tn_plan_set_timing_mask",
           "tn_plan_timings", "tn_adj_create", "tn_adj_forward", "tn_adj_backward", "tn_adj_loss_terms",
           "tn_adj_destroy", "tn_grid_params_default", "tn_grid_build", "tn_grid_get_info", "tn_grid_download",
           "tn_grid_destroy", "tn_seed_count_from_grid", "tn_seed_from_grid", "tn_plan_create_from_grid",
           "tn_plan_field_points", "tn_thin_nodes", "tn_protect_features"]


class DeviceError(RuntimeError):
    """A failed device call (status code and the library's message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class UnsupportedMode(DeviceError):
    pass


class NoDevice(DeviceError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"trussnet_amd library not built: {LIB_PATH} (python -m trussnet_amd.build)")
        # torch's wheel bundles a HIP runtime of the same soname as the one the library
        # links; loaded first, it serves both, where two runtimes in one process leave
        # torch with no device
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.tn_last_error.restype = C.c_char_p
        L.tn_relax_params_default.argtypes = [C.POINTER(RelaxParamsC)]
        L.tn_device_count.argtypes = [C.POINTER(C.c_int32)]
        L.tn_relax.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView),
                               C.POINTER(RelaxStatsC), C.c_int32]
        L.tn_relax_noalloc.argtypes = L.tn_relax.argtypes
        L.tn_relax_batch.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView),
                                     C.POINTER(RelaxStatsC), C.c_int32, C.c_int32]
        L.tn_plan_create.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.POINTER(NodesView), C.c_int32,
                                     C.POINTER(vp)]
        L.tn_plan_run.argtypes = [vp, C.POINTER(RelaxStatsC)]
        L.tn_plan_download.argtypes = [vp, C.POINTER(NodesView)]
        L.tn_plan_destroy.argtypes = [vp]
        L.tn_plan_hash_info.argtypes = [vp, C.POINTER(HashInfoC)]
        L.tn_plan_reset.argtypes = [vp]
        L.tn_plan_rebuild.argtypes = [vp] + [vp] * 5
        L.tn_plan_force.argtypes = [vp, vp]
        L.tn_plan_field_points.argtypes = [vp, C.c_int32, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp]
        L.tn_plan_step.argtypes = [vp, C.c_int32, C.POINTER(RelaxStatsC)]
        L.tn_thin_nodes.argtypes = [C.POINTER(GridView), C.POINTER(RelaxParamsC), C.c_int64] + [vp] * 6 + [C.c_int32]
        L.tn_protect_features.argtypes = [C.POINTER(GridView), vp, C.c_int64, C.c_int64] + [vp] * 6 + [C.c_int32]
        L.tn_plan_set_timing.argtypes = [vp, C.c_int32]
        L.tn_plan_set_timing_mask.argtypes = [vp, C.c_uint32]
        L.tn_plan_timings.argtypes = [vp, C.POINTER(KernelTimeC), C.c_int32, C.POINTER(C.c_int32)]
        L.tn_seed_count.argtypes = [C.POINTER(GridView), vp, C.POINTER(RelaxParamsC), C.c_int32,
                                    C.POINTER(C.c_int64)]
        L.tn_seed.argtypes = [C.POINTER(GridView), vp, C.POINTER(RelaxParamsC), C.c_int32, C.c_int64, vp, vp, vp,
                              vp, vp]
        for name in SYMBOLS:
            if name not in ("tn_last_error", "tn_relax_params_default", "tn_grid_params_default"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _raise(code: int):
    msg = lib().tn_last_error().decode(errors="replace")
    if code == TN_ERR_NODEVICE:
        raise NoDevice(code, msg)
    if code == TN_ERR_UNSUPPORTED:
        raise UnsupportedMode(code, msg)
    raise DeviceError(code, msg)


def _check(code: int):
    if code != TN_OK:
        _raise(code)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    c = C.c_int32(0)
    rc = lib().tn_device_count(C.byref(c))
    return int(c.value) if rc == TN_OK else 0


def default_params() -> RelaxParams:
    c = RelaxParamsC()
    lib().tn_relax_params_default(C.byref(c))
    return RelaxParams(**{name: (bool(getattr(c, name)) if isinstance(getattr(RelaxParams(), name), bool)
                                 else getattr(c, name)) for name, _ in RelaxParamsC._fields_})


def relax(grid: Grid, nodes: Nodes, prm: RelaxParams | None = None, device: int = -1) -> tuple[Nodes, dict]:
    """Relax a copy of `nodes` on HIP device `device` (-1: the first). Returns (nodes, stats)."""
    prm = prm or RelaxParams()
    out = nodes.copy()
    gv, pc, nv = grid.view(), prm.to_c(), out.view()
    st = RelaxStatsC()
    _check(lib().tn_relax(C.byref(gv), C.byref(pc), C.byref(nv), C.byref(st), device))
    return out, st.as_dict()


def relax_noalloc(grid: Grid, nodes: Nodes, prm: RelaxParams | None = None, device: int = -1) -> tuple[Nodes, dict]:
    """relax() through tn_relax_noalloc: the same result, with every device buffer carved
    from the library's static scratch array (no device allocation)."""
    prm = prm or RelaxParams()
    out = nodes.copy()
    gv, pc, nv = grid.view(), prm.to_c(), out.view()
    st = RelaxStatsC()
    _check(lib().tn_relax_noalloc(C.byref(gv), C.byref(pc), C.byref(nv), C.byref(st), device))
    return out, st.as_dict()


def relax_batch(grids: list, nodes: list, prms=None, device: int = -1) -> list:
    """Relax many volumes concurrently on one device (tn_relax_batch): [(nodes, stats)]
    per volume, each equal to relax(grids[k], nodes[k], prms[k], device)."""
    count = len(grids)
    if len(nodes) != count:
        raise ValueError("one node set per grid")
    if prms is None or isinstance(prms, RelaxParams):
        prms = [prms or RelaxParams()] * count
    outs = [nd.copy() for nd in nodes]
    gv = (GridView * max(count, 1))(*[g.view() for g in grids])
    pc = (RelaxParamsC * max(count, 1))(*[p.to_c() for p in prms])
    nv = (NodesView * max(count, 1))(*[o.view() for o in outs])
    st = (RelaxStatsC * max(count, 1))()
    _check(lib().tn_relax_batch(gv, pc, nv, st, count, device))
    return [(outs[k], st[k].as_dict()) for k in range(count)]


def seed(grid: Grid, prm: RelaxParams | None = None, device: int = -1) -> Nodes:
    """seed_cpu on the device (src/v2m_particles.cpp:625-726): the graded HCP lattice
    seeds, classified, then the junction-line seeds (``prm.jseed`` > 0) and the
    corner nodes (``prm.corners``), with part3 filled. Needs ``grid.grade``."""
    prm = prm or RelaxParams()
    if grid.grade is None:
        raise ValueError("seeding needs the grid's grades")
    g = grid.normalized() if grid.grade.dtype != np.uint8 else grid
    grade = np.ascontiguousarray(g.grade, np.uint8)
    gv, pc = g.view(), prm.to_c()
    n = C.c_int64(0)
    _check(lib().tn_seed_count(C.byref(gv), _p(grade), C.byref(pc), device, C.byref(n)))
    n = int(n.value)
    P = np.zeros((n, 3), np.float32)
    lab = np.zeros(n, np.uint16)
    typ = np.zeros(n, np.uint8)
    part = np.zeros((n, 2), np.uint16)
    part3 = np.zeros(n, np.uint16)
    _check(lib().tn_seed(C.byref(gv), _p(grade), C.byref(pc), device, n, _p(P), _p(lab), _p(typ), _p(part),
                         _p(part3)))
    return Nodes(P, lab, typ, part, part3)


def _compacted(fn, nodes: Nodes) -> Nodes:
    x = nodes.copy()
    part3 = x.part3.copy() if x.part3 is not None else None
    kept = C.c_int64(0)
    _check(fn(x.n, _p(x.P), _p(x.lab), _p(x.typ), _p(x.part), _p(part3) if part3 is not None else None,
              C.byref(kept)))
    m = int(kept.value)
    return Nodes(x.P[:m], x.lab[:m], x.typ[:m], x.part[:m], part3[:m] if part3 is not None else None)


def thin(grid: Grid, nodes: Nodes, thin: float, device: int = -1) -> Nodes:
    """thin_nodes on the device (src/v2m_particles.cpp:1223-1357): the nodes kept, in
    order, with factor ``thin`` (<= 0: all)."""
    gv, pc = grid.view(), RelaxParams(thin=thin).to_c()
    return _compacted(lambda *a: lib().tn_thin_nodes(C.byref(gv), C.byref(pc), *a, device), nodes)


def protect_features(grid: Grid, feat: np.ndarray, nodes: Nodes, device: int = -1) -> Nodes:
    """protect_features (src/v2m_particles.cpp:1002-1221) with the shape input's feature
    curves ``feat`` (Grid::feat): the nodes kept, in order."""
    gv = grid.view()
    feat = np.ascontiguousarray(feat, np.float32)
    return _compacted(lambda *a: lib().tn_protect_features(C.byref(gv), _p(feat), feat.size, *a, device), nodes)


class Plan:
    """A grid and its seed nodes resident on one device (tn_plan_* of the C ABI)."""

    def __init__(self, grid: Grid | None, seeds: Nodes, prm: RelaxParams | None = None, device: int = -1,
                 resident=None):
        """resident: a grid built on the device (trussnet_amd.gridbuild.Resident) to
        plan over instead of a host grid (tn_plan_create_from_grid)."""
        self.prm = prm or RelaxParams()
        self.grid = grid
        self.seeds = seeds
        self.n = seeds.n
        self._h = C.c_void_p()
        pc, nv = self.prm.to_c(), seeds.view()
        if resident is not None:
            L = lib()
            L.tn_plan_create_from_grid.argtypes = [C.c_void_p, C.POINTER(RelaxParamsC), C.POINTER(NodesView),
                                                   C.POINTER(C.c_void_p)]
            _check(L.tn_plan_create_from_grid(resident.h, C.byref(pc), C.byref(nv), C.byref(self._h)))
        else:
            gv = grid.view()
            _check(lib().tn_plan_create(C.byref(gv), C.byref(pc), C.byref(nv), device, C.byref(self._h)))
        self.hash = self.hash_info()

    def close(self):
        if self._h:
            lib().tn_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def hash_info(self) -> HashInfoC:
        hi = HashInfoC()
        _check(lib().tn_plan_hash_info(self._h, C.byref(hi)))
        return hi

    def run(self) -> dict:
        """Restore the seeds and relax them on the device (no host traffic but the control block)."""
        st = RelaxStatsC()
        _check(lib().tn_plan_run(self._h, C.byref(st)))
        return st.as_dict()

    def reset(self):
        _check(lib().tn_plan_reset(self._h))

    def step(self, iters: int) -> dict:
        st = RelaxStatsC()
        _check(lib().tn_plan_step(self._h, iters, C.byref(st)))
        return st.as_dict()

    def download(self) -> Nodes:
        out = self.seeds.copy()
        nv = out.view()
        _check(lib().tn_plan_download(self._h, C.byref(nv)))
        return out

    def rebuild(self) -> dict:
        n = self.n
        r = {"key": np.zeros(n, np.int32), "sorted": np.zeros(n, np.int32),
             "cstart": np.zeros(self.hash.nkeys + 1, np.int32), "nbr": np.zeros(n * K, np.int32),
             "nnb": np.zeros(n, np.int32)}
        _check(lib().tn_plan_rebuild(self._h, _p(r["key"]), _p(r["sorted"]), _p(r["cstart"]), _p(r["nbr"]),
                                     _p(r["nnb"])))
        return r

    def force(self) -> np.ndarray:
        F = np.zeros(self.n * 4, np.float32)
        _check(lib().tn_plan_force(self._h, _p(F)))
        return F

    def field_points(self, a: int, b: int, X: np.ndarray):
        """The device's field queries at points X (m x 3): phi_a, psi_ab, its gradient,
        the owner and the runner-up (tn_plan_field_points)."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 3)
        m = X.shape[0]
        phi, psi = np.zeros(m, np.float32), np.zeros(m, np.float32)
        grad = np.zeros((m, 3), np.float32)
        lab, sec = np.zeros(m, np.int32), np.zeros(m, np.int32)
        _check(lib().tn_plan_field_points(self._h, int(a), int(b), m, _p(X), _p(phi), _p(psi), _p(grad), _p(lab),
                                          _p(sec)))
        return phi, psi, grad, lab, sec

    # the kernel classes, in the order of their bits in a timing mask
    TIMING_CLASSES = ("keys", "scan", "radix_sort", "permute", "neighbors", "force", "move_easy", "move", "move_int",
                      "move_rest", "move_def", "stats", "refresh", "control")

    def set_timing(self, on=True):
        """Per-kernel device timing (HIP events on the stream of each launch); clears the
        totals. on: True / False for every class, or an iterable of class names to time
        only those (the others record no events)."""
        if isinstance(on, bool):
            _check(lib().tn_plan_set_timing(self._h, 1 if on else 0))
            return
        mask = 0
        for name in on:
            mask |= 1 << self.TIMING_CLASSES.index(name)
        _check(lib().tn_plan_set_timing_mask(self._h, mask))

    def timings(self) -> dict:
        """{kernel class: (launches, total ms, items)} since timing was switched on (items: the
        nodes those launches processed)."""
        buf = (KernelTimeC * 32)()
        cnt = C.c_int32(0)
        _check(lib().tn_plan_timings(self._h, buf, 32, C.byref(cnt)))
        return {buf[k].name.decode(): (int(buf[k].launches), float(buf[k].ms), int(buf[k].items))
                for k in range(min(cnt.value, 32))}
