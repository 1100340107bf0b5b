"""The grid build on the device (tn_grid_build of the C ABI): build_grid_cpu
(src/v2m_grid.cpp:47-722) from a label volume, a tissue-probability input or a
gray-scale input, to the Grid the relaxation and the seeding read.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import BL, SLOT, Grid, GridInfoC, GridParams, VolumeViewC
from .relax import _check, lib

_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        vp = C.c_void_p
        L.tn_grid_build.argtypes = [C.POINTER(VolumeViewC), vp, C.c_int32, C.POINTER(vp)]
        L.tn_grid_get_info.argtypes = [vp, C.POINTER(GridInfoC)]
        L.tn_grid_download.argtypes = [vp] + [vp] * 7
        L.tn_grid_destroy.argtypes = [vp]
        L.tn_grid_params_default.argtypes = [vp]
        L.tn_grid_params_default.restype = None
        for n in ("tn_grid_build", "tn_grid_get_info", "tn_grid_download", "tn_grid_destroy"):
            getattr(L, n).restype = C.c_int
        _bound = True
    return L


def volume_view(labels: np.ndarray, vs=(1.0, 1.0, 1.0), prob: np.ndarray | None = None, gray: np.ndarray | None = None,
                thresholds=None, hvox: np.ndarray | None = None) -> tuple[VolumeViewC, list]:
    """The C view of a (nz, ny, nx) volume and the arrays it points into."""
    lab = np.ascontiguousarray(labels, np.uint16)
    nz, ny, nx = lab.shape
    v = VolumeViewC()
    v.nx, v.ny, v.nz = nx, ny, nz
    v.vs[:] = [float(x) for x in vs]
    keep = [lab]
    v.L = lab.ctypes.data_as(C.POINTER(C.c_uint16))
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    if prob is not None:
        pr = np.ascontiguousarray(prob, np.float32)
        keep.append(pr)
        v.prob, v.nprob = fp(pr), pr.shape[0]
    if thresholds is not None and len(thresholds):
        gr = np.ascontiguousarray(gray, np.float32)
        th = np.ascontiguousarray(np.sort(np.asarray(thresholds, np.float32)))
        keep += [gr, th]
        v.gray, v.thresholds, v.nthr = fp(gr), fp(th), th.size
    if hvox is not None:
        hv = np.ascontiguousarray(hvox, np.float32).ravel()
        keep.append(hv)
        v.hvox = fp(hv)
    return v, keep


def build(labels: np.ndarray, vs=(1.0, 1.0, 1.0), prm: GridParams | None = None, *, prob: np.ndarray | None = None,
          gray: np.ndarray | None = None, thresholds=None, hvox: np.ndarray | None = None, device: int = -1,
          info: dict | None = None) -> Grid:
    """Build the grid of a (nz, ny, nx) label volume on the device. With prob
    (C, nz, ny, nx) the interface fields come from the probabilities; with gray and
    thresholds they are the intensity's memberships (labels: the number of
    thresholds <= the intensity). info, if given, receives the Grid's scalars."""
    L = _bind()
    prm = prm or GridParams()
    pc, pkeep = prm.to_c()
    v, vkeep = volume_view(labels, vs, prob, gray, thresholds, hvox)
    h = C.c_void_p()
    _check(L.tn_grid_build(C.byref(v), C.cast(C.byref(pc), C.c_void_p), device, C.byref(h)))
    try:
        gi = GridInfoC()
        _check(L.tn_grid_get_info(h, C.byref(gi)))
        nv = gi.nx * gi.ny * gi.nz
        nb = gi.nbx * gi.nby * gi.nbz
        bl_cnt = np.empty(nb, np.int32)
        bl_lab = np.empty(nb * BL, np.uint16)
        bl_slot = np.empty(nb, np.int32)
        phi = np.empty(gi.nslot * SLOT, np.float32)
        hh = np.empty(nv, np.float32)
        grade = np.empty(nv, np.uint8)
        gTW = np.empty(2 * gi.gm, np.float32)
        ptr = lambda a: a.ctypes.data if a.size else None
        _check(L.tn_grid_download(h, ptr(bl_cnt), ptr(bl_lab), ptr(bl_slot), ptr(phi), ptr(hh), ptr(grade), ptr(gTW)))
    finally:
        L.tn_grid_destroy(h)
    if info is not None:
        info.update({k: getattr(gi, k) for k, _ in GridInfoC._fields_})
    lab = np.ascontiguousarray(labels, np.uint16).ravel()
    gI = np.ascontiguousarray(gray, np.float32).ravel() if gi.gm else None
    return Grid(gi.nx, gi.ny, gi.nz, tuple(float(x) for x in vs), lab, bl_cnt, bl_lab, bl_slot, phi, hh,
                float(gi.hmin), float(gi.hmax), grade, gI, gTW if gi.gm else None, int(gi.gm))


class Resident:
    """A grid built on the device and kept there (a tn_grid handle)."""

    def __init__(self, handle: C.c_void_p):
        self.h = handle

    def info(self) -> dict:
        gi = GridInfoC()
        _check(_bind().tn_grid_get_info(self.h, C.byref(gi)))
        return {k: getattr(gi, k) for k, _ in GridInfoC._fields_}

    def seed(self, prm=None):
        """seed_cpu on this grid without leaving the device (tn_seed_from_grid)."""
        from ._abi import Nodes, RelaxParams, RelaxParamsC

        L = _bind()
        pc = (prm or RelaxParams()).to_c()
        vp = C.c_void_p
        L.tn_seed_count_from_grid.argtypes = [vp, C.POINTER(RelaxParamsC), C.POINTER(C.c_int64)]
        L.tn_seed_from_grid.argtypes = [vp, C.POINTER(RelaxParamsC), C.c_int64] + [vp] * 5
        n = C.c_int64(0)
        _check(L.tn_seed_count_from_grid(self.h, C.byref(pc), C.byref(n)))
        n = int(n.value)
        P = np.zeros((n, 3), np.float32)
        lab = np.zeros(n, np.uint16)
        typ = np.zeros(n, np.uint8)
        part = np.zeros((n, 2), np.uint16)
        part3 = np.zeros(n, np.uint16)
        ptr = lambda a: a.ctypes.data if a.size else None
        _check(L.tn_seed_from_grid(self.h, C.byref(pc), n, ptr(P), ptr(lab), ptr(typ), ptr(part), ptr(part3)))
        return Nodes(P, lab, typ, part, part3)

    def plan(self, seeds, prm=None):
        """A relaxation plan over this grid (tn_plan_create_from_grid)."""
        from .relax import Plan

        return Plan(None, seeds, prm, resident=self)

    def close(self):
        if self.h:
            _bind().tn_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_resident(labels: np.ndarray, vs=(1.0, 1.0, 1.0), prm: GridParams | None = None, *, prob=None, gray=None,
                   thresholds=None, hvox=None, device: int = -1) -> Resident:
    """Build the grid on the device and keep it there (no download)."""
    L = _bind()
    pc, pkeep = (prm or GridParams()).to_c()
    v, vkeep = volume_view(labels, vs, prob, gray, thresholds, hvox)
    h = C.c_void_p()
    _check(L.tn_grid_build(C.byref(v), C.cast(C.byref(pc), C.c_void_p), device, C.byref(h)))
    return Resident(h)
