"""ctypes mirrors of the C ABI structs in include/trussnet_amd.h.

The structs describe v2mesh's Grid (src/v2m_grid.h:45-81), RelaxParams
(src/v2m_particles.h:21-41), Nodes (src/v2m_particles.h:78-87) and RelaxStats
(src/v2m_particles.h:89-96) as raw arrays. ``Grid`` and ``Nodes`` below hold the
numpy arrays a view points into, so a view stays valid while its owner lives.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

K = 24
MAXLEV = 8
BS = 8
SLOT = 512
BL = 16
NOLAB = 0xFFFF
INTERIOR, INTERFACE, JUNCTION, CORNER = 0, 1, 2, 3

TN_OK, TN_ERR_ARG, TN_ERR_NODEVICE, TN_ERR_HIP, TN_ERR_UNSUPPORTED, TN_ERR_INTERNAL = range(6)

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)


class GridView(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nbx", C.c_int32), ("nby", C.c_int32), ("nbz", C.c_int32),
        ("vs", C.c_float * 3),
        ("L", _u16p), ("bl_cnt", _i32p), ("bl_lab", _u16p), ("bl_slot", _i32p),
        ("phi", _f32p), ("nslot", C.c_int64), ("h", _f32p),
        ("hmin", C.c_float), ("hmax", C.c_float),
        ("gI", _f32p), ("gTW", _f32p), ("ngTW", C.c_int64), ("gm", C.c_int32),
        ("feat", _f32p), ("nfeat", C.c_int64),
    ]


class RelaxParamsC(C.Structure):
    _fields_ = [
        ("nseed", C.c_int32), ("t", C.c_float), ("skin", C.c_float), ("fscale", C.c_float),
        ("fsurf", C.c_float), ("dt", C.c_float), ("maxstep", C.c_float), ("snap", C.c_float),
        ("max_iters", C.c_int32), ("dptol", C.c_float), ("jseed", C.c_float),
        ("corners", C.c_int32), ("voxel_trap", C.c_int32), ("thin", C.c_float),
        ("fire", C.c_int32), ("fire_dtmax", C.c_float), ("verbose", C.c_int32),
        ("rebuild_div", C.c_int32),
    ]


class NodesView(C.Structure):
    _fields_ = [("n", C.c_int64), ("P", _f32p), ("lab", _u16p), ("typ", _u8p), ("part", _u16p)]


class RelaxStatsC(C.Structure):
    _fields_ = [
        ("iters", C.c_int32), ("rebuilds", C.c_int32), ("last_move", C.c_float), ("last_p99", C.c_float),
        ("fire_resets", C.c_int32), ("fire_dt", C.c_float),
        ("n_interior", C.c_int64), ("n_interface", C.c_int64), ("n_junction", C.c_int64), ("n_corner", C.c_int64),
        ("ms_seed", C.c_double), ("ms_hash", C.c_double), ("ms_force", C.c_double), ("ms_move", C.c_double),
        ("power_fallbacks", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class HashInfoC(C.Structure):
    _fields_ = [
        ("b0", C.c_float), ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
        ("nlev", C.c_int32), ("dim", (C.c_int32 * 3) * MAXLEV), ("off", C.c_int32 * MAXLEV),
        ("nkeys", C.c_int32),
    ]


class KernelTimeC(C.Structure):
    _fields_ = [("name", C.c_char * 16), ("launches", C.c_int64), ("ms", C.c_double), ("items", C.c_int64)]


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


class GridParamsC(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("sigma", "sigma_thin", "thin_lo", "thin_hi", "thick", "thin_floor", "preserve",
                                          "sigma_curv", "hbase", "hmin", "hmax", "K", "g")] + [
        ("hlab", _f32p), ("nhlab", C.c_int32), ("iglob", C.c_float), ("ilab", _f32p), ("nilab", C.c_int32),
        ("ipair", _f32p), ("npair", C.c_int32)]


class VolumeViewC(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("vs", C.c_float * 3), ("L", _u16p),
                ("prob", _f32p), ("nprob", C.c_int32), ("gray", _f32p), ("thresholds", _f32p), ("nthr", C.c_int32),
                ("hvox", _f32p)]


class GridInfoC(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nbx", C.c_int32), ("nby", C.c_int32),
                ("nbz", C.c_int32), ("nslot", C.c_int64), ("hbase", C.c_float), ("hmin", C.c_float),
                ("hmax", C.c_float), ("hcurv", C.c_float), ("R", C.c_int32), ("limit_sweeps", C.c_int32),
                ("overflow_bricks", C.c_int32), ("gm", C.c_int32),
                ("prob_fields", C.c_int32)]


@dataclass
class GridParams:
    """GridParams (src/v2m_grid.h:21-43): smoothing, thin-layer blend, sizing and
    gradient limit of the grid build; hlab per-label sizes, iglob / ilab / ipair the
    interface sizes (InterfaceSizes, src/v2m_isize.h:25-28, ipair as (a, b, h) rows)."""
    sigma: float = 1.0
    sigma_thin: float = 0.35
    thin_lo: float = 2.0
    thin_hi: float = 4.0
    thick: float = 0.0
    thin_floor: float = 0.5
    preserve: float = 0.0
    sigma_curv: float = 2.0
    hbase: float = 0.0
    hmin: float = 0.0
    hmax: float = 0.0
    K: float = 3.0
    g: float = 0.3
    hlab: tuple = ()
    iglob: float = 0.0
    ilab: tuple = ()
    ipair: tuple = ()

    def to_c(self) -> tuple[GridParamsC, list]:
        """The C struct and the arrays it points into (keep them alive with it)."""
        c = GridParamsC()
        for n in ("sigma", "sigma_thin", "thin_lo", "thin_hi", "thick", "thin_floor", "preserve", "sigma_curv",
                  "hbase", "hmin", "hmax", "K", "g", "iglob"):
            setattr(c, n, float(getattr(self, n)))
        keep = []
        for n, cnt in (("hlab", "nhlab"), ("ilab", "nilab"), ("ipair", "npair")):
            a = np.ascontiguousarray(np.asarray(getattr(self, n), np.float32).ravel())
            keep.append(a)
            setattr(c, n, a.ctypes.data_as(_f32p) if a.size else None)
            setattr(c, cnt, a.size // 3 if n == "ipair" else a.size)
        return c, keep


@dataclass
class RelaxParams:
    """RelaxParams with the reference defaults (src/v2m_particles.h:22-39)."""
    nseed: int = 8
    t: float = 1.3
    skin: float = 0.3
    fscale: float = 1.2
    fsurf: float = 1.0
    dt: float = 0.5
    maxstep: float = 0.2
    snap: float = 0.5
    max_iters: int = 500
    dptol: float = 2e-3
    jseed: float = 0.8
    corners: bool = True
    voxel_trap: bool = False
    thin: float = 0.0
    fire: bool = True
    fire_dtmax: float = 2.0
    verbose: bool = False
    rebuild_div: int = 300

    def to_c(self) -> RelaxParamsC:
        c = RelaxParamsC()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, int(v) if isinstance(v, bool) else v)
        return c


@dataclass
class Grid:
    """The field the nodes move in (v2mesh Grid, src/v2m_grid.h:45-81), x fastest.

    ``L`` and ``h`` are flat arrays of nx*ny*nz; ``phi`` holds nslot*512 values.
    Gray-scale input (gm = m > 0): ``gI`` is the intensity per voxel and ``gTW`` the m
    ascending thresholds followed by their m membership widths (Grid::gI, gray_w).
    Shape input (gm < 0): ``gTW`` is the labels' compiled analytic-field program and
    ``feat`` the feature curves (Grid::feat) whose pinned nodes the seeding adds.
    """
    nx: int
    ny: int
    nz: int
    vs: tuple
    L: np.ndarray
    bl_cnt: np.ndarray
    bl_lab: np.ndarray
    bl_slot: np.ndarray
    phi: np.ndarray
    h: np.ndarray
    hmin: float
    hmax: float
    grade: np.ndarray | None = None
    gI: np.ndarray | None = None
    gTW: np.ndarray | None = None
    gm: int = 0
    feat: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nbx(self):
        return (self.nx + BS - 1) // BS

    @property
    def nby(self):
        return (self.ny + BS - 1) // BS

    @property
    def nbz(self):
        return (self.nz + BS - 1) // BS

    @property
    def nslot(self):
        return self.phi.size // SLOT

    def normalized(self) -> "Grid":
        """The same grid with every array contiguous and of the ABI dtype."""
        return Grid(self.nx, self.ny, self.nz, tuple(float(v) for v in self.vs),
                    np.ascontiguousarray(self.L, np.uint16).ravel(),
                    np.ascontiguousarray(self.bl_cnt, np.int32).ravel(),
                    np.ascontiguousarray(self.bl_lab, np.uint16).ravel(),
                    np.ascontiguousarray(self.bl_slot, np.int32).ravel(),
                    np.ascontiguousarray(self.phi, np.float32).ravel(),
                    np.ascontiguousarray(self.h, np.float32).ravel(),
                    float(self.hmin), float(self.hmax),
                    None if self.grade is None else np.ascontiguousarray(self.grade, np.uint8).ravel(),
                    None if self.gI is None else np.ascontiguousarray(self.gI, np.float32).ravel(),
                    None if self.gTW is None else np.ascontiguousarray(self.gTW, np.float32).ravel(), int(self.gm),
                    None if self.feat is None else np.ascontiguousarray(self.feat, np.float32).ravel())

    def check(self) -> None:
        nv = self.nx * self.ny * self.nz
        nb = self.nbx * self.nby * self.nbz
        if self.L.size != nv or self.h.size != nv:
            raise ValueError("grid: L and h need nx*ny*nz values")
        if self.bl_cnt.size != nb or self.bl_slot.size != nb or self.bl_lab.size != nb * BL:
            raise ValueError("grid: brick tables need nbx*nby*nbz entries")
        if self.phi.size % SLOT:
            raise ValueError("grid: phi must hold whole 512-value slots")

    def view(self) -> GridView:
        g = self.normalized() if not self._is_normal() else self
        g.check()
        v = GridView()
        v.nx, v.ny, v.nz = g.nx, g.ny, g.nz
        v.nbx, v.nby, v.nbz = g.nbx, g.nby, g.nbz
        v.vs[:] = [float(x) for x in g.vs]
        v.L = _ptr(g.L, C.c_uint16)
        v.bl_cnt = _ptr(g.bl_cnt, C.c_int32)
        v.bl_lab = _ptr(g.bl_lab, C.c_uint16)
        v.bl_slot = _ptr(g.bl_slot, C.c_int32)
        v.phi = _ptr(g.phi, C.c_float)
        v.nslot = g.nslot
        v.h = _ptr(g.h, C.c_float)
        v.hmin, v.hmax = g.hmin, g.hmax
        v.gI = _ptr(g.gI, C.c_float)
        v.gTW = _ptr(g.gTW, C.c_float)
        v.ngTW = 0 if g.gTW is None else g.gTW.size
        v.gm = int(g.gm)
        v.feat = _ptr(g.feat, C.c_float)
        v.nfeat = 0 if g.feat is None else g.feat.size
        if g is not self:
            self._keep.append(g)
        return v

    def _is_normal(self) -> bool:
        want = [(self.L, np.uint16), (self.bl_cnt, np.int32), (self.bl_lab, np.uint16), (self.bl_slot, np.int32),
                (self.phi, np.float32), (self.h, np.float32)]
        want += [(a, np.float32) for a in (self.gI, self.gTW, self.feat) if a is not None]
        return all(a.dtype == t and a.flags["C_CONTIGUOUS"] for a, t in want)

    def arrays(self) -> dict:
        d = {"dims": np.array([self.nx, self.ny, self.nz], np.int32), "vs": np.array(self.vs, np.float32),
             "hminmax": np.array([self.hmin, self.hmax], np.float32), "L": self.L, "bl_cnt": self.bl_cnt,
             "bl_lab": self.bl_lab, "bl_slot": self.bl_slot, "phi": self.phi, "h": self.h}
        if self.gm:
            d.update(gTW=self.gTW, gm=np.int32(self.gm))
        if self.gI is not None:
            d["gI"] = self.gI
        if self.feat is not None:
            d["feat"] = self.feat
        return d

    @staticmethod
    def from_arrays(d, prefix: str = "") -> "Grid":
        g = lambda k: d[prefix + k]
        nx, ny, nz = (int(x) for x in g("dims"))
        hmin, hmax = (float(x) for x in g("hminmax"))
        grade = d[prefix + "grade"] if (prefix + "grade") in d else None
        gm = int(d[prefix + "gm"]) if (prefix + "gm") in d else 0
        gI = d[prefix + "gI"] if gm > 0 else None
        feat = d[prefix + "feat"] if (prefix + "feat") in d else None
        return Grid(nx, ny, nz, tuple(float(x) for x in g("vs")), g("L"), g("bl_cnt"), g("bl_lab"), g("bl_slot"),
                    g("phi"), g("h"), hmin, hmax, grade, gI, g("gTW") if gm else None, gm, feat).normalized()


@dataclass
class Nodes:
    """Nodes (src/v2m_particles.h:78-87): P (n,3) float32, lab, typ, part (n,2), and
    optionally part3 (n,), the third partner of a CORNER node (not read by the
    relaxation, which keeps corners fixed)."""
    P: np.ndarray
    lab: np.ndarray
    typ: np.ndarray
    part: np.ndarray
    part3: np.ndarray | None = None

    def __post_init__(self):
        self.P = np.ascontiguousarray(self.P, np.float32).reshape(-1, 3)
        self.lab = np.ascontiguousarray(self.lab, np.uint16).ravel()
        self.typ = np.ascontiguousarray(self.typ, np.uint8).ravel()
        self.part = np.ascontiguousarray(self.part, np.uint16).reshape(-1, 2)
        if self.part3 is not None:
            self.part3 = np.ascontiguousarray(self.part3, np.uint16).ravel()
        n = self.lab.size
        if self.P.shape[0] != n or self.typ.size != n or self.part.shape[0] != n or (
                self.part3 is not None and self.part3.size != n):
            raise ValueError("nodes: P, lab, typ, part and part3 must describe the same node count")

    @property
    def n(self) -> int:
        return int(self.lab.size)

    def copy(self) -> "Nodes":
        return Nodes(self.P.copy(), self.lab.copy(), self.typ.copy(), self.part.copy(),
                     None if self.part3 is None else self.part3.copy())

    def view(self) -> NodesView:
        v = NodesView()
        v.n = self.n
        v.P = _ptr(self.P, C.c_float)
        v.lab = _ptr(self.lab, C.c_uint16)
        v.typ = _ptr(self.typ, C.c_uint8)
        v.part = _ptr(self.part, C.c_uint16)
        return v

    def equal(self, other: "Nodes") -> bool:
        return (np.array_equal(self.P.view(np.uint32), other.P.view(np.uint32)) and np.array_equal(self.typ, other.typ)
                and np.array_equal(self.part, other.part) and np.array_equal(self.lab, other.lab))
