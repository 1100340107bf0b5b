"""The node hand-off file between the relaxation and the step after it.

The reference's `--dump-nodes` file (dump_nodes, src/v2m_pipeline.cpp:58-73): one
JSON header line mapping each array name to [dtype, shape, byte offset], then the
raw little-endian arrays in that order -- P (float32, n x 3), Label (uint16, n),
Type (uint8, n), Partner (uint16, n x 2). The header is what nlohmann::json's
compact dump writes: keys in sorted order, no spaces. Files written here are read
by the reference's own tools (tools/v2mview.py) and vice versa.
"""
from __future__ import annotations

import json

import numpy as np

from ._abi import Nodes

_ORDER = [("P", "float32", 3), ("Label", "uint16", 1), ("Type", "uint8", 1), ("Partner", "uint16", 2)]


def header(n: int) -> bytes:
    """The header line of an n-node file (with its newline)."""
    off, spec = 0, {}
    for name, dt, width in _ORDER:
        shape = [n, width] if width > 1 else [n]
        spec[name] = [dt, shape, off]
        off += n * width * np.dtype(dt).itemsize
    return (json.dumps(spec, separators=(",", ":"), sort_keys=True) + "\n").encode()


def dump_nodes(path: str, nodes: Nodes) -> None:
    """Write nodes in the reference's hand-off format."""
    n = nodes.n
    with open(path, "wb") as f:
        f.write(header(n))
        f.write(np.ascontiguousarray(nodes.P, "<f4").tobytes())
        f.write(np.ascontiguousarray(nodes.lab, "<u2").tobytes())
        f.write(np.ascontiguousarray(nodes.typ, "u1").tobytes())
        f.write(np.ascontiguousarray(nodes.part, "<u2").tobytes())


def load_nodes(path: str) -> Nodes:
    """Read a hand-off file (the reference's or this module's)."""
    with open(path, "rb") as f:
        head = f.readline()
        body = f.read()
    spec = json.loads(head)
    arrays = {}
    for name, (dt, shape, off) in spec.items():
        count = int(np.prod(shape))
        arrays[name] = np.frombuffer(body, dtype=np.dtype(dt).newbyteorder("<"), count=count,
                                     offset=int(off)).reshape(shape)
    return Nodes(arrays["P"], arrays["Label"], arrays["Type"], arrays["Partner"])


def node_attributes(nodes: Nodes) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The points-mode node attributes (src/v2mesh.cpp:1094-1110): NodeLabel (int32),
    NodeType (int32), NodePartner (int32, n x 3): the partner labels a node's type
    carries (one for an interface node, two for a junction, three for a corner,
    the third from part3), -1 for none or an unset partner (0xFFFF)."""
    n = nodes.n
    typ = np.asarray(nodes.typ, np.int32)
    part = np.asarray(nodes.part, np.int32).reshape(n, 2)
    p3 = np.asarray(nodes.part3, np.int32) if nodes.part3 is not None else np.full(n, 0xFFFF, np.int32)
    lab = lambda v: np.where(v == 0xFFFF, -1, v)
    partners = np.full((n, 3), -1, np.int32)
    partners[:, 0] = np.where(typ >= 1, lab(part[:, 0]), -1)
    partners[:, 1] = np.where(typ >= 2, lab(part[:, 1]), -1)
    partners[:, 2] = np.where(typ >= 3, lab(p3), -1)
    return np.asarray(nodes.lab, np.int32), typ, partners


def nodes_to_world(P: np.ndarray, voxelsize, affine=None) -> np.ndarray:
    """Grid mm to world coordinates through the volume's 3 x 4 affine (row-major,
    voxel index -> world; identity scaled by nothing if None), as nodes_to_world
    (src/v2m_pipeline.cpp:754-767), in double precision."""
    P = np.asarray(P, np.float64).reshape(-1, 3)
    uvw = P / np.asarray(voxelsize, np.float64)
    if affine is None:
        return uvw * np.asarray(voxelsize, np.float64)
    A = np.asarray(affine, np.float64).reshape(-1)[:12].reshape(3, 4)
    return uvw @ A[:, :3].T + A[:, 3]


def _jdata(a: np.ndarray, dtype: str) -> dict:
    """A JData annotated array, zlib-compressed and base64-encoded (text JMesh)."""
    import base64
    import zlib

    a = np.ascontiguousarray(a)
    return {"_ArrayType_": dtype, "_ArraySize_": list(a.shape), "_ArrayZipType_": "zlib",
            "_ArrayZipSize_": [1, int(a.size)],
            "_ArrayZipData_": base64.b64encode(zlib.compress(a.tobytes(), 1)).decode()}


def write_jmesh_points(path: str, nodes: Nodes, voxelsize=(1.0, 1.0, 1.0), affine=None) -> None:
    """The points-mode JMesh output (src/v2mesh.cpp:1094-1112 through mesh_to_json,
    src/v2m_jmesh.cpp:77-161, text form): MeshNode (world coordinates, float64 n x 3),
    NodeLabel, NodeType, NodePartner (int32), keys in that order."""
    lab, typ, partners = node_attributes(nodes)
    doc = {"MeshNode": _jdata(nodes_to_world(nodes.P, voxelsize, affine), "double"),
           "NodeLabel": _jdata(lab, "int32"), "NodeType": _jdata(typ, "int32"),
           "NodePartner": _jdata(partners, "int32")}
    with open(path, "w") as f:
        json.dump(doc, f)


def read_jmesh_arrays(path: str) -> dict:
    """The annotated arrays of a text JMesh file (zlib or raw _ArrayData_), by key."""
    import base64
    import zlib

    with open(path) as f:
        doc = json.load(f)
    types = {"double": np.float64, "single": np.float32, "int32": np.int32, "uint32": np.uint32, "int16": np.int16,
             "uint16": np.uint16, "int8": np.int8, "uint8": np.uint8, "int64": np.int64}
    out = {}
    for k, v in doc.items():
        if not isinstance(v, dict) or "_ArrayType_" not in v:
            continue
        dt = types[v["_ArrayType_"]]
        if "_ArrayZipData_" in v:
            raw = zlib.decompress(base64.b64decode(v["_ArrayZipData_"]))
            a = np.frombuffer(raw, dt)
        else:
            a = np.asarray(v["_ArrayData_"], dt)
        out[k] = a.reshape(v["_ArraySize_"])
    return out
