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
