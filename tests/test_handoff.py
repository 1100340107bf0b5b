"""The node hand-off file (the reference's --dump-nodes format, dump_nodes in
src/v2m_pipeline.cpp:58-73): the header its compact JSON dump writes (sorted keys,
no spaces; offsets P 0, Label 12n, Type 14n, Partner 15n) and a byte-exact round
trip of a golden case's relaxed nodes."""
import os

import numpy as np

from conftest import Fixture
from trussnet_amd import handoff


def test_header_matches_the_reference_layout():
    n = 7
    want = ('{"Label":["uint16",[7],84],"P":["float32",[7,3],0],"Partner":["uint16",[7,2],105],'
            '"Type":["uint8",[7],98]}\n').encode()
    assert handoff.header(n) == want


def test_round_trip(tmp_path):
    fx = Fixture("sectors48")
    prm, res, _ = fx.runs()[2]
    path = os.path.join(tmp_path, "nodes.bin")
    handoff.dump_nodes(path, res)
    raw = open(path, "rb").read()
    head = handoff.header(res.n)
    assert raw.startswith(head) and len(raw) == len(head) + res.n * (12 + 2 + 1 + 4)
    back = handoff.load_nodes(path)
    assert back.equal(res)
    body = raw[len(head):]
    n = res.n
    assert body[:12 * n] == res.P.astype("<f4").tobytes()
    assert body[15 * n:] == res.part.astype("<u2").tobytes()
    assert np.array_equal(np.frombuffer(body[14 * n:15 * n], np.uint8), res.typ)


def test_points_attributes_follow_the_types():
    """NodePartner carries one partner per interface, two per junction, three per
    corner (the third from part3), -1 elsewhere and for unset partners
    (src/v2mesh.cpp:1094-1110); NodeLabel and NodeType are the labels and types."""
    from trussnet_amd._abi import Nodes

    P = np.zeros((5, 3), np.float32)
    lab = np.array([1, 2, 3, 4, 5], np.uint16)
    typ = np.array([0, 1, 2, 3, 1], np.uint8)
    part = np.array([[7, 8], [9, 10], [11, 12], [13, 14], [0xFFFF, 3]], np.uint16)
    part3 = np.array([0xFFFF, 0xFFFF, 0xFFFF, 15, 0xFFFF], np.uint16)
    L, T, Pt = handoff.node_attributes(Nodes(P, lab, typ, part, part3))
    assert L.tolist() == [1, 2, 3, 4, 5] and T.tolist() == [0, 1, 2, 3, 1]
    assert Pt.tolist() == [[-1, -1, -1], [9, -1, -1], [11, 12, -1], [13, 14, 15], [-1, -1, -1]]


def test_points_jmesh_round_trip(tmp_path):
    """The points-mode JMesh file: MeshNode in world coordinates (through the
    affine), NodeLabel / NodeType / NodePartner, zlib + base64 JData arrays."""
    fx = Fixture("tpm32")
    prm, res, _ = fx.runs()[1]
    path = os.path.join(tmp_path, "nodes.jmsh")
    vs = (0.8, 0.9, 1.1)
    affine = [[0.8, 0.0, 0.0, -10.0], [0.0, 0.0, 1.1, 5.0], [0.0, 0.9, 0.0, 2.5]]
    handoff.write_jmesh_points(path, res, vs, affine)
    got = handoff.read_jmesh_arrays(path)
    assert list(got) == ["MeshNode", "NodeLabel", "NodeType", "NodePartner"]
    P = res.P.astype(np.float64)
    u = P / np.array(vs)
    want = np.stack([0.8 * u[:, 0] - 10.0, 1.1 * u[:, 2] + 5.0, 0.9 * u[:, 1] + 2.5], axis=1)
    assert np.allclose(got["MeshNode"], want, rtol=0, atol=1e-12)
    L, T, Pt = handoff.node_attributes(res)
    assert np.array_equal(got["NodeLabel"], L) and np.array_equal(got["NodeType"], T)
    assert np.array_equal(got["NodePartner"], Pt) and got["NodePartner"].shape == (res.n, 3)
    assert np.all(Pt[res.typ == 0] == -1) and np.all(Pt[res.typ == 1][:, 1:] == -1)
