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
