"""Write the golden fixtures of node thinning and feature protection from the
compiled reference (oracle/_ref/libv2m_ref.so; make -C oracle ref):

    python tests/golden/make_thin.py

thin_nodes (src/v2m_particles.cpp:1223-1357) runs on each golden case's seeds
(seed_cpu's output, part3 included) at thin = 0.5, 0.8 and 1.1; protect_features
(src/v2m_particles.cpp:1002-1221) runs on the shape case's relaxed nodes (the
golden relaxations of shapes40) with that grid's feature curves. Both compact the
kept nodes in order, so each result is stored as the keep flags of the input
nodes (bit-packed), checked here to reproduce the reference's output arrays bit
for bit. File: tests/golden/thin.npz (numpy, compressed, no pickles).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import SHAPES40  # noqa: E402
from oracle import ref  # noqa: E402
from trussnet_amd._abi import GridParams, Grid, Nodes  # noqa: E402

CASES = ["spheres64", "sectors48", "tpm32", "ellipsoids56", "gray40", "shapes40"]
THIN = [0.5, 0.8, 1.1]


def load(name):
    d = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    grid = Grid.from_arrays(d, "grid_")
    seeds = Nodes(d["seed_P"], d["seed_lab"], d["seed_typ"], d["seed_part"], d["seed_part3"])
    return d, grid, seeds


def keep_of(inp: Nodes, out: Nodes) -> np.ndarray:
    """The keep flags that select `out` from `inp` (a subsequence, in order)."""
    keep = np.zeros(inp.n, bool)
    k = 0
    for i in range(inp.n):
        if k < out.n and (inp.P[i].view(np.uint32) == out.P[k].view(np.uint32)).all() and inp.lab[i] == out.lab[k] \
                and inp.typ[i] == out.typ[k] and (inp.part[i] == out.part[k]).all() and inp.part3[i] == out.part3[k]:
            keep[i] = True
            k += 1
    assert k == out.n, "the output is not a subsequence of the input"
    sel = Nodes(inp.P[keep], inp.lab[keep], inp.typ[keep], inp.part[keep], inp.part3[keep])
    assert sel.equal(out) and np.array_equal(sel.part3, out.part3)
    return keep


def main():
    out = {"thin_cases": np.array(CASES), "thin_values": np.array(THIN, np.float32)}
    for name in CASES:
        d, grid, seeds = load(name)
        rg = ref.RefGrid.from_grid(grid)
        for k, B in enumerate(THIN):
            res = rg.thin(seeds, B)
            keep = keep_of(seeds, res)
            out[f"{name}_thin{k}"] = np.packbits(keep)
            print(f"{name}: thin {B}: {seeds.n} -> {res.n}", flush=True)
    # protect_features on the shape case's relaxed nodes
    d, grid, seeds = load("shapes40")
    rg = ref.RefGrid.build_shapes(SHAPES40, 0.0, GridParams(hbase=3.0))
    g2 = rg.grid()
    assert np.array_equal(g2.h, grid.h) and np.array_equal(g2.phi, grid.phi), "the shape grid differs from the fixture"
    feat = rg.feat()
    out["protect_feat"] = feat
    r = 0
    while f"run{r}_P" in d:
        nodes = Nodes(d[f"run{r}_P"], seeds.lab, d[f"run{r}_typ"], d[f"run{r}_part"], seeds.part3)
        res = rg.protect_features(nodes)
        out[f"protect_run{r}"] = np.packbits(keep_of(nodes, res))
        print(f"shapes40 run {r}: protect_features {nodes.n} -> {res.n}", flush=True)
        r += 1
    np.savez_compressed(os.path.join(HERE, "thin.npz"), **out)


if __name__ == "__main__":
    main()
