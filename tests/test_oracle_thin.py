"""The CPU restatement of node thinning and feature protection (oracle/thin.c).

Pinned by the reference itself: tests/golden/thin.npz holds the nodes that the
reference's thin_nodes (src/v2m_particles.cpp:1223-1357) keeps from six golden
cases' seeds at three thin factors, and what its protect_features
(src/v2m_particles.cpp:1002-1221) keeps from the shape case's relaxed nodes. A
brute-force check of the greedy rule (no hash: every pair tested) pins the
restatement's search on two cases, and, where the reference library is built,
both are compared on seeds the fixtures do not hold.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, Fixture
from oracle import orc
from trussnet_amd._abi import Nodes

TH = np.load(os.path.join(GOLDEN, "thin.npz"), allow_pickle=False)
CASES = [str(c) for c in TH["thin_cases"]]
THIN = [float(b) for b in TH["thin_values"]]


def seeds_of(case):
    s = case.seeds
    return Nodes(s.P, s.lab, s.typ, s.part, case.d["seed_part3"])


def mask(key, n):
    return np.unpackbits(TH[key])[:n].astype(bool)


@pytest.mark.parametrize("name", CASES)
def test_thin_matches_reference_fixtures(oracle_lib, name):
    case = Fixture(name)
    s = seeds_of(case)
    for k, B in enumerate(THIN):
        keep = orc.thin_keep(case.grid, s, B)
        assert np.array_equal(keep, mask(f"{name}_thin{k}", s.n)), (name, B)


def test_thin_off_keeps_everything(oracle_lib):
    case = Fixture("spheres64")
    assert orc.thin_keep(case.grid, seeds_of(case), 0.0).all()


def test_protect_matches_reference_fixtures(oracle_lib):
    case = Fixture("shapes40")
    s = seeds_of(case)
    feat = TH["protect_feat"]
    for r, (prm, res, _) in enumerate(case.runs()):
        if f"protect_run{r}" not in TH:
            continue
        nodes = Nodes(res.P, s.lab, res.typ, res.part, s.part3)
        keep = orc.protect_keep(case.grid, feat, nodes)
        want = mask(f"protect_run{r}", nodes.n)
        assert np.array_equal(keep, want), r
        assert (~want).sum() > 0   # the fixture exercises the drop
        assert keep[nodes.typ == 3].all()   # pinned nodes stay


def _has(s, j, l):
    t = s.typ[j]
    return (s.lab[j] == l) | ((t != 0) & (s.part[j, 0] == l)) | ((t >= 2) & (s.part[j, 1] == l)) | \
        ((t == 3) & (s.part3[j] == l))


def brute_thin(grid, s, B):
    """thin_nodes's greedy rule with every pair tested (no hash)."""
    n = s.n
    vs = np.asarray(grid.vs, np.float32)
    idx = np.clip(np.floor(s.P / vs + np.float32(0.5)).astype(np.int64), 0, np.array([grid.nx, grid.ny, grid.nz]) - 1)
    h = grid.h.reshape(grid.nz, grid.ny, grid.nx)[idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float32)
    c0 = np.float32(max(np.float32(1e-6), np.float32(B) * np.float32(grid.hmin)))
    nlev = 1
    while nlev < 12 and c0 * np.float32(1 << (nlev - 1)) < np.float32(B) * np.float32(grid.hmax):
        nlev += 1
    kept = s.typ >= 2
    cand = [i for i in range(n) if s.typ[i] in (0, 1)]
    cand.sort(key=lambda i: (s.typ[i] == 0, h[i], i))
    j_all = np.arange(n)
    for i in cand:
        r = np.float32(B) * h[i]
        r2 = r * r
        L = 0
        while L + 1 < nlev and c0 * np.float32(1 << L) < r:
            L += 1
        c = c0 * np.float32(1 << L)
        cell = np.floor(s.P / c).astype(np.int64)
        near = (np.abs(cell - cell[i]) <= 1).all(axis=1)
        e = s.P - s.P[i]
        d2 = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
        ok = kept & near & (d2 < r2) & (j_all != i)
        if s.typ[i] == 1:
            a, b = s.lab[i], s.part[i, 0]
            ok &= (s.typ != 0) & _has(s, j_all, a) & _has(s, j_all, b)
        else:
            ok &= _has(s, j_all, s.lab[i])
        kept[i] = not ok.any()
    return kept


@pytest.mark.parametrize("name", ["sectors48", "tpm32"])
def test_thin_fixture_obeys_the_greedy_rule(name):
    case = Fixture(name)
    s = seeds_of(case)
    k = 1   # thin 0.8
    assert np.array_equal(brute_thin(case.grid, s, THIN[k]), mask(f"{name}_thin{k}", s.n))


def test_thin_against_reference_on_other_seeds(oracle_lib):
    from oracle import ref

    if not ref.available():
        pytest.skip("reference library not built (it is built where the reference sources are)")
    from trussnet_amd import workload
    from trussnet_amd._abi import RelaxParams

    grid = workload.plain_small()
    rg = ref.RefGrid.from_grid(grid)
    for jseed in (0.8, 0.0):
        s = rg.seed(RelaxParams(jseed=jseed))
        for B in (0.3, 0.9, 1.4):
            want = rg.thin(s, B)
            keep = orc.thin_keep(grid, s, B)
            got = Nodes(s.P[keep], s.lab[keep], s.typ[keep], s.part[keep], s.part3[keep])
            assert got.equal(want) and np.array_equal(got.part3, want.part3), (jseed, B)
