"""Node thinning and feature protection on the device (tn_thin_nodes,
tn_protect_features) through the C ABI, against the reference's results in
tests/golden/thin.npz (thin_nodes at three factors on six cases' seeds;
protect_features on the shape case's relaxed nodes) and against the oracle on
the large workload's two million device seeds, bit for bit.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, Fixture
from trussnet_amd._abi import Nodes, RelaxParams

pytestmark = pytest.mark.gpu

TH = np.load(os.path.join(GOLDEN, "thin.npz"), allow_pickle=False)
CASES = [str(c) for c in TH["thin_cases"]]
THIN = [float(b) for b in TH["thin_values"]]


@pytest.fixture(scope="module")
def tn():
    from trussnet_amd import relax

    if relax.device_count() < 1:
        pytest.fail("no HIP device visible")
    return relax


def seeds_of(case):
    s = case.seeds
    return Nodes(s.P, s.lab, s.typ, s.part, case.d["seed_part3"])


def selected(s, keep):
    return Nodes(s.P[keep], s.lab[keep], s.typ[keep], s.part[keep], s.part3[keep])


def same(got, want):
    return got.equal(want) and np.array_equal(got.part3, want.part3)


@pytest.mark.parametrize("name", CASES)
def test_thin_matches_reference(tn, name):
    case = Fixture(name)
    s = seeds_of(case)
    for k, B in enumerate(THIN):
        keep = np.unpackbits(TH[f"{name}_thin{k}"])[:s.n].astype(bool)
        got = tn.thin(case.grid, s, B, device=0)
        assert same(got, selected(s, keep)), (name, B, got.n, int(keep.sum()))
    assert same(tn.thin(case.grid, s, 0.0, device=0), s)


def test_protect_matches_reference(tn):
    case = Fixture("shapes40")
    s = seeds_of(case)
    feat = TH["protect_feat"]
    ran = 0
    for r, (prm, res, _) in enumerate(case.runs()):
        if f"protect_run{r}" not in TH:
            continue
        nodes = Nodes(res.P, s.lab, res.typ, res.part, s.part3)
        keep = np.unpackbits(TH[f"protect_run{r}"])[:nodes.n].astype(bool)
        got = tn.protect_features(case.grid, feat, nodes, device=0)
        assert same(got, selected(nodes, keep)), (r, got.n, int(keep.sum()))
        ran += 1
    assert ran > 0
    nodes = Nodes(case.runs()[0][1].P, s.lab, s.typ, s.part, s.part3)
    assert same(tn.protect_features(case.grid, np.zeros(0, np.float32), nodes, device=0), nodes)


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_thin_large_vs_oracle(tn, oracle_lib):
    from oracle import orc
    from trussnet_amd import workload

    grid = workload.plain_large()
    s = tn.seed(grid, RelaxParams(), device=0)
    print(f"[thin] plain_large: {s.n} seeds", flush=True)
    for B in (0.8, 1.2):
        got = tn.thin(grid, s, B, device=0)
        keep = orc.thin_keep(grid, s, B)
        assert same(got, selected(s, keep)), (B, got.n, int(keep.sum()))
        assert 0 < got.n < s.n
    # without third partners: the corners' part3 reads as no label
    s2 = Nodes(s.P, s.lab, s.typ, s.part, None)
    got = tn.thin(grid, s2, 1.2, device=0)
    keep = orc.thin_keep(grid, s2, 1.2)
    assert got.n == int(keep.sum()) and np.array_equal(got.P, s.P[keep])
