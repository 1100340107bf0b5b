"""tn_relax_noalloc (the relaxation with every device buffer in the library's static
scratch array) against tn_relax, through the C ABI.

Two calls in a row, the second with fewer nodes (the array keeps the first call's
contents), each bit-equal to tn_relax on the same grid, seeds and parameters, with
the same statistics; a relaxation larger than the array is refused with TN_ERR_ARG
and leaves the nodes untouched, and the entry point works again afterwards.
"""
import numpy as np
import pytest

from conftest import Fixture
from trussnet_amd._abi import Nodes, RelaxParams

pytestmark = pytest.mark.gpu

STAT_KEYS = ["iters", "rebuilds", "last_move", "last_p99", "fire_resets", "fire_dt", "n_interior", "n_interface",
             "n_junction", "n_corner", "power_fallbacks"]


@pytest.fixture(scope="module")
def tn():
    from trussnet_amd import relax

    if relax.device_count() < 1:
        pytest.fail("no HIP device visible")
    return relax


def _same(tn, case, prm):
    got, st = tn.relax_noalloc(case.grid, case.seeds, prm, device=0)
    want, sw = tn.relax(case.grid, case.seeds, prm, device=0)
    assert got.equal(want), (case.name, int((got.P != want.P).any(axis=1).sum()))
    for k in STAT_KEYS:
        assert np.float32(st[k]) == np.float32(sw[k]), (case.name, k, st[k], sw[k])
    return got


def test_noalloc_equals_tn_relax_twice(tn):
    big, small = Fixture("spheres64"), Fixture("jitter64")
    assert small.seeds.n < big.seeds.n
    prm = RelaxParams(max_iters=50, dptol=0.0)
    print(f"[noalloc] spheres64 ({big.seeds.n} nodes) then jitter64 ({small.seeds.n} nodes), 50 iterations",
          flush=True)
    got = _same(tn, big, prm)
    ref_prm, ref_nodes, _ = big.runs()[2]
    _same(tn, small, prm)
    # the golden run of the reference through the no-allocation path
    out, _ = tn.relax_noalloc(big.grid, big.seeds, ref_prm, device=0)
    assert out.equal(ref_nodes)
    assert got.n == big.seeds.n


def test_noalloc_other_modes(tn):
    for name in ("sectors48", "gray40"):
        case = Fixture(name)
        for prm, want, _ in case.runs()[:3]:
            out, _ = tn.relax_noalloc(case.grid, case.seeds, prm, device=0)
            assert out.equal(want), (name, prm.max_iters, prm.fire, prm.voxel_trap)


def test_noalloc_refuses_what_does_not_fit(tn):
    case = Fixture("spheres64")
    reps = 80   # about 240 k nodes: several hundred MB of node state, beyond the array
    s = case.seeds
    big = Nodes(np.tile(s.P, (reps, 1)), np.tile(s.lab, reps), np.tile(s.typ, reps), np.tile(s.part, (reps, 1)))
    before = big.copy()
    with pytest.raises(tn.DeviceError) as e:
        tn.relax_noalloc(case.grid, big, RelaxParams(max_iters=5), device=0)
    assert e.value.code == 1 and "scratch" in str(e.value)
    assert big.equal(before)
    _same(tn, Fixture("jitter64"), RelaxParams(max_iters=20, dptol=0.0))
