"""The HIP relaxation at the benchmark sizes against the oracle, through the C ABI.

The large graded six-label volume (256^3, about 2 M nodes) and the
probability-input volume (182x218x182x6, about 1.4 M nodes) are built here,
seeded on the device (the reference's seed_cpu, defaults: junction-line seeds
and corners included), relaxed for the benchmark's 100 iterations on the device
and by the CPU restatement, and compared bit for bit, with the statistics.
"""
import numpy as np
import pytest

from trussnet_amd._abi import RelaxParams

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

STAT_KEYS = ["iters", "rebuilds", "last_move", "last_p99", "fire_resets", "fire_dt", "n_interior", "n_interface",
             "n_junction", "n_corner"]


@pytest.fixture(scope="module")
def tn():
    from trussnet_amd import relax

    if relax.device_count() < 1:
        pytest.fail("no HIP device visible")
    return relax


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["plain_large", "prob_large"])
def test_full_size_vs_oracle(tn, oracle_lib, name):
    from trussnet_amd import workload

    grid = getattr(workload, name)()
    seeds = tn.seed(grid, RelaxParams(), device=0)
    assert seeds.n > 1_000_000
    prm = RelaxParams(max_iters=100, dptol=0.0)
    print(f"[large] {name}: {seeds.n} nodes, {prm.max_iters} iterations", flush=True)
    got, st = tn.relax(grid, seeds, prm, device=0)
    want, sw = oracle_lib.relax(grid, seeds, prm)
    assert got.equal(want), (name, int((got.P != want.P).any(axis=1).sum()))
    for k in STAT_KEYS:
        assert np.float32(st[k]) == np.float32(sw[k]), (name, k, st[k], sw[k])
    assert st["rebuilds"] >= 10 and st["n_junction"] > 0 and st["n_corner"] >= 0


@pytest.mark.timeout(600)
def test_full_size_voxel_trap_vs_oracle(tn, oracle_lib):
    """The voxel trap on the large volume (about 2 M nodes, a rebuild almost every
    iteration): the nodes snapped onto the voxel faces' lattice have up to hundreds
    of candidates within reach, so the K-nearest search cuts its buffers back many
    times a wave. 30 iterations against the oracle, bit for bit."""
    from trussnet_amd import workload

    grid = workload.plain_large()
    seeds = tn.seed(grid, RelaxParams(voxel_trap=True), device=0)
    assert seeds.n > 1_000_000
    prm = RelaxParams(max_iters=30, dptol=0.0, voxel_trap=True, fire=False)
    print(f"[large] voxel trap: {seeds.n} nodes, {prm.max_iters} iterations", flush=True)
    got, st = tn.relax(grid, seeds, prm, device=0)
    want, sw = oracle_lib.relax(grid, seeds, prm)
    assert got.equal(want), int((got.P != want.P).any(axis=1).sum())
    for k in STAT_KEYS:
        assert np.float32(st[k]) == np.float32(sw[k]), (k, st[k], sw[k])
    assert st["rebuilds"] >= 20
