"""The device path on the benchmark's other configurations and exact-path switches.

* The 64 uneven volumes of the batched workload (edges 32-160 voxels, 2-4 labels,
  uniform h 2-5 voxels, about 1.5 M nodes), seeded on the device and relaxed
  together by tn_relax_batch for 100 iterations: every volume bit-equal to the
  oracle's relaxation of the same grid and seeds, with its statistics.
* FIRE's power sign: with TN_SERIAL_POWER=1 every FIRE iteration takes the serial
  sum in node order (the CPU's own summation, src/v2m_particles.cpp:849-855); the
  golden FIRE runs must keep the reference's bits, and power_fallbacks must count
  every iteration.
* The device field queries (tn_plan_field_points) at the golden cases' query points:
  phi, psi, its gradient, the owner and the runner-up, bit for bit with the
  reference's values.
"""
import numpy as np
import pytest

from conftest import CASES, Fixture, bits
from trussnet_amd._abi import RelaxParams

pytestmark = pytest.mark.gpu

STAT_KEYS = ["iters", "rebuilds", "last_move", "last_p99", "fire_resets", "fire_dt", "n_interior", "n_interface",
             "n_junction", "n_corner"]


@pytest.fixture(scope="module")
def tn():
    from trussnet_amd import relax

    if relax.device_count() < 1:
        pytest.fail("no HIP device visible")
    return relax


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_batch_workload_vs_oracle(tn, oracle_lib):
    from trussnet_amd import workload

    grids = workload.batch_volumes(64, seed=0)
    seeds = [tn.seed(g, RelaxParams(), device=0) for g in grids]
    ns = [s.n for s in seeds]
    assert len(grids) == 64 and min(ns) > 0
    prm = RelaxParams(max_iters=100, dptol=0.0)
    print(f"[batch] 64 volumes, {sum(ns)} nodes ({min(ns)} .. {max(ns)}), 100 iterations", flush=True)
    out = tn.relax_batch(grids, seeds, [prm] * len(grids), device=0)
    print("[batch] device done; oracle per volume", flush=True)
    for k, (g, s, (got, st)) in enumerate(zip(grids, seeds, out)):
        want, sw = oracle_lib.relax(g, s, prm)
        assert got.equal(want), (k, s.n, int((got.P != want.P).any(axis=1).sum()))
        for key in STAT_KEYS:
            assert np.float32(st[key]) == np.float32(sw[key]), (k, key, st[key], sw[key])


@pytest.mark.parametrize("name", ["spheres64", "sectors48", "tpm32", "ellipsoids56"])
def test_serial_power_path_is_exact(tn, monkeypatch, name):
    monkeypatch.setenv("TN_SERIAL_POWER", "1")
    case = Fixture(name)
    ran = 0
    for prm, want, stats in case.runs():
        if not prm.fire or prm.voxel_trap:
            continue
        got, st = tn.relax(case.grid, case.seeds, prm, device=0)
        assert got.equal(want), (name, prm.max_iters, prm.dptol)
        for k, v in stats.items():
            assert np.float32(st[k]) == np.float32(v), (name, k, st[k], v)
        assert st["power_fallbacks"] == st["iters"] > 0
        ran += 1
    assert ran > 0
    monkeypatch.delenv("TN_SERIAL_POWER")
    prm = next(p for p, _, _ in case.runs() if p.fire and not p.voxel_trap)
    _, st = tn.relax(case.grid, case.seeds, prm, device=0)
    assert st["power_fallbacks"] < st["iters"]   # the tree sum certifies the sign almost always


@pytest.mark.parametrize("name", CASES)
def test_field_queries_match_reference(tn, name):
    case = Fixture(name)
    d = case.d
    a, b = (int(x) for x in d["fq_ab"])
    with tn.Plan(case.grid, case.seeds, RelaxParams(max_iters=1)) as plan:
        phi, psi, grad, lab, sec = plan.field_points(a, b, d["fq_X"])
    assert np.array_equal(bits(phi), bits(d["fq_phi"]))
    assert np.array_equal(bits(psi), bits(d["fq_psi"]))
    assert np.array_equal(bits(grad), bits(d["fq_grad"]))
    assert np.array_equal(lab, d["fq_label"]) and np.array_equal(sec, d["fq_second"])
