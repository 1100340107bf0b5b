"""The HIP relaxation against the reference's outputs, through the C ABI.

Each golden case (tests/golden, written by the compiled reference) is run on
the device stage by stage and whole: the hash geometry, keys, the stable sort,
cell starts, K-nearest lists and forces, then full relaxations. Everything is
compared bit for bit: the device path performs the reference's float
operations in the reference's order.
"""
import numpy as np
import pytest

from conftest import CASES, Fixture, bits
from trussnet_amd._abi import K, RelaxParams

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tn():
    from trussnet_amd import relax

    if relax.device_count() < 1:
        pytest.fail("no HIP device visible")
    return relax


@pytest.fixture(scope="module", params=CASES)
def case(request):
    return Fixture(request.param)


def test_hash_and_rebuild(tn, case):
    d = case.d
    with tn.Plan(case.grid, case.seeds, RelaxParams()) as plan:
        hi = plan.hash
        assert hi.nlev == int(d["hash_nlev"]) and hi.nkeys == int(d["hash_nkeys"])
        assert np.float32(hi.b0) == d["hash_b0"]
        assert [list(hi.dim[L]) for L in range(8)] == d["hash_dim"].tolist()
        rb = plan.rebuild()
        for k in ("key", "sorted", "cstart", "nnb"):
            assert np.array_equal(rb[k], d["rb_" + k]), k
        assert np.array_equal(rb["nbr"], d["rb_nbr"])


def test_forces(tn, case):
    with tn.Plan(case.grid, case.seeds, RelaxParams()) as plan:
        F = plan.force()
    assert np.array_equal(bits(F), bits(case.d["force_F"]))


def test_relax_runs(tn, case):
    """Whole relaxations: FIRE and Jacobi, fixed counts and the convergence stop."""
    for prm, want, stats in case.runs():
        got, st = tn.relax(case.grid, case.seeds, prm)
        assert got.equal(want), (case.name, prm.max_iters, prm.dptol, prm.fire,
                                 int((got.P != want.P).any(axis=1).sum()))
        for k, v in stats.items():
            assert np.float32(st[k]) == np.float32(v), (case.name, prm.max_iters, k, st[k], v)


@pytest.mark.parametrize("variant", ["lattice", "junctions", "default", "voxel"])
def test_seeding(tn, case, variant):
    """Device seeding = the reference's seed_cpu: the lattice alone; with junction-line
    seeds on a 0.4-spacing key grid; with the defaults (junction seeds at 0.8, corners);
    the lattice in voxel-trap mode (each point takes its voxel's label). A shape grid's
    seeds end with the pinned feature nodes of its curves (add_feature_nodes,
    src/v2m_particles.cpp:304-620)."""
    from trussnet_amd._abi import Nodes

    d = case.d
    if case.grid.gm < 0:   # shape grids: every variant ends with the feature nodes
        assert case.grid.feat is not None and case.grid.feat.size > 0
    hcp = "hcp_" if "hcp_P" in d else "seed_"   # seed_cpu's own nodes (a case may relax other seeds)
    prm, pre = {"lattice": (RelaxParams(jseed=0.0, corners=False), "lat_"),
                "junctions": (RelaxParams(jseed=0.4, corners=False), "js_"),
                "default": (RelaxParams(), hcp),
                "voxel": (RelaxParams(jseed=0.0, corners=False, voxel_trap=True), "vox_")}[variant]
    got = tn.seed(case.grid, prm, device=0)
    want = Nodes(d[pre + "P"], d[pre + "lab"], d[pre + "typ"], d[pre + "part"])
    assert got.n == want.n, (case.name, variant, got.n, want.n)
    assert got.equal(want), (case.name, variant, int((got.P != want.P).any(axis=1).sum()))
    if variant == "default":
        assert np.array_equal(got.part3, d[pre + "part3"])
    else:   # (no corner seeds; a shape grid's feature nodes may have a fourth label)
        plain = got.typ != 3 if case.grid.gm < 0 else np.ones(got.n, bool)
        assert (got.part3[plain] == 0xFFFF).all()


def test_generated_workload_vs_oracle(tn, oracle_lib):
    """A graded six-label workload built here, seeded on the device, relaxed on both sides."""
    from trussnet_amd import volumes, workload

    lab = volumes.nested_ellipsoids(96, nlab=6, seed=5, axes=(8.0, 45.0), grf_sigma=3.0, grf_amp=2.0)
    grid = workload.build_grid(lab, workload.graded_size(lab, 1.5, 6.0, 0.05), hmin=1.5, hmax=6.0)
    seeds = tn.seed(grid, RelaxParams(jseed=0.0, corners=False), device=0)
    assert seeds.n > 20000
    for prm in (RelaxParams(max_iters=30, dptol=0.0), RelaxParams(max_iters=300)):
        got, st = tn.relax(grid, seeds, prm)
        want, sw = oracle_lib.relax(grid, seeds, prm)
        assert got.equal(want), (prm.max_iters, int((got.P != want.P).any(axis=1).sum()))
        for k in ("iters", "rebuilds", "fire_resets", "n_interior", "n_interface", "n_junction", "n_corner"):
            assert st[k] == sw[k], k
        assert np.float32(st["last_move"]) == np.float32(sw["last_move"])


def test_plan_rerun_is_identical(tn):
    """A plan restores its seeds on every run: repeated runs give the same bits."""
    case = Fixture("sectors48")
    prm, want, _ = case.runs()[1]
    with tn.Plan(case.grid, case.seeds, prm) as plan:
        for _ in range(3):
            plan.run()
            assert plan.download().equal(want)


def test_timing_classes(tn):
    """Timing a subset of the kernel classes records only those, leaves the results
    unchanged, and the class names come back in the order their mask bits take."""
    case = Fixture("sectors48")
    prm, want, _ = case.runs()[1]
    with tn.Plan(case.grid, case.seeds, prm) as plan:
        plan.set_timing(("neighbors", "move_rest"))
        plan.run()
        assert plan.download().equal(want)
        t = plan.timings()
        assert tuple(t) == tn.Plan.TIMING_CLASSES
        assert t["neighbors"][0] > 0 and t["neighbors"][1] > 0.0
        for name in ("force", "move_easy", "stats", "keys"):
            assert t[name][0] == 0, name
        plan.set_timing(True)
        plan.run()
        assert plan.download().equal(want)
        t = plan.timings()
        assert t["force"][0] == t["stats"][0] > 0
        plan.set_timing(False)
        plan.run()
        assert all(v[0] == 0 for v in plan.timings().values())


def test_steps_compose(tn):
    """Iterations run one call at a time reach the same state as one whole run."""
    case = Fixture("spheres64")
    prm = RelaxParams(max_iters=12, dptol=0.0)
    with tn.Plan(case.grid, case.seeds, prm) as plan:
        plan.run()
        whole = plan.download()
        plan.reset()
        for _ in range(12):
            plan.step(1)
        assert plan.download().equal(whole)


@pytest.mark.parametrize("which", ["all", "labels"])
def test_batch_equals_the_reference_runs(tn, which):
    """tn_relax_batch: every golden case's runs relaxed together, each to the reference's
    bits -- all cases (gray-scale and shape grids among them: the general move kernels),
    and the label volumes alone (the move kernels compiled for label volumes)."""
    grids, seeds, prms, wants = [], [], [], []
    for name in CASES:
        fx = Fixture(name)
        if which == "labels" and fx.grid.gm != 0:
            continue
        for prm, want, stats in fx.runs():
            grids.append(fx.grid)
            seeds.append(fx.seeds)
            prms.append(prm)
            wants.append((name, want, stats))
    out = tn.relax_batch(grids, seeds, prms, device=0)
    assert len(out) == len(wants)
    for (got, st), (name, want, stats), prm in zip(out, wants, prms):
        assert got.equal(want), (name, prm.max_iters, prm.dptol, prm.fire)
        for k, v in stats.items():
            assert np.float32(st[k]) == np.float32(v), (name, prm.max_iters, k, st[k], v)


def test_shape_grid_seed_then_relax(tn):
    """A shape input from its grid alone on the device: seeding (its feature nodes
    included) then the relaxation, equal to the reference's seed_cpu + relax_cpu."""
    case = Fixture("shapes40")
    seeds = tn.seed(case.grid, RelaxParams(), device=0)
    prm, want, _ = case.runs()[0]
    got, _ = tn.relax(case.grid, seeds, prm)
    assert got.equal(want)
