"""Write the golden fixtures of the relaxation from the compiled reference.

Run in a container that holds the reference sources (it needs
oracle/_ref/libv2m_ref.so, built by ``make -C oracle ref``):

    python tests/golden/make_golden.py

For each case it builds the grid with the reference's build_grid_cpu
(src/v2m_grid.cpp:47), seeds it with seed_cpu (src/v2m_particles.cpp:625), and
stores, as data only:
  * the grid arrays and the seeds (the inputs);
  * the first rebuild: keys, sorted ids, cell starts, K-nearest lists, h per node
    (src/v2m_particles.cpp:767-820, through the reference's own bodies);
  * the forces of the seeds and one move of every seed from rest
    (v2m_force / v2m_move, src/opencl/v2m_particle_body.cl:161-184, 844-935);
  * field queries at points near the interfaces (phi, psi and its gradient, the
    owning label and the runner-up);
  * relax_cpu's result and stats for several (iterations, dptol, FIRE/Jacobi) runs.
Files: tests/golden/<case>.npz (numpy, compressed, no pickles).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref  # noqa: E402
from trussnet_amd import volumes  # noqa: E402
from trussnet_amd._abi import RelaxParams  # noqa: E402

STAT_KEYS = ["iters", "rebuilds", "last_move", "last_p99", "fire_resets", "fire_dt", "n_interior", "n_interface",
             "n_junction", "n_corner"]


def graded_size(lab: np.ndarray, lo: float, hi: float, slope: float) -> np.ndarray:
    """A user sizing field: lo at the interfaces, growing by `slope` per voxel of distance."""
    from scipy.ndimage import distance_transform_edt

    edge = np.zeros(lab.shape, bool)
    for ax in range(3):
        d = np.diff(lab.astype(np.int32), axis=ax) != 0
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[ax] = slice(0, -1)
        sl1[ax] = slice(1, None)
        edge[tuple(sl0)] |= d
        edge[tuple(sl1)] |= d
    dist = distance_transform_edt(~edge)
    return np.minimum(hi, lo + slope * dist).astype(np.float32)


SHAPES40 = """{"Shapes": [
  {"Grid":     {"Tag": 0, "Size": [40, 40, 40]}},
  {"Sphere":   {"O": [20, 20, 20], "R": 15, "Tag": 1}},
  {"Box":      {"O": [12, 12, 12], "Size": [16, 16, 10], "Tag": 2}},
  {"Cylinder": {"C0": [0, 20, 26], "C1": [40, 20, 26], "R": 5, "Tag": 3}}
]}"""

CASES = {
    # the small plain workload: 64^3, spheres r 24 / 12, uniform 3-voxel size; HCP seeds
    "spheres64": dict(vol=lambda: volumes.two_spheres(64, 24.0, 12.0), grid=dict(hbase=3.0, hmin=3.0, hmax=3.0),
                      runs=[(1, 0.0, True), (5, 0.0, True), (50, 0.0, True), (300, 2e-3, True), (40, 0.0, False)]),
    # the same volume seeded as the small plain benchmark states it: a 3 mm cubic grid
    # jittered by +-0.25 x 3 per axis (SplitMix64, seed 0), labelled and classified
    "jitter64": dict(vol=lambda: volumes.two_spheres(64, 24.0, 12.0), grid=dict(hbase=3.0, hmin=3.0, hmax=3.0),
                     points=lambda: volumes.jittered_points(64, 3.0, 0.25, 0),
                     runs=[(1, 0.0, True), (50, 0.0, True), (300, 2e-3, True), (50, 0.0, False)]),
    # a graded 4-label ball: junction lines, corners, a multi-level hash
    "sectors48": dict(vol=lambda: volumes.sectors(48), grid=dict(hbase=3.0),
                      runs=[(1, 0.0, True), (10, 0.0, True), (100, 0.0, True), (400, 2e-3, True), (40, 0.0, False),
                            (60, 0.0, False, True), (300, 2e-3, False, True)]),
    # nested perturbed ellipsoids with a user sizing field graded off the interfaces
    "ellipsoids56": dict(vol=lambda: volumes.nested_ellipsoids(56, nlab=4, seed=3, axes=(6.0, 24.0), grf_sigma=2.0,
                                                               grf_amp=1.0, centre_jitter=2.0),
                         grid=dict(hbase=4.0, hmin=1.5, hmax=4.0), hvox=(1.5, 4.0, 0.25),
                         runs=[(1, 0.0, True), (30, 0.0, True), (200, 2e-3, True)]),
    # a tissue-probability input: the softmax of six Gaussian random fields (channel 0
    # the exterior), labels by argmax, each label's field its smoothed probability
    "tpm32": dict(vol=lambda: volumes.tpm_softmax((32, 34, 30), 6, 3.0, 1)[1],
                  prob=lambda: volumes.tpm_softmax((32, 34, 30), 6, 3.0, 1)[0], grid=dict(hbase=3.0, hmin=2.0),
                  runs=[(1, 0.0, True), (30, 0.0, True), (200, 2e-3, True)]),
    # a gray-scale input: three thresholds cut nested perturbed shells out of a smooth
    # intensity; the fields are the intensity's memberships, psi of adjacent labels the
    # intensity itself
    "gray40": dict(vol=lambda: volumes.gray_ball(40), thr=volumes.GRAY_THRESHOLDS, grid=dict(hbase=2.5),
                   runs=[(1, 0.0, True), (30, 0.0, True), (200, 2e-3, True), (30, 0.0, False), (40, 0.0, False, True)]),
    # shape constructs: the labels' analytic fields (a compiled program, gm = -1) with a
    # box's edges and corners and a cylinder's rims pinned by the reference's seeding
    "shapes40": dict(shapes=SHAPES40, grid=dict(hbase=3.0),
                     runs=[(1, 0.0, True), (30, 0.0, True), (200, 2e-3, True), (30, 0.0, False)]),
    # anisotropic voxels (0.8 x 0.9 x 1.3 mm): every coordinate -> voxel division is inexact
    "aniso40": dict(vol=lambda: volumes.sectors(40), vs=(0.8, 0.9, 1.3), grid=dict(hbase=2.5),
                    runs=[(1, 0.0, True), (20, 0.0, True), (150, 2e-3, True), (30, 0.0, False), (40, 0.0, False, True)]),
}


def field_samples(G, nodes, rng, m=1500):
    """Points near interface nodes (and a few anywhere), with a label pair per case."""
    idx = np.flatnonzero(nodes.typ != 0)
    pick = rng.choice(idx, size=min(m, idx.size), replace=idx.size < m) if idx.size else np.zeros(0, int)
    X = nodes.P[pick] + rng.normal(scale=0.6, size=(pick.size, 3)).astype(np.float32)
    ext = np.array([G.nx * G.vs[0], G.ny * G.vs[1], G.nz * G.vs[2]], np.float32)
    Y = rng.uniform(-1.0, 1.0, size=(m // 5, 3)).astype(np.float32) * ext * 0.6 + ext * 0.5
    return np.concatenate([X, Y]).astype(np.float32), (int(nodes.lab[pick[0]]) if pick.size else 1,
                                                       int(nodes.part[pick[0], 0]) if pick.size else 0)


def make(name: str, spec: dict) -> str:
    lab = spec["vol"]() if "vol" in spec else None
    hvox = None
    if "hvox" in spec:
        lo, hi, slope = spec["hvox"]
        hvox = graded_size(lab, lo, hi, slope).ravel()
    prob = spec["prob"]() if "prob" in spec else None
    if "shapes" in spec:
        from trussnet_amd._abi import GridParams

        rg = ref.RefGrid.build_shapes(spec["shapes"], 0.0, GridParams(**spec["grid"]))
    else:
        rg = ref.RefGrid.build(lab, vs=spec.get("vs", (1.0, 1.0, 1.0)), hvox=hvox, prob=prob,
                               thresholds=spec.get("thr"), **spec["grid"])
    G = rg.grid()
    prm = RelaxParams()
    hcp = rg.seed(prm)
    # the relaxation's input: seed_cpu's nodes, or the case's own points classified
    seeds = rg.classify_points(spec["points"]()) if "points" in spec else hcp
    out = {}
    for k, v in G.arrays().items():
        out["grid_" + k] = v
    out["seed_P"], out["seed_lab"], out["seed_typ"], out["seed_part"] = seeds.P, seeds.lab, seeds.typ, seeds.part
    out["seed_part3"] = seeds.part3
    if "points" in spec:   # seed_cpu's own output, for the seeding tests
        out["hcp_P"], out["hcp_lab"], out["hcp_typ"], out["hcp_part"] = hcp.P, hcp.lab, hcp.typ, hcp.part
        out["hcp_part3"] = hcp.part3
    # the lattice seeds alone: seed_cpu without junction-line seeds and corners
    out["grid_grade"] = G.grade
    lat = rg.seed(RelaxParams(jseed=0.0, corners=False))
    out["lat_P"], out["lat_lab"], out["lat_typ"], out["lat_part"] = lat.P, lat.lab, lat.typ, lat.part
    # the lattice seeds in voxel mode (each point takes its voxel's label)
    vx = rg.seed(RelaxParams(jseed=0.0, corners=False, voxel_trap=True))
    out["vox_P"], out["vox_lab"], out["vox_typ"], out["vox_part"] = vx.P, vx.lab, vx.typ, vx.part
    # junction-line seeds on a finer key grid, no corners
    js = rg.seed(RelaxParams(jseed=0.4, corners=False))
    out["js_P"], out["js_lab"], out["js_typ"], out["js_part"] = js.P, js.lab, js.typ, js.part
    hi = rg.hash_info(prm)
    out["hash_nlev"] = np.int32(hi.nlev)
    out["hash_nkeys"] = np.int32(hi.nkeys)
    out["hash_b0"] = np.float32(hi.b0)
    out["hash_dim"] = np.array([list(hi.dim[L]) for L in range(8)], np.int32)
    out["hash_off"] = np.array(list(hi.off), np.int32)
    rb = rg.rebuild(seeds, prm)
    nnb = rb["nnb"]
    nbr = rb["nbr"].reshape(-1, 24).copy()
    nbr[np.arange(24)[None, :] >= nnb[:, None]] = -1  # only the first nnb entries are defined
    for k in ("key", "sorted", "cstart", "nnb", "hn"):
        out["rb_" + k] = rb[k]
    out["rb_nbr"] = nbr.ravel()
    F = rg.force(seeds, prm, rb["hn"], rb["nbr"], rb["nnb"])
    out["force_F"] = F
    fdt = np.float32(np.sqrt(np.float32(prm.dt)))
    V0 = np.zeros(seeds.n * 3, np.float32)
    m1, hn1, V1, pw1, mv1 = rg.move(seeds, prm, rb["hn"], F, V0, fdt, np.float32(0.1))
    out["move_fdt"] = fdt
    out["move_P"], out["move_typ"], out["move_part"] = m1.P, m1.typ, m1.part
    out["move_hn"], out["move_V"], out["move_pw"], out["move_mv"] = hn1, V1, pw1, mv1
    rng = np.random.default_rng(7)
    X, (a, b) = field_samples(G, seeds, rng)
    phi, psi, grad, labx, sec = rg.field_points(a, b, X)
    out["fq_X"], out["fq_ab"] = X, np.array([a, b], np.int32)
    out["fq_phi"], out["fq_psi"], out["fq_grad"], out["fq_label"], out["fq_second"] = phi, psi, grad, labx, sec
    runs = []
    for r, run in enumerate(spec["runs"]):
        iters, dptol, fire = run[:3]
        vox = len(run) > 3 and run[3]
        p = RelaxParams(max_iters=iters, dptol=dptol, fire=fire, voxel_trap=vox)
        res, st = rg.relax(seeds, p)
        out[f"run{r}_P"], out[f"run{r}_typ"], out[f"run{r}_part"] = res.P, res.typ, res.part
        out[f"run{r}_stats"] = np.array([st[k] for k in STAT_KEYS], np.float64)
        runs.append((iters, dptol, int(fire), int(vox)))
        print(f"  {name} run {r}: iters {st['iters']} rebuilds {st['rebuilds']} resets {st['fire_resets']} "
              f"types {[st['n_interior'], st['n_interface'], st['n_junction'], st['n_corner']]}")
    out["runs"] = np.array(runs, np.float64)
    out["stat_keys"] = np.array(STAT_KEYS)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {seeds.n} seeds, {G.nslot} slots, hash {hi.nlev} levels / {hi.nkeys} keys -> "
          f"{os.path.getsize(path) / 1e3:.0f} kB")
    return path


def main(argv):
    names = argv or list(CASES)
    for name in names:
        make(name, CASES[name])


if __name__ == "__main__":
    main(sys.argv[1:])
