#!/usr/bin/env python3
"""Throughput of the truss relaxation on one MI355X: node-iterations per second.

A step is one whole relaxation of the large plain workload on the device: the
seeds (already resident in HBM) are restored, the hash and K-nearest lists are
built, and 100 iterations run (max_iters = 100, dptol = 0, so every iteration
runs), rebuilds included. The workload: a 256^3 six-label volume of nested,
randomly oriented ellipsoids (semi-axes U(20, 120) voxels, seed 0) with
boundaries perturbed by a smoothed Gaussian random field (sigma 4, amplitude 3
voxels), sizes graded from 1.5 to 6 voxels by distance to the nearest interface,
seeded by the device seeding (graded HCP lattice + classification).

value = nodes x iterations x steps / wall time of the timed steps. Each step
ends with a device synchronisation (the plan's stream), so the host clock
brackets exactly the device work.

roofline: the kernel with the largest device time (the timed classes that launch
the same kernel merged: the general move pass runs twice an iteration). The warmup
times every class with HIP events on the stream of each launch and picks it; the
timed steps record events around its launches only. Its algorithmic bytes per
launch (the per-node figures below, times the nodes the launches processed) over
its average launch time in the timed steps, against 8 TB/s. traffic: its HBM bytes
per launch from two rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE) over a short run
of this script made before the timed run, each scaled by what the counter reads for
the plain device copy of the node positions (16 bytes per node) in the same pass.
rooflines_next: the next three kernels by device time, in the same terms, from the
warmup's events.
hbm_frac_loop: the whole loop's algorithmic bytes, iterations x n x (83 + 4 K)
+ rebuilds x n x (179 + 4 K), over the step time, against 8 TB/s.
cpu_baseline: the reference's own relax_cpu (oracle/_ref, compiled from its
sources) on this host's cores, on the same grid and seeds, for a bounded
number of iterations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload large|small|prob|voxel|gray|batch|adjoint|grid|pipeline]
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "node-iterations/sec at 2M nodes (plain vs batched/adjoint), % of HBM peak"
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per item (node, or cell for the scan) of each kernel class
# (K = mean list length). Neighbour and field gathers count as reuse, not traffic.
def item_bytes(kbar: float, radix_passes: int) -> dict:
    return {
        "force": 37 + 4 * kbar,       # P,h 16 + typ 1 + nnb 4 + nbr 4K read; F 16 written
        "move": 104,                  # queue 4, F 16, P,h 16, V 16, typ 1, part 4, lab 2 read; P,h 16, V 16, typ 1,
                                      # part 4, mv 4, pw 4 written
        "move_int": 104,
        "move_rest": 104,
        "move_def": 104,
        "stats": 36,                  # mv 4, pw 4, P,h 16, P0 12 read
        "refresh": 47 + 4 * kbar,     # list 4, sorted 4, P,h 16, typ 1, lab 2 read; nbr 4K, nnb 4, P0 16 written
        "neighbors": 23 + 4 * kbar,   # P,h 16, typ 1, lab 2 read; nbr 4K + nnb 4 written
        "keys": 24,                   # P 12 + h 4 read; key 4 written; count 4
        "scan": 8,                    # per cell: read + write
        "radix_sort": 20 * radix_passes,
        "permute": 134,               # sorted 4, X 16, V 16, typ 1, part 4, lab 2, id 4 read; the same 47 written,
                                      # Ps 16, P0 16, Ms 8, slot_of 4
    }


def launch_bytes(name: str, timings: dict, per_item: dict, queues: dict | None = None) -> float | None:
    """Algorithmic bytes of one average launch of a kernel class (queues: the timings the
    queue lengths come from, when the queue passes were not timed themselves)."""
    launches, _, items = timings[name]
    if not launches:
        return None
    if name == "move_easy":   # every node: typ 1; finished nodes 94 more; queued ones 4 (the queue)
        q = queues or timings
        queued = sum(q[k][2] / max(q[k][0], 1) for k in ("move", "move_int", "move_rest") if k in q)
        n = items / launches
        return 1 * n + 94 * (n - queued) + 4 * queued
    if name not in per_item:
        return None
    return per_item[name] * items / launches


# the kernels of each timed class (their function names in a rocprofv3 trace)
CLASS_KERNELS = {
    "force": ["k_force"], "move_easy": ["k_move_easy"], "move": ["k_move<1>"], "move_int": ["k_move<2>"],
    "move_rest": ["k_move<0>"], "move_def": ["k_move<0>"],
    "stats": ["k_stats"],
    "neighbors": ["k_truss"], "refresh": ["k_refresh_list", "k_truss_node"], "permute": ["k_permute"],
    "keys": ["k_fill_i32", "k_keys"],
    "control": ["k_control"],
}


def by_kernel(timings: dict, per_item: dict, queues: dict | None = None) -> dict:
    """The timed classes merged by the kernel they launch: {kernel: {classes, launches, ms,
    bytes, items}} (classes whose bytes per item are not known are left out)."""
    kernels: dict = {}
    for cls, (cnt, ms, items) in timings.items():
        b = launch_bytes(cls, timings, per_item, queues) if cnt else None
        if b is None:
            continue
        k = "+".join(CLASS_KERNELS.get(cls, [f"k_{cls}"]))
        a = kernels.setdefault(k, {"classes": [], "launches": 0, "ms": 0.0, "bytes": 0.0, "items": 0.0})
        a["classes"].append(cls)
        a["launches"] += cnt
        a["ms"] += ms
        a["bytes"] += b * cnt
        a["items"] += items
    return kernels


def kernel_key(full: str) -> str:
    """A trace's kernel name as CLASS_KERNELS writes it: no namespace, no arguments, and
    the label-volume flag of the move kernels' template dropped ("void
    tnk::k_move<0, true>(...)" -> "k_move<0>", "k_move_easy<true>" -> "k_move_easy")."""
    name = full.split("(")[0].split("::")[-1]
    for flag in ("true", "false"):
        if name.endswith(f", {flag}>"):
            name = name[: -len(f", {flag}>")] + ">"
        elif name.endswith(f"<{flag}>"):
            name = name[: -len(f"<{flag}>")]
    return name


def pmc_traffic(args) -> dict | None:
    """HBM bytes per launch of each kernel from two rocprofv3 PMC passes (FETCH_SIZE, then
    WRITE_SIZE) over a short run of this script, run before this process touches the GPU.
    Returns {kernel: (mean FETCH_SIZE KiB, mean WRITE_SIZE KiB)} and the same counters for
    the largest plain device copy of the run (the positions restored at a reset, 16 bytes
    per node), which calibrates them."""
    exe = shutil.which("rocprofv3")
    if not exe:
        log("rocprofv3 not found: no traffic measurement")
        return None
    out = tempfile.mkdtemp(prefix="tn_pmc_", dir="/tmp")
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-cpu", "--no-traffic",
             "--no-extras", "--iters", "10", "--workload", args.workload]
    vals: dict = {}
    copy = {}
    for k, counter in enumerate(("FETCH_SIZE", "WRITE_SIZE")):
        t = time.perf_counter()
        cmd = [exe, "--pmc", counter, "-f", "csv", "-d", out, "-o", counter.lower(), "--"] + child
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), capture_output=True, text=True,
                               timeout=300)
        except subprocess.TimeoutExpired:
            log(f"rocprofv3 --pmc {counter}: timed out")
            return None
        files = glob.glob(os.path.join(out, "**", f"{counter.lower()}_counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            log(f"rocprofv3 --pmc {counter}: rc {r.returncode}\n{r.stderr[-1500:]}")
            return None
        per: dict = {}
        for row in csv.DictReader(open(files[0])):
            name = kernel_key(row["Kernel_Name"])
            per.setdefault(name, []).append(float(row["Counter_Value"]))
        for name, v in per.items():
            vals.setdefault(name, [0.0, 0.0])[k] = sum(v) / len(v)
        copy[counter] = max(per.get("__amd_rocclr_copyBuffer", [0.0]))
        log(f"pmc {counter}: {len(per)} kernels ({time.perf_counter() - t:.0f} s)")
    shutil.rmtree(out, ignore_errors=True)
    return {"kernels": vals, "copy_kib": (copy["FETCH_SIZE"], copy["WRITE_SIZE"])}


def build_workload(name: str):
    from trussnet_amd import relax, workload
    from trussnet_amd._abi import RelaxParams

    t = time.perf_counter()
    if name == "large":
        grid = workload.plain_large()
        desc = ("plain_large: 256^3 six-label nested perturbed ellipsoids, h graded 1.5-6 voxels, "
                "device seeding (seed_cpu defaults), 100 iterations")
    elif name == "prob":
        grid = workload.prob_large()
        desc = ("prob_large: 182x218x182x6 softmax of six Gaussian random fields (sigma 6, seed 1), labels by "
                "argmax, fields = smoothed probabilities, uniform h 1.8, device seeding, 100 iterations")
    elif name == "voxel":
        grid = workload.plain_large()
        desc = ("plain_large in voxel-trap mode (the staircase move, voxel-labelled seeds, FIRE off): 256^3 "
                "six-label nested perturbed ellipsoids, h graded 1.5-6 voxels, 100 iterations")
    elif name == "gray":
        grid = workload.gray_large()
        desc = ("gray_large: 256^3 gray-scale ball (1 - r/115 + 0.06 x a Gaussian random field, sigma 3), "
                "thresholds 0.2/0.5/0.75, grid built on the device (hbase 4.5, hmin 1.5, hmax 6 voxels), "
                "device seeding, 100 iterations")
    else:
        grid = workload.plain_small()
        desc = "plain_small: 64^3 two spheres r 24/12, uniform h 3, device seeding, 50 iterations"
    log(f"grid {grid.nx}x{grid.ny}x{grid.nz}: {grid.nslot} slots ({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    seeds = relax.seed(grid, RelaxParams(voxel_trap=name == "voxel"), device=0)
    log(f"seeds: {seeds.n} ({time.perf_counter() - t:.1f} s)")
    return grid, seeds, desc


def cpu_baseline(grid, seeds, iters: int, voxel: bool = False) -> dict | None:
    from trussnet_amd._abi import RelaxParams

    prm = RelaxParams(max_iters=iters, dptol=0.0, voxel_trap=voxel)
    try:
        from oracle import ref

        kind = "reference"
        if not ref.available():
            raise RuntimeError("reference library not built")
        rg = ref.RefGrid.from_grid(grid)
        run = lambda: rg.relax(seeds, prm)
    except Exception as e:  # the restatement, if the reference library is absent
        log(f"reference CPU path unavailable ({e}); timing the CPU restatement")
        from oracle import orc

        kind = "port"
        run = lambda: orc.relax(grid, seeds, prm)
    t = time.perf_counter()
    _, st = run()
    dt = time.perf_counter() - t
    cores = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 64))
    return {"value": seeds.n * st["iters"] / dt, "unit": "node-iterations/s", "cores": cores, "kind": kind,
            "sample": f"relax_cpu on the same grid and {seeds.n} seeds, {st['iters']} iterations "
                      f"({st['rebuilds']} rebuilds), {dt:.1f} s"}


def bench_batch(args, emit: bool = True) -> dict:
    """Many small volumes: tn_relax_batch (all volumes' plans concurrent on their own
    streams) against a loop of tn_relax calls, one volume after another. A step is one
    relaxation of every volume, uploads and downloads included."""
    from trussnet_amd import relax, workload
    from trussnet_amd._abi import RelaxParams

    t = time.perf_counter()
    grids = workload.batch_volumes(64, seed=0)
    log(f"64 grids ({time.perf_counter() - t:.1f} s)")
    seeds = [relax.seed(g, RelaxParams(), device=0) for g in grids]
    ns = [sd.n for sd in seeds]
    total = sum(ns)
    log(f"seeds: {total} in 64 volumes ({min(ns)} .. {max(ns)})")
    iters = args.iters or 100
    prm = RelaxParams(max_iters=iters, dptol=0.0)
    for _ in range(args.warmup):
        relax.relax_batch(grids, seeds, prm, device=0)
    t = time.perf_counter()
    for _ in range(args.steps):
        out = relax.relax_batch(grids, seeds, prm, device=0)
    dt = time.perf_counter() - t
    t = time.perf_counter()
    loop = [relax.relax(g, sd, prm, device=0) for g, sd in zip(grids, seeds)]
    dt_loop = time.perf_counter() - t
    same = all(a[0].equal(b[0]) for a, b in zip(out, loop))
    value = total * iters * args.steps / dt
    log(f"batch {value:.4g} node-iterations/s; loop of tn_relax {total * iters / dt_loop:.4g}; identical {same}")
    cpu = None
    if not args.no_cpu:   # the reference on the first four volumes
        cpu = cpu_baseline_many(grids[:4], seeds[:4], args.cpu_iters or iters)
    line = {
        "metric": "node-iterations/sec, 64 volumes relaxed together (batched)", "value": value,
        "unit": "node-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch_volumes(64, seed 0): edges U(32,160), 2-4 nested ellipsoid labels, "
                               "uniform h U(2,5) voxels, device seeding (seed_cpu defaults), 100 iterations, uploads included",
                   "nodes": total, "volumes": 64, "iterations": iters, "parallelism": "single GPU"},
        "loop_value": total * iters / dt_loop, "identical_to_loop": same, "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def bench_adjoint(args, emit: bool = True) -> dict:
    """The differentiable relaxation: forward (iters Jacobi steps of the smoothed
    truss, then the energy + quality loss) and the hand-written adjoint (gradients
    with respect to the start positions and a 32^3 sizing grid, checkpoints every 8
    steps), on the 128^3 three-label volume. A step is one forward and one backward."""
    import numpy as np
    import torch

    from oracle import orc
    from trussnet_amd import adjoint, relax, workload
    from trussnet_amd._abi import RelaxParams

    t = time.perf_counter()
    grid = workload.adjoint_volume()
    seeds = relax.seed(grid, RelaxParams(), device=0)
    iters = args.iters or 64
    prm = adjoint.AdjParams(iters=iters, ckpt_every=8)
    pb, x0, theta = adjoint.build_problem(grid, seeds, prm, grid_pts=32, device=0)
    n = seeds.n
    log(f"adjoint problem: {n} nodes, {pb.tets.shape[0]} tets, theta {theta.shape} ({time.perf_counter() - t:.1f} s)")
    xd = torch.from_numpy(x0).cuda()
    td = torch.from_numpy(theta).cuda()
    gx = torch.empty_like(xd)
    gt = torch.empty_like(td)
    for _ in range(args.warmup):
        pb.forward_ptr(xd.data_ptr(), td.data_ptr())
        pb.backward_ptr(td.data_ptr(), gx.data_ptr(), gt.data_ptr())
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        loss = pb.forward_ptr(xd.data_ptr(), td.data_ptr())
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t) / args.steps
    t = time.perf_counter()
    for _ in range(args.steps):
        pb.backward_ptr(td.data_ptr(), gx.data_ptr(), gt.data_ptr())
    torch.cuda.synchronize()
    t_bwd = (time.perf_counter() - t) / args.steps
    log(f"forward {1e3 * t_fwd:.2f} ms, backward {1e3 * t_bwd:.2f} ms (x{t_bwd / t_fwd:.2f}), loss {loss:.6f}")
    # the same with every state kept (checkpoint every step: nothing recomputed)
    prm1 = adjoint.AdjParams(iters=iters, ckpt_every=1)
    pb1 = adjoint.Problem(pb.nbr, pb.nnb, pb.typ, pb.tets, theta.shape, pb.origin, pb.spacing, prm1, device=0)
    gx1, gt1 = torch.empty_like(xd), torch.empty_like(td)
    pb1.forward_ptr(xd.data_ptr(), td.data_ptr())
    pb1.backward_ptr(td.data_ptr(), gx1.data_ptr(), gt1.data_ptr())
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        pb1.forward_ptr(xd.data_ptr(), td.data_ptr())
    torch.cuda.synchronize()
    t_fwd1 = (time.perf_counter() - t) / args.steps
    t = time.perf_counter()
    for _ in range(args.steps):
        pb1.backward_ptr(td.data_ptr(), gx1.data_ptr(), gt1.data_ptr())
    torch.cuda.synchronize()
    t_bwd1 = (time.perf_counter() - t) / args.steps
    # both gradients are deterministic (the grid's is summed in fixed point)
    same = bool(torch.equal(gx1, gx)) and bool(torch.equal(gt1, gt))
    log(f"every state kept: forward {1e3 * t_fwd1:.2f} ms, backward {1e3 * t_bwd1:.2f} ms (x{t_bwd1 / t_fwd1:.2f}), "
        f"same gradient as the checkpointed run: {same}")
    pb1.close()
    check = None
    if not args.no_cpu:
        # the FP64 CPU gradient and central differences (eps 1e-3 h) of the FP64 forward
        pc, keep = adjoint.problem_struct(pb.nbr, pb.nnb, pb.typ, pb.tets, theta.shape, pb.origin, pb.spacing, prm)
        t = time.perf_counter()
        L64, gx64, gt64 = orc.adj_grad(pc, x0, theta)
        t_cpu = time.perf_counter() - t
        g1, g2 = gx.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64)
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        # central differences of the FP64 forward, component by component: the sizing
        # grid's two largest gradient entries at eps = 1e-3 h, and the start positions'
        # largest entry at 1e-3 h and 1e-4 h (the 64-step trajectory bends sharply in the
        # positions: there 1e-3 h is too coarse a step for a central difference)
        hbar = float(np.mean(theta))

        def central(ex, et, eps_h):
            eps = eps_h * hbar
            lp = orc.adj_loss(pc, x0 + eps * ex, theta + eps * et)[0]
            lm = orc.adj_loss(pc, x0 - eps * ex, theta - eps * et)[0]
            return (lp - lm) / (2 * eps)

        fd = []
        for c in np.argsort(-np.abs(g2.ravel()))[:2]:
            e = np.zeros(theta.size)
            e[c] = 1.0
            want = central(0.0 * x0, e.reshape(theta.shape), 1e-3)
            fd.append({"component": f"theta[{int(c)}]", "eps_over_h": 1e-3, "fd": want, "adjoint": float(g2.ravel()[c]),
                       "rel": abs(float(g2.ravel()[c]) - want) / abs(want)})
        i, a_ = np.unravel_index(int(np.argmax(np.abs(g1))), g1.shape)
        e = np.zeros(x0.shape)
        e[i, a_] = 1.0
        for eps_h in (1e-3, 1e-4):
            want = central(e, 0.0 * theta, eps_h)
            fd.append({"component": f"x0[{int(i)},{int(a_)}]", "eps_over_h": eps_h, "fd": want,
                       "adjoint": float(g1[i, a_]), "rel": abs(float(g1[i, a_]) - want) / abs(want)})
        check = {"loss_fp64": L64, "loss_rel": abs(loss - L64) / abs(L64), "grad_x_rel": rel(g1, gx64),
                 "grad_theta_rel": rel(g2, gt64), "fd": fd, "cpu_grad_s": t_cpu}
        log(f"check: {check}")
    line = {
        "metric": "node-iterations/sec of the differentiable relaxation forward; adjoint cost ratio",
        "value": n * iters / t_fwd, "unit": "node-iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * (t_fwd + t_bwd), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "adjoint_volume: 128^3 three nested perturbed ellipsoid labels, h 1.5, device seeding; "
                               "64 smoothed Jacobi steps, checkpoints every 8, 32^3 sizing grid (trilinear, smoothstep fractions), "
                               "loss = mean truss energy + mean (1 - smoothed Joe-Liu) over a Delaunay of the seeds",
                   "nodes": n, "tets": int(pb.tets.shape[0]), "iterations": iters, "parallelism": "single GPU"},
        "forward_ms": 1e3 * t_fwd, "backward_ms": 1e3 * t_bwd, "adjoint_cost_ratio": t_bwd / t_fwd,
        "every_state_kept": {"forward_ms": 1e3 * t_fwd1, "backward_ms": 1e3 * t_bwd1,
                             "adjoint_cost_ratio": t_bwd1 / t_fwd1, "identical_gradient": same},
        "loss": loss, "check": check,
    }
    if emit:
        print(json.dumps(line), flush=True)
    pb.close()
    return line


def bench_grid(args) -> int:
    """The grid build on the device (tn_grid_build) of the large workload's 256^3
    label volume with the reference's default GridParams (hbase 4.5, hmin 1.5, hmax
    6 voxels): bricks, slots, curvature and interface fields with the thin-layer
    blend, sizing, gradient limit, grades. A step is one build from the host labels
    (their 32 MB upload included; the grid stays on the device). Against the
    reference's build_grid_cpu on this host's cores."""
    from trussnet_amd import gridbuild, volumes
    from trussnet_amd._abi import GridParams

    lab = volumes.nested_ellipsoids(256, nlab=6, seed=0, axes=(20.0, 120.0), grf_sigma=4.0, grf_amp=3.0)
    prm = GridParams(hbase=4.5, hmin=1.5, hmax=6.0)
    nv = lab.size
    info = {}
    for _ in range(args.warmup):
        gridbuild.build(lab, prm=prm, device=0, info=info)
    t = time.perf_counter()
    for _ in range(args.steps):
        G = gridbuild.build_resident(lab, prm=prm, device=0)
        G.close()
    dt = time.perf_counter() - t
    value = nv * args.steps / dt
    log(f"grid build: {1e3 * dt / args.steps:.1f} ms per build, {info.get('nslot')} slots, "
        f"{info.get('limit_sweeps')} limit sweeps")
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref

            if not ref.available():
                raise RuntimeError("reference library not built")
            t = time.perf_counter()
            r = ref.RefGrid.build_full(lab, prm=prm)
            dtc = time.perf_counter() - t
            sc = r.scalars()
            cores = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 64))
            cpu = {"value": nv / dtc, "unit": "voxels/s", "cores": cores, "kind": "reference",
                   "sample": f"build_grid_cpu on the same volume and parameters, one build, {dtc:.2f} s "
                             f"({sc['nslot']} slots, {sc['limit_sweeps']} limit sweeps)"}
        except Exception as e:
            log(f"reference CPU path unavailable ({e})")
    line = {
        "metric": "voxels/s of the grid build (build_grid_cpu on the device)", "value": value, "unit": "voxels/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain_large's volume: 256^3 six nested perturbed ellipsoid labels (seed 0), "
                               "GridParams hbase 4.5, hmin 1.5, hmax 6 voxels, defaults otherwise; label upload included",
                   "voxels": nv, "slots": info.get("nslot"), "limit_sweeps": info.get("limit_sweeps"),
                   "parallelism": "single GPU"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_pipeline(args) -> int:
    """Labels to relaxed nodes on the device: tn_grid_build of the large workload's
    256^3 volume (hbase 4.5, hmin 1.5, hmax 6 voxels, the reference's defaults
    otherwise), tn_seed_from_grid, tn_plan_create_from_grid, 100 iterations, the
    nodes downloaded. A step is the whole chain; the reference's build_grid_cpu,
    seed_cpu and relax_cpu (the same 100 iterations unless --cpu-iters says otherwise,
    then scaled) time the same chain."""
    from trussnet_amd import gridbuild, volumes
    from trussnet_amd._abi import GridParams, RelaxParams

    lab = volumes.nested_ellipsoids(256, nlab=6, seed=0, axes=(20.0, 120.0), grf_sigma=4.0, grf_amp=3.0)
    gprm = GridParams(hbase=4.5, hmin=1.5, hmax=6.0)
    iters = args.iters or 100
    rprm = RelaxParams(max_iters=iters, dptol=0.0)

    def chain():
        t0 = time.perf_counter()
        res = gridbuild.build_resident(lab, prm=gprm, device=0)
        t1 = time.perf_counter()
        seeds = res.seed(RelaxParams())
        t2 = time.perf_counter()
        with res.plan(seeds, rprm) as plan:
            res.close()
            st = plan.run()
            out = plan.download()
        t3 = time.perf_counter()
        return seeds.n, st, (t1 - t0, t2 - t1, t3 - t2)

    for _ in range(args.warmup):
        chain()
    t = time.perf_counter()
    for _ in range(args.steps):
        n, st, parts = chain()
    dt = (time.perf_counter() - t) / args.steps
    log(f"pipeline: {n} nodes; grid {1e3 * parts[0]:.1f} ms, seeds {1e3 * parts[1]:.1f} ms, "
        f"plan + {iters} iterations + download {1e3 * parts[2]:.1f} ms")
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import ref

            if not ref.available():
                raise RuntimeError("reference library not built")
            t0 = time.perf_counter()
            r = ref.RefGrid.build_full(lab, prm=gprm)
            t1 = time.perf_counter()
            sd = r.seed(RelaxParams())
            t2 = time.perf_counter()
            _, stc = r.relax(sd, RelaxParams(max_iters=args.cpu_iters or iters, dptol=0.0))
            t3 = time.perf_counter()
            total = (t1 - t0) + (t2 - t1) + (t3 - t2) * iters / max(1, stc["iters"])
            cores = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 64))
            cpu = {"value": total, "unit": "s", "cores": cores, "kind": "reference",
                   "sample": f"build_grid_cpu {t1 - t0:.2f} s + seed_cpu {t2 - t1:.2f} s ({sd.n} nodes) + "
                             f"relax_cpu {stc['iters']} iterations {t3 - t2:.2f} s scaled to {iters}"}
        except Exception as e:
            log(f"reference CPU path unavailable ({e})")
    line = {
        "metric": "seconds from a 256^3 label volume to relaxed nodes (grid build, seeding, relaxation)",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain_large's volume with the reference's curvature sizing (GridParams hbase 4.5, "
                               f"hmin 1.5, hmax 6), seed_cpu defaults, {iters} iterations, labels up and nodes down "
                               "over PCIe included", "nodes": n, "iterations": iters,
                   "stage_ms": [round(1e3 * x, 2) for x in parts], "parallelism": "single GPU"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline_many(grids, seeds, iters: int) -> dict | None:
    from trussnet_amd._abi import RelaxParams

    try:
        from oracle import ref

        if not ref.available():
            raise RuntimeError("reference library not built")
    except Exception as e:
        log(f"reference CPU path unavailable ({e})")
        return None
    prm = RelaxParams(max_iters=iters, dptol=0.0)
    nit, secs = 0, 0.0
    for g, sd in zip(grids, seeds):
        rg = ref.RefGrid.from_grid(g)
        t = time.perf_counter()
        _, st = rg.relax(sd, prm)
        secs += time.perf_counter() - t
        nit += sd.n * st["iters"]
    cores = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 64))
    return {"value": nit / secs, "unit": "node-iterations/s", "cores": cores, "kind": "reference",
            "sample": f"relax_cpu on {len(grids)} of the volumes, {iters} iterations each, {secs:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["large", "small", "prob", "voxel", "gray", "batch", "adjoint", "grid",
                                           "pipeline"], default="large")
    ap.add_argument("--iters", type=int, default=0, help="iterations per step (default: the workload's)")
    ap.add_argument("--cpu-iters", type=int, default=0, help="iterations of the CPU baseline (default: the GPU's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 PMC passes")
    ap.add_argument("--no-extras", action="store_true",
                    help="large workload: skip the batched and adjoint figures added to its line")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs the relaxation on one GPU; --gpus {args.gpus} "
              f"(world size {world}) is not supported", file=sys.stderr)
        return 2

    if args.workload == "batch":
        bench_batch(args)
        return 0
    if args.workload == "adjoint":
        bench_adjoint(args)
        return 0
    if args.workload == "grid":
        return bench_grid(args)
    if args.workload == "pipeline":
        return bench_pipeline(args)
    pmc = None if args.no_traffic else pmc_traffic(args)   # before this process initialises the GPU

    from trussnet_amd import relax
    from trussnet_amd._abi import RelaxParams

    grid, seeds, desc = build_workload(args.workload)
    iters = args.iters or (50 if args.workload == "small" else 100)
    voxel = args.workload == "voxel"
    prm = RelaxParams(max_iters=iters, dptol=0.0, voxel_trap=voxel)
    t = time.perf_counter()
    plan = relax.Plan(grid, seeds, prm, device=0)   # grid and seeds up, first rebuild
    plan_ms = 1e3 * (time.perf_counter() - t)
    log(f"plan: {seeds.n} nodes, hash {plan.hash.nlev} levels / {plan.hash.nkeys} cells ({plan_ms:.1f} ms)")
    # The warmup times every kernel class, to find the dominant kernel; the timed steps
    # record events around that kernel's launches only (every class's events together
    # cost about 4% of the step).
    passes = 0 if plan.hash.nkeys <= 1 else (int(plan.hash.nkeys - 1).bit_length() + 7) // 8
    plan.set_timing(True)
    for w in range(args.warmup):
        st = plan.run()
        log(f"warmup {w}: {st['iters']} iterations, {st['rebuilds']} rebuilds")
    warm = plan.timings() if args.warmup else {}
    if warm:
        agg = by_kernel(warm, item_bytes(24.0, passes))
        top = max(agg, key=lambda k: agg[k]["ms"])
        plan.set_timing(tuple(agg[top]["classes"]))
        log(f"timed steps: events around {top} only (the dominant kernel of the warmup)")
    else:
        plan.set_timing(True)
    t = time.perf_counter()
    for k in range(args.steps):
        st = plan.run()
    dt = time.perf_counter() - t
    timings = plan.timings()
    if st["iters"] != iters:
        raise RuntimeError(f"a step ran {st['iters']} iterations, not {iters}")
    t = time.perf_counter()
    plan.download()   # the relaxed nodes back to the host
    d2h_ms = 1e3 * (time.perf_counter() - t)
    rb = plan.rebuild()
    kbar = float(rb["nnb"].mean())
    n = seeds.n
    value = n * iters * args.steps / dt
    ms_step = 1e3 * dt / args.steps
    log(f"timed: {args.steps} steps in {dt:.3f} s -> {value:.4g} node-iterations/s; "
        f"{st['rebuilds']} rebuilds/step, mean list {kbar:.2f}")
    for label, tm, steps in (("timed", timings, args.steps), ("warmup", warm, args.warmup)):
        for name, (cnt, ms, items) in sorted(tm.items(), key=lambda kv: -kv[1][1]):
            if cnt:
                log(f"  {label} {name:<11} {cnt:6d} launches {ms:9.2f} ms total {1e3 * ms / cnt:9.1f} us/launch "
                    f"{items / cnt:11.0f} items/launch")
    # roofline of the dominant kernel (the timed classes merged by kernel: the general
    # move pass runs twice an iteration, as two classes), from the timed steps' events
    per_item = item_bytes(kbar, passes)

    def kernel_roofline(kernels: dict, kname: str, steps: int) -> dict:
        a = kernels[kname]
        avg_s = a["ms"] / a["launches"] / 1e3
        nbytes = a["bytes"] / a["launches"]
        achieved = nbytes / avg_s
        traffic = None
        if pmc and pmc["copy_kib"][0] > 0 and pmc["copy_kib"][1] > 0:
            # the copy moves 16 n bytes each way: bytes = counter KiB x 1024 x factor
            fr = 16.0 * n / (pmc["copy_kib"][0] * 1024.0)
            fw = 16.0 * n / (pmc["copy_kib"][1] * 1024.0)
            ks = [pmc["kernels"].get(k) for k in kname.split("+")]
            if ks and all(ks):
                rd = sum(v[0] for v in ks) * 1024.0 * fr
                wr = sum(v[1] for v in ks) * 1024.0 * fw
                traffic = rd + wr
                log(f"traffic of {kname}: {rd / 1e6:.1f} MB read + {wr / 1e6:.1f} MB written per launch "
                    f"(calibration: read x{fr:.3f}, write x{fw:.3f})")
        return {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic, "bytes_per_launch": nbytes,
                "avg_launch_us": avg_s * 1e6, "items_per_launch": a["items"] / a["launches"],
                "ms_per_step": a["ms"] / max(steps, 1), "classes": a["classes"]}

    timed_k = by_kernel(timings, per_item, warm or None)
    dom = max(timed_k, key=lambda k: timed_k[k]["ms"])
    roofline = kernel_roofline(timed_k, dom, args.steps)
    # the next kernels by device time, in the same terms, from the warmup's events
    others = []
    if warm:
        warm_k = by_kernel(warm, per_item)
        for k in sorted(warm_k, key=lambda k: -warm_k[k]["ms"]):
            if k != dom and len(others) < 3:
                others.append(dict(kernel_roofline(warm_k, k, args.warmup), timed="warmup"))
    loop_bytes = iters * n * (83 + 4 * kbar) + st["rebuilds"] * n * (179 + 4 * kbar)
    cpu = None if args.no_cpu else cpu_baseline(grid, seeds, args.cpu_iters or iters, voxel)
    line = {
        "metric": METRIC, "value": value, "unit": "node-iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "nodes": n, "iterations": iters, "rebuilds_per_step": st["rebuilds"],
                   "mean_list": round(kbar, 3), "node_types": [st["n_interior"], st["n_interface"],
                                                                st["n_junction"], st["n_corner"]],
                   "parallelism": "single GPU"},
        "hbm_frac_loop": loop_bytes / (dt / args.steps) / HBM_PEAK,
        "roofline": roofline,
        "rooflines_next": others,
        "cpu_baseline": cpu,
        # the host boundary, outside `value`: plan creation (grid and seeds up, the first
        # rebuild) and the relaxed nodes down
        "host_transfer": {"plan_create_ms": plan_ms, "download_ms": d2h_ms,
                          "pcie_inclusive_value": n * iters / (dt / args.steps + 1e-3 * (plan_ms + d2h_ms))},
    }
    plan.close()
    if args.workload == "large" and not args.no_extras:
        # the two added patterns on the same GPU: 64 volumes relaxed together against a
        # loop of single calls, and the adjoint's cost and accuracy
        sub = argparse.Namespace(**vars(args))
        sub.iters, sub.no_cpu, sub.steps, sub.warmup = 0, True, 1, 1
        b = bench_batch(sub, emit=False)
        line["batched"] = {"value": b["value"], "loop_value": b["loop_value"], "ratio": b["value"] / b["loop_value"],
                           "identical_to_loop": b["identical_to_loop"], "ms_per_step": b["ms_per_step"],
                           "workload": b["config"]["workload"], "nodes": b["config"]["nodes"], "volumes": 64}
        sub.steps, sub.no_cpu = 3, False
        a = bench_adjoint(sub, emit=False)
        ck = a.get("check") or {}
        line["adjoint"] = {"forward_ms": a["forward_ms"], "backward_ms": a["backward_ms"],
                           "adjoint_cost_ratio": a["adjoint_cost_ratio"], "nodes": a["config"]["nodes"],
                           "iterations": a["config"]["iterations"], "checkpoint_every": 8,
                           "grad_x_rel_fp64": ck.get("grad_x_rel"), "grad_theta_rel_fp64": ck.get("grad_theta_rel"),
                           "fd": ck.get("fd"), "workload": a["config"]["workload"]}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
