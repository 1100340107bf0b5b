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

roofline: the kernel with the largest device time in the timed region (HIP
events on the plan's stream), its algorithmic bytes per launch (per-node
figures below, times the nodes) over its average launch time, against 8 TB/s.
hbm_frac_loop: the whole loop's algorithmic bytes, iterations x n x (83 + 4 K)
+ rebuilds x n x (179 + 4 K), over the step time, against 8 TB/s.
cpu_baseline: the reference's own relax_cpu (oracle/_ref, compiled from its
sources) on this host's cores, on the same grid and seeds, for a bounded
number of iterations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|small]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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
        "stats": 36,                  # mv 4, pw 4, P,h 16, P0 12 read
        "refresh": 47 + 4 * kbar,     # list 4, sorted 4, P,h 16, typ 1, lab 2 read; nbr 4K, nnb 4, P0 16 written
        "neighbors": 23 + 4 * kbar,   # P,h 16, typ 1, lab 2 read; nbr 4K + nnb 4 written
        "keys": 24,                   # P 12 + h 4 read; key 4 written; count 4
        "scan": 8,                    # per cell: read + write
        "radix_sort": 20 * radix_passes,
        "gather": 68,                 # sorted 4 + X 16 read, Ps 16 written; P0 16 + 16
        "order": 6,
    }


def launch_bytes(name: str, timings: dict, per_item: dict) -> float | None:
    """Algorithmic bytes of one average launch of a kernel class."""
    launches, _, items = timings[name]
    if not launches:
        return None
    if name == "move_easy":   # every node: order 4 + typ 1; finished nodes 94 more; queued ones 4 (the queue)
        queued = timings["move"][2] / max(timings["move"][0], 1)
        n = items / launches
        return 5 * n + 94 * (n - queued) + 4 * queued
    if name not in per_item:
        return None
    return per_item[name] * items / launches


def build_workload(name: str):
    from trussnet_amd import relax, workload
    from trussnet_amd._abi import RelaxParams

    t = time.perf_counter()
    if name == "large":
        grid = workload.plain_large()
        desc = ("plain_large: 256^3 six-label nested perturbed ellipsoids, h graded 1.5-6 voxels, "
                "device HCP seeding, 100 iterations")
    else:
        grid = workload.plain_small()
        desc = "plain_small: 64^3 two spheres r 24/12, uniform h 3, device HCP seeding, 50 iterations"
    log(f"grid {grid.nx}x{grid.ny}x{grid.nz}: {grid.nslot} slots ({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    seeds = relax.seed(grid, RelaxParams(jseed=0.0, corners=False), device=0)
    log(f"seeds: {seeds.n} ({time.perf_counter() - t:.1f} s)")
    return grid, seeds, desc


def cpu_baseline(grid, seeds, iters: int) -> dict | None:
    from trussnet_amd._abi import RelaxParams

    prm = RelaxParams(max_iters=iters, dptol=0.0)
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["large", "small"], default="large")
    ap.add_argument("--iters", type=int, default=0, help="iterations per step (default: the workload's)")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs the relaxation on one GPU; --gpus {args.gpus} "
              f"(world size {world}) is not supported", file=sys.stderr)
        return 2

    from trussnet_amd import relax
    from trussnet_amd._abi import RelaxParams

    grid, seeds, desc = build_workload(args.workload)
    iters = args.iters or (100 if args.workload == "large" else 50)
    prm = RelaxParams(max_iters=iters, dptol=0.0)
    plan = relax.Plan(grid, seeds, prm, device=0)
    log(f"plan: {seeds.n} nodes, hash {plan.hash.nlev} levels / {plan.hash.nkeys} cells")
    for w in range(args.warmup):
        st = plan.run()
        log(f"warmup {w}: {st['iters']} iterations, {st['rebuilds']} rebuilds")
    plan.set_timing(True)
    t = time.perf_counter()
    for k in range(args.steps):
        st = plan.run()
    dt = time.perf_counter() - t
    timings = plan.timings()
    if st["iters"] != iters:
        raise RuntimeError(f"a step ran {st['iters']} iterations, not {iters}")
    rb = plan.rebuild()
    kbar = float(rb["nnb"].mean())
    n = seeds.n
    value = n * iters * args.steps / dt
    ms_step = 1e3 * dt / args.steps
    log(f"timed: {args.steps} steps in {dt:.3f} s -> {value:.4g} node-iterations/s; "
        f"{st['rebuilds']} rebuilds/step, mean list {kbar:.2f}")
    for name, (cnt, ms, items) in sorted(timings.items(), key=lambda kv: -kv[1][1]):
        if cnt:
            log(f"  {name:<11} {cnt:6d} launches {ms:9.2f} ms total {1e3 * ms / cnt:9.1f} us/launch "
                f"{items / cnt:11.0f} items/launch")
    # roofline of the dominant kernel
    passes = 0 if plan.hash.nkeys <= 1 else (int(plan.hash.nkeys - 1).bit_length() + 7) // 8
    per_item = item_bytes(kbar, passes)
    ranked = [k for k, v in sorted(timings.items(), key=lambda kv: -kv[1][1])
              if v[0] and launch_bytes(k, timings, per_item) is not None]
    dom = ranked[0]
    cnt, ms, _ = timings[dom]
    avg_s = ms / cnt / 1e3
    nbytes = launch_bytes(dom, timings, per_item)
    achieved = nbytes / avg_s
    roofline = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                "bytes_per_launch": nbytes, "avg_launch_us": avg_s * 1e6,
                "items_per_launch": timings[dom][2] / cnt}
    loop_bytes = iters * n * (83 + 4 * kbar) + st["rebuilds"] * n * (179 + 4 * kbar)
    cpu = None if args.no_cpu else cpu_baseline(grid, seeds, args.cpu_iters)
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
