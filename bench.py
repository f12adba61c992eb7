"""Benchmark of the neighbourhood-search hot path on one MI355X.

Workload (one step): the 3D weakly-compressible-SPH neighbour step on 100^3 = 1,000,000
jittered lattice points (spacing 0.01, jitter U[+-0.001], Float32), search radius
2 * (1.2 * 0.01) (~58 neighbours per point): displace every point by a seeded
U[+-0.0005], update! (counting-sort cell-list rebuild), then foreach_point_neighbor
counting the neighbours of every point.  Inputs are resident in HBM when timing starts.

Prints one JSON line.  `value` is neighbour pairs (ordered, self included) per second of
whole steps.  `roofline` is the traversal kernel's algorithmic bytes per launch over its
average launch time (HIP events on the launch stream), against 8 TB/s.  `cpu_baseline`
times the oracle's threaded restatement of the reference CPU path on a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pointneighbors.jl_amd"))

HBM_PEAK_GBS = 8000.0
SHAPE = (100, 100, 100)
SPACING = 0.01
JITTER = 0.001
DISP = 0.0005


def traversal_bytes(n_points: int, n_nonempty: int, d: int = 3, s: int = 4, f_in: int = 0,
                    f_own: int = 0, f_out: int = 4) -> int:
    """Algorithmic bytes of one traversal launch (cell-tile model): every query cell
    stages its 3^d stencil (coordinates + per-neighbour fields), reads the permutation
    and its own fields, writes its outputs, and reads 3^(d-1) row ranges (8 B each)."""
    return n_points * (3**d * (d * s + f_in) + 4 + f_own + f_out) + n_nonempty * 3 ** (d - 1) * 8


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="approximate CPU time budget of the cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design (z-slab sharding with halo exchange), not built", file=sys.stderr)
        sys.exit(2)

    import torch
    import pointneighbors as pn
    from pointneighbors import synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    pts = synthetic.jittered_lattice(SHAPE, SPACING, JITTER, dtype=np.float32, seed=1)
    n = pts.shape[0]
    r = np.float32(2.0) * (np.float32(1.2) * np.float32(SPACING))
    lo = pts.min(axis=0) - np.float32(3) * r
    hi = pts.max(axis=0) + np.float32(3) * r
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    x = synthetic.to_device_coords(pts, dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234)
    disp = torch.empty((n, 3), dtype=torch.float32, device=dev)
    f = pn.CountNeighbors(counts, accumulate=False)
    stream = torch.cuda.current_stream(dev)

    def step(time_kernels=None):
        torch.rand((n, 3), generator=gen, device=dev, out=disp)
        disp.mul_(2 * DISP).sub_(DISP)
        x.add_(disp.t())
        pn.update(nhs, x, x)
        if time_kernels is not None:
            time_kernels[0].record(stream)
        pn.foreach_point_neighbor(f, x, x, nhs)
        if time_kernels is not None:
            time_kernels[1].record(stream)

    pn.initialize(nhs, x, x)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: exactly K whole steps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for k in range(args.steps):
        step(ev[k])
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = t_start.elapsed_time(t_end)
    ms_step = ms_total / args.steps
    trav_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    pairs = int(counts.sum().item())
    n_binned, n_nonempty, max_count = nhs.stats()

    algo = traversal_bytes(n, n_nonempty)
    achieved = algo / (trav_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traverse_count_1m_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pts, lo, hi, r, args.cpu_seconds)

    line = {
        "metric": "neighbour pairs/s (1M 3D Float32 points, update! + count per step)",
        "value": pairs / (ms_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 jittered lattice, torch Philox displacements)",
        "config": {"workload": "wcsph-1m-3d-f32: 100^3 lattice, spacing 0.01, jitter 0.001, "
                               "r = 2*(1.2*0.01), step = displace U[+-0.0005] + update! + "
                               "count neighbours",
                   "n_points": n, "search_radius": float(r), "n_cells": int(np.prod(cl.n_cells)),
                   "nonempty_cells": int(n_nonempty), "max_cell_population": int(max_count)},
        "pairs_per_step": pairs,
        "traversal_ms": trav_ms,
        "traversal_pairs_per_s": pairs / (trav_ms * 1e-3),
        "update_and_displace_ms": ms_step - trav_ms,
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_traverse<3,float,float,CountF>",
                     "algorithmic_bytes_per_launch": algo},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def cpu_baseline(pts, lo, hi, r, budget_s):
    """The oracle's threaded restatement of the reference CPU path (ParallelUpdate build
    with atomic pushes + point-major traversal statically chunked over points) on the
    same 1M-point workload, repeated until about `budget_s` seconds of work."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    import pointneighbors as pn
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    g = orc.Grid(3, cl.min_corner.astype(np.float64), cl.n_cells, float(r), coord_dtype=np.float32,
                 radius_dtype=np.float32, cell_dtype=np.float32, max_points_per_cell=100)
    reps, t_build, t_trav, pairs = 0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    while True:
        st, counts, tb, tt = g.threaded_count_f32(pts, cores)
        if st != 0:
            return None
        reps += 1
        t_build += tb
        t_trav += tt
        pairs = int(counts.sum())
        if time.perf_counter() - t0 > budget_s or reps >= 50:
            break
    per_step = (t_build + t_trav) / reps
    return {"value": pairs / per_step, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"{reps} x (update! + count) of the same 1M-point Float32 workload "
                      f"(one fixed configuration, no displacement); build {t_build / reps * 1e3:.1f} ms, "
                      f"traversal {t_trav / reps * 1e3:.1f} ms per step",
            "traversal_pairs_per_s": pairs / (t_trav / reps)}


if __name__ == "__main__":
    main()
