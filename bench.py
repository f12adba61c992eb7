"""Benchmark of the neighbourhood-search hot path on one MI355X.

Workload: 100^3 = 1,000,000 jittered lattice points (spacing 0.01, jitter U[+-0.001],
Float32), smoothing length h = 1.2 * 0.01, search radius 2h (~58 neighbours per point).
One step (default `--workload wcsph`, the weakly-compressible SPH step):
    displace every point by a seeded U[+-0.0005]; update! (counting-sort rebuild);
    cubic-spline density summation; Tait equation of state; pressure acceleration.
`--workload count` replaces the physics by one neighbour-counting traversal.
Inputs are resident in HBM when timing starts.

Prints one JSON line.  `value` is neighbour pairs (ordered, self included) per second of
whole steps.  `roofline` is the dominant traversal kernel's algorithmic bytes per launch
over its average launch time (HIP events the library records on its launch stream right
around each kernel, tuning key "time_launches") against 8 TB/s.
`cpu_baseline` times the oracle's threaded restatement of the reference CPU path on the
same workload, on a bounded sample.
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
RHO0, SOUND, GAMMA = 1000.0, 10.0, 7.0


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
    ap.add_argument("--workload", choices=["wcsph", "count", "batch", "large", "csr", "tlsph", "hash",
                                           "trixi-wcsph", "nbody", "update"], default="wcsph")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="approximate CPU time budget of the cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the traversal launch and build strategy on this workload first")
    ap.add_argument("--tuning-cache", default=None, help="JSON cache for --autotune results")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a tuning knob (algorithm, persist_blocks, chunk, canonical_order) "
                         "before timing; repeatable")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, copy 256 MiB device-to-device (a known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design (z-slab sharding with halo exchange), not built", file=sys.stderr)
        sys.exit(2)

    import torch
    import pointneighbors as pn
    from pointneighbors import sph, synthetic

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, dev)))
        return
    if args.workload == "large":
        print(json.dumps(bench_large(args, dev)))
        return
    if args.workload == "csr":
        print(json.dumps(bench_csr(args, dev)))
        return
    if args.workload == "tlsph":
        print(json.dumps(bench_tlsph(args, dev)))
        return
    if args.workload == "hash":
        print(json.dumps(bench_hash(args, dev)))
        return
    if args.workload == "trixi-wcsph":
        print(json.dumps(bench_trixi_wcsph(args, dev)))
        return
    if args.workload == "nbody":
        print(json.dumps(bench_nbody(args, dev)))
        return
    if args.workload == "update":
        print(json.dumps(bench_update(args, dev)))
        return
    pts = synthetic.jittered_lattice(SHAPE, SPACING, JITTER, dtype=np.float32, seed=1)
    n = pts.shape[0]
    h = np.float32(1.2) * np.float32(SPACING)
    r = np.float32(2.0) * h
    lo = pts.min(axis=0) - np.float32(3) * r
    hi = pts.max(axis=0) + np.float32(3) * r
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    x = synthetic.to_device_coords(pts, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234)
    disp = torch.empty((n, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    mass = torch.full((n,), RHO0 * SPACING**3, dtype=torch.float32, device=dev)
    rho = torch.zeros(n, dtype=torch.float32, device=dev)
    P = torch.zeros(n, dtype=torch.float32, device=dev)
    dv = torch.zeros((n, 3), dtype=torch.float32, device=dev).t()
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    f_count = pn.CountNeighbors(counts, accumulate=False)
    f_rho = sph.SPHDensity(rho, mass, float(h), "cubic_spline", accumulate=False)
    f_dv = sph.SPHPressureAcceleration(dv, mass, P, float(h), "cubic_spline", accumulate=False)

    def step():
        disp.uniform_(-DISP, DISP, generator=gen)
        x.add_(disp.t())
        pn.update(nhs, x, x)
        if args.workload == "count":
            pn.foreach_point_neighbor(f_count, x, x, nhs)
            return
        pn.foreach_point_neighbor(f_rho, x, x, nhs)
        sph.tait_pressure_term(rho, RHO0, SOUND, GAMMA, out=P)
        pn.foreach_point_neighbor(f_dv, x, x, nhs)

    pn.initialize(nhs, x, x)
    tuned = None
    tune_table = None
    if args.autotune:
        from pointneighbors import autotune
        key = autotune.signature(nhs, x, args.workload)
        tuned, tune_table = autotune.autotune(nhs, step, reps=5, warmup=2, cache=args.tuning_cache, key=key)
    if args.tune:
        tuned = dict(tuned or {})
        for kv in args.tune:
            k, v = kv.split("=", 1)
            nhs.set_tuning(k, int(v))
            tuned[k] = int(v)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    # each traversal kernel's duration: HIP events the library records on its launch stream
    # right around the kernel
    nhs.set_tuning("time_launches", 1)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for _ in range(args.steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_step = t_start.elapsed_time(t_end) / args.steps
    launch_ms = nhs.launch_times()
    nhs.set_tuning("time_launches", 0)
    per_step = 1 if args.workload == "count" else 2
    if len(launch_ms) != per_step * args.steps:
        raise RuntimeError(f"expected {per_step * args.steps} timed traversal launches, got {len(launch_ms)}")

    # pairs of the final configuration (outside the timed region)
    pn.foreach_point_neighbor(f_count, x, x, nhs)
    torch.cuda.synchronize()
    pairs = int(counts.sum().item())
    n_binned, n_nonempty, max_count = nhs.stats()

    extra = {}
    if args.workload == "count":
        trav_ms = float(np.mean(launch_ms))
        kernel = "CountF"
        algo = traversal_bytes(n, n_nonempty)
        extra["count_traversal_ms"] = trav_ms
    else:
        dens_ms = float(np.mean(launch_ms[0::2]))
        # the equation of state alone, after the timed steps
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(20):
            sph.tait_pressure_term(rho, RHO0, SOUND, GAMMA, out=P)
        e1.record(stream)
        torch.cuda.synchronize()
        eos_ms = e0.elapsed_time(e1) / 20
        pres_ms = float(np.mean(launch_ms[1::2]))
        extra.update({"density_traversal_ms": dens_ms, "eos_ms": eos_ms, "pressure_traversal_ms": pres_ms,
                      "density_pairs_per_s": pairs / (dens_ms * 1e-3),
                      "pressure_pairs_per_s": pairs / (pres_ms * 1e-3)})
        # the dominant (slower) traversal of the step
        if pres_ms >= dens_ms:
            trav_ms, kernel = pres_ms, "PressureAccelF"
            # per neighbour: coordinates + m_j + P_j; per query: P_i; out: 3 floats
            algo = traversal_bytes(n, n_nonempty, f_in=8, f_own=4, f_out=12)
        else:
            trav_ms, kernel = dens_ms, "DensityF"
            # per neighbour: coordinates + m_j; out: rho_i
            algo = traversal_bytes(n, n_nonempty, f_in=4, f_own=0, f_out=4)
    kernel = f"k_cells<3,float,float,{kernel}>"
    achieved = algo / (trav_ms * 1e-3) / 1e9

    # measured HBM bytes per launch of this kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_1m.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pts, lo, hi, r, h, args.workload, args.cpu_seconds)
    copy_gbs = streaming_copy_gbs(dev)
    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        extra["copy_probe_bytes"] = src.numel() * 4

    line = {
        "metric": "neighbour pairs/s at 1M 3D Float32 points (whole step) + traversal % of HBM peak",
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
        "config": {"workload": ("wcsph-1m-3d-f32: 100^3 lattice, spacing 0.01, jitter 0.001, h = 1.2*0.01, "
                                "r = 2h; step = displace U[+-0.0005] + update! + cubic-spline density + "
                                "Tait EOS + pressure acceleration") if args.workload == "wcsph" else
                               ("count-1m-3d-f32: 100^3 lattice, spacing 0.01, jitter 0.001, r = 2*(1.2*0.01); "
                                "step = displace U[+-0.0005] + update! + count neighbours"),
                   "n_points": n, "search_radius": float(r), "n_cells": int(np.prod(cl.n_cells)),
                   "nonempty_cells": int(n_nonempty), "max_cell_population": int(max_count)},
        "pairs_per_step": pairs,
        "dominant_traversal_ms": trav_ms,
        "dominant_kernel": kernel,
        **extra,
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "algorithmic_bytes_per_launch": algo, "measured_copy_gbs": copy_gbs,
                     "frac_of_measured_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "tuning": tuned,
        "autotune_table": tune_table,
    }
    print(json.dumps(line))


def streaming_copy_gbs(dev, mib=1024, reps=10):
    """The box's streaming bandwidth: a device-to-device copy of `mib` MiB (read + write bytes
    over the best of `reps` timed copies), reported beside the nominal 8 TB/s."""
    import torch
    src = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    best = float("inf")
    for _ in range(reps + 2):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    del src, dst
    return 2 * mib * 1024 * 1024 / (best * 1e-3) / 1e9


BATCH_CLOUDS, BATCH_MIN, BATCH_MAX, BATCH_R = 512, 1000, 65536, 2.4


def batch_clouds(seed=1):
    """512 jittered lattices (spacing 1, jitter U[+-0.25], Float32), sizes log-uniform in
    [1000, 65536], each in its own box."""
    from pointneighbors import synthetic
    u = synthetic.uniform(seed, BATCH_CLOUDS)
    sizes = np.round(np.exp(np.log(BATCH_MIN) + u * (np.log(BATCH_MAX) - np.log(BATCH_MIN)))).astype(int)
    clouds = []
    for c, s_ in enumerate(sizes):
        side = max(2, int(round(s_ ** (1 / 3))))
        shape = (side, side, max(1, int(round(s_ / side / side))))
        clouds.append(synthetic.jittered_lattice(shape, 1.0, 0.25, dtype=np.float32, seed=seed * 100003 + c))
    return clouds


def bench_batch(args, dev):
    """512 uneven clouds, 100-step loop of displace U[+-0.05] + update! + count: per-cloud
    launches (one GridNeighborhoodSearch per cloud, no synchronization inside the step)
    against one batched search whose step is captured in a graph."""
    import torch
    import pointneighbors as pn
    from pointneighbors import batch, synthetic
    clouds = batch_clouds()
    r = np.float32(BATCH_R)
    margin = np.float32(3 * BATCH_R)
    cls = [pn.FullGridCellList(min_corner=p.min(axis=0) - margin, max_corner=p.max(axis=0) + margin,
                               search_radius=r) for p in clouds]
    offs = np.concatenate([[0], np.cumsum([len(p) for p in clouds])])
    n = int(offs[-1])
    allpts = np.concatenate(clouds)
    steps = max(args.steps, 1)
    stream = torch.cuda.current_stream(dev)

    # per-cloud launches
    y1 = synthetic.to_device_coords(allpts, dev)
    c1 = torch.zeros(n, dtype=torch.int32, device=dev)
    step1 = torch.zeros(1, dtype=torch.int64, device=dev)
    nhss, views, outs = [], [], []
    for c in range(len(clouds)):
        v = y1[:, offs[c]:offs[c + 1]]
        nh = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=len(clouds[c]), cell_list=cls[c])
        pn.initialize(nh, v, v)
        nhss.append(nh)
        views.append(v)
        outs.append(pn.CountNeighbors(c1[offs[c]:offs[c + 1]], accumulate=False))

    def per_cloud_step():
        batch.displace_uniform(y1, 0.05, 7, step1)
        for c in range(len(clouds)):
            pn.update(nhss[c], views[c], views[c], check_errors=False)
            pn.foreach_point_neighbor(outs[c], views[c], views[c], nhss[c])

    # one batched search, step captured in a graph
    nb = batch.BatchedGridNeighborhoodSearch(3, search_radius=r, cell_lists=cls, n_points=[len(p) for p in clouds])
    y2 = synthetic.to_device_coords(allpts, dev)
    c2 = torch.zeros(n, dtype=torch.int32, device=dev)
    step2 = torch.zeros(1, dtype=torch.int64, device=dev)

    def batched_step():
        batch.displace_uniform(y2, 0.05, 7, step2)
        nb.update(y2, check_errors=False)
        nb.count_neighbors(c2)

    def timed(fn, k):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    per_cloud_ms = timed(per_cloud_step, min(steps, 20))
    for nh in nhss:
        nh.check()
    eager_ms = timed(batched_step, steps)
    nb.check()
    s = torch.cuda.Stream(dev)
    s.wait_stream(stream)
    with torch.cuda.stream(s):
        batched_step()
    stream.wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batched_step()
    graph_ms = timed(g.replay, steps)
    nb.check()
    pairs = int(c2.sum().item())
    return {
        "metric": "batched update! + count of 512 uneven clouds: speedup over per-cloud launches",
        "value": per_cloud_ms / graph_ms, "unit": "x", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": graph_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded SplitMix64 jittered lattices and displacements)",
        "config": {"workload": "batch-512-clouds: sizes log-uniform 1000..65536, spacing 1, jitter 0.25, "
                               "r = 2.4; step = displace U[+-0.05] + update! + count",
                   "n_clouds": len(clouds), "n_points": n},
        "per_cloud_ms_per_step": per_cloud_ms, "batched_eager_ms_per_step": eager_ms,
        "batched_graph_ms_per_step": graph_ms, "pairs_per_step": pairs,
        "pairs_per_s": pairs / (graph_ms * 1e-3),
    }


def _timed(fn, k, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(k):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / k


def bench_large(args, dev):
    """256^3 = 16,777,216 jittered lattice points (spacing 1, jitter 0.1), Float32,
    h = 1.2, r = 2h: WCSPH forward (update! + density + EOS + pressure acceleration)
    against forward + adjoint of L = sum |a_i|^2 with respect to the positions (reverse
    mode through the pressure and density adjoint kernels; no stored pair lists)."""
    import torch
    import pointneighbors as pn
    from pointneighbors import sph, synthetic
    pts = synthetic.jittered_lattice((256, 256, 256), 1.0, 0.1, dtype=np.float32, seed=1)
    n = len(pts)
    h = np.float32(1.2)
    r = np.float32(2) * h
    lo, hi = pts.min(axis=0) - 3 * r, pts.max(axis=0) + 3 * r
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    x = synthetic.to_device_coords(pts, dev)
    del pts
    mass = torch.full((n,), 1000.0, dtype=torch.float32, device=dev)
    rho = torch.zeros(n, dtype=torch.float32, device=dev)
    P = torch.zeros(n, dtype=torch.float32, device=dev)
    dv = torch.zeros((n, 3), dtype=torch.float32, device=dev).t()
    f_rho = sph.SPHDensity(rho, mass, float(h), accumulate=False)
    f_dv = sph.SPHPressureAcceleration(dv, mass, P, float(h), accumulate=False)
    B = RHO0 * SOUND**2 / GAMMA
    pn.initialize(nhs, x, x)

    def forward():
        pn.update(nhs, x, x)
        pn.foreach_point_neighbor(f_rho, x, x, nhs)
        sph.tait_pressure_term(rho, RHO0, SOUND, GAMMA, out=P)
        pn.foreach_point_neighbor(f_dv, x, x, nhs)

    xg = x.detach().requires_grad_(True)

    def forward_adjoint():
        pn.update(nhs, x, x)
        rh = sph.sph_density(xg, mass, nhs, float(h))
        pp = B * ((rh / RHO0) ** GAMMA - 1.0)
        a = sph.sph_pressure_acceleration(xg, mass, pp / (rh * rh), nhs, float(h))
        loss = (a * a).sum()
        (g,) = torch.autograd.grad(loss, xg)
        return g

    k = max(2, min(args.steps, 10))
    fwd_ms = _timed(forward, k)
    adj_ms = _timed(forward_adjoint, k)
    g = forward_adjoint()
    torch.cuda.synchronize()
    c = torch.zeros(n, dtype=torch.int32, device=dev)
    pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
    pairs = int(c.sum().item())
    return {
        "metric": "forward + adjoint time / forward time (16.8M 3D Float32 WCSPH step)",
        "value": adj_ms / fwd_ms, "unit": "ratio", "n_gpus": 1, "steps": k, "warmup": 2,
        "ms_per_step": adj_ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded SplitMix64 jittered lattice)",
        "config": {"workload": "wcsph-16m-3d-f32: 256^3 lattice, spacing 1, jitter 0.1, h = 1.2, r = 2h; "
                               "loss = sum |a|^2, adjoint w.r.t. positions", "n_points": n},
        "forward_ms": fwd_ms, "forward_adjoint_ms": adj_ms, "pairs_per_step": pairs,
        "forward_pairs_per_s": pairs / (fwd_ms * 1e-3),
        "grad_finite": bool(torch.isfinite(g).all().item()),
    }


def bench_csr(args, dev):
    """PrecomputedNeighborhoodSearch on an undeformed 128^3 lattice (spacing 1, Float32),
    r = 3: CSR build (count, scan, fill, sort rows) and a TLSPH-style density sum over
    the stored lists."""
    import torch
    import pointneighbors as pn
    from pointneighbors import sph, synthetic
    pts = synthetic.jittered_lattice((128, 128, 128), 1.0, 0.0, dtype=np.float32, seed=1)
    n = len(pts)
    r = np.float32(3.0)
    lo, hi = pts.min(axis=0) - r, pts.max(axis=0) + r
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    inner = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    pre = pn.PrecomputedNeighborhoodSearch[3](search_radius=r, n_points=n, update_neighborhood_search=inner)
    x = synthetic.to_device_coords(pts, dev)
    k = max(2, min(args.steps, 10))
    build_ms = _timed(lambda: pn.initialize(pre, x, x), k)
    nnz = int(pre.neighbor_lists.col_idx.numel())
    lens = pre.neighbor_lists.lengths()
    mass = torch.ones(n, dtype=torch.float32, device=dev)
    rho = torch.zeros(n, dtype=torch.float32, device=dev)
    f = sph.SPHDensity(rho, mass, 1.5, "wendland_c2", accumulate=False)
    frozen = pn.freeze_neighborhood_search(pre)
    sum_ms = _timed(lambda: pn.foreach_point_neighbor(f, x, x, frozen), k)
    return {
        "metric": "CSR neighbour-list build time (128^3 lattice, r = 3)", "value": build_ms, "unit": "ms",
        "n_gpus": 1, "steps": k, "warmup": 2, "ms_per_step": build_ms, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (undeformed lattice)",
        "config": {"workload": "csr-128cube-3d-f32: 128^3 lattice, spacing 1, r = 3; build + list density",
                   "n_points": n},
        "nnz": nnz, "max_row": int(lens.max().item()), "interior_row": int(lens[64 + 128 * 64 + 128 * 128 * 64].item()),
        "build_ms": build_ms, "list_density_ms": sum_ms, "list_pairs_per_s": nnz / (sum_ms * 1e-3),
    }


def reference_cloud(side, search_radius_factor, seed=1):
    """The reference's benchmark cloud (test/point_cloud.jl:4-57 as driven by
    benchmarks/run_benchmarks.jl:60-113): a side^3 unit lattice (1-based), jittered by
    N(0, 0.05^2), stably sorted by the cells (size search_radius_factor) of that first jitter
    with the first dimension slowest, jittered again by N(0, 0.05^2); float64, lattice units.
    Julia's RNG stream cannot be reproduced; the jitter comes from numpy's seeded generator."""
    from pointneighbors import synthetic
    rng = np.random.default_rng(seed)
    X = synthetic.jittered_lattice((side,) * 3, 1.0, 0.0, dtype=np.float64) + 1.0
    X = X + 0.05 * rng.standard_normal(X.shape)
    cells = np.floor(X / float(search_radius_factor)).astype(np.int64)
    X = X + 0.05 * rng.standard_normal(X.shape)
    return X[np.lexsort((cells[:, 2], cells[:, 1], cells[:, 0]))]


H100_WCSPH_MS = {"f32": 3.407714, "f64": 6.047292}  # benchmarks/plot_benchmarks.jl:67-68 (1,030,301 points)


def bench_trixi_wcsph(args, dev):
    """TrixiParticles' WCSPH interact! on the cloud of the reference's published WCSPH runs
    (benchmark_wcsph, benchmarks/smoothed_particle_hydrodynamics.jl:44-102, run by
    run_benchmark_full_grid(benchmark_wcsph, (10, 10, 10), 9, search_radius_factor = 3.0f0),
    benchmarks/plot_benchmarks.jl:53-56): 101^3 = 1,030,301 points rescaled by the domain size
    102, r = 3/102, FullGridCellList over [0, 1]^3, WendlandC2 with h = r/2, density
    1000 + U[0, 1), mass 0.1 * spacing, velocity 0, StateEquationCole(c = 10, exponent 1),
    ArtificialViscosityMonaghan(alpha = 0.02), DensityDiffusionMolteniColagrossi(delta = 0.1).
    Times one interact! per call (dv accumulates, the pressure precomputed), as the reference
    does, in Float32 (the headline) and Float64."""
    import torch
    import pointneighbors as pn
    from pointneighbors import synthetic
    side = 101
    base = reference_cloud(side, 3.0)
    n = len(base)
    k = max(5, min(args.steps, 50))
    gen = np.random.default_rng(2)
    rand = gen.random(n)

    def run(dt, tdt):
        domain = dt(side + 1)
        X = (base.astype(dt) / domain).astype(dt)
        r = dt(3.0) / domain
        spacing = r / dt(3)
        h = float(r) / 2
        cl = pn.FullGridCellList(min_corner=np.zeros(3, dtype=dt), max_corner=np.ones(3, dtype=dt), search_radius=r)
        nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
        x = synthetic.to_device_coords(X, dev)
        pn.initialize(nhs, x, x)
        rho = torch.as_tensor((dt(1000) + rand.astype(dt)).astype(dt), device=dev)
        mass = torch.full((n,), float(dt(0.1) * spacing), dtype=tdt, device=dev)
        v = torch.zeros((n, 4), dtype=tdt, device=dev)
        v[:, 3] = rho
        v = v.t()
        p = pn.state_equation_cole(rho, 1000.0, 10.0, 1.0)
        dv = torch.zeros_like(v)
        f = pn.WCSPHInteraction(dv, v, p, mass, h, 10.0)
        ms = _timed(lambda: pn.foreach_point_neighbor(f, x, x, nhs), k)
        c = torch.zeros(n, dtype=torch.int32, device=dev)
        pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
        pairs = int(c.sum().item())
        return {"interact_ms": ms, "pairs": pairs, "pairs_per_s": pairs / (ms * 1e-3),
                "finite": bool(torch.isfinite(dv).all().item()), "speedup_vs_h100": H100_WCSPH_MS[str(dt.__name__).replace("float", "f")] / ms}

    r32 = run(np.float32, torch.float32)
    r64 = run(np.float64, torch.float64)
    return {
        "metric": "WCSPH interact! time (benchmark_wcsph, 1,030,301 points)", "value": r32["interact_ms"],
        "unit": "ms", "n_gpus": 1, "steps": k, "warmup": 2, "ms_per_step": r32["interact_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": r32["interact_ms"] / H100_WCSPH_MS["f32"],
        "dtype": "f32", "data": "synthetic (the reference's jittered, cell-sorted lattice; numpy RNG)",
        "config": {"workload": "trixi-wcsph-101cube-3d: 101^3 lattice / 102, r = 3/102, h = r/2, WendlandC2, "
                               "ContinuityDensity, Cole EOS (c 10, exponent 1), Monaghan alpha 0.02, "
                               "Molteni-Colagrossi delta 0.1", "n_points": n},
        "f32": r32, "f64": r64,
        "h100_ms": H100_WCSPH_MS,
    }


def bench_nbody(args, dev):
    """benchmark_n_body (benchmarks/n_body.jl:34-51) on the reference's published cloud
    (101^3 / 102, r = 3/102, Float32, FullGridCellList over [0, 1]^3): dv .= 0, then the n-body
    closure over all neighbours; mass 1e10 (1 + U[0, 1)), G = 6.6743e-11."""
    import torch
    import pointneighbors as pn
    from pointneighbors import synthetic
    side = 101
    X = (reference_cloud(side, 3.0).astype(np.float32) / np.float32(side + 1)).astype(np.float32)
    n = len(X)
    r = np.float32(3.0) / np.float32(side + 1)
    cl = pn.FullGridCellList(min_corner=np.zeros(3, np.float32), max_corner=np.ones(3, np.float32), search_radius=r)
    nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    x = synthetic.to_device_coords(X, dev)
    pn.initialize(nhs, x, x)
    mass = torch.as_tensor((1e10 * (np.random.default_rng(3).random(n) + 1)).astype(np.float32), device=dev)
    dv = torch.zeros((n, 3), dtype=torch.float32, device=dev).t()
    f = pn.NBodyAcceleration(dv, mass, 6.6743e-11, accumulate=False)  # dv .= 0 then +=
    k = max(5, min(args.steps, 50))
    ms = _timed(lambda: pn.foreach_point_neighbor(f, x, x, nhs), k)
    c = torch.zeros(n, dtype=torch.int32, device=dev)
    pn.foreach_point_neighbor(pn.CountNeighbors(c), x, x, nhs)
    pairs = int(c.sum().item())
    return {
        "metric": "n-body closure time (benchmark_n_body, 1,030,301 points)", "value": ms, "unit": "ms",
        "n_gpus": 1, "steps": k, "warmup": 2, "ms_per_step": ms, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (the reference's jittered, cell-sorted lattice)",
        "config": {"workload": "nbody-101cube-3d-f32: 101^3 lattice / 102, r = 3/102", "n_points": n},
        "pairs": pairs, "pairs_per_s": pairs / (ms * 1e-3), "finite": bool(torch.isfinite(dv).all().item()),
    }


def bench_update(args, dev):
    """benchmark_update_alternating (benchmarks/update.jl:28-49) on the 1M-point WCSPH cloud:
    coordinates2 = coordinates perturbed by N(0, (4e-4 r)^2); each call does update! with
    coordinates then update! with coordinates2 (each a full counting-sort rebuild, checked for
    the reference's errors); reported per update!.  initialize! (benchmark_initialize) beside it."""
    import torch
    import pointneighbors as pn
    from pointneighbors import synthetic
    pts = synthetic.jittered_lattice(SHAPE, SPACING, JITTER, dtype=np.float32, seed=1)
    n = pts.shape[0]
    h = np.float32(1.2) * np.float32(SPACING)
    r = np.float32(2.0) * h
    lo = pts.min(axis=0) - np.float32(3) * r
    hi = pts.max(axis=0) + np.float32(3) * r
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    nhs = pn.GridNeighborhoodSearch[3](search_radius=r, n_points=n, cell_list=cl)
    x = synthetic.to_device_coords(pts, dev)
    pts2 = (pts + np.random.default_rng(4).standard_normal(pts.shape) * (4e-4 * float(r))).astype(np.float32)
    x2 = synthetic.to_device_coords(pts2, dev)
    pn.initialize(nhs, x, x)
    k = max(5, min(args.steps, 50))

    def alternating():
        pn.update(nhs, x, x)
        pn.update(nhs, x, x2)

    pair_ms = _timed(alternating, k)
    init_ms = _timed(lambda: pn.initialize(nhs, x, x), k)
    s1, i1 = nhs.cell_table()
    pn.update(nhs, x, x2)
    s2, _ = nhs.cell_table()
    moved = int(np.abs(np.diff(s1) - np.diff(s2)).sum() // 2)
    return {
        "metric": "update! time (benchmark_update_alternating, 1M 3D Float32)", "value": 0.5 * pair_ms,
        "unit": "ms", "n_gpus": 1, "steps": k, "warmup": 2, "ms_per_step": 0.5 * pair_ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded jittered lattice, N(0, (4e-4 r)^2) perturbation)",
        "config": {"workload": "update-1m-3d-f32: 100^3 lattice, spacing 0.01, r = 0.024, alternating clouds",
                   "n_points": n},
        "update_ms": 0.5 * pair_ms, "initialize_ms": init_ms, "points_per_s": n / (0.5 * pair_ms * 1e-3),
        "cell_changes_lower_bound": moved,
    }


def bench_tlsph(args, dev):
    """TLSPH on the cloud of the reference's published TLSPH runs: 101^3 = 1,030,301 points, a
    unit lattice jittered twice by N(0, 0.05^2) and stably sorted by the cells (size 3) of the
    first jitter, first dimension slowest (test/point_cloud.jl:4-57 as driven by
    benchmarks/run_benchmarks.jl:60-113; lattice units, not rescaled to a unit box), search radius 3, h = r/2, Wendland C2, the material of
    setup_tlsph (benchmarks/smoothed_particle_hydrodynamics.jl:146-187: rho 1000, m 0.1,
    E 1.4e6, nu 0.4, Ganzenmueller penalty alpha 0.1), Float32, precomputed lists of the initial
    coordinates.  Times the two loops the reference times: interact_structure_structure!
    (benchmark_tlsph, accumulating into dv) and calc_deformation_grad!
    (benchmark_tlsph_deformation_grad), both per call with the lists resident.  A Float64 pass
    of the same cloud is reported beside the Float32 headline."""
    import torch
    import pointneighbors as pn
    from pointneighbors import sph, synthetic
    side = 101
    r = np.float32(3.0)
    X = reference_cloud(side, 3.0).astype(np.float32)
    n = len(X)
    h = float(r) / 2
    rho0, m0, E, nu, alpha = 1000.0, 0.1, 1.4e6, 0.4, 0.1
    k = max(2, min(args.steps, 10))

    def run(pts, radius, tdt):
        pre = pn.PrecomputedNeighborhoodSearch[3](search_radius=radius, n_points=n)
        Xd = synthetic.to_device_coords(pts, dev)
        pn.initialize(pre, Xd, Xd)
        mass = torch.full((n,), m0, dtype=tdt, device=dev)
        rho = torch.full((n,), rho0, dtype=tdt, device=dev)
        vol = mass / rho
        res = {"nnz": int(pre.neighbor_lists.col_idx.numel())}
        res["correction_matrix_ms"] = _timed(lambda: sph.tlsph_correction_matrix(pre, Xd, vol, h), k)
        L = sph.tlsph_correction_matrix(pre, Xd, vol, h)
        x = Xd * 1.01 + 0.002 * torch.sin(Xd)
        res["deformation_grad_ms"] = _timed(lambda: sph.tlsph_deformation_gradient(pre, Xd, x, vol, h, correction=L), k)
        F = sph.tlsph_deformation_gradient(pre, Xd, x, vol, h, correction=L)
        res["pk1_corrected_ms"] = _timed(lambda: sph.tlsph_pk1_corrected(F, rho, E, nu, correction=L), k)
        Pr = sph.tlsph_pk1_corrected(F, rho, E, nu, correction=L)
        dv = sph.tlsph_interaction(pre, Xd, x, mass, rho, Pr, h, F=F, young_modulus=E, penalty_alpha=alpha)
        res["interaction_ms"] = _timed(lambda: sph.tlsph_interaction(pre, Xd, x, mass, rho, Pr, h, F=F, young_modulus=E,
                                                                     penalty_alpha=alpha, dv=dv), k)
        res["finite"] = bool(torch.isfinite(F).all().item() and torch.isfinite(dv).all().item())
        del pre
        return res

    r32 = run(X, r, torch.float32)
    r64 = run(X.astype(np.float64), np.float64(r), torch.float64)
    inter_ms, grad_ms, nnz = r32["interaction_ms"], r32["deformation_grad_ms"], r32["nnz"]
    h100_interact_ms, h100_grad_ms = 3.801, 0.825  # benchmarks/plot_benchmarks.jl:76,84
    return {
        "metric": "TLSPH interaction time (benchmark_tlsph, 1,030,301 points)", "value": inter_ms, "unit": "ms",
        "n_gpus": 1, "steps": k, "warmup": 2, "ms_per_step": inter_ms, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": inter_ms / h100_interact_ms, "dtype": "f32",
        "data": "synthetic (jittered lattice, smooth deformation)",
        "config": {"workload": "tlsph-101cube-3d-f32: 101^3 lattice, 2 x N(0, 0.05^2) jitter, cell-sorted, r = 3, h = 1.5, "
                               "Wendland C2, rho 1000, m 0.1, E 1.4e6, nu 0.4, penalty alpha 0.1",
                   "n_points": n},
        "nnz": nnz, "interaction_ms": inter_ms, "interaction_pairs_per_s": nnz / (inter_ms * 1e-3),
        "deformation_grad_ms": grad_ms, "deformation_grad_pairs_per_s": nnz / (grad_ms * 1e-3),
        "speedup_vs_h100": {"interaction": h100_interact_ms / inter_ms, "deformation_grad": h100_grad_ms / grad_ms},
        "correction_matrix_ms": r32["correction_matrix_ms"], "pk1_corrected_ms": r32["pk1_corrected_ms"],
        "finite": r32["finite"],
        "f64": dict(r64, speedup_vs_h100={"interaction": 6.187 / r64["interaction_ms"],  # plot_benchmarks.jl:75,83
                                          "deformation_grad": 1.423 / r64["deformation_grad_ms"]}),
    }


def bench_hash(args, dev):
    """update! + neighbour count on the 1M-point Float32 WCSPH cloud (100^3 lattice, spacing
    0.01, r = 0.024) with a SpatialHashingCellList of list_size 2N (the reference's
    recommendation), beside the same steps with the FullGridCellList; counts must agree."""
    import torch
    import pointneighbors as pn
    from pointneighbors import synthetic
    pts = synthetic.jittered_lattice(SHAPE, SPACING, JITTER, dtype=np.float32, seed=1)
    n = pts.shape[0]
    r = np.float32(2.0) * (np.float32(1.2) * np.float32(SPACING))
    x = synthetic.to_device_coords(pts, dev)
    lo, hi = pts.min(axis=0) - np.float32(3) * r, pts.max(axis=0) + np.float32(3) * r
    grids = {
        "spatial_hashing": pn.GridNeighborhoodSearch[3](
            search_radius=r, n_points=n, cell_list=pn.SpatialHashingCellList[3](list_size=2 * n)),
        "full_grid": pn.GridNeighborhoodSearch[3](
            search_radius=r, n_points=n, cell_list=pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)),
    }
    k = max(2, min(args.steps, 20))
    res, counts = {}, {}
    for name, nhs in grids.items():
        c = torch.zeros(n, dtype=torch.int32, device=dev)
        f = pn.CountNeighbors(c, accumulate=False)
        pn.initialize(nhs, x, x)
        res[name + "_update_ms"] = _timed(lambda: pn.update(nhs, x, x), k)
        res[name + "_count_ms"] = _timed(lambda: pn.foreach_point_neighbor(f, x, x, nhs), k)
        counts[name] = c
    pairs = int(counts["full_grid"].sum().item())
    step = res["spatial_hashing_update_ms"] + res["spatial_hashing_count_ms"]
    return {
        "metric": "neighbour pairs/s, update! + count with a SpatialHashingCellList (1M 3D Float32)",
        "value": pairs / (step * 1e-3), "unit": "pairs/s", "n_gpus": 1, "steps": k, "warmup": 2,
        "ms_per_step": step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 jittered lattice)",
        "config": {"workload": "hash-1m-3d-f32: 100^3 lattice, spacing 0.01, r = 0.024, list_size 2N", "n_points": n},
        "pairs": pairs, "counts_equal_full_grid": bool(torch.equal(counts["spatial_hashing"], counts["full_grid"])),
        **res,
    }


def cpu_baseline(pts, lo, hi, r, h, workload, budget_s):
    """The oracle's threaded restatement of the reference CPU path (ParallelUpdate build
    with atomic pushes + point-major traversals statically chunked over points) on the
    same 1M-point workload, repeated until about `budget_s` seconds of work."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    import pointneighbors as pn
    cl = pn.FullGridCellList(min_corner=lo, max_corner=hi, search_radius=r)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    g = orc.Grid(3, cl.min_corner.astype(np.float64), cl.n_cells, float(r), coord_dtype=np.float32,
                 radius_dtype=np.float32, cell_dtype=np.float32, max_points_per_cell=100)
    reps, t_build, t_work, pairs = 0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    while True:
        if workload == "count":
            st, counts, tb, tt = g.threaded_count_f32(pts, cores)
            pairs = int(counts.sum()) if st == 0 else 0
        else:
            st, rho, dv, tb, tt = g.threaded_wcsph_f32(pts, cores, RHO0 * SPACING**3, float(h), RHO0, SOUND, GAMMA)
            if st == 0 and pairs == 0:
                _, counts, _, _ = g.threaded_count_f32(pts, cores)
                pairs = int(counts.sum())
        if st != 0:
            return None
        reps += 1
        t_build += tb
        t_work += tt
        if time.perf_counter() - t0 > budget_s or reps >= 50:
            break
    per_step = (t_build + t_work) / reps
    return {"value": pairs / per_step, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"{reps} steps (update! + {'count' if workload == 'count' else 'density + EOS + pressure'}) "
                      f"of the same 1M-point Float32 configuration, no displacement; build "
                      f"{t_build / reps * 1e3:.1f} ms, traversals {t_work / reps * 1e3:.1f} ms per step"}


if __name__ == "__main__":
    main()
