#!/usr/bin/env python3
"""Throughput of octree ray traversal on one MI355X.

Workload (default "caves"): a 512^3 octree with brick dim 16 (seeded porous
noise terrain with worm caves; 40% Empty / 10% Solid / rest Parted brick
slots), 1920x1080 primary rays from each of 8 seeded orbit poses. One step
renders all 8 poses on the device (rays generated in the kernel, 24 bytes of
output per ray: hit, voxel xyz, albedo, distance); the tree is resident in
HBM before timing starts. Prints one JSON line.

By default a step is one batched launch (svo_batch_run: the 8 render jobs in
one persistent kernel), whose outputs are checked equal to 8 separate
svo_render launches; `--launch per-frame` times the 8 separate launches
instead. A single launch ends when its slowest waves end (a few tiles take
~10x the mean number of traversal steps), so one launch per frame leaves the
GPU partly idle at the end of every frame; the batched launch overlaps those
tails with other frames' tiles. Both figures are reported.

The roofline object prices the timed launch against HBM bandwidth with the
algorithmic bytes per ray counted by the instrumented CPU restatement on a
sample of the same rays; the cpu_baseline object times that restatement on
the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Mrays/s at 1080p into 512³; p50/p99 latency batched vs per-call at fixed QPS")
HBM_PEAK_GBS = 8000.0
# VALU issue ceiling of an MI355X in G wave64 instructions per second: each
# of a CU's 4 SIMDs is 16 lanes wide and issues one wave64 VALU instruction
# every 4 cycles, so a CU completes one per cycle; 256 CUs at the 2.4 GHz
# peak engine clock.
VALU_PEAK_GIPS = 256 * 1.0 * 2.4
# Algorithmic bytes of one logical event of the reference traversal
EVENT_BYTES = {"stack_iterations": 12, "child_reads": 4, "brick_reads": 4, "voxel_tests": 8,
               "data_palette_reads": 4, "hits": 4}
OUT_BYTES_PER_RAY = 24  # hit + voxel xyz + albedo + distance


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def step_parity(otree, rays, outs, extra=None) -> dict:
    """Every ray of the measured step against the oracle: `outs` are the
    step's device outputs (hit, albedo, voxel, distance), one dict per render
    job in ray order, and `extra` = (rays, outputs) of a cast job. Reports the
    rays whose fields differ and the boundary hits (impact within 1e-6 of a
    second face of the hit cell, where the hit cell is decided by an exact
    f32 tie), and how many of those differ."""
    import numpy as np

    cores = min(16, os.cpu_count() or 1)
    groups = [(rays, outs)] + ([extra] if extra is not None else [])
    checked = bad = boundary = bad_boundary = 0
    for rr, oo in groups:
        exp = otree.cast(rr, threads=cores)
        got = {k: np.concatenate([o[k].cpu().numpy().reshape(-1) for o in oo]) for k in
               ("hit", "albedo", "voxel", "distance")}
        n = len(exp)
        differ = np.zeros(n, bool)
        differ |= got["hit"].view(np.uint32) != exp["hit"]
        differ |= got["albedo"].view(np.uint32) != exp["albedo"]
        differ |= (got["voxel"].view(np.uint32).reshape(n, 3) != exp["voxel"]).any(axis=1)
        differ |= got["distance"].view(np.uint32) != exp["distance"].view(np.uint32)
        b = exp["boundary"] == 1
        checked += n
        bad += int(differ.sum())
        boundary += int(b.sum())
        bad_boundary += int((differ & b).sum())
    return {"rays_checked": checked, "fields": "hit, albedo, voxel xyz, distance (bits)",
            "mismatching_rays": bad, "boundary_hits": boundary,
            "boundary_fraction": boundary / max(1, checked),
            "boundary_mismatches": bad_boundary}


def cpu_baseline(cfg, tree, vps, seconds: float, extra_rays=None):
    """Times the CPU restatement (single thread, then all cores) on a bounded
    sample of the same rays and returns (cpu_baseline, per-ray event counts of
    the step's rays: the primary rays plus `extra_rays` if given, the loaded
    oracle tree, the primary rays)."""
    import numpy as np

    import oracle

    t0 = time.time()
    otree = oracle.OracleTree(tree.export_model())
    log(f"cpu baseline: oracle tree loaded in {time.time() - t0:.1f}s")
    rays = np.concatenate([oracle.viewport_rays(v.origin, v.direction, v.frustum, v.fov,
                                                cfg.width, cfg.height) for v in vps])
    # event counts on a strided sample (instrumented run)
    _, cnt = otree.cast(rays[::97], counters=True)
    per_ray = {k: getattr(cnt, k) / cnt.rays for k in EVENT_BYTES}
    if extra_rays is not None:
        _, ce = otree.cast(extra_rays[::97], counters=True)
        wp, we = len(rays), len(extra_rays)
        per_ray = {k: (per_ray[k] * wp + getattr(ce, k) / ce.rays * we) / (wp + we)
                   for k in EVENT_BYTES}
    # single thread: a strided sample sized to about `seconds / 2`
    probe = rays[::4001]
    t0 = time.perf_counter()
    otree.cast(probe, threads=1)
    rate1 = len(probe) / max(time.perf_counter() - t0, 1e-9)
    n1 = int(min(len(rays), max(len(probe), rate1 * seconds / 2)))
    stride = max(1, len(rays) // n1)
    sample = np.ascontiguousarray(rays[::stride])
    t0 = time.perf_counter()
    otree.cast(sample, threads=1)
    t1 = time.perf_counter() - t0
    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    otree.cast(sample, threads=cores)
    tn = time.perf_counter() - t0
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    base = {
        "value": round(len(sample) / t1 / 1e6, 3), "unit": "Mrays/s", "cores": 1, "kind": "port",
        "sample": (f"{len(sample)} of the {len(rays)} rays of the 8 poses (every {stride}th), "
                   f"CPU restatement of get_by_ray, g++ -O3 -ffp-contract=off"),
        "all_cores": {"value": round(len(sample) / tn / 1e6, 3), "cores": cores},
        "cpu_model": cpu_model,
    }
    return base, per_ray, otree, rays


def copy_calibration(args) -> int:
    """A plain device copy of 256 MiB per step, the reference point for
    rocprofv3 FETCH_SIZE / WRITE_SIZE (bytes read and written are known)."""
    import torch

    n_bytes = 256 << 20
    src = torch.empty(n_bytes // 4, dtype=torch.int32, device="cuda:0").fill_(7)
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"mode": "copy", "bytes_read_per_step": n_bytes,
                      "bytes_written_per_step": n_bytes, "steps": args.steps,
                      "GBps": round(2 * n_bytes * args.steps / el / 1e9, 1)}), flush=True)
    return 0


def batched_bench(args, render_variant: str) -> dict:
    """Batched mode: trees of uneven size (log-uniform 64..max), content of the
    plain heightfield (built bottom-up) or the caves by seed, brick dims
    4/8/16, each with 1-4 viewports of resolution drawn from {128^2, 512^2,
    720p, 1080p}; all rays in one persistent launch per batch kernel variant
    (a replayed hipGraph), against one plain render launch per viewport.
    Outputs of the two forms are compared for equality."""
    import torch

    from shocovox_amd import (Batch, DeviceTree, Octree, alloc_hits, batch_variants,
                              render_variants, select_batch_variant, select_render_variant,
                              workloads as W)

    t0 = time.time()
    specs, views = W.batched_pool_specs(args.seed, args.batched_trees, args.batched_max_size)
    trees, dtrees = [], []
    kinds = {2: 0, 4: 0}
    for kind, size, dim, seed in specs:
        kinds[kind] += 1
        t = Octree.scene(kind, size, dim, seed)
        trees.append(t)
        dtrees.append(DeviceTree(t))
    rays = sum(w * h for _, _, w, h in views)
    pool_bytes = sum(d.device_bytes for d in dtrees)
    sizes = [t.get_size() for t in trees]
    log(f"batched: {len(trees)} trees ({pool_bytes / 2**30:.2f} GiB), {len(views)} viewports, "
        f"{rays} rays, built in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    fields = ("hit", "albedo", "voxel", "distance")
    plain_bufs = [alloc_hits(w * h, dev, fields) for _, _, w, h in views]
    batch_bufs = [alloc_hits(w * h, dev, fields) for _, _, w, h in views]
    stream = torch.cuda.current_stream()
    reps = max(1, args.batched_reps)

    def plain_pass():
        for (ti, vp, w, h), b in zip(views, plain_bufs):
            dtrees[ti].render_into(vp, w, h, b, stream)

    # the per-viewport baseline runs the render variant fastest on this pool,
    # among the headline's pick and the two forms of the one-tile kernel
    rnames = render_variants()
    cands = []
    for name in (render_variant, "w5-glut-8x8", "w5-glut-8x8-ww-g1"):
        if name in rnames and name not in cands:
            cands.append(name)
    plain_ms = {}
    for name in cands:
        select_render_variant(rnames.index(name))
        plain_pass()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            plain_pass()
        torch.cuda.synchronize()
        plain_ms[name] = (time.perf_counter() - t1) / reps * 1e3
        log(f"batched pool, per-viewport launches with {name}: {plain_ms[name]:.3f} ms")
    plain_variant = min(plain_ms, key=plain_ms.get)
    select_render_variant(rnames.index(plain_variant))
    plain_pass()  # the reference outputs for the equality check
    torch.cuda.synchronize()
    plain_s = plain_ms[plain_variant] / 1e3
    out = {"workload": (f"{len(trees)} trees (sizes {min(sizes)}..{max(sizes)} log-uniform, "
                        f"{sizes.count(1024)} of 1024^3; brick dims 4/8/16; {kinds[2]} caves, "
                        f"{kinds[4]} heightfield; {pool_bytes / 2**30:.2f} GiB resident), "
                        f"{len(views)} viewports of 128^2/512^2/720p/1080p, {rays} rays"),
           "trees": len(trees), "tree_sizes": sorted(sizes), "pool_bytes": pool_bytes,
           "viewports": len(views), "rays": rays,
           "plain": {"ms": round(plain_s * 1e3, 3), "mrays_per_s": round(rays / plain_s / 1e6, 1),
                     "launches": len(views), "render_variant": plain_variant,
                     "variants_ms": {k: round(v, 3) for k, v in plain_ms.items()}},
           "batched": {}}
    names = batch_variants()
    best = None
    for vi, name in enumerate(names):
        if name.startswith("r1-") and not args.batched_all_variants:
            continue  # lane refill: measured 2-3x slower; timed with --batched-all-variants
        select_batch_variant(vi)
        batch = Batch(len(views))
        jobs = [Batch.render_job(dtrees[ti], vp, w, h, b)
                for (ti, vp, w, h), b in zip(views, batch_bufs)]
        batch.run(jobs, stream)
        torch.cuda.synchronize()
        same = all(torch.equal(a[k], b[k]) for a, b in zip(plain_bufs, batch_bufs) for k in fields)
        t1 = time.perf_counter()
        for _ in range(reps):
            batch.run(jobs, stream)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t1) / reps
        out["batched"][name] = {"ms": round(el * 1e3, 3), "mrays_per_s": round(rays / el / 1e6, 1),
                                "equal_to_plain": bool(same)}
        log(f"batched {name}: {el * 1e3:.3f} ms, {rays / el / 1e6:.1f} Mrays/s, equal {same}")
        for b in batch_bufs:
            for k in fields:
                b[k].zero_()
        if same and (best is None or el < out["batched"][best]["ms"] / 1e3):
            best = name
        del batch
    select_batch_variant(names.index(best) if best else 0)
    out["best_variant"] = best
    if best:
        out["speedup_vs_plain"] = round(plain_s / (out["batched"][best]["ms"] / 1e3), 3)
    del dtrees, trees
    torch.cuda.synchronize()
    return out


def served_pool(args):
    """(specs, trees): the served pool's trees, built on the host. Sizes are
    log-uniform over 64..max (stratified), brick dims 4/8/16 below 1024^3 and
    16 from 1024^3 on; content by seed: caves or the plain heightfield below
    2048^3, spheres over a heightfield or the heightfield at 2048^3, any of
    the three at 1024^3."""
    from shocovox_amd import Octree, workloads as W

    specs = W.served_pool_specs(args.seed, args.served_trees, args.served_max_size)
    t0 = time.time()
    trees = []
    for k, (kind, size, dim, seed) in enumerate(specs):
        trees.append(Octree.scene(kind, size, dim, seed))
        if size >= 1024:
            log(f"served pool: tree {k} ({size}^3, kind {kind}) built at {time.time() - t0:.1f}s")
    return specs, trees


def admit_pool(trees, budget: int):
    """Admits the trees in order while their device bytes fit the pool budget
    (itself at most the free device memory less the loop's reserve); returns
    the device trees (None where refused) and the bytes admitted."""
    from shocovox_amd import serve

    dtrees, total = [], 0
    for t in trees:
        if budget - total <= 0:
            log(f"served: tree {t.get_size()}^3 refused: the pool budget is spent")
            dtrees.append(None)
            continue
        try:
            d = serve.admit(t, budget_bytes=budget - total, reserve_bytes=0)
        except Exception as e:  # admission refused: the tree stays on the host
            log(f"served: tree {t.get_size()}^3 refused: {e}")
            dtrees.append(None)
            continue
        dtrees.append(d)
        total += d.device_bytes
    return dtrees, total


def served_bench(args, capacity_mrays: float) -> dict:
    """Served mode over a pool of resident trees admitted by free HBM (and a
    budget): the same seeded Poisson plans (Zipf(1.1) over trees and
    resolutions, 5% sparse edits of 1-1000 voxels) served by per-request
    plain launches and by continuous batching, each mode on its own copy of
    the pool (sessions edit their trees). Per mode: p50/p99 at fixed
    fractions of a nominal capacity, then the largest offered load whose p99
    stays within a latency target (the plain mode's p99 at the target load),
    found by search; the throughput ratio of the modes is taken at that
    equal p99."""
    import torch

    from shocovox_amd import (batch_variants, current_batch_variant, current_render_variant,
                              render_variants, serve)

    t0 = time.time()
    specs, pristine = served_pool(args)
    log(f"served: pool of {len(pristine)} trees built in {time.time() - t0:.1f}s")
    # device memory kept free for the loop's output pools and a margin
    cfg0 = serve.ServeConfig(qps=1.0, duration_s=1.0, seed=args.seed,
                             max_inflight=args.served_inflight,
                             max_batch_rays=args.served_batch_rays)
    reserve = (cfg0.max_inflight * max(cfg0.max_batch_rays, 3840 * 2160) * 24) + (2 << 30)
    free, total = serve.device_memory()
    budget = min(int(args.served_budget_gb * 2 ** 30), free - reserve)
    modes = ("plain", "batched")
    if args.served_kernel_mode:
        modes += ("batched-kernel",)
    results = {}
    admitted = refused = pool_bytes = None
    levels = {}
    for mi, mode in enumerate(modes):
        # every mode but the last runs on a copy of the pristine pool
        t1 = time.time()
        ptrees = [t.clone() for t in pristine] if mi < len(modes) - 1 else pristine
        dtrees, nbytes = admit_pool(ptrees, budget)
        keep = [i for i, d in enumerate(dtrees) if d is not None]
        if admitted is None:
            admitted, refused, pool_bytes = keep, len(ptrees) - len(keep), nbytes
        elif keep != admitted:
            raise SystemExit("served: admission differs between modes")
        ptrees = [ptrees[i] for i in keep]
        dtrees = [dtrees[i] for i in keep]
        log(f"served {mode}: {len(dtrees)} trees admitted ({nbytes / 2**30:.2f} GiB, "
            f"{len(specs) - len(keep)} refused) in {time.time() - t1:.1f}s")
        sizes = [t.get_size() for t in ptrees]
        probe = serve.plan(serve.ServeConfig(qps=1000.0, duration_s=20.0, seed=args.seed), sizes)
        mean_rays = sum(r.width * r.height for r in probe if r.kind == 0) / len(probe)
        qps_cap = capacity_mrays * 1e6 / mean_rays

        def session(load, seconds, seed=args.seed):
            cfg = serve.ServeConfig(qps=round(load * qps_cap, 1), duration_s=seconds, seed=seed,
                                    mode=mode, max_inflight=args.served_inflight,
                                    max_batch_rays=args.served_batch_rays)
            r = serve.run(ptrees, dtrees, cfg)
            return {"load": load, "qps": cfg.qps, "mrays_per_s": round(r["mrays_per_s"], 1),
                    "p50_ms": round(r["p50_ms"], 3), "p99_ms": round(r["p99_ms"], 3),
                    "render_p99_ms": round(r["render_p99_ms"], 3),
                    "edit_p99_ms": round(r["edit_p99_ms"], 3),
                    "edit_phase_p99_ms": {k: round(r[f"edit_{k}_p99_ms"], 3)
                                          for k in ("wait", "apply", "upload")},
                    "requests": int(r["requests"]), "edits": int(r["edits"]),
                    "launches": int(r["launches"]),
                    "jobs_per_launch": round(r["mean_jobs_per_launch"], 2),
                    "single_launches": int(r["single_launches"]),
                    "elapsed_s": round(r["elapsed_s"], 3)}

        # untimed warm-up: the first edits of a tree grow its device arrays
        session(0.6, min(1.0, args.served_seconds), seed=args.seed + 7)
        rows = []
        for frac in args.served_loads:
            row = session(frac, args.served_seconds)
            log(f"served {mode} load {frac}: {row}")
            rows.append(row)
            levels.setdefault(frac, {})[mode] = row
        if mode == "plain":
            ref = min(args.served_loads, key=lambda f: abs(f - args.served_slo_load))
            slo = levels[ref]["plain"]["p99_ms"]
        # the largest load whose p99 stays within the target: grow by 1.3x
        # until it does not, then two bisection steps
        lo, hi, probes = None, None, []
        load = 0.5
        while load <= args.served_max_load:
            row = session(load, args.served_seconds)
            probes.append(row)
            log(f"served {mode} search load {load:.3f}: p99 {row['p99_ms']} ms (target {slo})")
            if row["p99_ms"] <= slo:
                lo = row
                load *= 1.3
            else:
                hi = row
                break
        for _ in range(2 if lo is not None and hi is not None else 0):
            load = math.sqrt(lo["load"] * hi["load"])
            row = session(load, args.served_seconds)
            probes.append(row)
            log(f"served {mode} search load {load:.3f}: p99 {row['p99_ms']} ms (target {slo})")
            if row["p99_ms"] <= slo:
                lo = row
            else:
                hi = row
        results[mode] = {"rows": rows, "search": probes, "at_target": lo,
                         "capacity_qps": lo["qps"] if lo else None,
                         "capacity_mrays_per_s": lo["mrays_per_s"] if lo else None,
                         "capped": lo is not None and hi is None}
        del dtrees
        torch.cuda.synchronize()
        if mi < len(modes) - 1:
            del ptrees
    p, b = results["plain"], results["batched"]
    ratio = (round(b["capacity_mrays_per_s"] / p["capacity_mrays_per_s"], 3)
             if p["capacity_mrays_per_s"] and b["capacity_mrays_per_s"] else None)
    fixed = []
    for frac in args.served_loads:
        row = {"load": frac, **{m: levels[frac][m] for m in modes}}
        row["throughput_ratio"] = round(levels[frac]["batched"]["mrays_per_s"] /
                                        max(levels[frac]["plain"]["mrays_per_s"], 1e-9), 3)
        fixed.append(row)
    sizes = [specs[i][1] for i in admitted]
    return {
        "workload": (f"{len(admitted)} resident trees of {len(specs)} (sizes "
                     f"{min(sizes)}..{max(sizes)} log-uniform, {sizes.count(2048)} of 2048^3; "
                     f"{pool_bytes / 2**30:.2f} GiB admitted by free HBM within a "
                     f"{args.served_budget_gb:g} GiB budget, {refused} refused), Poisson "
                     "arrivals, Zipf(1.1) over tree and resolution {256^2, 720p, 1080p, 4K}, "
                     "5% edits of 1..1000 voxels"),
        "trees": len(admitted), "refused": refused, "pool_bytes": pool_bytes,
        "tree_specs": [list(specs[i]) for i in admitted],
        "device_total_bytes": total,
        "max_inflight": args.served_inflight, "max_batch_rays": args.served_batch_rays,
        "batch_variant": batch_variants()[current_batch_variant()],
        "render_variant": render_variants()[current_render_variant()],
        "arrival_seconds": args.served_seconds,
        "nominal_capacity_qps": round(qps_cap, 1),
        "p99_target_ms": slo, "p99_target_from": f"plain mode's p99 at load {ref}",
        "capacity_at_target": {m: {"qps": results[m]["capacity_qps"],
                                   "mrays_per_s": results[m]["capacity_mrays_per_s"],
                                   "p99_ms": results[m]["at_target"]["p99_ms"]
                                   if results[m]["at_target"] else None,
                                   "search_capped": results[m]["capped"]} for m in modes},
        "throughput_ratio_at_equal_p99": ratio,
        "levels": fixed,
        "search": {m: results[m]["search"] for m in modes},
        "warmup_session": "1 s at load 0.6 per mode, each mode on its own copy of the pool",
    }


def lod_mode(args) -> int:
    """get_by_ray_at_lod on the caves workload with MIP maps on: per step the
    8 poses rendered by render_lod_kernel and 2M incoherent rays cast by
    cast_lod_kernel, at viewing distance --lod-distance (prints one JSON line;
    the figures of the LOD kernels' profiles)."""
    import torch

    from shocovox_amd import DeviceTree, alloc_hits, workloads as W

    cfg = W.CAVES
    t0 = time.time()
    tree = W.build_scene(cfg, seed=args.seed)
    tree.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    d = DeviceTree(tree)
    vps = W.viewports(cfg, seed=args.seed)
    inc = torch.from_numpy(W.incoherent_rays(cfg.size, 2_000_000, seed=args.seed + 4)).to("cuda:0")
    log(f"lod: caves with MIP maps resident in {time.time() - t0:.1f}s")
    vd = args.lod_distance
    fields = ("hit", "albedo", "voxel", "distance")
    outs = [alloc_hits(cfg.width * cfg.height, "cuda:0", fields) for _ in vps]

    def step():
        for vp, o in zip(vps, outs):
            o.update(d.render_at_lod(vp, cfg.width, cfg.height, vd, fields))
        d.cast_at_lod(inc, vd, fields)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t1) / args.steps
    rays = cfg.width * cfg.height * len(vps) + inc.shape[0]
    hits = float(torch.stack([o["hit"].float().mean() for o in outs]).mean())
    print(json.dumps({"mode": "lod", "viewing_distance": vd, "rays_per_step": rays,
                      "ms_per_step": round(el * 1e3, 3), "mrays_per_s": round(rays / el / 1e6, 1),
                      "hit_rate": round(hits, 4)}), flush=True)
    return 0


def spheres_block(args) -> dict:
    """The 2048^3 spheres workload (brick dim 32): one step is a 3840x2160
    render plus 8M incoherent rays (origins uniform in the volume,
    directions uniform on the sphere), in one batched launch (the fastest of
    a few batch variants whose outputs equal the first one's) and as a render
    launch plus a cast launch."""
    import torch

    from shocovox_amd import (Batch, DeviceTree, alloc_hits, batch_variant_scratch,
                              batch_variants, select_batch_variant, workloads as W)

    cfg = W.SPHERES
    t0 = time.time()
    tree = W.build_scene(cfg, seed=args.seed)
    d = DeviceTree(tree)
    del tree
    vp = W.viewports(cfg, seed=args.seed)[0]
    inc = torch.from_numpy(W.incoherent_rays(cfg.size, 8_000_000, seed=args.seed + 6)).to("cuda:0")
    rays = cfg.width * cfg.height + inc.shape[0]
    log(f"spheres: built and resident ({d.device_bytes / 2**30:.2f} GiB) in {time.time() - t0:.1f}s")
    fields = ("hit", "albedo", "voxel", "distance")
    pr, pc = alloc_hits(cfg.width * cfg.height, "cuda:0", fields), alloc_hits(inc.shape[0],
                                                                            "cuda:0", fields)
    stream = torch.cuda.current_stream()
    steps = max(3, args.steps // 4)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / steps

    def launches():
        d.render_into(vp, cfg.width, cfg.height, pr, stream)
        d.cast_into(inc, inc.shape[0], pc, stream)

    per_launch = timed(launches)
    ref = [{k: b[k].clone() for k in fields} for b in (pr, pc)]
    names = batch_variants()
    tuning = {}
    for name in ("classic-w5-g5", "classic-w6-g6", "ww-w5-g5", "ww-w6-g6", "classic-w5-g5-n73"):
        if name not in names:
            continue
        select_batch_variant(names.index(name))
        bt = Batch(2)
        jobs = [Batch.render_job(d, vp, cfg.width, cfg.height, pr), Batch.cast_job(d, inc, pc)]
        el = timed(lambda: bt.run(jobs, stream))
        same = all(torch.equal(b[k], r[k]) for b, r in zip((pr, pc), ref) for k in fields)
        tuning[name] = {"ms": round(el * 1e3, 3), "equal_to_launches": bool(same),
                        "scratch_bytes_per_lane": batch_variant_scratch(names.index(name))}
        log(f"spheres batched {name}: {el * 1e3:.3f} ms, equal {same}")
        del bt
    best = min((n for n in tuning if tuning[n]["equal_to_launches"]),
               key=lambda n: tuning[n]["ms"], default=None)
    out = {"workload": (f"spheres: {cfg.size}^3 octree, brick dim {cfg.brick_dim}, "
                        f"{cfg.width}x{cfg.height} primary rays + 8M incoherent rays per step"),
           "rays_per_step": rays, "device_bytes": d.device_bytes, "steps": steps,
           "launches": {"ms_per_step": round(per_launch * 1e3, 3),
                        "mrays_per_s": round(rays / per_launch / 1e6, 1)},
           "batched": tuning, "best_variant": best}
    if best:
        ms = tuning[best]["ms"]
        out["ms_per_step"] = ms
        out["mrays_per_s"] = round(rays / (ms * 1e-3) / 1e6, 1)
    del d
    torch.cuda.synchronize()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="caves", choices=["caves", "plain", "spheres"],
                    help="caves: the 512^3 1080p workload (default); spheres: the 2048^3 "
                         "workload, one 3840x2160 frame plus 8M incoherent rays per step")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--launch", default="batched", choices=["batched", "per-frame"],
                    help="one batched launch per step (default) or one render launch per pose")
    ap.add_argument("--batch-variant", default="auto",
                    help="batch kernel variant of the batched step (name or index; default: "
                         "autotune)")
    ap.add_argument("--variant", default="auto",
                    help="render launch variant name or index; 'auto' times each and keeps "
                         "the fastest before the timed region")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-mode comparison")
    ap.add_argument("--batched-trees", type=int, default=64)
    ap.add_argument("--batched-max-size", type=int, default=1024)
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--batched-all-variants", action="store_true",
                    help="time the lane-refill batch variants too")
    ap.add_argument("--no-served", action="store_true", help="skip the served-mode sessions")
    ap.add_argument("--no-spheres", action="store_true",
                    help="skip the 2048^3 spheres block (4K + 8M incoherent rays per step)")
    ap.add_argument("--served-trees", type=int, default=32)
    ap.add_argument("--served-max-size", type=int, default=2048)
    ap.add_argument("--served-budget-gb", type=float, default=40.0,
                    help="device memory the served pool may take (admission stops there)")
    ap.add_argument("--served-seconds", type=float, default=1.5)
    ap.add_argument("--served-inflight", type=int, default=2)
    # launches of 8M rays: a request of up to 4K completes in one launch, so a
    # popular tree goes idle often enough for its edits to start promptly
    ap.add_argument("--served-batch-rays", type=int, default=1 << 23)
    ap.add_argument("--served-loads", type=float, nargs="+", default=[0.3, 0.6, 0.9, 1.2],
                    help="offered load as fractions of the measured render capacity")
    ap.add_argument("--served-slo-load", type=float, default=0.6,
                    help="the p99 target is the plain mode's p99 at this load")
    ap.add_argument("--served-max-load", type=float, default=4.0,
                    help="upper end of the capacity search, as a fraction of capacity")
    ap.add_argument("--served-kernel-mode", action="store_true",
                    help="also serve with the persistent kernel for every launch")
    ap.add_argument("--mode", default="render", choices=["render", "copy", "lod"],
                    help="'copy' runs a plain device-to-device copy of known size (for "
                         "calibrating rocprofv3 byte counters); 'lod' times get_by_ray_at_lod "
                         "on the caves workload with MIP maps on")
    ap.add_argument("--lod-distance", type=float, default=64.0,
                    help="viewing distance of --mode lod")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: every box here has "
              "a single MI355X and this build runs no multi-GPU path (rays or trees would shard "
              "one per GPU with no exchange on the hot path).", file=sys.stderr)
        return 2

    import torch

    if args.mode == "copy":
        return copy_calibration(args)
    if args.mode == "lod":
        return lod_mode(args)

    from shocovox_amd import (Batch, DeviceTree, alloc_hits, batch_variant_scratch,
                              batch_variants, collect_kernel_ms, current_batch_variant,
                              render_variant_scratch, render_variants, select_batch_variant,
                              select_render_variant, set_timing, workloads as W)

    cfg = W.CONFIGS[args.workload]
    t0 = time.time()
    tree = W.build_scene(cfg, seed=args.seed)
    vps = W.viewports(cfg, seed=args.seed)
    log(f"scene {cfg.name} built in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    dtree = DeviceTree(tree)
    rays_per_frame = cfg.width * cfg.height
    rays_per_step = rays_per_frame * len(vps)
    fields = ("hit", "albedo", "voxel", "distance")
    bufs = [alloc_hits(rays_per_frame, dev, fields) for _ in vps]
    stream = torch.cuda.current_stream()

    incoherent = None
    inc_host = None
    if cfg.name == "spheres":
        # the 2048^3 workload adds 8M incoherent rays (origins in the volume,
        # directions uniform on the sphere) to every step
        inc_host = W.incoherent_rays(cfg.size, 8_000_000, seed=args.seed + 6)
        incoherent = (torch.from_numpy(inc_host).to(dev), alloc_hits(len(inc_host), dev, fields))
        rays_per_step += len(inc_host)

    def per_frame_step():
        for vp, b in zip(vps, bufs):
            dtree.render_into(vp, cfg.width, cfg.height, b, stream)
        if incoherent is not None:
            dtree.cast_into(incoherent[0], incoherent[0].shape[0], incoherent[1], stream)

    # one launch per pose: autotune the render kernel's launch variants
    # (builds that spill registers to scratch are not candidates)
    names = render_variants()
    render_scratch = {n: render_variant_scratch(i) for i, n in enumerate(names)}
    tuning = {}
    if args.variant == "auto":
        set_timing(True)
        for i, name in enumerate(names):
            if render_scratch[name]:
                continue
            select_render_variant(i)
            per_frame_step()
            collect_kernel_ms()
            per_frame_step()
            per_frame_step()
            ms = collect_kernel_ms()
            if incoherent is not None:
                ms = ms[0::2]  # the render launches; the casts do not depend on the variant
            tuning[name] = round(sum(ms) / max(1, len(ms)), 4)
        set_timing(False)
        best = min(tuning, key=tuning.get)
        log(f"autotune (avg ms per frame): {tuning} -> {best}")
    else:
        best = names[int(args.variant)] if args.variant.isdigit() else args.variant
    select_render_variant(names.index(best))

    def time_steps(step, n):
        """Runs `step` n times between events on the launch stream; returns
        (wall seconds, average event-timed ms per step)."""
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(n)]
        t0 = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            step()
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, sum(a.elapsed_time(b) for a, b in evs) / max(1, n)

    for _ in range(args.warmup):
        per_frame_step()
    pf_wall, pf_ms = time_steps(per_frame_step, max(3, args.steps // 2))
    pf_steps = max(3, args.steps // 2)
    per_frame = {"render_variant": best, "ms_per_step": round(pf_wall / pf_steps * 1e3, 4),
                 "mrays_per_s": round(rays_per_step * pf_steps / pf_wall / 1e6, 1),
                 "launches_per_step": len(vps) + (incoherent is not None)}
    log(f"per-frame launches: {per_frame}")
    # reference outputs of the step, for the batched launch's equality check
    per_frame_out = [{k: b[k].clone() for k in fields} for b in bufs]
    if incoherent is not None:
        per_frame_out.append({k: incoherent[1][k].clone() for k in fields})

    batch = None
    batch_tuning = {}
    bnames = batch_variants()
    batch_scratch = {n: batch_variant_scratch(i) for i, n in enumerate(bnames)}
    if args.launch == "batched":
        outs = bufs + ([incoherent[1]] if incoherent is not None else [])

        def make(vi):
            select_batch_variant(vi)
            bt = Batch(len(outs))
            jobs = [Batch.render_job(dtree, vp, cfg.width, cfg.height, b)
                    for vp, b in zip(vps, bufs)]
            if incoherent is not None:
                jobs.append(Batch.cast_job(dtree, incoherent[0], incoherent[1]))
            return bt, jobs

        if args.batch_variant == "auto":
            for vi, name in enumerate(bnames):
                if name.startswith("r1-"):
                    continue  # lane refill: measured 2x slower on coherent frames
                if batch_scratch[name]:
                    continue  # spills registers to scratch
                bt, jobs = make(vi)
                bt.run(jobs, stream)
                _, ms = time_steps(lambda: bt.run(jobs, stream), 3)
                batch_tuning[name] = round(ms, 4)
                del bt
            bvariant = min(batch_tuning, key=batch_tuning.get)
            log(f"batch autotune (ms per step): {batch_tuning} -> {bvariant}")
        else:
            bvariant = (bnames[int(args.batch_variant)] if args.batch_variant.isdigit()
                        else args.batch_variant)
        batch, jobs = make(bnames.index(bvariant))

        def step():
            batch.run(jobs, stream)
    else:
        bvariant = None
        step = per_frame_step

    for _ in range(args.warmup):
        step()
    elapsed, avg_launch_ms = time_steps(step, args.steps)
    ms_per_step = elapsed / args.steps * 1e3
    value = rays_per_step * args.steps / elapsed / 1e6
    equal = None
    if batch is not None:
        outs = bufs + ([incoherent[1]] if incoherent is not None else [])
        equal = all(torch.equal(o[k], r[k]) for o, r in zip(outs, per_frame_out) for k in fields)
        if not equal:
            raise SystemExit("batched launch outputs differ from the per-frame launches")
    hits = float(torch.stack([b["hit"].float().mean() for b in bufs]).mean())
    launch_desc = (f"one batched launch per step ({bvariant}: {len(vps)} render jobs"
                   f"{' + 1 cast job' if incoherent is not None else ''})" if batch is not None
                   else f"{per_frame['launches_per_step']} launches per step ({best})")
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} Mrays/s, "
        f"{launch_desc}, event-timed {avg_launch_ms:.4f} ms per step, hit rate {hits:.3f}")

    roofline = None
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, per_ray, otree, prays = cpu_baseline(cfg, tree, vps, args.cpu_seconds, inc_host)
        b_ray = sum(EVENT_BYTES[k] * v for k, v in per_ray.items()) + OUT_BYTES_PER_RAY
        # the timed unit: one batched launch (all rays of the step), or one
        # per-frame render launch (the render rays of one pose)
        if batch is not None:
            rays_per_launch, launch_ms, kname = rays_per_step, avg_launch_ms, "batch_kernel"
        else:
            rays_per_launch = rays_per_frame
            launch_ms = tuning.get(best) if tuning else pf_ms / len(vps)
            kname = "render_kernel"
        achieved = b_ray * rays_per_launch / (launch_ms * 1e-3) / 1e9
        traffic = None
        counters = {}
        issue = None
        measured_frac = None
        # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
        # workload, scaled by the same counters' reading of a copy of known size;
        # VALU instructions per launch from a SQ_INSTS_VALU run
        tf = ROOT / "profiles" / f"{kname}_{cfg.name}_{cfg.height}p_traffic.json"
        if tf.exists():
            tj = json.loads(tf.read_text())
            if tj.get("units_per_launch") == rays_per_launch:
                traffic = tj.get("bytes_per_launch")
                counters = {k: tj[k] for k in ("l2_hit_rate", "valu_lane_utilization")
                            if k in tj}
                if traffic:
                    measured_frac = traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                if tj.get("valu_insts_per_launch"):
                    gips = tj["valu_insts_per_launch"] / (launch_ms * 1e-3) / 1e9
                    issue = {"achieved": round(gips, 1), "peak": VALU_PEAK_GIPS,
                             "unit": "G wave64 VALU instructions/s",
                             "frac": round(gips / VALU_PEAK_GIPS, 4),
                             "valu_insts_per_launch": tj["valu_insts_per_launch"],
                             "counters_from": tf.name}
        # the limiter named by the counters: whichever ceiling the kernel is
        # closest to, VALU issue or measured HBM bytes
        limiter = None
        if issue is not None or measured_frac is not None:
            limiter = ("valu-issue" if (issue["frac"] if issue else 0.0) >= (measured_frac or 0.0)
                       else "hbm")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "kernel": kname,
                    "frac_kind": "logical: algorithmic bytes of the reference's event "
                                 "sequence per ray x rays / launch time",
                    "measured_hbm_frac": (round(measured_frac, 4) if measured_frac is not None
                                          else None),
                    "limiter": limiter, "issue": issue,
                    "bytes_per_ray": round(b_ray, 1),
                    "events_per_ray": {k: round(v, 3) for k, v in per_ray.items()},
                    "avg_launch_ms": round(launch_ms, 4),
                    "rays_per_launch": rays_per_launch, **counters}
        log(f"cpu baseline {cpu}")
        t0 = time.time()
        parity = step_parity(otree, prays, bufs,
                             (inc_host, [incoherent[1]]) if incoherent is not None else None)
        log(f"step parity against the oracle ({time.time() - t0:.1f}s): {parity}")
        del otree

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {cfg.size}^3 octree, brick dim {cfg.brick_dim}, "
                               f"{cfg.width}x{cfg.height} primary rays x {len(vps)} poses"
                               + (", + 8M incoherent rays" if incoherent is not None else ""),
                   "octree_size": cfg.size, "brick_dim": cfg.brick_dim, "width": cfg.width,
                   "height": cfg.height, "poses": len(vps), "rays_per_step": rays_per_step,
                   "seed": args.seed, "hit_rate": round(hits, 4),
                   "device_bytes": dtree.device_bytes, "launch": launch_desc,
                   "batch_variant": bvariant, "batch_autotune_ms": batch_tuning or None,
                   "outputs_equal_to_per_frame": equal, "per_frame": per_frame,
                   "launch_variant": best, "autotune_ms": tuning or None,
                   "scratch_bytes_per_lane": {
                       "batch": {n: b for n, b in batch_scratch.items() if b},
                       "render": {n: b for n, b in render_scratch.items() if b}},
                   "tree": W.tree_statistics(tree)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if not args.no_batched:
        line["batched"] = batched_bench(args, best)
        # the served sessions' plain launches use the render variant that
        # was fastest on the batched workload's mix of trees
        select_render_variant(render_variants().index(line["batched"]["plain"]["render_variant"]))
    if not args.no_served:
        # offered loads as fractions of the one-launch-per-frame capacity
        line["served"] = served_bench(args, per_frame["mrays_per_s"])
    if not args.no_spheres and args.workload == "caves":
        prev = current_batch_variant()
        line["spheres"] = spheres_block(args)
        select_batch_variant(prev)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
