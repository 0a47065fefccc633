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
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Mrays/s at 1080p into 512³; p50/p99 latency batched vs per-call at fixed QPS")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one logical event of the reference traversal
EVENT_BYTES = {"stack_iterations": 12, "child_reads": 4, "brick_reads": 4, "voxel_tests": 8,
               "data_palette_reads": 4, "hits": 4}
OUT_BYTES_PER_RAY = 24  # hit + voxel xyz + albedo + distance


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cfg, tree, vps, seconds: float, extra_rays=None) -> tuple[dict, dict]:
    """Times the CPU restatement (single thread, then all cores) on a bounded
    sample of the same rays and returns (cpu_baseline, per-ray event counts of
    the step's rays: the primary rays plus `extra_rays` if given)."""
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
    return base, per_ray


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


BATCH_RESOLUTIONS = ((128, 128), (512, 512), (1280, 720), (1920, 1080))


def batched_bench(args, render_variant: str) -> dict:
    """Batched mode: trees of uneven size (log-uniform 64..max), each with 1-4
    viewports of resolution drawn from {128^2, 512^2, 720p, 1080p}; all rays in
    one persistent launch per batch kernel variant (a replayed hipGraph),
    against one plain render launch per viewport. Outputs of the two forms are
    compared for equality."""
    import math

    import numpy as np
    import torch

    from shocovox_amd import (Batch, DeviceTree, Octree, alloc_hits, batch_variants,
                              select_batch_variant, workloads as W)

    rng = np.random.default_rng(args.seed + 700)
    max_exp = int(np.log2(args.batched_max_size))
    t0 = time.time()
    trees, dtrees, views = [], [], []
    for i in range(args.batched_trees):
        size = int(2 ** rng.integers(6, max_exp + 1))
        dim = int(rng.choice([4, 8, 16]))
        kind = int(rng.integers(1, 3)) if size <= 128 else 2
        t = Octree.scene(kind, size, dim, args.seed * 1000 + 500 + i)
        trees.append(t)
        dtrees.append(DeviceTree(t))
        c = size / 2.0
        for _ in range(int(rng.integers(1, 5))):
            w, h = BATCH_RESOLUTIONS[int(rng.integers(0, len(BATCH_RESOLUTIONS)))]
            az = float(rng.uniform(0, 2 * math.pi))
            eye = (c + 1.3 * size * math.cos(az), 0.9 * size, c + 1.3 * size * math.sin(az))
            views.append((i, W.look_at_viewport(eye, (c, c / 2, c), w, h), w, h))
    rays = sum(w * h for _, _, w, h in views)
    log(f"batched: {len(trees)} trees, {len(views)} viewports, {rays} rays, built in "
        f"{time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    fields = ("hit", "albedo", "voxel", "distance")
    plain_bufs = [alloc_hits(w * h, dev, fields) for _, _, w, h in views]
    batch_bufs = [alloc_hits(w * h, dev, fields) for _, _, w, h in views]
    stream = torch.cuda.current_stream()
    reps = max(1, args.batched_reps)

    def plain_pass():
        for (ti, vp, w, h), b in zip(views, plain_bufs):
            dtrees[ti].render_into(vp, w, h, b, stream)

    plain_pass()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t1) / reps
    out = {"workload": (f"{len(trees)} trees (sizes 64..{args.batched_max_size}, brick dims "
                        f"4/8/16), {len(views)} viewports of 128^2/512^2/720p/1080p"),
           "trees": len(trees), "viewports": len(views), "rays": rays,
           "plain": {"ms": round(plain_s * 1e3, 3), "mrays_per_s": round(rays / plain_s / 1e6, 1),
                     "launches": len(views), "render_variant": render_variant},
           "batched": {}}
    names = batch_variants()
    best = None
    for vi, name in enumerate(names):
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
    return out


def served_bench(args, capacity_mrays: float) -> dict:
    """Served mode: a pool of resident trees admitted by free HBM, then the
    same seeded Poisson plan (Zipf(1.1) over trees and resolutions, 5% sparse
    edits of 1-1000 voxels) served by continuous batching (mode "batched",
    which runs a launch holding one whole request as a plain render, and
    "batched-kernel", which never does) and by per-request plain launches, at
    fixed fractions of the measured render capacity."""
    import numpy as np

    from shocovox_amd import Octree, batch_variants, current_batch_variant, serve

    rng = np.random.default_rng(args.seed + 500)
    max_exp = int(np.log2(args.served_max_size))
    specs = []
    for i in range(args.served_trees):
        size = int(2 ** rng.integers(6, max_exp + 1))  # log-uniform over 64..max
        dim = int(rng.choice([4, 8, 16]))
        kind = int(rng.integers(1, 3)) if size <= 128 else 2
        specs.append((kind, size, dim, int(args.seed * 1000 + i)))
    cfg0 = serve.ServeConfig(qps=1.0, duration_s=1.0, seed=args.seed,
                             max_inflight=args.served_inflight,
                             max_batch_rays=args.served_batch_rays)
    modes = ("batched", "plain", "batched-kernel")
    # device memory kept free for the loop's output pools and a margin; every
    # mode gets its own copy of the pool, so each starts from the same trees
    # and sees the same edits in the same order (sessions mutate the trees)
    reserve = (cfg0.max_inflight * max(cfg0.max_batch_rays, 3840 * 2160) * 24) + (1 << 30)
    t0 = time.time()
    trees, dtrees, refused, pool_bytes = [], [], 0, 0
    for kind, size, dim, seed in specs:
        t = Octree.scene(kind, size, dim, seed)
        try:
            d = serve.admit(t, reserve_bytes=reserve)
        except Exception as e:  # admission refused: the tree stays on the host
            log(f"served: tree {size}^3 refused: {e}")
            refused += 1
            continue
        trees.append(t)
        dtrees.append(d)
        pool_bytes += d.device_bytes
    log(f"served: {len(trees)} trees resident ({pool_bytes / 2**30:.2f} GiB) in "
        f"{time.time() - t0:.1f}s, {refused} refused")
    pools = {modes[0]: (trees, dtrees)}
    for mode in modes[1:]:
        copies = [t.clone() for t in trees]
        pools[mode] = (copies, [serve.admit(t, reserve_bytes=reserve) for t in copies])
    sizes = [t.get_size() for t in trees]
    probe = serve.plan(serve.ServeConfig(qps=1000.0, duration_s=20.0, seed=args.seed), sizes)
    mean_rays = sum(r.width * r.height for r in probe if r.kind == 0) / len(probe)
    qps_cap = capacity_mrays * 1e6 / mean_rays
    # Untimed warm-up session per pool (the served counterpart of --warmup):
    # the first edits of a tree grow its device arrays and fault in its host
    # mirror, a one-time cost that would otherwise land on one measured session.
    t0 = time.time()
    for mode, (ptrees, pdtrees) in pools.items():
        serve.run(ptrees, pdtrees, serve.ServeConfig(
            qps=round(0.6 * qps_cap, 1), duration_s=min(1.0, args.served_seconds),
            seed=args.seed + 7, mode=mode, max_inflight=args.served_inflight,
            max_batch_rays=args.served_batch_rays))
    log(f"served: warm-up sessions in {time.time() - t0:.1f}s")
    levels = []
    for frac in args.served_loads:
        qps = round(frac * qps_cap, 1)
        row = {"load": frac, "qps": qps}
        for mode in modes:
            ptrees, pdtrees = pools[mode]
            cfg = serve.ServeConfig(qps=qps, duration_s=args.served_seconds, seed=args.seed,
                                    mode=mode, max_inflight=args.served_inflight,
                                    max_batch_rays=args.served_batch_rays)
            # a mode's sessions run back to back on its pool; their edits accumulate
            r = serve.run(ptrees, pdtrees, cfg)
            row[mode] = {"mrays_per_s": round(r["mrays_per_s"], 1),
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
            log(f"served load {frac} ({qps} req/s) {mode}: {row[mode]}")
        b, p = row["batched"], row["plain"]
        row["throughput_ratio"] = round(b["mrays_per_s"] / max(p["mrays_per_s"], 1e-9), 3)
        levels.append(row)
    return {
        "workload": (f"{len(trees)} resident trees (sizes {min(sizes)}..{max(sizes)}, brick dims "
                     "4/8/16), Poisson arrivals, Zipf(1.1) over tree and resolution "
                     "{256^2, 720p, 1080p, 4K}, 5% edits of 1..1000 voxels"),
        "trees": len(trees), "refused": refused, "pool_bytes": pool_bytes,
        "mean_rays_per_request": round(mean_rays), "capacity_qps": round(qps_cap, 1),
        "max_inflight": args.served_inflight, "max_batch_rays": args.served_batch_rays,
        "batch_variant": batch_variants()[current_batch_variant()],
        "arrival_seconds": args.served_seconds,
        "warmup_session": "1 s at load 0.6 per mode, each mode on its own copy of the pool",
        "levels": levels,
    }


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
    ap.add_argument("--batched-max-size", type=int, default=512)
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--no-served", action="store_true", help="skip the served-mode sessions")
    ap.add_argument("--served-trees", type=int, default=32)
    ap.add_argument("--served-max-size", type=int, default=512)
    ap.add_argument("--served-seconds", type=float, default=2.0)
    ap.add_argument("--served-inflight", type=int, default=2)
    # launches of 8M rays: a request of up to 4K completes in one launch, so a
    # popular tree goes idle often enough for its edits to start promptly
    ap.add_argument("--served-batch-rays", type=int, default=1 << 23)
    ap.add_argument("--served-loads", type=float, nargs="+", default=[0.3, 0.6, 0.9, 1.2],
                    help="offered load as fractions of the measured render capacity")
    ap.add_argument("--mode", default="render", choices=["render", "copy"],
                    help="'copy' runs a plain device-to-device copy of known size (for "
                         "calibrating rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: every box here has "
              "a single MI355X and this build runs no multi-GPU path (rays or trees would shard "
              "one per GPU with no exchange on the hot path).", file=sys.stderr)
        return 2

    import torch

    if args.mode == "copy":
        return copy_calibration(args)

    from shocovox_amd import (Batch, DeviceTree, alloc_hits, batch_variants, collect_kernel_ms,
                              render_variants, select_batch_variant, select_render_variant,
                              set_timing, workloads as W)

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
    names = render_variants()
    tuning = {}
    if args.variant == "auto":
        set_timing(True)
        for i, name in enumerate(names):
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
    if not args.no_cpu_baseline:
        cpu, per_ray = cpu_baseline(cfg, tree, vps, args.cpu_seconds, inc_host)
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
        # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
        # workload, scaled by the same counters' reading of a copy of known size
        tf = ROOT / "profiles" / f"{kname}_{cfg.name}_{cfg.height}p_traffic.json"
        if tf.exists():
            tj = json.loads(tf.read_text())
            if tj.get("units_per_launch") == rays_per_launch:
                traffic = tj.get("bytes_per_launch")
                counters = {k: tj[k] for k in ("l2_hit_rate", "valu_lane_utilization")
                            if k in tj}
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "kernel": kname,
                    "bytes_per_ray": round(b_ray, 1),
                    "events_per_ray": {k: round(v, 3) for k, v in per_ray.items()},
                    "avg_launch_ms": round(launch_ms, 4),
                    "rays_per_launch": rays_per_launch, **counters}
        log(f"cpu baseline {cpu}")

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
                   "launch_variant": best, "autotune_ms": tuning or None},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.no_batched:
        line["batched"] = batched_bench(args, best)
    if not args.no_served:
        # offered loads as fractions of the one-launch-per-frame capacity
        line["served"] = served_bench(args, per_frame["mrays_per_s"])
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
