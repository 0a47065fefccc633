#!/usr/bin/env python3
"""Throughput of octree ray traversal on one MI355X.

Workload (default "caves"): a 512^3 octree with brick dim 16 (seeded porous
noise terrain with worm caves; 40% Empty / 10% Solid / rest Parted brick
slots), 1920x1080 primary rays from each of 8 seeded orbit poses. One step
renders all 8 poses on the device (rays generated in the kernel, 24 bytes of
output per ray: hit, voxel xyz, albedo, distance); the tree is resident in
HBM before timing starts. Prints one JSON line.

The roofline object prices the traversal kernel against HBM bandwidth with
the algorithmic bytes per ray counted by the instrumented CPU restatement on
a sample of the same rays; the cpu_baseline object times that restatement
on the host cores.
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


def cpu_baseline(cfg, tree, vps, seconds: float) -> tuple[dict, dict]:
    """Times the CPU restatement (single thread, then all cores) on a bounded
    sample of the same rays and returns (cpu_baseline, per-ray event counts)."""
    import numpy as np

    import oracle

    t0 = time.time()
    otree = oracle.OracleTree(tree.export_model())
    log(f"cpu baseline: oracle tree loaded in {time.time() - t0:.1f}s")
    rays = np.concatenate([oracle.viewport_rays(v.origin, v.direction, v.frustum, v.fov,
                                                cfg.width, cfg.height) for v in vps])
    # event counts on a strided sample (instrumented run)
    _, cnt = otree.cast(rays[::97], counters=True)
    n = cnt.rays
    per_ray = {k: getattr(cnt, k) / n for k in EVENT_BYTES}
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="caves", choices=["caves", "plain"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", default="auto",
                    help="render launch variant name or index; 'auto' times each and keeps "
                         "the fastest before the timed region")
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

    from shocovox_amd import (DeviceTree, alloc_hits, collect_kernel_ms, render_variants,
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

    def step():
        for vp, b in zip(vps, bufs):
            dtree.render_into(vp, cfg.width, cfg.height, b, stream)

    names = render_variants()
    tuning = {}
    if args.variant == "auto":
        # autotune: time every launch variant on the same frames, keep the fastest
        set_timing(True)
        for i, name in enumerate(names):
            select_render_variant(i)
            step()
            collect_kernel_ms()
            step()
            step()
            ms = collect_kernel_ms()
            tuning[name] = round(sum(ms) / max(1, len(ms)), 4)
        set_timing(False)
        best = min(tuning, key=tuning.get)
        log(f"autotune (avg ms per frame): {tuning} -> {best}")
    else:
        best = names[int(args.variant)] if args.variant.isdigit() else args.variant
    select_render_variant(names.index(best))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    set_timing(True)
    collect_kernel_ms()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    set_timing(False)
    kernel_ms = collect_kernel_ms(cap=args.steps * len(vps) + 16)
    ms_per_step = elapsed / args.steps * 1e3
    value = rays_per_step * args.steps / elapsed / 1e6
    avg_launch_ms = sum(kernel_ms) / max(1, len(kernel_ms))
    hits = float(torch.stack([b["hit"].float().mean() for b in bufs]).mean())
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} Mrays/s, "
        f"avg launch {avg_launch_ms:.4f} ms over {len(kernel_ms)} launches, hit rate {hits:.3f}")

    roofline = None
    cpu = None
    if not args.no_cpu_baseline:
        cpu, per_ray = cpu_baseline(cfg, tree, vps, args.cpu_seconds)
        b_ray = sum(EVENT_BYTES[k] * v for k, v in per_ray.items()) + OUT_BYTES_PER_RAY
        achieved = b_ray * rays_per_frame / (avg_launch_ms * 1e-3) / 1e9
        traffic = None
        # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
        # workload, scaled by the same counters' reading of a copy of known size
        tf = ROOT / "profiles" / f"render_kernel_{cfg.name}_{cfg.height}p_traffic.json"
        if tf.exists():
            tj = json.loads(tf.read_text())
            if tj.get("units_per_launch") == rays_per_frame:
                traffic = tj.get("bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "kernel": "render_kernel",
                    "bytes_per_ray": round(b_ray, 1),
                    "events_per_ray": {k: round(v, 3) for k, v in per_ray.items()},
                    "avg_launch_ms": round(avg_launch_ms, 4),
                    "rays_per_launch": rays_per_frame}
        log(f"cpu baseline {cpu}")

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {cfg.size}^3 octree, brick dim {cfg.brick_dim}, "
                               f"{cfg.width}x{cfg.height} primary rays x {len(vps)} poses",
                   "octree_size": cfg.size, "brick_dim": cfg.brick_dim, "width": cfg.width,
                   "height": cfg.height, "poses": len(vps), "rays_per_step": rays_per_step,
                   "seed": args.seed, "hit_rate": round(hits, 4),
                   "device_bytes": dtree.device_bytes, "launch_variant": best,
                   "autotune_ms": tuning or None},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
