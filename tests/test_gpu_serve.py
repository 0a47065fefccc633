"""Served mode on the GPU.

Each render request's result is the same alone and in the mix: a session
served by continuous batching (chunked, tile-major batched launches) and a
session of one plain launch per request, run on separate copies of the same
trees with the same plan, produce equal per-request digests; a host replay of
the plan through the oracle (edits applied in arrival order) produces the
same digests too, and all three copies end in the same octree.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import oracle
from parity import result_digest
from shocovox_amd import DeviceTree, Octree, serve

pytestmark = pytest.mark.gpu

# (scene kind, size, brick dim, seed)
TREES = [(2, 64, 4, 5), (1, 64, 8, 7), (2, 128, 8, 6), (2, 32, 2, 8)]
SMALL_RES = ((64, 64), (96, 80), (128, 128), (200, 120))


def _say(msg: str) -> None:
    print(msg, flush=True)


def _trees():
    return [Octree.scene(k, s, d, seed) for k, s, d, seed in TREES]


def _models_equal(a: Octree, b: Octree) -> bool:
    ma, mb = a.export_model(), b.export_model()
    return ma.keys() == mb.keys() and all(
        np.array_equal(np.asarray(ma[k]), np.asarray(mb[k])) for k in ma)


def _oracle_digests(cfg: serve.ServeConfig, trees: list) -> list:
    """Replays the plan on host octrees: edits in arrival order, every render
    cast by the oracle against the tree state at its arrival."""
    plan = serve.plan(cfg, [t.get_size() for t in trees])
    cache = {}
    out = []
    for r in plan:
        if r.kind == 1:
            serve.apply_edit(trees[r.tree], r)
            cache.pop(r.tree, None)
            out.append(0)
            continue
        if r.tree not in cache:
            cache[r.tree] = oracle.OracleTree(trees[r.tree].export_model())
        v = r.viewport
        rays = oracle.viewport_rays(v.origin, v.direction, v.frustum, v.fov, r.width, r.height)
        out.append(result_digest(cache[r.tree].cast(rays, threads=8)))
    return out


@pytest.mark.parametrize("load", ["steady", "burst"])
def test_served_results_same_alone_and_in_mix(gpu_device, load):
    # steady: arrivals rarely overlap; burst: ~400 requests within 4 ms, so
    # launches mix chunks of many requests and trees
    qps, dur = (400, 0.5) if load == "steady" else (100000, 0.004)
    cfg = serve.ServeConfig(qps=qps, duration_s=dur, seed=9, max_inflight=3,
                            max_batch_rays=8192, max_batch_jobs=64, edit_fraction=0.15,
                            max_edit_voxels=300, resolutions=SMALL_RES)
    plan = serve.plan(cfg, [s for _, s, _, _ in TREES])
    n_edit = sum(r.kind for r in plan)
    _say(f"serve plan ({load}): {len(plan)} requests, {n_edit} edits")
    assert n_edit > 5 and len(plan) > 100

    a, b, c, k = _trees(), _trees(), _trees(), _trees()
    _say("serve: batched session")
    ra = serve.run(a, [DeviceTree(t) for t in a], cfg, digests=True)
    _say(f"batched: {ra['launches']} launches, {ra['mean_jobs_per_launch']:.2f} chunks/launch, "
         f"p50 {ra['p50_ms']:.2f} ms p99 {ra['p99_ms']:.2f} ms")
    _say("serve: plain session")
    rb = serve.run(b, [DeviceTree(t) for t in b], dataclasses.replace(cfg, mode="plain"),
                   digests=True)
    _say(f"plain: {rb['launches']} launches, p50 {rb['p50_ms']:.2f} ms p99 {rb['p99_ms']:.2f} ms")
    _say("serve: batched-kernel session")
    rk = serve.run(k, [DeviceTree(t) for t in k], dataclasses.replace(cfg, mode="batched-kernel"),
                   digests=True)
    _say(f"batched-kernel: {rk['launches']} launches, single {ra['single_launches']} in mode 0")
    assert rk["single_launches"] == 0 and rb["single_launches"] == 0
    if load == "steady":
        assert ra["single_launches"] > 0
    for r in (ra, rb, rk):
        assert r["requests"] == len(plan) and r["rejected"] == 0
        assert r["renders"] + r["edits"] == len(plan) and r["edits"] == n_edit
        assert r["rays"] == sum(q.width * q.height for q in plan if q.kind == 0)
        assert (r["latencies_ms"] >= 0).all()
    # requests larger than a launch's budget were cut into several chunks
    chunks = ra["mean_jobs_per_launch"] * ra["launches"]
    assert chunks > ra["renders"] + 0.5
    if load == "burst":
        assert ra["mean_jobs_per_launch"] > 1.5
    assert np.array_equal(ra["digests"], rb["digests"])
    assert np.array_equal(ra["digests"], rk["digests"])
    assert all(int(d) != 0 for d, q in zip(ra["digests"], plan) if q.kind == 0)

    _say("serve: oracle replay")
    expect = _oracle_digests(cfg, c)
    got = [int(x) for x in ra["digests"]]
    bad = [i for i, (e, g) in enumerate(zip(expect, got)) if e != g]
    assert not bad, f"{len(bad)} requests differ from the oracle replay, first {bad[:5]}"
    for i in range(len(TREES)):
        assert _models_equal(a[i], c[i]) and _models_equal(b[i], c[i]), i
        assert _models_equal(k[i], c[i]), i


def test_served_queue_limit_rejects_and_accounts(gpu_device):
    """A burst far above capacity with a queue limit of 2: arrivals that find
    the queue full are rejected (latency -1), the rest complete."""
    cfg = serve.ServeConfig(qps=20000, duration_s=0.05, seed=4, max_inflight=1,
                            max_batch_rays=4096, max_batch_jobs=8, max_queue=2,
                            edit_fraction=0.0, resolutions=((256, 256),))
    trees = _trees()
    _say("serve: queue-limit burst")
    r = serve.run(trees, [DeviceTree(t) for t in trees], cfg)
    lat = r["latencies_ms"]
    assert r["rejected"] > 0 and r["rejected"] == int((lat < 0).sum())
    assert r["renders"] + r["rejected"] == r["requests"]
    assert r["rays"] == r["renders"] * 256 * 256


def test_admission_by_free_device_memory(gpu_device):
    free, total = serve.device_memory()
    assert 0 < free <= total
    t = Octree.scene(2, 128, 8, 2)
    d = serve.admit(t, reserve_bytes=1 << 20)
    assert d.device_bytes > 0
    with pytest.raises(Exception) as e:
        serve.admit(t, reserve_bytes=total)
    assert "device bytes" in str(e.value)


def test_served_deadline_drops_late_requests(gpu_device):
    """Requests not started within the deadline are dropped (latency -2) and
    counted; the rest complete."""
    cfg = serve.ServeConfig(qps=20000, duration_s=0.05, seed=6, max_inflight=1,
                            max_batch_rays=1 << 16, max_batch_jobs=8, deadline_ms=1.0,
                            edit_fraction=0.2, max_edit_voxels=50, resolutions=((256, 256),))
    trees = _trees()
    _say("serve: deadline burst")
    r = serve.run(trees, [DeviceTree(t) for t in trees], cfg)
    lat = r["latencies_ms"]
    assert r["timed_out"] > 0 and r["timed_out"] == int((lat == -2).sum())
    assert r["renders"] + r["edits"] + r["timed_out"] + r["rejected"] == r["requests"]
    assert (lat[lat >= 0] >= 0).all() and r["renders"] > 0


def test_served_with_a_non_8x8_batch_variant_selected(gpu_device):
    """The loop's chunked launches are laid out in 8x8 tiles: with a variant of
    16x4 tiles selected, serving falls back to an 8x8 variant and its results
    still equal per-request launches."""
    from shocovox_amd import batch_variants, current_batch_variant, select_batch_variant
    names = batch_variants()
    prev = current_batch_variant()
    cfg = serve.ServeConfig(qps=100000, duration_s=0.003, seed=12, max_inflight=2,
                            max_batch_rays=8192, max_batch_jobs=64, edit_fraction=0.1,
                            max_edit_voxels=100, resolutions=SMALL_RES)
    a, b = _trees(), _trees()
    _say("serve: burst with classic-w5-g5-16x4 selected")
    select_batch_variant(names.index("classic-w5-g5-16x4"))
    try:
        ra = serve.run(a, [DeviceTree(t) for t in a], cfg, digests=True)
    finally:
        select_batch_variant(prev)
    rb = serve.run(b, [DeviceTree(t) for t in b], dataclasses.replace(cfg, mode="plain"),
                   digests=True)
    assert ra["mean_jobs_per_launch"] > 1.5
    assert np.array_equal(ra["digests"], rb["digests"])
