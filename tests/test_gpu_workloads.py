"""GPU parity on the benchmark workloads at their configured shapes, and
re-uploads that change a tree's brick dimension.

Every test prints what it is about to run before touching the GPU.
"""
from __future__ import annotations

import pytest

import oracle
from parity import assert_parity, gpu_to_numpy
from shocovox_amd import DeviceTree, Octree, workloads as W

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def test_all_eight_caves_poses_against_the_oracle(gpu_device):
    """Every pose of the 512^3 / dim 16 headline step, rendered on the device,
    against the oracle on a strided sample of each frame's rays (every 7th
    pixel, 296,229 rays per pose)."""
    import torch
    cfg = W.CAVES
    tree = W.build_scene(cfg, seed=1)
    otree = oracle.OracleTree(tree.export_model())
    d = DeviceTree(tree)
    for k, vp in enumerate(W.viewports(cfg, seed=1)):
        _say(f"caves pose {k}: {cfg.width}x{cfg.height}, every 7th ray against the oracle")
        bufs = d.render(vp, cfg.width, cfg.height)
        torch.cuda.synchronize()
        got = {f: v[::7] for f, v in gpu_to_numpy(bufs).items()}
        rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                    cfg.height)[::7]
        exp = otree.cast(rays, threads=16)
        res = assert_parity(exp, got, f"caves pose {k}")
        assert exp["hit"].sum() > 0.2 * len(rays)
        _say(f"caves pose {k}: {res}")


def test_update_that_changes_the_brick_dimension(gpu_device):
    """Trees created from a description of one brick dimension, then brought
    to an octree of another by svo_tree_update (a full re-upload): the brick
    arrays are reallocated by bytes, not by slots, and every render equals
    the oracle on the octree the tree was updated to."""
    import torch
    s = 128.0
    vp = W.look_at_viewport((1.5 * s, 1.2 * s, 1.3 * s), (s / 2, s / 3, s / 2), 320, 200)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 320, 200)
    for (size0, dim0), (size1, dim1) in (((128, 4), (128, 16)), ((128, 16), (128, 4)),
                                         ((64, 4), (256, 16))):
        first = Octree.scene(2, size0, dim0, 11)
        flat = first.flatten()
        d = DeviceTree(desc=flat.desc())
        tree = Octree.scene(2, size1, dim1, 12)
        _say(f"created from {size0}^3 dim {dim0}, updated to {size1}^3 dim {dim1}")
        d.update(tree)
        bufs = d.render(vp, 320, 200)
        torch.cuda.synchronize()
        exp = oracle.OracleTree(tree.export_model()).cast(rays)
        res = assert_parity(exp, gpu_to_numpy(bufs), f"dim {dim0} -> {dim1}")
        assert exp["hit"].sum() > 1000
        _say(f"dim {dim0} -> {dim1}: {res}")
        del d, flat


def test_occupancy_beyond_the_32_bit_range_is_refused(gpu_device):
    """The traversal reads occupancy words through 32-bit byte offsets, so a
    tree whose occupancy sidecar would exceed 4 GiB is refused at upload,
    before anything is allocated or copied (the description below claims
    2^24 bricks of 16^3: 8 GiB of occupancy; its arrays are never read)."""
    from shocovox_amd import OctreeError
    flat = Octree.scene(2, 64, 16, 3).flatten()
    d = flat.desc()
    d.brick_count = 1 << 24
    _say("upload of a description with 2^24 bricks of 16^3 (8 GiB of occupancy)")
    with pytest.raises(OctreeError, match="4 GiB"):
        DeviceTree(desc=d)
    d = flat.desc()
    ok = DeviceTree(desc=d)  # the same tree at its real size is accepted
    del ok, flat


def test_batched_pool_at_configured_size(gpu_device):
    """The batched workload at its configured size: 64 trees log-uniform over
    64^3..1024^3 (brick dims 4/8/16, heightfield or caves content), 1-4
    viewports each of 128^2 / 512^2 / 720p / 1080p, all in one persistent
    launch; every job's outputs on a strided sample of its rays (every 61st)
    against the oracle, and the whole launch equal to per-viewport launches."""
    import torch
    from shocovox_amd import Batch, alloc_hits
    specs, views = W.batched_pool_specs(1, 64, 1024)
    _say(f"batched pool: {len(specs)} trees, {len(views)} viewports, "
         f"{sum(w * h for _, _, w, h in views)} rays")
    trees = [Octree.scene(*s) for s in specs]
    dtrees = [DeviceTree(t) for t in trees]
    fields = ("hit", "pix", "albedo", "voxel", "distance", "normal", "flags")
    bufs = [alloc_hits(w * h, gpu_device, fields) for _, _, w, h in views]
    plain = [alloc_hits(w * h, gpu_device, fields) for _, _, w, h in views]
    _say("batched pool: one batched launch, then one render launch per viewport")
    batch = Batch(len(views))
    batch.run([Batch.render_job(dtrees[ti], vp, w, h, b)
               for (ti, vp, w, h), b in zip(views, bufs)])
    for (ti, vp, w, h), b in zip(views, plain):
        dtrees[ti].render_into(vp, w, h, b)
    torch.cuda.synchronize()
    for a, b in zip(bufs, plain):
        for k in fields:
            assert torch.equal(a[k], b[k]), k
    checked = 0
    for ti, spec in enumerate(specs):
        jobs = [(k, v) for k, v in enumerate(views) if v[0] == ti]
        if not jobs:
            continue
        otree = oracle.OracleTree(trees[ti].export_model())
        for k, (_, vp, w, h) in jobs:
            rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w, h)[::61]
            got = {f: v[::61] for f, v in gpu_to_numpy(bufs[k]).items()}
            assert_parity(otree.cast(rays, threads=16), got, f"tree {spec} job {k}")
            checked += len(rays)
        del otree
    _say(f"batched pool: {checked} sampled rays equal to the oracle")


def test_served_pool_up_to_2048_with_4k_requests(gpu_device):
    """A served session shaped like the served workload: trees up to 2048^3
    (the 2048^3 spheres at brick dim 32 among them), requests of 256^2 to 4K
    by Zipf(1.1), 5% edits of 1-1000 voxels applied in place. The digest of
    every render equals an oracle replay of the plan (edits applied in
    arrival order) for every request whose tree the replay has loaded in its
    current state: all renders of trees below 2048^3, and those of the
    2048^3 trees that no edit changed since the replay last loaded them."""
    from shocovox_amd import serve
    from parity import result_digest
    specs = [(4, 64, 8, 21), (2, 128, 4, 22), (2, 256, 8, 23), (4, 512, 16, 24),
             (2, 1024, 16, 25), (3, 2048, 32, 26), (4, 2048, 16, 27)]
    _say(f"served pool: {[s[1] for s in specs]}")
    live = [Octree.scene(*s) for s in specs]
    replay = [t.clone() for t in live]
    dtrees = [serve.admit(t) for t in live]
    cfg = serve.ServeConfig(qps=500, duration_s=0.5, seed=31, max_inflight=2)
    plan = serve.plan(cfg, [s[1] for s in specs])
    n_edit = sum(r.kind for r in plan)
    big = sum(1 for r in plan if r.kind == 0 and r.width == 3840)
    _say(f"served plan: {len(plan)} requests, {n_edit} edits, {big} 4K renders")
    assert n_edit >= 3 and big >= 3
    r = serve.run(live, dtrees, cfg, digests=True)
    _say(f"served: {r['launches']} launches, p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms")
    assert r["requests"] == len(plan) and r["rejected"] == 0 and r["timed_out"] == 0
    cache, loads, checked, skipped = {}, {}, 0, 0
    for i, q in enumerate(plan):
        if q.kind == 1:
            serve.apply_edit(replay[q.tree], q)
            cache.pop(q.tree, None)
            continue
        if q.tree not in cache:
            # a 2048^3 tree is loaded into the oracle once (its export and
            # load take seconds); its renders after a later edit are skipped
            if specs[q.tree][1] >= 2048 and loads.get(q.tree, 0) >= 1:
                skipped += 1
                continue
            cache[q.tree] = oracle.OracleTree(replay[q.tree].export_model())
            loads[q.tree] = loads.get(q.tree, 0) + 1
        v = q.viewport
        rays = oracle.viewport_rays(v.origin, v.direction, v.frustum, v.fov, q.width, q.height)
        exp = result_digest(cache[q.tree].cast(rays, threads=16))
        assert int(r["digests"][i]) == exp, f"request {i} (tree {specs[q.tree]}, {q.width}x{q.height})"
        checked += 1
    _say(f"served: {checked} render digests equal to the oracle replay, {skipped} skipped")
    assert checked >= 0.7 * (len(plan) - n_edit)
