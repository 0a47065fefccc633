"""GPU parity: the HIP traversal against the oracle, through the C ABI.

Every test prints what it is about to run (workload, seed, ray count) before
touching the GPU, so a failing case is named even if the run dies.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from kat_trees import fixed_kats, make_edge_ray_point_to, make_ray_point_to, random_fill
from parity import assert_parity, gpu_to_numpy
from shocovox_amd import DeviceTree, Octree, batch_variants, workloads as W

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def _cast(tree: Octree, rays: np.ndarray, dev) -> dict:
    import torch
    d = DeviceTree(tree)
    t = torch.from_numpy(np.ascontiguousarray(rays, np.float32).reshape(-1, 6)).to(dev)
    bufs = d.cast(t)
    torch.cuda.synchronize()
    return gpu_to_numpy(bufs)


def _oracle(tree: Octree) -> oracle.OracleTree:
    return oracle.OracleTree(tree.export_model())


def test_kat_fixed_rays(gpu_device):
    """The 17 fixed-ray known-answer tests of raytracing/tests.rs on the GPU."""
    for name, (builder, r, expect) in fixed_kats().items():
        _say(f"KAT {name}")
        tree = builder()
        data = tree.export_model()["data"]
        got = _cast(tree, r[None, :], gpu_device)
        exp = _oracle(tree).cast(r[None, :])
        assert_parity(exp, got, name)
        hit = bool(got["hit"][0])
        kind = expect[0]
        if kind == "some":
            assert hit, name
        elif kind == "none":
            assert not hit, name
        elif kind == "data":
            # voxel_data!(&5) is Informative: no colour index, data index set
            assert hit and got["albedo"][0] == 0, name
            assert (got["pix"][0] & 0xFFFF) == 0xFFFF and (got["pix"][0] >> 16) != 0xFFFF, name
            assert int(data[int(got["pix"][0]) >> 16]) == expect[1], name
        elif kind == "albedo":
            assert hit and got["albedo"][0] == expect[1], name
        elif kind == "albedo_normal":
            assert hit and got["albedo"][0] == expect[1], name
            assert tuple(got["normal"][0]) == expect[2], name
        elif kind == "albedo_normal_len":
            assert hit and got["albedo"][0] == expect[1], name
            assert np.linalg.norm(got["normal"][0]) < expect[2], name
        elif kind == "none_or_data":
            assert (not hit) or ((got["pix"][0] & 0xFFFF) == 0xFFFF and
                                 int(data[int(got["pix"][0]) >> 16]) == expect[1]), name


@pytest.mark.parametrize("seed", range(8))
def test_kat_random_trees(gpu_device, seed):
    """test_get_by_ray_from_outside / _where_dim_is_2 / _from_edge / _from_inside
    (raytracing/tests.rs:144-251) with seeded generators."""
    rng = np.random.default_rng(seed)
    cases = [
        ("from_outside", *random_fill(4, 1, rng), make_ray_point_to, 0.0),
        ("from_outside_dim2", *random_fill(4, 2, rng), make_ray_point_to, 0.0),
        ("from_edge", *random_fill(8, 1, rng, zs=(1, 2, 3)), make_edge_ray_point_to, 0.1),
        ("from_inside", *random_fill(16, 1, rng, zs=(1, 2, 3)), make_ray_point_to, 0.0),
    ]
    for name, tree, filled, maker, off in cases:
        if not filled:
            continue
        rays = np.stack([maker(np.array(p, np.float32) + np.float32(off), rng) for p in filled])
        _say(f"random KAT {name} seed={seed} rays={len(rays)}")
        got = _cast(tree, rays, gpu_device)
        exp = _oracle(tree).cast(rays)
        assert_parity(exp, got, f"{name}/{seed}")
        assert got["hit"].all(), f"{name}/{seed}: every ray at a filled voxel must hit"
        assert ((got["pix"] & 0xFFFF) == 0xFFFF).all()
        data = tree.export_model()["data"]
        assert (data[(got["pix"] >> 16).astype(np.int64)] == 5).all(), f"{name}/{seed}: data 5"


def test_plain_render_all_rays(gpu_device):
    """Every primary ray of the 128^3 / dim 8 workload, generated on the GPU."""
    import torch
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    _say(f"plain render {cfg.width}x{cfg.height}")
    d = DeviceTree(tree)
    bufs = d.render(vp, cfg.width, cfg.height)
    torch.cuda.synchronize()
    got = gpu_to_numpy(bufs)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    exp = _oracle(tree).cast(rays)
    res = assert_parity(exp, got, "plain render")
    assert exp["hit"].sum() > 1000
    _say(f"plain render parity {res}")


def test_plain_cast_matches_render(gpu_device):
    """svo_cast on host-generated rays equals svo_render on the same viewport."""
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=2)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    _say("plain cast")
    got = _cast(tree, rays, gpu_device)
    exp = _oracle(tree).cast(rays)
    assert_parity(exp, got, "plain cast")


def test_caves_pose_render(gpu_device):
    """One full 1920x1080 pose of the 512^3 / dim 16 workload."""
    import torch
    cfg = W.CAVES
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[3]
    _say(f"caves render pose 3 {cfg.width}x{cfg.height}")
    d = DeviceTree(tree)
    bufs = d.render(vp, cfg.width, cfg.height)
    torch.cuda.synchronize()
    got = gpu_to_numpy(bufs)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    exp = _oracle(tree).cast(rays, threads=16)
    res = assert_parity(exp, got, "caves pose")
    _say(f"caves parity {res}")


def test_caves_incoherent(gpu_device):
    """Incoherent rays (origins in the volume, directions on S^2) into 512^3."""
    cfg = W.CAVES
    tree = W.build_scene(cfg, seed=3)
    rays = W.incoherent_rays(cfg.size, 200_000, seed=5)
    _say("caves incoherent 200000")
    got = _cast(tree, rays, gpu_device)
    exp = _oracle(tree).cast(rays, threads=16)
    assert_parity(exp, got, "caves incoherent")


def test_cast_in_ray_order_equals_oracle(gpu_device):
    """svo_cast's traversal order for large ray sets (origin cell, direction
    octant, by a radix sort on the stream) changes no result: the same rays,
    with the order forced on (threshold 1) and off, each equal the oracle ray
    for ray, including rays with NaN, infinite and outside origins."""
    from shocovox_amd import set_cast_order_threshold
    tree = W.build_scene(W.CAVES, seed=3)
    rays = W.incoherent_rays(W.CAVES.size, 120_000, seed=9)
    rays[:7, :3] = [[np.nan, 1, 1], [np.inf, 5, 5], [-50, 20, 30], [600, 600, 600],
                    [-np.inf, 0, 0], [1e30, -1e30, 3], [0, 0, 0]]
    exp = _oracle(tree).cast(rays, threads=16)
    import torch
    from shocovox_amd import Batch, alloc_hits
    d = DeviceTree(tree)
    t = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(gpu_device)
    try:
        for threshold in (1, 0):
            set_cast_order_threshold(threshold)
            _say(f"ordered cast, threshold {threshold}, {len(rays)} rays")
            assert_parity(exp, _cast(tree, rays, gpu_device), f"order threshold {threshold}")
            # a batched run's cast job takes the same order
            _say(f"batched cast job, threshold {threshold}")
            out = alloc_hits(len(rays), gpu_device)
            bt = Batch(2)
            bt.run([Batch.cast_job(d, t, out)])
            torch.cuda.synchronize()
            assert_parity(exp, gpu_to_numpy(out), f"batched, order threshold {threshold}")
            del bt
    finally:
        set_cast_order_threshold(1 << 20)


def test_axis_aligned_and_degenerate_rays(gpu_device):
    """Rays along axes and planes (inf/NaN scale factors, signed zeros)."""
    tree = W.build_scene(W.PLAIN, seed=1)
    rays = []
    for o in [(64.5, 200.0, 64.5), (-10.0, 20.0, 30.0), (64.0, 64.0, -5.0), (0.0, 0.0, 0.0),
              (128.0, 40.0, 128.0), (20.0, 127.99, 20.0)]:
        for d in [(0, -1, 0), (1, 0, 0), (0, 0, 1), (-1, 0, 0), (0, 1, 0), (0, 0, -1),
                  (0.70710677, -0.70710677, 0), (0, -0.70710677, 0.70710677),
                  (-0.0, -1.0, 0.0), (0.57735026, -0.57735026, 0.57735026)]:
            rays.append([*o, *d])
    rays = np.array(rays, np.float32)
    _say(f"axis-aligned rays {len(rays)}")
    got = _cast(tree, rays, gpu_device)
    exp = _oracle(tree).cast(rays)
    assert_parity(exp, got, "axis aligned")


def test_extreme_rays(gpu_device):
    """NaN, infinite, huge, subnormal and zero ray components: the kernels'
    hardware float-to-int conversions (saturating, NaN -> 0) and the power-of-
    two arithmetic give the oracle's results, flags included, on every ray;
    also towards a tree of brick dimension 1 and 32."""
    nan, inf = float("nan"), float("inf")
    origins = [(64.5, 200.0, 64.5), (30.0, 40.0, 50.0), (-1e30, 64.0, 64.0), (1e30, 1e30, 1e30),
               (inf, 10.0, 10.0), (nan, 10.0, 10.0), (64.0, 64.0, 64.0), (-5.0, -5.0, -5.0),
               (127.99999, 0.0, 64.0), (3e9, 5.0, 5.0)]
    dirs = [(0, 0, 0), (nan, 0, -1), (0, nan, 0), (nan, nan, nan), (1e20, -1e20, 1e20),
            (1e-30, -1.0, 1e-30), (1e-40, 1e-40, -1.0), (inf, -1, 0), (-inf, -inf, inf),
            (0.6, -0.8, 0.0), (1, 1, 1), (-0.0, -0.0, 1.0), (-1.0, 0.0, 0.0)]
    rays = np.array([[*o, *d] for o in origins for d in dirs], np.float32)
    trees = [W.build_scene(W.PLAIN, seed=1), Octree.scene(1, 64, 1, 3), Octree.scene(2, 128, 32, 5)]
    for i, tree in enumerate(trees):
        _say(f"extreme rays, tree {i}")
        got = _cast(tree, rays, gpu_device)
        assert_parity(_oracle(tree).cast(rays), got, f"extreme rays, tree {i}")


def test_empty_inputs(gpu_device):
    """Zero rays and an empty tree are valid inputs."""
    import torch
    tree = Octree.new(16, 2)
    d = DeviceTree(tree)
    bufs = d.cast(torch.zeros((0, 6), dtype=torch.float32, device=gpu_device))
    assert bufs["hit"].numel() == 0
    rays = W.incoherent_rays(16, 1000, seed=9)
    got = _cast(tree, rays, gpu_device)
    assert not got["hit"].any()
    exp = _oracle(tree).cast(rays)
    assert_parity(exp, got, "empty tree")


def test_every_render_variant_matches_oracle(gpu_device):
    """All launch variants of the render kernel compute the same results."""
    import torch
    from shocovox_amd import render_variants, select_render_variant, current_render_variant
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=4)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    exp = _oracle(tree).cast(rays)
    d = DeviceTree(tree)
    before = current_render_variant()
    try:
        for i, name in enumerate(render_variants()):
            _say(f"render variant {name}")
            select_render_variant(i)
            bufs = d.render(vp, cfg.width, cfg.height)
            torch.cuda.synchronize()
            assert_parity(exp, gpu_to_numpy(bufs), name)
    finally:
        select_render_variant(before)


def test_kernel_builds_report_their_scratch(gpu_device):
    """Every variant reports its scratch bytes per lane; the initial render
    variant and the batched variant the serving loop uses do not spill, and
    most builds do not either (the autotuners consider only those)."""
    from shocovox_amd import (batch_variant_scratch, render_variant_scratch,
                              render_variants)
    _say("scratch bytes of every kernel variant")
    rn, bn = render_variants(), batch_variants()
    rs = {n: render_variant_scratch(i) for i, n in enumerate(rn)}
    bs = {n: batch_variant_scratch(i) for i, n in enumerate(bn)}
    _say(f"render {rs}\nbatch {bs}")
    assert rs["w6-slut-8x8-ww"] == 0 and rs["w6-slut-8x8-ww-g1"] == 0
    assert bs["classic-w5-g5"] == 0
    assert sum(v == 0 for v in rs.values()) >= len(rs) // 2
    assert sum(v == 0 for v in bs.values()) >= len(bs) // 2


def test_in_place_edits_on_device(gpu_device):
    """Sparse edits reach the device through svo_tree_update (changed records and
    bricks only); each round's render equals the oracle on the edited tree."""
    import torch
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=6)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    d = DeviceTree(tree)
    rng = np.random.default_rng(17)
    for rnd in range(4):
        n = int(rng.integers(1, 1000))
        xyz = rng.integers(0, cfg.size, size=(n, 3)).astype(np.uint32)
        ents = np.zeros((n, 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = 0xFF2020FF
        if rnd % 2:
            tree.clear_many(xyz)
        else:
            tree.insert_many(xyz, ents)
        recs, slots = d.update(tree)
        _say(f"edit round {rnd}: {n} voxels -> {recs} records, {slots} bricks uploaded")
        bufs = d.render(vp, cfg.width, cfg.height)
        torch.cuda.synchronize()
        exp = _oracle(tree).cast(rays)
        assert_parity(exp, gpu_to_numpy(bufs), f"edit round {rnd}")


def test_patches_on_streams_destroyed_after_each_update(gpu_device):
    """svo_tree_update keeps nothing on the caller's stream: every patch below
    is queued on a fresh HIP stream that is destroyed right after the call,
    and the next patch (which reuses the staging of the one two calls back)
    still runs; the final render equals the oracle on the edited tree."""
    import ctypes
    import torch
    torch.cuda.synchronize()
    # the HIP runtime already in the process (the one libsvo and torch use)
    with open("/proc/self/maps") as f:
        paths = sorted({ln.split()[-1] for ln in f if "libamdhip64.so" in ln})
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths[0])
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=8)
    vp = W.viewports(cfg)[0]
    d = DeviceTree(tree)
    rng = np.random.default_rng(23)
    for rnd in range(5):
        xyz = rng.integers(0, cfg.size, size=(200, 3)).astype(np.uint32)
        ents = np.zeros((200, 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = 0xFF20A0FF
        if rnd % 2:
            tree.clear_many(xyz)
        else:
            tree.insert_many(xyz, ents)
        s = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
        _say(f"patch {rnd} on a new stream, then the stream is destroyed")
        d.update(tree, stream=s.value)
        assert hip.hipStreamDestroy(s) == 0
    torch.cuda.synchronize()
    bufs = d.render(vp, cfg.width, cfg.height)
    torch.cuda.synchronize()
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    assert_parity(_oracle(tree).cast(rays), gpu_to_numpy(bufs), "patches on destroyed streams")


@pytest.mark.parametrize("size,dim,mips", [(32, 2, False), (64, 4, True), (128, 8, False),
                                           (256, 16, False)])
def test_patch_forms_on_device(gpu_device, size, dim, mips):
    """In-place patches in all three brick forms reach the device intact:
    word writes (a few voxels of an existing brick), slot fills plus words
    (new bricks, Solid bricks split by a clear) and whole slots (bricks of
    dimension 2 with three or more changed words, whole-node edits), with
    MIP bricks patched as well on one tree. Every round's render and a set of
    incoherent rays equal the oracle on the edited tree. (At 256^3 and dim 16
    a round's bricks hold over 2^20 values: the layout encodes them on
    several threads.)"""
    import torch
    from shocovox_amd import OctreeEntry
    rng = np.random.default_rng(size + dim)
    tree = Octree.new(size, dim)
    if mips:
        tree.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    xyz = rng.integers(0, size, size=(size * size, 3)).astype(np.uint32)
    ents = np.zeros((len(xyz), 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = (rng.integers(1, 9, len(xyz)).astype(np.uint32) << 24) | 0xFF
    tree.insert_many(xyz, ents)
    d = DeviceTree(tree)
    s = float(size)
    vp = W.look_at_viewport((1.7 * s, 1.2 * s, 1.5 * s), (s / 2, s / 2, s / 2), 256, 160)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 256, 160)
    inc = W.incoherent_rays(size, 4000, seed=size + 3)
    for rnd in range(6):
        kind = rnd % 3
        if kind == 0:  # sparse voxels: word writes, new bricks
            n = int(rng.integers(50, 600))
            xyz = rng.integers(0, size, size=(n, 3)).astype(np.uint32)
            ents = np.zeros((n, 3), np.uint32)
            ents[:, 0] = 1
            ents[:, 1] = (rng.integers(1, 9, n).astype(np.uint32) << 24) | 0x7F
            tree.insert_many(xyz, ents)
            tree.clear_many(rng.integers(0, size, size=(n // 2, 3)).astype(np.uint32))
        elif kind == 1:  # whole nodes: solid regions appear and split
            for _ in range(4):
                p = tuple(int(v) for v in rng.integers(0, size, size=3))
                tree.insert_at_lod(p, int(rng.choice([2, 4, 8])), OctreeEntry.visual(0x3366AAFF))
            tree.clear_many(rng.integers(0, size, size=(200, 3)).astype(np.uint32))
        else:
            for _ in range(3):
                p = tuple(int(v) for v in rng.integers(0, size, size=3))
                tree.clear_at_lod(p, int(rng.choice([2, 4])))
        recs, slots = d.update(tree)
        _say(f"patch forms {size}^3 dim {dim} round {rnd}: {recs} records, {slots} slots")
        got = gpu_to_numpy(d.render(vp, 256, 160))
        torch.cuda.synchronize()
        o = _oracle(tree)
        assert_parity(o.cast(rays), got, f"render after round {rnd}")
        assert_parity(o.cast(inc), _cast_tree(d, inc, gpu_device), f"rays after round {rnd}")


def _cast_tree(d, rays, dev) -> dict:
    import torch
    t = torch.from_numpy(np.ascontiguousarray(rays, np.float32).reshape(-1, 6)).to(dev)
    bufs = d.cast(t)
    torch.cuda.synchronize()
    return gpu_to_numpy(bufs)


def test_patch_that_outgrows_the_device_arrays(gpu_device):
    """Edits that need more records, brick slots and palette entries than the
    device arrays hold grow them on the device (old contents copied device to
    device, old buffers kept until the tree is destroyed) instead of
    re-sending the tree; renders equal the oracle after every round, and a
    render queued on another stream before the growing update still sees the
    tree as it was."""
    import torch
    from shocovox_amd import OctreeEntry
    tree = Octree.new(64, 4)
    tree.insert((3, 3, 3), OctreeEntry.visual(0x102030FF))
    d = DeviceTree(tree)
    s = float(tree.get_size())
    vp = W.look_at_viewport((1.6 * s, 1.1 * s, 1.4 * s), (s / 2, s / 2, s / 2), 320, 200)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 320, 200)
    rng = np.random.default_rng(29)
    side = torch.cuda.Stream()
    for rnd in range(3):
        before = d.device_bytes
        exp_prev = _oracle(tree).cast(rays)
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            queued = d.render(vp, 320, 200, stream=side)
        n = 3000 * (rnd + 1)
        xyz = rng.integers(0, 64, size=(n, 3)).astype(np.uint32)
        ents = np.zeros((n, 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = (rng.integers(1, 1 << 24, n).astype(np.uint32) << 8) | 0xFF
        tree.insert_many(xyz, ents)
        if rnd > 0:
            side.synchronize()  # without growth a patch rewrites the arrays in place
        recs, slots = d.update(tree)
        _say(f"growth round {rnd}: {recs} records, {slots} bricks, "
             f"{before} -> {d.device_bytes} device bytes")
        side.synchronize()
        if rnd == 0:
            assert d.device_bytes > before
        assert_parity(exp_prev, gpu_to_numpy(queued), f"render queued before round {rnd}")
        got = gpu_to_numpy(d.render(vp, 320, 200))
        torch.cuda.synchronize()
        assert_parity(_oracle(tree).cast(rays), got, f"growth round {rnd}")


@pytest.mark.parametrize("variant", range(len(batch_variants())))
def test_batched_mixed_jobs_and_replays(gpu_device, variant):
    """Many trees, uneven viewports and ray arrays in one persistent launch;
    the graph is replayed with different job sets (with eager work in
    between) and every job of every replay equals the oracle - for every
    batch kernel variant."""
    import torch
    from shocovox_amd import (Batch, Octree, alloc_hits, batch_variants, current_batch_variant,
                              select_batch_variant)
    names = batch_variants()
    _say(f"batch variant {names[variant]}")
    rng = np.random.default_rng(23)
    trees = [W.build_scene(W.PLAIN, seed=s) for s in (1, 2)]
    small = Octree.new(64, 4)
    xyz = rng.integers(0, 64, size=(5000, 3)).astype(np.uint32)
    ents = np.zeros((5000, 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = 0x3355AAFF
    small.insert_many(xyz, ents)
    trees.append(small)
    dtrees = [DeviceTree(t) for t in trees]
    oracles = [_oracle(t) for t in trees]
    prev = current_batch_variant()
    select_batch_variant(variant)
    try:
        batch = Batch(16)
    finally:
        select_batch_variant(prev)
    for replay in range(3):
        jobs, checks = [], []
        for k in range(int(rng.integers(3, 9))):
            ti = int(rng.integers(0, len(trees)))
            size = trees[ti].get_size()
            if rng.random() < 0.6:
                w, h = int(rng.integers(1, 200)), int(rng.integers(1, 150))
                c = size / 2
                vp = W.look_at_viewport((c + size * 1.3 * np.cos(k), size * 0.9, c + size * np.sin(k)),
                                        (c, c / 2, c), w, h)
                bufs = alloc_hits(w * h, gpu_device)
                jobs.append(Batch.render_job(dtrees[ti], vp, w, h, bufs))
                rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w, h)
            else:
                n = int(rng.integers(1, 3000))
                rays = W.incoherent_rays(size, n, seed=int(rng.integers(0, 1 << 30)))
                t = torch.from_numpy(rays).to(gpu_device)
                bufs = alloc_hits(n, gpu_device)
                jobs.append(Batch.cast_job(dtrees[ti], t, bufs))
                bufs["_rays"] = t
            checks.append((ti, rays, bufs))
        _say(f"batched replay {replay}: {len(jobs)} jobs")
        batch.run(jobs)
        torch.cuda.synchronize()
        # unrelated eager work between replays
        torch.ones(1 << 20, device=gpu_device).sum().item()
        for ti, rays, bufs in checks:
            got = gpu_to_numpy({k: v for k, v in bufs.items() if k != "_rays"})
            assert_parity(oracles[ti].cast(rays), got, f"replay {replay} tree {ti}")
    # an empty batch is valid
    batch.run([])
    torch.cuda.synchronize()


def test_batched_runs_back_to_back_on_two_streams(gpu_device):
    """Runs issued without any host wait in between, alternating between two
    streams, on one context: each run's table comes from the other pinned
    staging, and a run waits on its stream for the previous one, so every
    run's outputs (separate buffers per run) equal the oracle."""
    import torch
    from shocovox_amd import Batch, alloc_hits
    tree = W.build_scene(W.PLAIN, seed=3)
    d = DeviceTree(tree)
    o = _oracle(tree)
    streams = [torch.cuda.Stream(device=gpu_device), torch.cuda.Stream(device=gpu_device)]
    batch = Batch(4)
    runs = []
    for k in range(6):
        w, h = 96 + 16 * k, 64 + 8 * k
        vp = W.look_at_viewport((200.0 - 10 * k, 150.0, 180.0 + 5 * k), (64.0, 40.0, 64.0), w, h)
        bufs = alloc_hits(w * h, gpu_device)
        rays = W.incoherent_rays(128, 500 + 100 * k, seed=40 + k)
        t = torch.from_numpy(rays).to(gpu_device)
        cbufs = alloc_hits(len(rays), gpu_device)
        runs.append((vp, w, h, bufs, rays, t, cbufs))
    torch.cuda.synchronize()
    _say("back-to-back runs on two streams")
    for k, (vp, w, h, bufs, rays, t, cbufs) in enumerate(runs):
        batch.run([Batch.render_job(d, vp, w, h, bufs), Batch.cast_job(d, t, cbufs)],
                  streams[k % 2])
    torch.cuda.synchronize()
    for k, (vp, w, h, bufs, rays, t, cbufs) in enumerate(runs):
        prim = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w, h)
        assert_parity(o.cast(prim), gpu_to_numpy(bufs), f"run {k} render")
        assert_parity(o.cast(rays), gpu_to_numpy(cbufs), f"run {k} cast")


def test_spheres_2048_full_size(gpu_device):
    """The 2048^3 / dim 32 spheres workload at full size: 3840x2160 primary
    rays and 8M incoherent rays on the GPU. Every ray: no guard flag, every
    hit's voxel is filled under the host octree's get() and carries the hit's
    albedo, distance = |impact - origin|. A strided sample of each ray set
    equals the oracle bit for bit."""
    import torch
    cfg = W.SPHERES
    _say("spheres 2048^3 build")
    tree = W.build_scene(cfg, seed=1)
    d = DeviceTree(tree)
    vp = W.viewports(cfg)[0]
    _say(f"spheres render {cfg.width}x{cfg.height}")
    fields = ("hit", "pix", "albedo", "voxel", "distance", "impact", "flags")
    bufs = d.render(vp, cfg.width, cfg.height, fields=fields)
    torch.cuda.synchronize()
    prim = gpu_to_numpy(bufs)
    del bufs
    inc_rays = W.incoherent_rays(cfg.size, 8_000_000, seed=7)
    _say("spheres incoherent 8000000")
    t = torch.from_numpy(inc_rays).to(gpu_device)
    bufs = d.cast(t, fields=fields)
    torch.cuda.synchronize()
    inc = gpu_to_numpy(bufs)
    del bufs, t
    prim_rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                     cfg.height)
    for name, got, rays in (("primary", prim, prim_rays), ("incoherent", inc, inc_rays)):
        assert (got["flags"] == 0).all(), name
        hit = got["hit"] == 1
        _say(f"{name}: {int(hit.sum())} hits of {len(hit)}")
        assert hit.sum() > 0.05 * len(hit)
        vox = got["voxel"].reshape(-1, 3)[hit]
        ent = tree.get_many(vox)
        assert (ent[:, 0] != 0).all(), f"{name}: a hit voxel is empty"
        assert (ent[:, 1] == got["albedo"][hit]).all(), f"{name}: albedo differs from get()"
        dist = np.linalg.norm(got["impact"].reshape(-1, 3)[hit].astype(np.float64)
                              - rays[hit, :3].astype(np.float64), axis=1)
        assert np.allclose(dist, got["distance"][hit], rtol=1e-5, atol=1e-3), name
    _say("spheres oracle sample")
    otree = _oracle(tree)
    for name, got, rays, step in (("primary", prim, prim_rays, 83),
                                  ("incoherent", inc, inc_rays, 79)):
        sample = np.ascontiguousarray(rays[::step])
        exp = otree.cast(sample, threads=16)
        sub = {k: v.reshape(len(rays), -1)[::step].reshape((-1,) + v.shape[1:])
               for k, v in got.items()}
        assert_parity(exp, sub, f"spheres {name} sample")


def test_saved_tree_loaded_straight_to_device(gpu_device, tmp_path):
    """Octree::save, then svo_tree_load into device memory: renders equal the
    oracle on the original tree; with the host copy kept, an edit followed by
    svo_tree_update equals the oracle on the edited tree."""
    import torch
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=5)
    path = tmp_path / "plain.bencode"
    tree.save(path)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    _say("loaded tree render")
    d = DeviceTree.load(path)
    got = gpu_to_numpy(d.render(vp, cfg.width, cfg.height))
    torch.cuda.synchronize()
    assert_parity(_oracle(tree).cast(rays), got, "loaded tree")
    d2, host = DeviceTree.load(path, keep_host=True)
    host.clear_at_lod((64, 64, 64), 32)
    d2.update(host)
    _say("loaded tree edited")
    got = gpu_to_numpy(d2.render(vp, cfg.width, cfg.height))
    torch.cuda.synchronize()
    assert_parity(_oracle(host).cast(rays), got, "loaded tree after edit")


def test_forced_iteration_guard_on_device(gpu_device):
    """svo_tree_set_step_limit: the device reports guard-ended rays exactly as
    the host build of the same traversal does, and the default bound restores
    oracle parity."""
    import hostsim
    import torch
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    d = DeviceTree(tree)
    a = tree.flatten().arrays()
    t = torch.from_numpy(rays).to(gpu_device)
    for limit in (3, 0):
        _say(f"step limit {limit}")
        d.set_step_limit(limit)
        got = gpu_to_numpy(d.cast(t))
        torch.cuda.synchronize()
        exp = hostsim.cast(a, rays, max_steps=limit)
        assert_parity(exp, got, f"step limit {limit}")
        if limit:
            assert (got["flags"] & 1).any()
        else:
            assert (got["flags"] == 0).all()
            assert_parity(_oracle(tree).cast(rays), got, "default bound")


@pytest.mark.parametrize("name,dim", [("navigate", 8), ("navigate_x", 2), ("navigate_z", 32)])
def test_vox_scene_renders_like_the_oracle(gpu_device, name, dim):
    """A MagicaVoxel asset imported with Octree::load_vox_file, flattened and
    uploaded: a look-at render and 100k incoherent rays equal the oracle."""
    import torch
    from pathlib import Path
    path = Path(__file__).resolve().parent / "golden" / "vox" / f"{name}.vox"
    from oracle import vox as OV
    tree = Octree.load_vox_file(path, dim)
    s = float(tree.get_size())
    pts = np.array([p[:3] for p in OV.load(path.read_bytes(), dim)[1]], np.float64)
    lo, hi = pts.min(0), pts.max(0) + 1.0
    c, ext = (lo + hi) / 2.0, float((hi - lo).max())
    vp = W.look_at_viewport(tuple(c + np.array([0.9, 0.6, 1.1]) * ext), tuple(c), 640, 360,
                            vfov_deg=28.0)
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 640, 360)
    _say(f"vox {name} dim {dim} render")
    d = DeviceTree(tree)
    got = gpu_to_numpy(d.render(vp, 640, 360))
    torch.cuda.synchronize()
    exp = _oracle(tree).cast(rays)
    assert exp["hit"].sum() > 1000
    assert_parity(exp, got, f"vox {name} render")
    inc = W.incoherent_rays(int(s), 100_000, seed=17)
    _say(f"vox {name} dim {dim} incoherent")
    got = gpu_to_numpy(d.cast(torch.from_numpy(inc).to(gpu_device)))
    torch.cuda.synchronize()
    assert_parity(_oracle(tree).cast(inc), got, f"vox {name} incoherent")


@pytest.mark.parametrize("vd", [0.5, 8.0, 1e6])
def test_lod_traversal_on_device(gpu_device, vd):
    """get_by_ray_at_lod on the device (svo_cast_at_lod, svo_render_at_lod)
    with MIPs enabled equals the oracle; after in-place edits (MIP bricks
    patched with the records) it still does."""
    import torch
    tree = Octree.scene(2, 256, 8, 12)
    tree.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    d = DeviceTree(tree)
    vp = W.look_at_viewport((330.0, 240.0, 300.0), (128.0, 128.0, 128.0), 320, 200)
    prim = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 320, 200)
    inc = W.incoherent_rays(256, 100_000, seed=31)
    rng = np.random.default_rng(7)
    for rnd in range(2):
        _say(f"lod vd={vd} round {rnd} render")
        got = gpu_to_numpy(d.render_at_lod(vp, 320, 200, vd))
        torch.cuda.synchronize()
        o = _oracle(tree)
        assert_parity(o.cast(prim, viewing_distance=vd), got, f"lod render vd={vd} r{rnd}")
        exp = o.cast(inc, viewing_distance=vd)
        from shocovox_amd import set_cast_order_threshold
        try:
            # the second round casts in the sorted ray order (forced on)
            set_cast_order_threshold(1 if rnd else 0)
            _say(f"lod vd={vd} round {rnd} cast, ray order {'on' if rnd else 'off'}")
            got = gpu_to_numpy(d.cast_at_lod(torch.from_numpy(inc).to(gpu_device), vd))
            torch.cuda.synchronize()
        finally:
            set_cast_order_threshold(1 << 20)
        assert_parity(exp, got, f"lod cast vd={vd} r{rnd}")
        xyz = rng.integers(0, 256, size=(500, 3)).astype(np.uint32)
        ents = np.zeros((500, 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = 0x8844CCFF
        tree.insert_many(xyz, ents)
        d.update(tree)


def test_batched_lod_jobs(gpu_device):
    """get_by_ray_at_lod in the batched mode: render and cast jobs with
    viewing distances (and plain jobs beside them, in one run) on trees with
    MIPs enabled equal the oracle; a plain run on the same context afterwards
    still does; a negative viewing distance is refused."""
    import torch
    from shocovox_amd import Batch, alloc_hits
    from shocovox_amd._native import SvoError
    trees = [Octree.scene(2, 256, 8, 12), Octree.scene(1, 64, 4, 3)]
    for t in trees:
        t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    dtrees = [DeviceTree(t) for t in trees]
    oracles = [_oracle(t) for t in trees]
    fields = ("hit", "pix", "albedo", "voxel", "distance", "impact", "normal", "flags")
    vp = W.look_at_viewport((330.0, 240.0, 300.0), (128.0, 128.0, 128.0), 200, 136)
    prim = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 200, 136)
    inc = W.incoherent_rays(64, 30_000, seed=5)
    inc_dev = torch.from_numpy(inc).to(gpu_device)
    batch = Batch(8)
    cases = [("render", 0, 2.0), ("render", 0, 0.0), ("cast", 1, 0.5), ("cast", 1, 0.0),
             ("render", 0, 40.0)]
    bufs, jobs = [], []
    for kind, ti, vd in cases:
        if kind == "render":
            b = alloc_hits(200 * 136, gpu_device, fields)
            jobs.append(Batch.render_job(dtrees[ti], vp, 200, 136, b, viewing_distance=vd))
        else:
            b = alloc_hits(len(inc), gpu_device, fields)
            jobs.append(Batch.cast_job(dtrees[ti], inc_dev, b, viewing_distance=vd))
        bufs.append(b)
    for rnd, sel in enumerate((list(range(len(cases))), [1, 3])):
        _say(f"batched lod run {rnd}")
        for b in bufs:
            for k in fields:
                b[k].zero_()
        batch.run([jobs[i] for i in sel])
        torch.cuda.synchronize()
        for i in sel:
            kind, ti, vd = cases[i]
            rays = prim if kind == "render" else inc
            exp = oracles[ti].cast(rays, viewing_distance=vd) if vd > 0 else oracles[ti].cast(rays)
            assert_parity(exp, gpu_to_numpy(bufs[i]), f"batched {kind} vd={vd} run {rnd}")
    bad = Batch.cast_job(dtrees[1], inc_dev, bufs[3], viewing_distance=-1.0)
    with pytest.raises(SvoError):
        batch.run([bad])


def test_rays_that_leave_the_content_behind(gpu_device):
    """Rays with nothing ahead of them at the root (an empty tree, rays beside
    the content of a sparse scene): the batch kernel replays only the
    reference's restart nudges (crawl_out_of_root), the while-while form used
    by svo_render and svo_cast runs the full restarts; both equal the oracle,
    in one batched run and as plain launches."""
    import torch
    from shocovox_amd import Batch, alloc_hits
    fields = ("hit", "pix", "albedo", "voxel", "distance", "impact", "normal", "flags")
    w, h = 160, 96
    batch = Batch(8)
    trees = (Octree.new(512, 16), Octree.scene(1, 256, 16, 3), Octree.new(64, 4))
    for i, tree in enumerate(trees):
        size = tree.get_size()
        vp = W.look_at_viewport((1.3 * size, 0.9 * size, 1.2 * size), (size / 2,) * 3, w, h)
        prim = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w, h)
        inc = W.incoherent_rays(size, 20_000, seed=3 + i)
        o = _oracle(tree)
        exp_p, exp_i = o.cast(prim), o.cast(inc)
        d = DeviceTree(tree)
        _say(f"leave-behind tree {i} (size {size}): batched run")
        bp = alloc_hits(w * h, gpu_device, fields)
        bi = alloc_hits(len(inc), gpu_device, fields)
        inc_dev = torch.from_numpy(inc).to(gpu_device)
        batch.run([Batch.render_job(d, vp, w, h, bp), Batch.cast_job(d, inc_dev, bi)])
        torch.cuda.synchronize()
        assert_parity(exp_p, gpu_to_numpy(bp), f"batched render, tree {i}")
        assert_parity(exp_i, gpu_to_numpy(bi), f"batched cast, tree {i}")
        _say(f"leave-behind tree {i}: plain launches")
        got = d.render(vp, w, h)
        torch.cuda.synchronize()
        assert_parity(exp_p, gpu_to_numpy(got), f"plain render, tree {i}")
        assert_parity(exp_i, _cast(tree, inc, gpu_device), f"plain cast, tree {i}")
