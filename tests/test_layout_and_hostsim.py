"""CPU-side checks of the product path that need no GPU:

* the C ABI library loads and exports every function include/svo.h declares;
* the flattened device layout answers get() exactly like the host octree;
* the kernel's traversal source (traverse.hpp), compiled for the host,
  agrees bit for bit with the oracle (the GPU tests then check the gfx950
  build of the same source);
* invariants of the oracle's hits and misses.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import hostsim
import oracle
from kat_trees import fixed_kats
from parity import canonical_layout, compare
from shocovox_amd import Octree, OctreeEntry, _native, workloads as W

ROOT = Path(__file__).resolve().parent.parent


def test_c_abi_exports_every_declared_function():
    header = (ROOT / "include" / "svo.h").read_text()
    declared = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(svo_\w+)\s*\(", header, re.M))
    declared = {d for d in declared if not d.endswith("_t")}
    lib = C.CDLL(str(_native.LIB_PATH))
    missing = [d for d in sorted(declared) if not hasattr(lib, d)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(_native.exported_symbols()) <= declared
    assert len(declared) >= 30


def test_status_codes_and_errors():
    with pytest.raises(_native.SvoError) as e:
        Octree.new(12, 3)
    assert e.value.status == 2 and "brick" in str(e.value)


def _all_positions(n):
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
    return g.reshape(-1, 3).astype(np.uint32)


def test_flat_layout_roundtrip_plain():
    tree = W.build_scene(W.PLAIN, seed=1)
    flat = tree.flatten()
    o = oracle.OracleTree(tree.export_model())
    xyz = _all_positions(128)[::7]
    host = tree.get_many(xyz)
    ora = o.get_many(xyz)
    # oracle get() returns the palette index value; host get() the entry
    assert ((ora == 0xFFFFFFFF) == (host[:, 0] == 0)).all()
    for p in xyz[::997]:
        assert flat.get(tuple(int(v) for v in p)) == tree.get(tuple(int(v) for v in p))


def test_flat_layout_structure_plain():
    tree = W.build_scene(W.PLAIN, seed=1)
    a = tree.flatten().arrays()
    nodes = a["nodes"]
    kinds = nodes["meta"] & 3
    assert kinds[0] in (1, 2, 3)  # the root record comes first
    internal = nodes[kinds == 1]
    kids = internal["child"][internal["child"] != 0xFFFFFFFF]
    assert (kids < len(nodes)).all() and len(np.unique(kids)) == len(kids)
    # breadth-first numbering: children come after their parent
    for i in np.nonzero(kinds == 1)[0][:50]:
        c = nodes["child"][i]
        assert (c[c != 0xFFFFFFFF] > i).all()
    # occupancy sidecar equals the emptiness of the palette values
    n3 = a["brick_dim"] ** 3
    pix = a["brick_pix"].reshape(-1, n3)
    occ = a["brick_occ"].reshape(len(pix), -1)
    bits = np.unpackbits(occ.view(np.uint8), bitorder="little").reshape(len(pix), -1)[:, :n3]
    assert np.array_equal(bits.astype(bool), pix != 0xFFFFFFFF)


def test_flat_layout_roundtrip_kat_trees():
    for name, (builder, _, _) in fixed_kats().items():
        if name == "edge_case_brick_boundary_error":
            continue
        tree = builder()
        flat = tree.flatten()
        n = tree.get_size()
        step = max(1, n // 16)
        for x in range(0, n, step):
            for y in range(0, n, step):
                for z in range(0, n, step):
                    assert flat.get((x, y, z)) == tree.get((x, y, z)), (name, x, y, z)


def test_hostsim_equals_oracle_plain_all_rays():
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    # the kernel's ray generation compiled for the host equals the oracle's
    assert np.array_equal(hostsim.viewport_rays(vp, cfg.width, cfg.height).view(np.uint32),
                          rays.view(np.uint32))
    exp = oracle.OracleTree(tree.export_model()).cast(rays)
    got = hostsim.cast(tree.flatten().arrays(), rays)
    res = compare(exp, {k: got[k] for k in got.dtype.names})
    assert all(v == 0 for k, v in res.items() if k != "rays"), res
    assert (exp["flags"] == 0).all()


def test_hostsim_equals_oracle_kat_trees():
    for name, (builder, r, _) in fixed_kats().items():
        tree = builder()
        exp = oracle.OracleTree(tree.export_model()).cast(r[None, :])
        got = hostsim.cast(tree.flatten().arrays(), r[None, :])
        res = compare(exp, {k: got[k] for k in got.dtype.names})
        assert all(v == 0 for k, v in res.items() if k != "rays"), (name, res)


def test_hostsim_equals_oracle_incoherent_small_trees():
    rng = np.random.default_rng(21)
    for size, dim in [(8, 1), (16, 2), (32, 4), (64, 8)]:
        t = Octree.new(size, dim)
        xyz = rng.integers(0, size, size=(size * size, 3)).astype(np.uint32)
        ents = np.zeros((len(xyz), 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = (rng.integers(1, 8, size=len(xyz)).astype(np.uint32) << 24) | 0xFF
        t.insert_many(xyz, ents)
        rays = W.incoherent_rays(size, 4000, seed=size)
        # plus rays from outside the bounds aimed at the centre
        outside = rng.uniform(-size, 2 * size, size=(2000, 3)).astype(np.float32)
        d = (np.float32(size / 2) - outside).astype(np.float32)
        ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)
        d = (d / ln[:, None]).astype(np.float32)
        rays = np.concatenate([rays, np.concatenate([outside, d], 1)]).astype(np.float32)
        exp = oracle.OracleTree(t.export_model()).cast(rays)
        got = hostsim.cast(t.flatten().arrays(), rays)
        res = compare(exp, {k: got[k] for k in got.dtype.names})
        assert all(v == 0 for k, v in res.items() if k != "rays"), (size, dim, res)


def test_hostsim_equals_oracle_caves_sample():
    cfg = W.CAVES
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                cfg.height)[::41]
    exp = oracle.OracleTree(tree.export_model()).cast(rays, threads=4)
    got = hostsim.cast(tree.flatten().arrays(), rays)
    res = compare(exp, {k: got[k] for k in got.dtype.names})
    assert all(v == 0 for k, v in res.items() if k != "rays"), res


def test_hit_invariants_plain():
    """Every hit's voxel is non-empty under the reference's get() and holds the
    hit's palette value; hit distance equals |impact - origin|."""
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width, cfg.height)
    o = oracle.OracleTree(tree.export_model())
    h = o.cast(rays)
    hit = h["hit"] == 1
    vox = h["voxel"][hit]
    assert (o.get_many(vox) == h["pix"][hit]).all()
    assert (h["flags"] == 0).all()
    d = np.linalg.norm(h["impact"][hit].astype(np.float64) - rays[hit, :3], axis=1)
    assert np.allclose(d, h["distance"][hit], atol=1e-3)


def test_miss_invariant_brute_force():
    """A miss never passes through a non-empty voxel, checked by an f64 march
    away from cell boundaries (the reference's 0.1 restart nudge can skip
    thin corners, so rays within 0.11 of a boundary crossing are excluded)."""
    rng = np.random.default_rng(5)
    t = Octree.new(32, 2)
    xyz = rng.integers(0, 32, size=(300, 3)).astype(np.uint32)
    ents = np.zeros((300, 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = 0xFFFFFFFF
    t.insert_many(xyz, ents)
    o = oracle.OracleTree(t.export_model())
    solid = set(map(tuple, xyz.tolist()))
    rays = W.incoherent_rays(32, 3000, seed=8)
    h = o.cast(rays)
    checked = bad = 0
    for r, hh in zip(rays, h):
        if hh["hit"]:
            continue
        org, d = r[:3].astype(np.float64), r[3:].astype(np.float64)
        ts = np.arange(0.0, 60.0, 0.01)
        pts = org[None, :] + ts[:, None] * d[None, :]
        inside = ((pts >= 0) & (pts < 32)).all(1)
        frac = np.abs(pts - np.round(pts))
        near_edge = (frac < 0.11).sum(1) >= 2
        cells = np.floor(pts[inside & ~near_edge]).astype(int)
        checked += 1
        if any(tuple(c) in solid for c in cells):
            bad += 1
    assert checked > 1000
    assert bad <= 0.01 * checked


def _random_edits(tree, rng, n_insert, n_clear, size, colors=6):
    xyz = rng.integers(0, size, size=(n_insert, 3)).astype(np.uint32)
    ents = np.zeros((n_insert, 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = (rng.integers(1, colors, size=n_insert).astype(np.uint32) << 16) | 0xFF
    tree.insert_many(xyz, ents)
    tree.clear_many(rng.integers(0, size, size=(n_clear, 3)).astype(np.uint32))


@pytest.mark.parametrize("seed", range(5))
def test_incremental_layout_matches_fresh_tree(seed):
    """Edits applied in place to the layout (svo_flat_sync) answer casts and
    point queries exactly like the edited tree (oracle) after every round.
    (256^3 at dim 16: a round's bricks hold over 2^20 values, so their
    encoding runs on several threads.)"""
    rng = np.random.default_rng(seed)
    size, dim = [(32, 2), (64, 4), (128, 8), (64, 1), (256, 16)][seed]
    tree = Octree.new(size, dim)
    _random_edits(tree, rng, 2000, 0, size)
    flat = tree.flatten()
    flat.sync(tree)  # drains the build's log
    rays = W.incoherent_rays(size, 3000, seed=seed + 40)
    for _ in range(6):
        _random_edits(tree, rng, int(rng.integers(1, 400)), int(rng.integers(0, 400)), size)
        if rng.random() < 0.3:  # whole-node edits as well
            p = tuple(int(v) for v in rng.integers(0, size, size=3))
            s = int(rng.choice([2, 4, 8]))
            if rng.random() < 0.5:
                tree.insert_at_lod(p, s, OctreeEntry.visual(0x77AA33FF))
            else:
                tree.clear_at_lod(p, s)
        flat.sync(tree)
        a = flat.arrays()
        # word for word what a fresh flatten of the edited tree holds
        assert canonical_layout(a) == canonical_layout(tree.flatten().arrays())
        exp = oracle.OracleTree(tree.export_model()).cast(rays)
        got = hostsim.cast(a, rays)
        res = compare(exp, {k: got[k] for k in got.dtype.names})
        assert all(v == 0 for k, v in res.items() if k != "rays"), res
        for p in rng.integers(0, size, size=(300, 3)):
            p = tuple(int(v) for v in p)
            assert flat.get(p) == tree.get(p), p


def test_resumable_traversal_equals_cast_ray():
    """ray_begin / ray_step (the refill kernels' form) give the same result,
    bit for bit, as cast_ray and the oracle on the plain workload, a caves
    sample, the KAT trees and incoherent rays into small random trees."""
    cases = []
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    cases.append((tree, oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov,
                                             cfg.width, cfg.height)))
    caves = W.build_scene(W.CAVES, seed=1)
    vp = W.viewports(W.CAVES)[3]
    cases.append((caves, oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov,
                                              W.CAVES.width, W.CAVES.height)[::53]))
    for name, (builder, r, _) in fixed_kats().items():
        cases.append((builder(), r[None, :]))
    rng = np.random.default_rng(77)
    for size, dim in [(16, 2), (64, 4)]:
        t = Octree.new(size, dim)
        xyz = rng.integers(0, size, size=(size * size, 3)).astype(np.uint32)
        ents = np.zeros((len(xyz), 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = 0xFF00FFFF
        t.insert_many(xyz, ents)
        cases.append((t, W.incoherent_rays(size, 3000, seed=size)))
    for t, rays in cases:
        a = t.flatten().arrays()
        base = hostsim.cast(a, rays)
        res = hostsim.cast(a, rays, resumable=True)
        assert base.tobytes() == res.tobytes()
        # the while-while arrangement (cast_ray_ww) too
        assert base.tobytes() == hostsim.cast(a, rays, regrouped=True).tobytes()
        assert base.tobytes() == hostsim.cast(a, rays, regrouped="nocrawl").tobytes()
        exp = oracle.OracleTree(t.export_model()).cast(rays)
        r = compare(exp, {k: res[k] for k in res.dtype.names})
        assert all(v == 0 for k, v in r.items() if k != "rays"), r


def test_p2_and_general_forms_agree():
    """The kernels' P2 form (divisions by bounds sizes as multiplications by
    exact reciprocals, the POP's fmod as floor * size) and the general form
    give the same results, bit for bit, as each other and the oracle, for
    plain and LOD traversal, on trees of every brick dimension 1..32."""
    rng = np.random.default_rng(5)
    for size, dim in [(8, 1), (16, 2), (64, 4), (128, 8), (128, 16), (256, 32)]:
        t = Octree.new(size, dim)
        n = min(size * size, 6000)
        xyz = rng.integers(0, size, size=(n, 3)).astype(np.uint32)
        ents = np.zeros((n, 3), np.uint32)
        ents[:, 0] = 1
        ents[:, 1] = (rng.integers(1, 8, size=n).astype(np.uint32) << 24) | 0xFF
        t.insert_many(xyz, ents)
        t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
        a = t.flatten().arrays()
        rays = W.incoherent_rays(size, 3000, seed=size + dim)
        o = oracle.OracleTree(t.export_model())
        for vd in (None, 2.0):
            p2 = hostsim.cast(a, rays, viewing_distance=vd)
            gen = hostsim.cast(a, rays, viewing_distance=vd, general=True)
            assert p2.tobytes() == gen.tobytes(), (size, dim, vd)
            exp = o.cast(rays) if vd is None else o.cast(rays, viewing_distance=vd)
            r = compare(exp, {k: p2[k] for k in p2.dtype.names})
            assert all(v == 0 for k, v in r.items() if k != "rays"), (size, dim, vd, r)


def test_extreme_rays_on_the_host_build():
    """NaN, infinite, huge, subnormal and zero ray components: the host build
    of the kernels' traversal (both arrangements) equals the oracle on every
    ray, towards trees of brick dimension 8, 1 and 32."""
    nan, inf = float("nan"), float("inf")
    origins = [(64.5, 200.0, 64.5), (30.0, 40.0, 50.0), (-1e30, 64.0, 64.0), (1e30, 1e30, 1e30),
               (inf, 10.0, 10.0), (nan, 10.0, 10.0), (64.0, 64.0, 64.0), (-5.0, -5.0, -5.0),
               (127.99999, 0.0, 64.0), (3e9, 5.0, 5.0)]
    dirs = [(0, 0, 0), (nan, 0, -1), (0, nan, 0), (nan, nan, nan), (1e20, -1e20, 1e20),
            (1e-30, -1.0, 1e-30), (1e-40, 1e-40, -1.0), (inf, -1, 0), (-inf, -inf, inf),
            (0.6, -0.8, 0.0), (1, 1, 1), (-0.0, -0.0, 1.0), (-1.0, 0.0, 0.0)]
    rays = np.array([[*o, *d] for o in origins for d in dirs], np.float32)
    for tree in (W.build_scene(W.PLAIN, seed=1), Octree.scene(1, 64, 1, 3),
                 Octree.scene(2, 128, 32, 5)):
        a = tree.flatten().arrays()
        exp = oracle.OracleTree(tree.export_model()).cast(rays)
        for regrouped in (False, True, "nocrawl"):
            got = hostsim.cast(a, rays, regrouped=regrouped)
            res = compare(exp, {k: got[k] for k in got.dtype.names})
            assert all(v == 0 for k, v in res.items() if k != "rays"), (regrouped, res)


def test_rays_that_leave_the_content_behind():
    """Rays with nothing ahead of them at the root level (an empty tree, rays
    passing beside the content of a sparse scene) make the reference restart
    from the root at p + 0.1 d until they leave the tree; the kernels replay
    only those nudges and the iteration count (crawl_out_of_root). Both
    arrangements equal the oracle, which runs the full restarts, and agree
    with each other under lowered iteration bounds (where the replayed count
    decides the guard flag)."""
    for tree in (Octree.new(64, 4), Octree.new(512, 16), Octree.scene(1, 256, 16, 3)):
        size = tree.get_size()
        a = tree.flatten().arrays()
        rays = W.incoherent_rays(size, 1500, seed=3)
        vp = W.look_at_viewport((1.3 * size, 0.9 * size, 1.2 * size), (size / 2,) * 3, 48, 32)
        prim = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, 48, 32)
        o = oracle.OracleTree(tree.export_model())
        for r in (rays, prim):
            exp = o.cast(r)
            for regrouped in (False, True, "nocrawl"):
                got = hostsim.cast(a, r, regrouped=regrouped)
                res = compare(exp, {k: got[k] for k in got.dtype.names})
                assert all(v == 0 for k, v in res.items() if k != "rays"), (size, regrouped, res)
        for limit in (7, 300, 3000):
            g1 = hostsim.cast(a, rays, max_steps=limit)
            for form in (True, "nocrawl"):
                g2 = hostsim.cast(a, rays, max_steps=limit, regrouped=form)
                assert g1.tobytes() == g2.tobytes(), (size, limit, form)


def test_regrouped_traversal_under_forced_guard():
    """cast_ray_ww counts iterations exactly as cast_ray (an iteration
    resumed after its probe is not counted twice): with lowered bounds both
    end the same rays."""
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                cfg.height)[::5]
    a = tree.flatten().arrays()
    for limit in (1, 2, 3, 5, 8, 13):
        base = hostsim.cast(a, rays, max_steps=limit)
        assert base.tobytes() == hostsim.cast(a, rays, max_steps=limit, regrouped=True).tobytes()
        assert base.tobytes() == hostsim.cast(a, rays, max_steps=limit,
                                              regrouped="nocrawl").tobytes()


def test_forced_iteration_guard_reports_and_ends_rays():
    """A lowered iteration bound (the guard the reference lacks) ends long rays
    as misses with flag bit 0 and leaves every other ray exactly as the
    oracle has it."""
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                cfg.height)[::7]
    exp = oracle.OracleTree(tree.export_model()).cast(rays)
    a = tree.flatten().arrays()
    for limit in (1, 3, 6):
        got = hostsim.cast(a, rays, max_steps=limit)
        guarded = (got["flags"] & 1) == 1
        assert guarded.any() and not guarded.all(), limit
        assert (got["hit"][guarded] == 0).all()
        ok = ~guarded
        res = compare(exp[ok], {k: got[k][ok] for k in got.dtype.names})
        assert all(v == 0 for k, v in res.items() if k != "rays"), (limit, res)
    # the default bound is never reached on this workload
    assert (hostsim.cast(a, rays)["flags"] == 0).all()


def test_parity_comparison_itself():
    """The comparison used by every parity test: a NaN equals any NaN (sign
    and payload of generated NaNs are platform-defined), while a NaN against a
    number and a one-ulp difference both count as mismatches."""
    dt = np.dtype([("hit", "<u4"), ("impact", "<f4", (3,))])
    e = np.zeros(3, dt)
    e["hit"] = 1
    e["impact"] = [[np.nan, 1, 2], [1, 2, 3], [np.nan, 0, 0]]
    g = {"hit": e["hit"].copy(),
         "impact": np.array([[-np.nan, 1, 2], [1, 2, np.nextafter(np.float32(3), np.float32(4))],
                             [1.0, 0, 0]], np.float32)}
    assert compare(e, g) == {"rays": 3, "hit": 0, "impact": 2}
