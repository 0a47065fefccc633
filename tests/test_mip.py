"""MIP maps of the host octree (octree/mipmap.rs), mirroring the reference's
mipmap_tests (octree/tests.rs:155-588), plus the strategy API, the default
strategy, incremental-versus-recalculated agreement and persistence."""
from __future__ import annotations

import math

import numpy as np
import pytest

from shocovox_amd import MIPResamplingMethods as M, Octree, OctreeEntry

OOB = 8
RED, GREEN, BLUE = 0xFF0000FF, 0x00FF00FF, 0x0000FFFF


def _v(c: int) -> OctreeEntry:
    return OctreeEntry.visual(c)


def _mix(n: int, channels) -> int:
    """((255^2 / n).sqrt() as u32) in the given byte positions, alpha 255
    (the expected values of the reference tests)."""
    f = np.float32(255.0) * np.float32(255.0) / np.float32(n)
    v = int(np.sqrt(np.float32(f)))
    out = 0xFF
    for shift in channels:
        out |= v << shift
    return out


RG_MIX = _mix(2, (16, 24))
RGB_MIX = _mix(3, (8, 16, 24))


def _albedo(e: OctreeEntry):
    return e.albedo if e.kind in (1, 3) else None


def _six(tree, base=(0, 0, 0)):
    x, y, z = base
    for p, c in (((0, 0, 0), RED), ((0, 0, 1), GREEN), ((0, 1, 0), RED), ((0, 1, 1), GREEN),
                 ((1, 0, 0), RED), ((1, 0, 1), GREEN)):
        tree.insert((x + p[0], y + p[1], z + p[2]), _v(c))


def test_mixed_mip_lvl1():
    """test_mixed_mip_lvl1 (tests.rs:161-204)."""
    t = Octree.new(2, 1)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter)
    _six(t)
    s = t.albedo_mip_map_resampling_strategy()
    assert _albedo(s.sample_root_mip(OOB, (0, 0, 0))) == RG_MIX


def test_simple_solid_mip_lvl2_where_dim_is_2():
    """(tests.rs:206-276)"""
    t = Octree.new(4, 2)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter)
    for p in ((0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1)):
        t.insert(p, _v(RED))
    s = t.albedo_mip_map_resampling_strategy()
    assert _albedo(s.sample_root_mip(OOB, (0, 0, 0))) == RED
    for p in ((0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
        assert _albedo(s.sample_root_mip(OOB, p)) is None, p


def test_mixed_mip_lvl2_where_dim_is_2():
    """(tests.rs:278-356)"""
    t = Octree.new(4, 2)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter)
    _six(t)
    s = t.albedo_mip_map_resampling_strategy()
    assert _albedo(s.sample_root_mip(OOB, (0, 0, 0))) == RG_MIX
    for p in ((0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
        assert _albedo(s.sample_root_mip(OOB, p)) is None, p


def _twelve(t):
    _six(t)
    for p, c in (((8, 0, 0), RED), ((8, 0, 1), GREEN), ((8, 1, 0), BLUE), ((8, 1, 1), GREEN),
                 ((9, 1, 0), RED), ((9, 0, 1), BLUE)):
        t.insert(p, _v(c))


def _check_lvl2_dim4(t):
    s = t.albedo_mip_map_resampling_strategy()
    assert _albedo(s.sample_root_mip(0, (0, 0, 0))) == RG_MIX
    assert _albedo(s.sample_root_mip(1, (0, 0, 0))) == RGB_MIX
    assert _albedo(s.sample_root_mip(OOB, (0, 0, 0))) == RG_MIX
    assert _albedo(s.sample_root_mip(OOB, (2, 0, 0))) == RGB_MIX


def test_mixed_mip_lvl2_where_dim_is_4():
    """(tests.rs:358-471)"""
    t = Octree.new(16, 4)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter).set_method_at(2, M.BoxFilter)
    _twelve(t)
    _check_lvl2_dim4(t)


def test_mixed_mip_regeneration_lvl2_where_dim_is_4():
    """(tests.rs:473-587): MIPs switched on after the inserts are recalculated."""
    t = Octree.new(16, 4)
    t.set_auto_simplify(False)
    _twelve(t)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter).set_method_at(2, M.BoxFilter).recalculate_mips()
    _check_lvl2_dim4(t)


def test_expected_mixes_by_hand():
    """The reference's expected colours, from the definition of the box
    filter: sqrt(mean of squares) per channel, truncated."""
    assert RG_MIX == (180 << 24) | (180 << 16) | 0xFF
    assert RGB_MIX == (147 << 24) | (147 << 16) | (147 << 8) | 0xFF
    assert int(math.sqrt(255 * 255 / 2)) == 180 and int(math.sqrt(255 * 255 / 3)) == 147


def test_default_strategy_and_setters():
    """MIPMapStrategy::default (mipmap.rs:591-604) and the clamping setters."""
    s = Octree.new(8, 2).albedo_mip_map_resampling_strategy()
    assert not s.is_enabled()
    assert s.get_method_at(1) == M.PointFilter
    assert all(s.get_method_at(k) == M.BoxFilter for k in (0, 2, 3, 4, 5, 9))
    assert [s.get_new_color_similarity_at(k) for k in (1, 2, 3, 4, 5)] == pytest.approx(
        [0.0, 0.1, 0.05, 0.02, 0.0])
    s.set_method_at(3, M.Posterize(1.7)).set_color_similarity_thr_at(5, -2.0)
    assert s.get_method_at(3) == M.Posterize(1.0)
    assert s.get_new_color_similarity_at(5) == 0.0
    s.set_method({6: M.PointFilterBD}).set_color_similarity_thr({6: 0.25})
    assert s.get_method_at(6) == M.PointFilterBD and s.get_new_color_similarity_at(6) == 0.25
    s.reset()
    assert s.get_method_at(6) == M.BoxFilter and s.get_method_at(1) == M.PointFilter


def test_point_filter_takes_the_majority():
    """PointFilter (mipmap.rs:181-200): the most frequent colour of the 2^3
    cell (3 red, 2 green here)."""
    t = Octree.new(2, 1)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    for p, c in (((0, 0, 0), RED), ((0, 0, 1), GREEN), ((1, 1, 0), RED), ((1, 0, 1), GREEN),
                 ((0, 1, 1), RED)):
        t.insert(p, _v(c))
    assert _albedo(t.albedo_mip_map_resampling_strategy().sample_root_mip(OOB, (0, 0, 0))) == RED


def test_posterize_averages_the_largest_group():
    """Posterize(thr) (mipmap.rs:201-268): colours within thr*255 of a
    group's gamma mean join it; the largest group's mean is the MIP. (The
    reference's u32 channel differences wrap below zero in a release build,
    so a group absorbs only colours no brighter than its mean; the brighter
    colour is sampled first here.)"""
    a, b, far = 0xD20000FF, 0xC80000FF, 0x0000FFFF  # 210, then 200 red; blue far away
    t = Octree.new(2, 1)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.Posterize(0.1))
    for p, c in (((0, 0, 0), a), ((0, 0, 1), b), ((0, 1, 0), far)):
        t.insert(p, _v(c))
    got = _albedo(t.albedo_mip_map_resampling_strategy().sample_root_mip(OOB, (0, 0, 0)))
    r = round(math.sqrt(round((210 * 210 + 200 * 200) / 2)))
    assert r == 205
    assert got == (r << 24) | 0xFF


def test_color_threshold_reuses_a_palette_colour():
    """With a similarity threshold at the level, a sampled colour within
    thr*255 of a palette colour becomes the first such colour of the palette
    (mipmap.rs:527-560): the red/green mix lies 195 from red, under 204."""
    t = Octree.new(4, 2)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter).set_color_similarity_thr_at(1, 0.8)
    _six(t)
    assert _albedo(t.albedo_mip_map_resampling_strategy().sample_root_mip(OOB, (0, 0, 0))) == RED


@pytest.mark.parametrize("size,dim", [(16, 2), (32, 4), (64, 8)])
def test_incremental_mips_equal_recalculated(size, dim):
    """Inserting with MIPs on gives the same MIPs as inserting first and
    recalculating (BoxFilter everywhere: no ties to break)."""
    rng = np.random.default_rng(size + dim)
    n = 400
    xyz = rng.integers(0, size, size=(n, 3)).astype(np.uint32)
    cols = (rng.integers(1, 6, n).astype(np.uint32) * 0x28000000) | 0x3030FF

    def build(mips_first: bool) -> Octree:
        t = Octree.new(size, dim)
        t.set_auto_simplify(False)
        st = t.albedo_mip_map_resampling_strategy().set_method_at(1, M.BoxFilter)
        st.set_color_similarity_thr({2: 0.0, 3: 0.0, 4: 0.0})
        if mips_first:
            st.switch_albedo_mip_maps(True)
        for p, c in zip(xyz, cols):
            t.insert(tuple(int(v) for v in p), _v(int(c)))
        if not mips_first:
            st.switch_albedo_mip_maps(True)
        return t

    a, b = build(True), build(False)
    sa, sb = a.albedo_mip_map_resampling_strategy(), b.albedo_mip_map_resampling_strategy()
    for o in range(9):
        for x in range(dim):
            for y in range(dim):
                for z in range(dim):
                    assert _albedo(sa.sample_root_mip(o, (x, y, z))) == \
                        _albedo(sb.sample_root_mip(o, (x, y, z))), (o, x, y, z)


def test_mips_survive_save_and_load(tmp_path):
    t = Octree.new(16, 4)
    t.set_auto_simplify(False)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True).set_method_at(
        1, M.BoxFilter).set_method_at(2, M.Posterize(0.25))
    _twelve(t)
    c = Octree.from_bytes(t.to_bytes())
    s, u = t.albedo_mip_map_resampling_strategy(), c.albedo_mip_map_resampling_strategy()
    assert u.is_enabled() and u.get_method_at(2) == M.Posterize(0.25)
    assert u.get_method_at(1) == M.BoxFilter
    for o in range(9):
        for p in ((0, 0, 0), (2, 0, 0), (1, 1, 1), (3, 3, 3)):
            assert s.sample_root_mip(o, p) == u.sample_root_mip(o, p)
    assert c.to_bytes() == t.to_bytes()
