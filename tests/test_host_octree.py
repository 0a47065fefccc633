"""Host octree (the builder feeding the device layout): the reference's edit
known-answer tests (src/octree/update/tests.rs, object_pool.rs tests) run
against the product's insert / update / clear / get, plus the construction
errors of Octree::new (src/octree/mod.rs:173-205)."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from shocovox_amd import EMPTY, Octree, OctreeEntry, OctreeError, voxel_data
from shocovox_amd._native import SvoError

RED = OctreeEntry.visual(0xFF0000FF)
GREEN = OctreeEntry.visual(0x00FF00FF)
BLUE = OctreeEntry.visual(0x0000FFFF)
ALBEDO = OctreeEntry.visual(0xFFAAEEFF)
EMPTY_ENTRY = OctreeEntry()
# OCTANT_OFFSET_REGION_LUT as integer offsets (lut.rs:156-197)
OFFSETS = [(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1), (0, 1, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1)]


def cube(n, start=0):
    return itertools.product(range(start, n), repeat=3)


def count_hits(tree, n, expect=None):
    hits = 0
    for p in cube(n):
        h = tree.get(p)
        if h != EMPTY_ENTRY:
            if expect is not None:
                assert h == expect, p
            hits += 1
    return hits


def new(size, dim, simplify=True):
    t = Octree.new(size, dim)
    t.set_auto_simplify(simplify)
    return t


# --- Octree::new validation (octree/mod.rs:173-205) ------------------------

@pytest.mark.parametrize("size,dim,status", [
    (0, 1, 2), (8, 3, 2), (8, 16, 1), (12, 2, 1), (4, 4, 3), (2, 1, 0), (128, 8, 0),
    (2048, 32, 0), (8192, 1, 1),  # f32 log2(8192) is 12.999999 in the reference's check
])
def test_new_validation(size, dim, status):
    if status == 0:
        t = Octree.new(size, dim)
        assert t.get_size() == size and t.brick_dim == dim
    else:
        with pytest.raises(OctreeError) as e:
            Octree.new(size, dim)
        assert e.value.status == status


def test_insert_out_of_bounds_is_invalid_position():
    t = Octree.new(4, 1)
    with pytest.raises(SvoError) as e:
        t.insert((4, 0, 0), RED)
    assert e.value.status == 4
    with pytest.raises(SvoError):
        t.clear((0, 5, 0))
    assert t.get((9, 9, 9)) == EMPTY_ENTRY


# --- update/tests.rs --------------------------------------------------------

def test_simplest_insert_and_get():  # :9-17
    red = OctreeEntry.visual(0xFF000001)
    t = new(2, 1, False)
    t.insert((0, 0, 0), red)
    assert t.get((0, 0, 0)) == red


@pytest.mark.parametrize("size,dim", [(2, 1), (4, 2)])
def test_simple_insert_and_get(size, dim):  # :20-46, :113-140
    t = new(size, dim, False)
    t.insert((1, 0, 0), RED)
    t.insert((0, 1, 0), GREEN)
    t.insert((0, 0, 1), BLUE)
    assert t.get((1, 0, 0)) == RED and t.get((0, 1, 0)) == GREEN and t.get((0, 0, 1)) == BLUE
    assert t.get((1, 1, 1)) == EMPTY_ENTRY
    t.insert((1, 0, 0), GREEN)
    assert t.get((1, 0, 0)) == GREEN and t.get((0, 1, 0)) == GREEN
    assert t.get((0, 0, 1)) == BLUE and t.get((1, 1, 1)) == EMPTY_ENTRY


def test_insert_empty():  # :49-54
    t = new(2, 1, False)
    t.insert((0, 0, 0), EMPTY_ENTRY)
    assert t.get((0, 0, 0)) == EMPTY_ENTRY


def test_complex_insert_and_get():  # :57-110
    red, green = 0xFF0000FF, 0x00FF00FF
    t = new(2, 1, False)
    t.insert((1, 0, 0), OctreeEntry.complex(red, 3))
    t.insert((0, 1, 0), OctreeEntry.complex(green, 1))
    t.insert((0, 0, 1), voxel_data(2))
    assert t.get((1, 0, 0)) == OctreeEntry.complex(red, 3)
    assert t.get((0, 1, 0)) == OctreeEntry.complex(green, 1)
    assert t.get((0, 0, 1)) == voxel_data(2)
    assert t.get((1, 1, 1)) == EMPTY_ENTRY
    t.insert((1, 0, 0), voxel_data(3))
    assert t.get((1, 0, 0)) == voxel_data(3)
    assert t.get((0, 1, 0)) == OctreeEntry.complex(green, 1)


@pytest.mark.parametrize("dim", [1, 2])
def test_insert_at_lod(dim):  # :143-242
    t = new(8, dim, False)
    t.insert_at_lod((0, 0, 0), 2, RED)
    for p in cube(2):
        assert t.get(p) == RED
    t.insert_at_lod((0, 0, 0), 4, GREEN)
    assert count_hits(t, 4, GREEN) == 64


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_simplified_insert_separated_by_clear(dim):  # :245-359
    t = Octree.new(8, dim)
    for p in cube(8):
        t.insert(p, RED)
    assert t.get((3, 3, 3)) == RED
    t.clear((3, 3, 3))
    assert t.get((3, 3, 3)) == EMPTY_ENTRY
    assert count_hits(t, 8, RED) == 511


def test_update_color_and_data():  # :361-402
    t = new(2, 1, False)
    t.insert((0, 0, 0), OctreeEntry.complex(0xFF0000FF, 3))
    t.update((0, 0, 0), OctreeEntry.visual(0x00FF00FF))
    assert t.get((0, 0, 0)) == OctreeEntry.complex(0x00FF00FF, 3)
    t.update((0, 0, 0), voxel_data(4))
    assert t.get((0, 0, 0)) == OctreeEntry.complex(0x00FF00FF, 4)
    t.update((0, 0, 0), EMPTY_ENTRY)
    assert t.get((0, 0, 0)) == OctreeEntry.complex(0x00FF00FF, 4)


@pytest.mark.parametrize("separate_by", ["clear", "insert"])
def test_uniform_solid_leaf_separated_dim1(separate_by):  # :404-461
    t = Octree.new(2, 1)
    base = OctreeEntry.visual(0xFFFF00FF)
    for off in OFFSETS:
        t.insert(off, base)
    assert t.get((0, 0, 0)) == base
    if separate_by == "clear":
        t.clear((0, 0, 0))
        assert t.get((0, 0, 0)) == EMPTY_ENTRY
    else:
        t.insert((0, 0, 0), base)
        assert t.get((0, 0, 0)) == base
    for off in OFFSETS[1:]:
        assert t.get(off) == base


@pytest.mark.parametrize("separate_by", ["clear", "insert"])
def test_uniform_solid_leaf_separated_dim4(separate_by):  # :463-568
    dim = 4
    t = Octree.new(8, dim)
    base = 0xFFFF00AA
    for o, off in enumerate(OFFSETS):
        for p in cube(dim // 2):
            pos = tuple(off[i] * (dim // 2) + p[i] for i in range(3))
            t.insert(pos, OctreeEntry.visual(base + o))
    assert t.get((0, 0, 0)) == OctreeEntry.visual(base)
    modified = OctreeEntry.visual(0x000000FF)
    if separate_by == "clear":
        t.clear((0, 0, 0))
        assert t.get((0, 0, 0)) == EMPTY_ENTRY
    else:
        t.insert((0, 0, 0), modified)
        assert t.get((0, 0, 0)) == modified
    for o, off in enumerate(OFFSETS):
        for p in cube(dim // 2):
            if p == (0, 0, 0) and o == 0:
                continue
            pos = tuple(off[i] * (dim // 2) + p[i] for i in range(3))
            assert t.get(pos) == OctreeEntry.visual(base + o)


def _block_color(x, y, z, dim):
    h = dim // 2
    r, g, b = x - x % h, y - y % h, z - z % h
    return OctreeEntry.visual((r << 24) | (g << 16) | (b << 8) | 255)


@pytest.mark.parametrize("separate_by", ["clear", "insert"])
def test_uniform_parted_brick_leaf_separated_dim4(separate_by):  # :570-641, :695-772
    dim = 4
    t = Octree.new(8, dim)
    for o, off in enumerate(OFFSETS):
        for p in cube(dim):
            pos = tuple(off[i] * dim + p[i] for i in range(3))
            t.insert(pos, _block_color(*p, dim))
            assert t.get(pos) == _block_color(*p, dim)
    assert t.get((0, 0, 0)) == OctreeEntry.visual(0x000000FF)
    modified = OctreeEntry.visual(0xFF0000FF)
    if separate_by == "clear":
        t.clear((1, 1, 1))
        assert t.get((1, 1, 1)).is_none()
    else:
        t.insert((1, 1, 1), modified)
        assert t.get((1, 1, 1)) == modified
        assert t.get((0, 0, 0)) == OctreeEntry.visual(0x000000FF)
    for o, off in enumerate(OFFSETS):
        for p in cube(dim):
            pos = tuple(off[i] * dim + p[i] for i in range(3))
            if p == (1, 1, 1) and o == 0:
                continue
            assert t.get(pos) == _block_color(*p, dim), pos


def test_simple_uniform_parted_brick_leaf_separated_by_insert():  # :643-693
    dim = 2
    t = Octree.new(4, dim)
    base = 0xF00000FF
    for o, off in enumerate(OFFSETS):
        t.insert_at_lod(tuple(v * dim for v in off), dim, OctreeEntry.visual(base + 2 * o))
    assert t.get((0, 0, 0)) == OctreeEntry.visual(base)
    modified = OctreeEntry.visual(0x000000FF)
    t.insert((0, 0, 0), modified)
    assert t.get((0, 0, 0)) == modified
    for p in cube(dim):
        for o, off in enumerate(OFFSETS):
            pos = tuple(off[i] * dim + p[i] for i in range(3))
            if p == (0, 0, 0) and o == 0:
                assert t.get(pos) == modified
            else:
                assert t.get(pos) == OctreeEntry.visual(base + 2 * o)


@pytest.mark.parametrize("dim,pos,size,hits", [
    (4, (1, 1, 1), 4, 27),   # :774-812
    (1, (2, 2, 2), 3, 8),    # :814-844
    (1, (3, 3, 3), 3, 1),    # :846-877
    (4, (1, 1, 1), 3, 27),   # :879-910
])
def test_insert_at_lod_unaligned(dim, pos, size, hits):
    t = new(8, dim, False)
    t.insert_at_lod(pos, size, RED)
    assert count_hits(t, 8, RED) == hits


def test_insert_at_lod_with_simplify():  # :912-966
    t = Octree.new(8, 1)
    t.insert_at_lod((4, 0, 0), 2, RED)
    for p in cube(2):
        assert t.get((4 + p[0], p[1], p[2])) == RED
    t.insert_at_lod((0, 0, 0), 4, GREEN)
    hits = count_hits(t, 4, GREEN)
    for p in itertools.product(range(4, 6), range(2), range(2)):
        if t.get(p) != EMPTY_ENTRY:
            assert t.get(p) == RED
            hits += 1
    assert hits == 64 + 8


@pytest.mark.parametrize("size,dim", [(2, 1), (4, 2)])
def test_simplifyable_insert_and_get(size, dim):  # :968-1028
    t = Octree.new(size, dim)
    for p in cube(size):
        t.insert(p, RED)
    t.insert((0, 0, 0), GREEN)
    assert t.get((0, 0, 0)) == GREEN
    for p in cube(size, 1):
        assert t.get(p) == RED


@pytest.mark.parametrize("size,dim", [(2, 1), (4, 2)])
def test_simple_clear(size, dim):  # :1030-1070
    t = new(size, dim, False)
    t.insert((1, 0, 0), RED)
    t.insert((0, 1, 0), GREEN)
    t.insert((0, 0, 1), BLUE)
    t.clear((0, 0, 1))
    assert t.get((1, 0, 0)) == RED and t.get((0, 1, 0)) == GREEN
    assert t.get((0, 0, 1)) == EMPTY_ENTRY and t.get((1, 1, 1)) == EMPTY_ENTRY


def test_ocbits_resolution_cases():  # :1072-1119
    t = Octree.new(64, 8)
    t.insert((0, 1, 1), RED)
    t.insert((1, 0, 0), RED)
    t.clear((1, 0, 0))
    assert t.get((1, 0, 0)) == EMPTY_ENTRY and t.get((0, 1, 1)) == RED
    t = Octree.new(64, 8)
    t.insert_at_lod((33, 33, 33), 2, RED)
    assert t.get((33, 33, 33)) == RED
    t.clear((33, 33, 33))
    assert t.get((33, 33, 33)) == EMPTY_ENTRY
    t = Octree.new(64, 8)
    t.insert((31, 31, 31), RED)
    t.clear((31, 31, 31))
    assert t.get((31, 31, 31)) == EMPTY_ENTRY


def test_double_clear():  # :1121-1137
    t = new(2, 1, False)
    black, white = OctreeEntry.visual(0x000000FF), OctreeEntry.visual(0xFFFFFFFF)
    t.insert((1, 0, 0), black)
    t.insert((0, 1, 0), white)
    t.insert((0, 0, 1), white)
    t.clear((0, 0, 1))
    t.clear((0, 0, 1))
    assert t.get((1, 0, 0)) == black and t.get((0, 1, 0)) == white
    assert t.get((0, 0, 1)) == EMPTY_ENTRY


@pytest.mark.parametrize("size,dim", [(2, 1), (4, 2)])
def test_simplifyable_clear(size, dim):  # :1139-1197
    t = Octree.new(size, dim)
    for p in cube(size):
        t.insert(p, ALBEDO)
    t.clear((0, 0, 0))
    assert t.get((0, 0, 0)) == EMPTY_ENTRY
    for p in cube(size, 1):
        assert t.get(p) == ALBEDO


def test_clear_to_nothing():  # :1199-1227
    t = Octree.new(4, 1)
    for p in cube(2):
        t.insert(p, ALBEDO)
    t.clear_at_lod((0, 0, 0), 2)
    for p in cube(2):
        assert t.get(p) == EMPTY_ENTRY


def test_clear_edge_case():  # :1229-1279
    t = Octree.new(64, 16)
    t.update((1, 0, 0), voxel_data(0xFACEFEED))
    t.insert_at_lod((0, 0, 0), 32, RED)
    t.clear_at_lod((5, 5, 5), 8)
    for p in cube(8, 5):
        assert t.get(p) == EMPTY_ENTRY
    for p in cube(5):
        assert t.get(p) == RED, p
    t.clear_at_lod((0, 0, 0), 32)
    xyz = np.array(list(cube(32)), np.uint32)
    assert (t.get_many(xyz)[:, 0] == EMPTY).all()


@pytest.mark.parametrize("dim,size,start,clear_size,n,expected", [
    (1, 8, (0, 0, 0), 2, 4, 64 - 8),          # :1281-1315
    (2, 8, (0, 0, 0), 2, 4, 64 - 8),          # :1317-1351
    (1, 8, (1, 1, 1), 2, 4, 64 - 8),          # :1353-1406
    (1, 8, (0, 0, 0), 3, 4, 64 - 8),          # :1466-1497
])
def test_clear_at_lod(dim, size, start, clear_size, n, expected):
    t = Octree.new(size, dim)
    t.insert_at_lod((0, 0, 0), 4, ALBEDO)
    t.clear_at_lod(start, clear_size)
    assert count_hits(t, n, ALBEDO) == expected


def test_clear_at_lod_with_unaligned_position_dim1_cleared_cells():  # :1353-1406
    t = Octree.new(8, 1)
    t.insert_at_lod((0, 0, 0), 4, ALBEDO)
    t.clear_at_lod((1, 1, 1), 2)
    for p in cube(2):
        assert t.get(p) == EMPTY_ENTRY
    for p in [(0, 0, 2), (0, 2, 0), (0, 2, 2), (2, 0, 0), (2, 0, 2), (2, 2, 0), (2, 2, 2)]:
        assert t.get(p).is_some()


def test_clear_at_lod_with_unaligned_position_where_dim_is_4():  # :1408-1464
    t = Octree.new(16, 4)
    t.insert_at_lod((0, 0, 0), 8, ALBEDO)
    assert count_hits(t, 8, ALBEDO) == 512
    t.clear_at_lod((1, 1, 1), 4)
    assert count_hits(t, 8, ALBEDO) == 512 - 27


def test_clear_at_lod_with_unaligned_size_where_dim_is_4():  # :1499-1548
    t = Octree.new(8, 4)
    t.insert_at_lod((0, 0, 0), 4, ALBEDO)
    assert count_hits(t, 8, ALBEDO) == 64
    t.clear_at_lod((0, 0, 0), 3)
    assert count_hits(t, 8, ALBEDO) == 64 - 27


def test_clear_whole_nodes_where_dim_is_4():  # :1550-1599
    t = Octree.new(16, 4)
    t.insert_at_lod((0, 0, 0), 8, ALBEDO)
    assert count_hits(t, 8, ALBEDO) == 512
    t.clear_at_lod((0, 0, 0), 5)
    assert count_hits(t, 8, ALBEDO) == 512 - 64


def test_overwrite_whole_nodes_where_dim_is_4():  # :1601-1656
    t = Octree.new(16, 4)
    t.insert_at_lod((0, 0, 0), 8, RED)
    assert count_hits(t, 8, RED) == 512
    t.insert_at_lod((0, 0, 0), 5, BLUE)
    red = blue = 0
    for p in cube(8):
        h = t.get(p)
        assert h != EMPTY_ENTRY
        red += h == RED
        blue += h == BLUE
    assert (red, blue) == (512 - 64, 64)


def test_edge_case_octree_set():  # :1658-1674
    t = Octree.new(8, 1)
    for p in cube(8, 6):
        e = OctreeEntry.visual(sum(p))
        t.insert(p, e)
        assert t.get(p) == e


def test_case_inserting_empty():  # :1676-1687
    t = Octree.new(4, 1)
    t.insert((3, 0, 0), OctreeEntry.visual(0))
    assert t.get((3, 0, 0)) == EMPTY_ENTRY


# --- batched entry points ---------------------------------------------------

def test_insert_many_matches_single_inserts():
    rng = np.random.default_rng(4)
    xyz = rng.integers(0, 32, size=(3000, 3)).astype(np.uint32)
    ents = np.zeros((3000, 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = (rng.integers(1, 6, size=3000) << 8) | 0xFF
    a = Octree.new(32, 4)
    a.insert_many(xyz, ents)
    b = Octree.new(32, 4)
    for p, e in zip(xyz, ents):
        b.insert(tuple(int(v) for v in p), OctreeEntry(int(e[0]), int(e[1]), int(e[2])))
    allp = np.array(list(cube(32)), np.uint32)
    assert np.array_equal(a.get_many(allp), b.get_many(allp))
    a.clear_many(xyz[:500])
    for p in xyz[:500]:
        b.clear(tuple(int(v) for v in p))
    assert np.array_equal(a.get_many(allp), b.get_many(allp))


def test_clone_is_equal_and_independent():
    """Octree::clone: the same model and bytes; edits of the copy leave the
    original as it was."""
    t = Octree.scene(2, 64, 4, 5)
    t.albedo_mip_map_resampling_strategy().switch_albedo_mip_maps(True)
    c = t.clone()
    assert c.to_bytes() == t.to_bytes()
    before = t.to_bytes()
    c.insert((1, 2, 3), OctreeEntry.visual(0x12345678))
    c.clear((40, 40, 40))
    assert t.to_bytes() == before and c.to_bytes() != before
