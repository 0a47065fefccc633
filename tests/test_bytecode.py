"""Persisted octrees in the reference's bencode format (Octree::save / load /
to_bytes / from_bytes), mirroring src/convert/bytecode_tests.rs, plus the
layout of the encoding and malformed-input handling.

No file written by the reference itself exists to pin the format against
(parity unpinned for reading reference-written files); the checks below pin
the restatement by round trips of the reference's own test trees, by the
byte layout the reference's encoders define for small known trees, and by
equality of the device layout of a loaded tree."""
from __future__ import annotations

import numpy as np
import pytest

from shocovox_amd import Octree, OctreeEntry, OctreeError, workloads as W


def _rgba(v: int) -> OctreeEntry:
    return OctreeEntry.visual(v & 0xFFFFFFFF)


def _models_equal(a: Octree, b: Octree) -> bool:
    ma, mb = a.export_model(), b.export_model()
    return ma.keys() == mb.keys() and all(np.array_equal(np.asarray(ma[k]), np.asarray(mb[k]))
                                          for k in ma)


def test_octree_file_io(tmp_path):
    """test_octree_file_io (bytecode_tests.rs:181-222)."""
    red = 0xFF0000FF
    tree = Octree.new(8, 1)
    tree.insert_at_lod((0, 0, 0), 4, _rgba(red))
    tree.clear_at_lod((0, 0, 0), 2)
    path = tmp_path / "test_junk_octree"
    tree.save(path)
    copy = Octree.load(path)
    hits = 0
    for x in range(4):
        for y in range(4):
            for z in range(4):
                assert tree.get((x, y, z)) == copy.get((x, y, z))
                e = tree.get((x, y, z))
                if e.kind != 0:
                    assert e == _rgba(red)
                    hits += 1
    assert hits == 64 - 8
    assert _models_equal(tree, copy)


@pytest.mark.parametrize("size,dim,lo", [(128, 1, 100), (128, 2, 100)])
def test_big_octree_serialize(size, dim, lo):
    """test_big_octree_serialize (:224-250) and ..._where_dim_is_2 (:336-371)."""
    tree = Octree.new(size, dim)
    r = np.arange(lo, size, dtype=np.uint32)
    xyz = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    ents = np.zeros((len(xyz), 3), np.uint32)
    ents[:, 0] = 1
    if dim == 1:
        ents[:, 1] = xyz.sum(1)  # Albedo::from(x + y + z)
    else:
        ents[:, 1] = (xyz[:, 0] << 24) + (xyz[:, 1] << 16) + (xyz[:, 2] << 8) + 0xFF
    tree.insert_many(xyz, ents)
    copy = Octree.from_bytes(tree.to_bytes())
    got = copy.get_many(xyz)
    exp = tree.get_many(xyz)
    assert np.array_equal(got, exp)
    assert _models_equal(tree, copy)


def test_small_octree_serialize_where_dim_is_1():
    """(:252-267)"""
    tree = Octree.new(2, 1)
    tree.insert((0, 0, 0), _rgba(1))
    copy = Octree.from_bytes(tree.to_bytes())
    assert copy.get((0, 0, 0)) == tree.get((0, 0, 0))
    assert copy.get((0, 0, 0)).kind != 0


@pytest.mark.parametrize("dim", [1, 2])
def test_octree_serialize_where_dim_is_1_and_2(dim):
    """(:269-301, :303-334)"""
    tree = Octree.new(4, dim)
    for x in range(4):
        for y in range(4):
            for z in range(4):
                tree.insert((x, y, z), _rgba((x << 24) + (y << 16) + (z << 8) + 0xFF))
    copy = Octree.from_bytes(tree.to_bytes())
    for x in range(4):
        for y in range(4):
            for z in range(4):
                assert copy.get((x, y, z)) == _rgba((x << 24) + (y << 16) + (z << 8) + 0xFF)


def test_encoding_layout_of_a_fresh_tree():
    """Octree::new(2, 1) holds one Nothing root with no children and empty
    palettes; its encoding, by the encoders of bytecode.rs:582-596,
    object_pool.rs:25-35, 97-107 and the markers of bytecode.rs:206-234,
    352-372: auto_simplify 1, size 2, dim 1, pool [first_available, [[1, "#"]]]
    with first_available 0 (ObjectPool::allocate, object_pool.rs:177-201, leaves
    it at 0 after the first push), children ["##x##"], MIPs ["#b"], two empty
    palettes, then the MIP strategy."""
    b = Octree.new(2, 1).to_bytes()
    assert b.startswith(b"li1ei2ei1eli0elli1e1:#eeel5:##x##el2:#belele")
    assert b.endswith(b"e")
    # reading it back gives the same tree
    assert Octree.from_bytes(b).to_bytes() == b


def test_encoding_markers_of_every_node_form():
    """Internal, Leaf (Empty / Solid / Parted bricks) and UniformLeaf nodes and
    Children / OccupancyBitmap / NoChildren connections all round trip, and
    their markers appear as the reference's encoders write them."""
    t = Octree.new(16, 2)
    t.insert((1, 1, 1), _rgba(0x11223344))       # Parted brick in a leaf
    t.insert_at_lod((8, 8, 8), 4, _rgba(0x55667788))  # Solid bricks
    b = t.to_bytes()
    for marker in (b"2:##i", b"3:###", b"4:##b#", b"3:#b#", b"2:#b", b"5:##c##", b"5:##b##"):
        assert marker in b, marker
    u = Octree.from_bytes(b)
    assert _models_equal(t, u) and u.to_bytes() == b
    # a whole-tree solid fill simplifies to a UniformLeaf root
    s = Octree.new(4, 2)
    s.insert_at_lod((0, 0, 0), 4, _rgba(0x01020304))
    sb = s.to_bytes()
    assert b"4:##u#" in sb
    assert _models_equal(s, Octree.from_bytes(sb))


def test_caves_tree_round_trip_and_layout(tmp_path):
    """A 128^3 caves tree: save, load, identical model and device layout."""
    tree = Octree.scene(2, 128, 8, 3)
    path = tmp_path / "caves.bencode"
    tree.save(path)
    copy = Octree.load(path)
    assert _models_equal(tree, copy)
    fa, fb = tree.flatten().arrays(), copy.flatten().arrays()
    for k in ("nodes", "brick_pix", "brick_occ", "colors"):
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("bad", [b"", b"l", b"li1ei8ei1e", b"le", b"i5e",
                                 b"li1ei8ei1eli0elli1e2:#?eeeeee", b"li2ei8ei1e",
                                 b"li1ei12ei1e"])
def test_malformed_input_is_refused(bad):
    with pytest.raises(OctreeError) as e:
        Octree.from_bytes(bad)
    assert e.value.status == 3  # SVO_INVALID_STRUCTURE


def test_truncated_real_file_is_refused():
    b = W.build_scene(W.PLAIN, seed=1).to_bytes()
    for cut in (10, len(b) // 3, len(b) - 1):
        with pytest.raises(OctreeError):
            Octree.from_bytes(b[:cut])


def test_load_missing_file(tmp_path):
    with pytest.raises(OctreeError):
        Octree.load(tmp_path / "nope")
