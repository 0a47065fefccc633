"""MagicaVoxel import (Octree::load_vox_file, convert/magicavoxel.rs:259-395).

The oracle (oracle/vox.py) restates the import; the product (libsvo,
svo_octree_load_vox) must build the same tree. The oracle itself is pinned
by the reference's rotation known-answer test (magicavoxel.rs:398-414), by
the group structure of the rotation encoding, by hand-derived positions of
small synthetic files, by numpy rotations of dense grids, and by the scene
structure of the reference's asset files (navigate.vox holds the parts of
navigate_x/y/z.vox). The asset files are copied under tests/golden/vox by
tests/golden/vox/copy_assets.py.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import vox_writer as VW
from oracle import vox as OV
from shocovox_amd import Octree, OctreeError

ASSETS = Path(__file__).resolve().parent / "golden" / "vox"
NAMES = ("navigate", "navigate_x", "navigate_y", "navigate_z")
PAL = VW.palette_ramp()


def _pack(c) -> int:
    r, g, b, a = c
    return (r << 24) | (g << 16) | (b << 8) | a


def _voxels(tree: Octree) -> dict:
    """Every non-empty voxel of a small tree: position -> albedo."""
    s = tree.get_size()
    r = np.arange(s, dtype=np.uint32)
    xyz = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    e = tree.get_many(xyz)
    keep = e[:, 0] != 0
    return {tuple(int(v) for v in p): int(a) for p, a in zip(xyz[keep], e[keep, 1])}


def _both(data: bytes, dim: int = 1):
    """(product tree, oracle size, oracle final voxels)."""
    size, ins = OV.load(data, dim)
    return Octree.from_vox_bytes(data, dim), size, OV.final_voxels(ins)


def _models_equal(a: Octree, b: Octree) -> bool:
    ma, mb = a.export_model(), b.export_model()
    return ma.keys() == mb.keys() and all(np.array_equal(np.asarray(ma[k]), np.asarray(mb[k]))
                                          for k in ma)


# ---------------------------------------------------------------- rotations

def test_matrix_parse():
    """test_matrix_parse (magicavoxel.rs:398-414)."""
    assert OV.parse_rotation_matrix(4) == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    ex = OV.parse_rotation_matrix((1 << 0) | (2 << 2) | (0 << 4) | (1 << 5) | (1 << 6))
    assert ex == [[0, 1, 0], [0, 0, -1], [-1, 0, 0]]


def test_rotation_bytes_cover_the_signed_permutations():
    """The valid bytes (bit 7 unused) name each of the 48 signed 3x3
    permutation matrices exactly twice; every other byte is refused."""
    seen = {}
    for b in range(256):
        try:
            m = OV.parse_rotation_matrix(b)
        except OV.VoxError:
            continue
        a = np.array(m)
        assert (np.abs(a).sum(0) == 1).all() and (np.abs(a).sum(1) == 1).all()
        seen.setdefault(tuple(map(tuple, m)), []).append(b)
    assert len(seen) == 48 and all(len(v) == 2 for v in seen.values())
    assert all(v[1] == v[0] | 0x80 for v in seen.values())


def _rotated_grid(dense: np.ndarray, m) -> np.ndarray:
    """numpy placement of a model under rotation m: output axis r is input
    axis c_r, reversed where the sign is negative, starting at 0."""
    cols = [int(np.nonzero(row)[0][0]) for row in m]
    out = np.transpose(dense, cols)
    for r, row in enumerate(m):
        if row[cols[r]] < 0:
            out = np.flip(out, r)
    return out


def _grid_voxels_lyup(grid: np.ndarray) -> dict:
    """Rzup grid -> {Lyup position: albedo}."""
    out = {}
    for x, y, z in zip(*np.nonzero(grid)):
        out[(int(x), int(z), int(y))] = _pack(PAL[int(grid[x, y, z]) - 1])
    return out


def _dense_model(size, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros(size, np.int64)
    mask = rng.random(size) < 0.3
    g[mask] = rng.integers(1, 256, int(mask.sum()))
    vox = [(int(x), int(y), int(z), int(g[x, y, z])) for x, y, z in zip(*np.nonzero(g))]
    return g, vox


# ------------------------------------------------------- synthetic placements

def test_single_model_positions():
    """One 4x6x8 model: its lowest corner lands on the origin, y and z swap
    (Rzup -> Lyup) and the edge is 2^ceil(log2(8)) = 8; file colour index i
    takes palette entry i - 1."""
    vox = [(0, 0, 0, 1), (3, 5, 7, 2), (1, 2, 3, 255)]
    data = VW.write([((4, 6, 8), vox)], VW.placed([0], [(0, {"_t": "7 -3 11"})]), PAL)
    tree, size, fv = _both(data)
    exp = {(0, 0, 0): _pack(PAL[0]), (3, 7, 5): _pack(PAL[1]), (1, 3, 2): _pack(PAL[254])}
    assert size == 8 and fv == exp
    assert tree.get_size() == 8 and _voxels(tree) == exp


def test_translated_pair():
    """Two 4^3 copies at (0,0,0) and (10,3,-5): the scene minimum is
    (-2,-2,-7), so the copies start at (0,0,5) and (10,3,0) (Rzup), i.e.
    (0,5,0) and (10,0,3) in the tree; extent 14 -> edge 16."""
    data = VW.write([((4, 4, 4), [(0, 0, 0, 1), (3, 3, 3, 2)])],
                    VW.placed([0, 0], [(0, {"_t": "0 0 0"}), (0, {"_t": "10 3 -5"})]), PAL)
    tree, size, fv = _both(data)
    exp = {(0, 5, 0): _pack(PAL[0]), (3, 8, 3): _pack(PAL[1]),
           (10, 0, 3): _pack(PAL[0]), (13, 3, 6): _pack(PAL[1])}
    assert size == 16 and fv == exp
    assert _voxels(tree) == exp


@pytest.mark.parametrize("seed", range(6))
def test_rotated_model_matches_numpy(seed):
    """Every one of the 48 orientations (8 per seed) of a 2x4x6 model equals
    numpy's transpose/flip of the dense grid."""
    g, vox = _dense_model((2, 4, 6), seed)
    mats = {}
    for b in range(128):
        try:
            mats.setdefault(tuple(map(tuple, OV.parse_rotation_matrix(b))), b)
        except OV.VoxError:
            pass
    items = sorted(mats.items())[8 * seed:8 * seed + 8]
    for m, b in items:
        data = VW.write([((2, 4, 6), vox)], VW.placed([0], [(0, {"_r": str(b)})]), PAL)
        exp = _grid_voxels_lyup(_rotated_grid(g, m))
        tree, size, fv = _both(data)
        assert size == 8
        assert fv == exp, b
        assert _voxels(tree) == exp, b


def test_nested_rotations_compose():
    """transform(_r a) -> group -> transform(_r b) -> shape places the model
    under a.b (magicavoxel.rs:142-150): b is applied first."""
    g, vox = _dense_model((2, 4, 6), 11)
    a, b = 0b0100001, 0b1011000 | 0b0001
    ma, mb = OV.parse_rotation_matrix(a), OV.parse_rotation_matrix(b)
    scenes = [VW.transform(1), VW.group([2]), VW.transform(3, {"_r": str(a)}), VW.group([4]),
              VW.transform(5, {"_r": str(b)}), VW.shape([0])]
    data = VW.write([((2, 4, 6), vox)], scenes, PAL)
    exp = _grid_voxels_lyup(_rotated_grid(_rotated_grid(g, mb), ma))
    tree, _, fv = _both(data)
    assert fv == exp and _voxels(tree) == exp


def test_transform_without_rotation_resets_orientation():
    """A transform with no _r places its subtree unrotated even under a
    rotated parent (the `else Matrix3::identity()` of magicavoxel.rs:149)."""
    g, vox = _dense_model((2, 4, 6), 12)
    scenes = [VW.transform(1), VW.group([2]), VW.transform(3, {"_r": "33"}), VW.group([4]),
              VW.transform(5, {"_t": "1 1 1"}), VW.shape([0])]
    data = VW.write([((2, 4, 6), vox)], scenes, PAL)
    exp = _grid_voxels_lyup(g)
    tree, _, fv = _both(data)
    assert fv == exp and _voxels(tree) == exp


def test_frames_and_frame_attributes():
    """Frame 0 only: a transform uses its first frame; shape models whose _f
    is not 0 are skipped (an absent _f counts as 0, "+0" parses as 0)."""
    vox0, vox1 = [(0, 0, 0, 1)], [(1, 1, 1, 2)]
    scenes = [VW.transform(1), VW.group([2, 4]),
              VW.transform(3, frames=[{"_t": "0 0 0"}, {"_t": "40 0 0"}]),
              VW.shape([(0, {"_f": "+0"}), (1, {"_f": "1"})]),
              VW.transform(5, {"_t": "4 0 0"}), VW.shape([(1, {})])]
    data = VW.write([((2, 2, 2), vox0), ((2, 2, 2), vox1)], scenes, PAL)
    # minimum (-1,-1,-1): model 0 at 0, model 1 (second shape) at x + 4
    exp = {(0, 0, 0): _pack(PAL[0]), (5, 1, 1): _pack(PAL[1])}
    tree, size, fv = _both(data)
    assert size == 8 and fv == exp and _voxels(tree) == exp


def test_later_insert_wins_and_zero_index():
    """Overlapping placements: the one visited later keeps the voxel; a
    stored colour index of 0 reads palette entry 0 (saturating)."""
    data = VW.write([((2, 2, 2), [(0, 0, 0, 3)]), ((2, 2, 2), [(0, 0, 0, 0)])],
                    VW.placed([0, 1], [(0, {}), (1, {})]), PAL)
    tree, _, fv = _both(data)
    assert fv == {(0, 0, 0): _pack(PAL[0])} and _voxels(tree) == fv


def test_extra_chunks_are_skipped():
    junk = VW._chunk(b"LAYR", b"\x00" * 12) + VW._chunk(b"MATL", b"\x01" * 20) + \
        VW._chunk(b"rOBJ", b"\x00" * 8, VW._chunk(b"NOTE", b"\x00" * 4))
    data = VW.write([((2, 2, 2), [(1, 0, 1, 9)])], VW.placed([0], [(0, {})]), PAL,
                    extra_chunks=junk)
    tree, _, fv = _both(data)
    assert fv == {(1, 1, 0): _pack(PAL[8])} and _voxels(tree) == fv


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_brick_dimensions(dim):
    g, vox = _dense_model((8, 6, 4), 3)
    data = VW.write([((8, 6, 4), vox)], VW.placed([0, 0], [(0, {}), (0, {"_t": "5 9 -2"})]),
                    PAL)
    tree, size, fv = _both(data, dim)
    assert tree.get_size() == size and tree.brick_dim == dim
    assert _voxels(tree) == fv


# ------------------------------------------------------------------ refusals

def _status(data: bytes, dim: int = 1) -> int:
    with pytest.raises(OctreeError) as e:
        Octree.from_vox_bytes(data, dim)
    return e.value.status


_M = [((2, 2, 2), [(0, 0, 0, 1)])]


@pytest.mark.parametrize("data", [
    b"", b"VOX ", b"VOX \xc8\x00\x00\x00MAIN", b"XOV " + b"\x00" * 40,
    VW.write(_M, VW.placed([0], [(0, {})]), PAL)[:-10],              # truncated
    VW.write(_M, VW.placed([0], [(0, {})]), None),                   # no palette
    VW.write(_M, [VW.group([1]), VW.transform(2), VW.shape([0])], PAL),  # root not a transform
    VW.write(_M, VW.placed([0], [(0, {"_r": "3"})]), PAL),           # column index 3
    VW.write(_M, VW.placed([0], [(0, {"_r": "5"})]), PAL),           # repeated column
    VW.write(_M, VW.placed([0], [(0, {"_r": "-4"})]), PAL),          # not a u8
    VW.write(_M, VW.placed([0], [(0, {"_t": "1 2"})]), PAL),         # two values
    VW.write(_M, VW.placed([0], [(0, {"_t": "1  2 3"})]), PAL),      # empty field
    VW.write(_M, VW.placed([0], [(0, {"_t": "1 x 3"})]), PAL),
    VW.write(_M, VW.placed([0], [(1, {})]), PAL),                    # missing model
    VW.write(_M, [VW.transform(1), VW.group([1])], PAL),             # cycle
    VW.write(_M, [VW.transform(1), VW.group([])], PAL),              # nothing placed
    VW.write(_M, [VW.transform(1), VW.group([2]), VW.transform(3, frames=[]),
                  VW.shape([0])], PAL),                              # no frames
])
def test_malformed_files_are_refused(data):
    assert _status(data) == 3  # SVO_INVALID_STRUCTURE


def test_odd_model_overflowing_the_tree_is_refused():
    """A 5^3 model spans [-2, 2] around its centre, so the edge is 4 while its
    voxel at 4 lies outside: an InvalidPosition (a panic in the reference)."""
    data = VW.write([((5, 5, 5), [(0, 0, 0, 1), (4, 4, 4, 2)])], VW.placed([0], [(0, {})]), PAL)
    size, ins = OV.load(data, 1)
    assert size == 4 and max(max(p[:3]) for p in ins) == 4
    assert _status(data) == 4  # SVO_INVALID_POSITION
    # without the far voxel the same model loads
    ok = VW.write([((5, 5, 5), [(0, 0, 0, 1)])], VW.placed([0], [(0, {})]), PAL)
    assert _voxels(Octree.from_vox_bytes(ok, 1)) == {(0, 0, 0): _pack(PAL[0])}


def test_brick_dimension_larger_than_the_tree_is_refused():
    data = VW.write(_M, VW.placed([0], [(0, {})]), PAL)  # extent 2 -> edge 2
    with pytest.raises(OctreeError) as e:
        Octree.new(2, 4)
    assert _status(data, 4) == e.value.status


def test_load_from_path_and_missing_file(tmp_path):
    data = VW.write([((2, 2, 2), [(1, 1, 1, 7)])], VW.placed([0], [(0, {})]), PAL)
    p = tmp_path / "m.vox"
    p.write_bytes(data)
    assert _voxels(Octree.load_vox_file(p, 1)) == {(1, 1, 1): _pack(PAL[6])}
    with pytest.raises(OctreeError):
        Octree.load_vox_file(tmp_path / "nope.vox", 1)


# ------------------------------------------------------------- asset files

# edges of the asset scenes (2^ceil(log2(extent))) and voxel counts
ASSET_SIZES = {"navigate": (512, 12452), "navigate_x": (256, 4772),
               "navigate_y": (256, 3092), "navigate_z": (512, 5252)}


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("dim", [1, 2, 4, 8, 32])
def test_asset_files_match_the_oracle(name, dim):
    """The product's tree equals a tree built by inserting the oracle's
    voxels in the oracle's order (same nodes, bricks and palettes), and at
    brick dimensions up to 4 every voxel reads back with its colour.

    Above 4 the trees are lossy in the reference too: a leaf whose bricks
    halve exactly is united into one half-resolution brick (simplify,
    update/mod.rs:859-990), and a later insert there splits it again with
    dilute_brick_data (update/mod.rs:597-623), which reads each child's
    region at offset 2 instead of brick_dim / 2."""
    data = (ASSETS / f"{name}.vox").read_bytes()
    size, ins = OV.load(data, dim)
    fv = OV.final_voxels(ins)
    assert (size, len(fv)) == ASSET_SIZES[name]
    tree = Octree.load_vox_file(ASSETS / f"{name}.vox", dim)
    ref = Octree.new(size, dim)
    a = np.array(ins, np.int64)
    ents = np.zeros((len(a), 3), np.uint32)
    ents[:, 0] = 1
    ents[:, 1] = a[:, 3]
    ref.insert_many(a[:, :3].astype(np.uint32), ents)
    assert _models_equal(tree, ref)
    if dim > 4:
        return
    xyz = np.array(list(fv.keys()), np.uint32)
    got = tree.get_many(xyz)
    assert (got[:, 0] == 1).all()
    assert np.array_equal(got[:, 1], np.array(list(fv.values()), np.uint32))


def test_navigate_is_the_union_of_its_parts():
    """navigate.vox holds the three arrow groups and the marker that
    navigate_x/y/z.vox hold one group each of: each part, moved by the offset
    that aligns its own colour, lands inside the whole, and the moved parts
    cover it."""
    full = OV.final_voxels(OV.load((ASSETS / "navigate.vox").read_bytes(), 1)[1])
    cover = set()
    for part in ("navigate_x", "navigate_y", "navigate_z"):
        pv = OV.final_voxels(OV.load((ASSETS / f"{part}.vox").read_bytes(), 1)[1])
        colours = set(pv.values())
        own = [c for c in colours if sum(1 for v in full.values() if v == c) ==
               sum(1 for v in pv.values() if v == c)]
        assert own
        c = own[0]
        lo_f = np.min([p for p, v in full.items() if v == c], 0)
        lo_p = np.min([p for p, v in pv.items() if v == c], 0)
        off = lo_f - lo_p
        moved = {tuple(int(k) for k in np.array(p) + off): v for p, v in pv.items()}
        for p, v in moved.items():
            assert full.get(p) == v
        cover |= set(moved)
    assert cover == set(full)


def test_asset_points_by_hand():
    """navigate_y.vox: three 40^3 arrows at z = 131-48, 131, 131+48 (x = y =
    0) and a 32^3 marker at (0, 0, 16). The scene minimum is (-20, -20, 0)
    (arrow half-size 20, marker bottom at 16 - 16), so a marker voxel
    (x, y, z) starts from (0,0,16) - (16,16,16) + (20,20,0) = (4, 4, 0) and
    lands at (x + 4, z, y + 4) once y and z swap."""
    data = (ASSETS / "navigate_y.vox").read_bytes()
    v = OV.parse(data)
    (_, marker), = [m for m in v["models"] if tuple(m[0]) == (32, 32, 32)]
    tree = Octree.from_vox_bytes(data, 1)
    pal = v["palette"]
    for x, y, z, i in marker:
        e = tree.get((x + 4, z, y + 4))
        assert e.kind == 1 and e.albedo == _pack(pal[i])


@pytest.mark.parametrize("dim,kept", [(4, 9), (8, 1), (16, 1)])
def test_split_of_a_united_leaf_keeps_the_reference_offsets(dim, kept):
    """A full 2x2x2 block in brick octant 1 of an otherwise empty leaf lets
    the leaf unite into a half-resolution brick; the next insert splits it
    with the reference's dilute_brick_data, whose child offset of 2 matches
    brick_dim / 2 only for dimension 4, so at 8 and 16 the block is lost (as
    in the reference)."""
    from shocovox_amd import OctreeEntry
    t = Octree.new(2 * dim, dim)
    c = OctreeEntry.visual(0x222222FF)
    pts = [(dim + i, j, k) for k in range(2) for j in range(2) for i in range(2)]
    for p in pts:
        t.insert(p, c)
    t.insert((dim + 1, 3, 1), c)
    assert sum(t.get(q).kind != 0 for q in pts + [(dim + 1, 3, 1)]) == kept
