"""Pins of the inputs that oracle and kernel share, checked on the CPU.

Both the oracle and the kernel traverse the trees the scene builders make and
the rays the viewport generator makes; a mistake in either would pass every
parity test. So:

* the benchmark trees (caves 512^3 / dim 16, spheres at dim 32, the
  insert-built plain tree) satisfy the reference's structural invariants on
  every node: an occupancy bit is set exactly when its (size/4)^3 sub-region
  holds a non-empty voxel (insert.rs:344-366, detail.rs:524-544, math/mod.rs:
  114-162), Solid bricks are never empty, a Leaf's octant mask is full for a
  Solid brick and zero for an Empty one, and the tree is a fixed point of the
  reference's simplify (update/mod.rs:689-1068, node.rs:316-330, 554-575),
  i.e. a tree the reference's own edits could have left;
* the oracle's pinhole rays (examples/cpu_render.rs:78-114 through pixel
  centres) are pinned by closed forms: unit length, the centre pixel of an
  odd-sized image looks along the view direction, the middle row's edge
  pixels make the angle atan((glass/2 - pixel/2) / fov) with it, the middle
  column's the angle of dir * fov + up * offset (the glass is spanned by
  right and the world's +y), +x of the image is normalize(up x direction)
  and +y is up;
* boundary hits (impact within 1e-6 of a second face of the hit cell) are
  flagged exactly where a ray meets a cell edge, counted on the workloads,
  and the kernel's traversal source equals the oracle on every one of them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import hostsim
import oracle
from kat_trees import normalized, ray
from shocovox_amd import Octree, OctreeEntry, workloads as W

F = np.float32
OCT = [(o & 1, (o >> 2) & 1, (o >> 1) & 1) for o in range(8)]  # octant = x + 2z + 4y


# --- trees --------------------------------------------------------------------

def _empty(m, values):
    """pix_points_to_empty (node.rs:405-427) for an array of palette index values."""
    v = np.asarray(values, np.uint32)
    ci, di = (v & 0xFFFF).astype(np.int64), (v >> 16).astype(np.int64)
    cnone = ci == 0xFFFF
    dnone = di == 0xFFFF
    c_empty = cnone | ((m["colors"][np.where(cnone, 0, ci)] & 0xFF) == 0)
    d_empty = dnone | (m["data"][np.where(dnone, 0, di)] == 0)
    return c_empty & d_empty


def _walk(m):
    """[(key, (x, y, z) min, size)] of every node reachable from the root."""
    n = m["pool_len"]
    kids = m["children"].reshape(-1, 8)
    out, todo = [], [(0, (0, 0, 0), int(m["octree_size"]))]
    while todo:
        k, mn, s = todo.pop()
        out.append((k, mn, s))
        if m["content_kind"][k] == 1 and m["children_kind"][k] == 1:
            for o, (ox, oy, oz) in enumerate(OCT):
                c = int(kids[k, o])
                if c < n and m["reserved"][c]:
                    h = s // 2
                    todo.append((c, (mn[0] + ox * h, mn[1] + oy * h, mn[2] + oz * h), h))
    return out


def _brick_grid(m, form, value, span):
    """Occupancy [x, y, z] of one brick covering span^3 voxels."""
    dim = int(m["brick_dim"])
    if form == 0:
        return np.zeros((span,) * 3, bool)
    if form == 1:
        return np.full((span,) * 3, not _empty(m, [value])[0])
    n3 = dim ** 3
    vals = m["voxels"][int(value) * n3:(int(value) + 1) * n3]
    occ = (~_empty(m, vals)).reshape(dim, dim, dim).transpose(2, 1, 0)  # flat x + y d + z d^2
    u = span // dim
    return occ.repeat(u, 0).repeat(u, 1).repeat(u, 2)


def _dense(m, nodes):
    s = int(m["octree_size"])
    g = np.zeros((s, s, s), bool)
    bk = m["brick_kind"].reshape(-1, 8)
    bv = m["brick_value"].reshape(-1, 8)
    for k, (x, y, z), size in nodes:
        kind = m["content_kind"][k]
        if kind == 2:
            h = size // 2
            for o, (ox, oy, oz) in enumerate(OCT):
                g[x + ox * h:x + ox * h + h, y + oy * h:y + oy * h + h,
                  z + oz * h:z + oz * h + h] = _brick_grid(m, bk[k, o], bv[k, o], h)
        elif kind == 3:
            g[x:x + size, y:y + size, z:z + size] = _brick_grid(m, bk[k, 0], bv[k, 0], size)
    return g


def _bits(region):
    s = region.shape[0] // 4
    cells = region.reshape(4, s, 4, s, 4, s).any(axis=(1, 3, 5))
    bits = 0
    for cx in range(4):
        for cy in range(4):
            for cz in range(4):
                if cells[cx, cy, cz]:
                    bits |= 1 << (cx + 4 * cy + 16 * cz)
    return bits


def _stored_bits(m, k):
    kind = m["content_kind"][k]
    if kind == 1:
        return int(m["internal_bits"][k])
    if kind in (2, 3):
        return int(m["children_bitmap"][k]) if m["children_kind"][k] == 2 else 0
    return 0


def _mask(o):
    ox, oy, oz = OCT[o]
    return 0x0000000000330033 << (ox * 2 + oy * 8 + oz * 32)


def _brick_eq(m, fa, va, fb, vb):
    """BrickData equality (Parted compares the voxel vectors)."""
    if fa != fb:
        return False
    if fa == 1:
        return va == vb
    if fa == 2:
        n3 = int(m["brick_dim"]) ** 3
        return np.array_equal(m["voxels"][va * n3:(va + 1) * n3], m["voxels"][vb * n3:(vb + 1) * n3])
    return True


def _simplify_would_change(m, k):
    """Whether the reference's simplify(k) (update/mod.rs:689-1068) would change
    node k itself (the recursion into children is covered by checking every
    node)."""
    kind = int(m["content_kind"][k])
    dim = int(m["brick_dim"])
    n3 = dim ** 3
    bk = m["brick_kind"].reshape(-1, 8)
    bv = m["brick_value"].reshape(-1, 8)

    def brick_simplifies(form, value):  # BrickData::simplify (node.rs:316-330)
        if form == 1:
            return True  # a Solid brick is homogeneous: it would be re-set (to itself unless empty)
        if form == 2:
            v = m["voxels"][int(value) * n3:(int(value) + 1) * n3]
            return bool((v == v[0]).all())
        return False

    if kind == 3:
        f, v = int(bk[k, 0]), int(bv[k, 0])
        return (f == 1 and bool(_empty(m, [v])[0])) or (f == 2 and brick_simplifies(f, v))
    if kind == 2:
        forms, vals = [int(f) for f in bk[k]], [int(v) for v in bv[k]]
        # bricks that simplify change the node (a Solid brick re-set to the
        # same value does not: only empty Solid values or homogeneous Parted ones)
        for f, v in zip(forms, vals):
            if f == 1 and _empty(m, [v])[0]:
                return True
            if f == 2 and brick_simplifies(f, v):
                return True
        if all(f == 1 for f in forms) and len(set(vals)) == 1:
            return True  # unites into UniformLeaf(Solid)
        half = dim // 2
        uniform = True
        for o in range(8):
            f, v = forms[o], vals[o]
            if f in (0, 1):
                uniform &= _brick_eq(m, f, v, forms[0], vals[0])
            else:
                b = m["voxels"][v * n3:(v + 1) * n3].reshape(dim, dim, dim)  # [z, y, x]
                if dim < 2:
                    uniform = False
                else:
                    base = b[0::2, 0::2, 0::2]
                    for dz in (0, 1):
                        for dy in (0, 1):
                            for dx in (0, 1):
                                uniform &= bool((b[dz::2, dy::2, dx::2] == base).all())
            if not uniform:
                break
        return uniform and half >= 1  # would become a UniformLeaf of a half-resolution brick
    if kind == 1:
        if int(m["internal_bits"][k]) == 0 or m["children_kind"][k] != 1:
            return True
        kids = [int(c) for c in m["children"].reshape(-1, 8)[k]]
        n = m["pool_len"]
        if not all(c < n and m["reserved"][c] for c in kids):
            return False
        c0 = kids[0]
        k0 = int(m["content_kind"][c0])
        for c in kids[1:]:
            kc = int(m["content_kind"][c])
            if k0 != kc or k0 == 1:
                return False
            if k0 == 0:
                continue
            nb = 1 if k0 == 3 else 8
            if not all(_brick_eq(m, int(bk[c0, o]), int(bv[c0, o]), int(bk[c, o]), int(bv[c, o]))
                       for o in range(nb)):
                return False
        return True  # all children equal: the node would take child 0's place
    return False


TREES = {
    "caves-512-d16-seed1": lambda: W.build_scene(W.CAVES, seed=1),
    "spheres-512-d32-seed1": lambda: Octree.scene(3, 512, 32, 1),
    "spheres-256-d8-seed4": lambda: Octree.scene(3, 256, 8, 4),
    "plain-128-d8-seed1": lambda: W.build_scene(W.PLAIN, seed=1),
}


@pytest.mark.parametrize("name", list(TREES))
def test_benchmark_tree_structure(name):
    tree = TREES[name]()
    m = tree.export_model()
    nodes = _walk(m)
    g = _dense(m, nodes)
    # the dense grid of the test's own walk agrees with the oracle's get()
    rng = np.random.default_rng(3)
    s = int(m["octree_size"])
    xyz = rng.integers(0, s, size=(200_000, 3)).astype(np.uint32)
    got = oracle.OracleTree(m).get_many(xyz) != 0xFFFFFFFF
    assert np.array_equal(got, g[xyz[:, 0], xyz[:, 1], xyz[:, 2]])
    assert g.any()
    bk = m["brick_kind"].reshape(-1, 8)
    bv = m["brick_value"].reshape(-1, 8)
    for k, (x, y, z), size in nodes:
        kind = m["content_kind"][k]
        region = g[x:x + size, y:y + size, z:z + size]
        assert _stored_bits(m, k) == _bits(region), f"{name}: occupancy bits of node {k}"
        if kind in (2, 3):
            for o in range(8 if kind == 2 else 1):
                if bk[k, o] == 1:
                    assert not _empty(m, [bv[k, o]])[0], f"{name}: empty Solid brick in node {k}"
        if kind == 2:
            bits = _stored_bits(m, k)
            for o in range(8):
                if bk[k, o] == 1:
                    assert bits & _mask(o) == _mask(o), f"{name}: Solid octant {o} of node {k}"
                elif bk[k, o] == 0:
                    assert bits & _mask(o) == 0, f"{name}: Empty octant {o} of node {k}"
        assert not _simplify_would_change(m, k), f"{name}: node {k} is not simplified"
    # the statistics reported beside benchmark results count the same voxels
    st = W.tree_statistics(tree)
    assert st["voxels"] == int(g.sum())


def test_structure_checks_catch_planted_mistakes():
    """The checks above fail on trees with a wrong occupancy bit, an empty
    Solid brick and an unsimplified leaf (mistakes planted in the dump)."""
    tree = Octree.new(16, 2)
    for p in [(0, 0, 0), (1, 1, 1), (5, 2, 9), (12, 12, 12)]:
        tree.insert(p, OctreeEntry.visual(0x336699FF))
    m = tree.export_model()
    nodes = _walk(m)
    g = _dense(m, nodes)
    assert all(_stored_bits(m, k) == _bits(g[x:x + s, y:y + s, z:z + s])
               for k, (x, y, z), s in nodes)
    leaf = next(k for k, _, _ in nodes if m["content_kind"][k] == 2)
    bad = dict(m, children_bitmap=m["children_bitmap"].copy())
    bad["children_bitmap"][leaf] ^= 1 << 63
    assert any(_stored_bits(bad, k) != _bits(g[x:x + s, y:y + s, z:z + s])
               for k, (x, y, z), s in nodes)
    bad = dict(m, brick_kind=m["brick_kind"].copy(), brick_value=m["brick_value"].copy())
    bad["brick_kind"][leaf * 8] = 1
    bad["brick_value"][leaf * 8] = 0xFFFFFFFF
    assert _simplify_would_change(bad, leaf)
    filled = m["voxels"][~_empty(m, m["voxels"])][0]  # the inserted colour's index value
    bad = dict(m, brick_kind=m["brick_kind"].copy(), brick_value=m["brick_value"].copy())
    bad["brick_kind"][leaf * 8:leaf * 8 + 8] = 1
    bad["brick_value"][leaf * 8:leaf * 8 + 8] = filled
    assert _simplify_would_change(bad, leaf)  # eight equal Solid bricks unite


# --- the ray generator ----------------------------------------------------------

def _angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
    return math.acos(min(1.0, max(-1.0, c)))


@pytest.mark.parametrize("eye,target,w,h,fov", [
    ((700.0, 400.0, -100.0), (256.0, 256.0, 256.0), 33, 21, 1.0),
    ((-50.0, 20.0, 64.0), (64.0, 64.0, 64.0), 129, 65, 2.5),
    ((10.0, 300.0, 11.0), (12.0, 0.0, 9.0), 7, 7, 0.75),
])
def test_viewport_rays_closed_forms(eye, target, w, h, fov):
    vp = W.look_at_viewport(eye, target, w, h)
    gw, gh = vp.frustum[0] * fov, vp.frustum[1] * fov  # same angular extent at glass distance fov
    rays = oracle.viewport_rays(vp.origin, vp.direction, (gw, gh, 1e9), fov, w, h)
    o, d = rays[:, :3], rays[:, 3:]
    assert (o == np.asarray(vp.origin, F)).all()
    assert np.abs(np.linalg.norm(d.astype(np.float64), axis=1) - 1.0).max() < 4e-7
    dirv = np.asarray(vp.direction, np.float64)
    # the glass point is formed in f32 at the eye's magnitude, fov away from
    # it: a few units in the last place of the eye, seen from fov, bound the
    # angle error (half a pixel, the size of a plausible mistake, is > 1e-3)
    tol = 4 * 2.0 ** -23 * float(np.abs(np.asarray(eye)).max()) / fov + 1e-6
    cx, cy = (w - 1) // 2, (h - 1) // 2
    assert _angle(d[cy * w + cx], dirv) < tol
    # pixel centres of the middle row and column: atan((glass/2 - pixel/2) / fov)
    for x in (0, w - 1):
        exp = math.atan((gw / 2 - gw / w / 2) / fov)
        assert abs(_angle(d[cy * w + x], dirv) - exp) < tol
    # the glass is spanned by right and the world's up (+y), which is not
    # perpendicular to a pitched view direction: the middle column's ray
    # through row y is dir * fov + up * ((y + 1/2) ph - gh / 2)
    up = np.array([0.0, 1.0, 0.0])
    for y in (0, h - 1):
        v = (y + 0.5) * (gh / h) - gh / 2
        exp = _angle(dirv * fov + up * v, dirv)
        assert abs(_angle(d[y * w + cx], dirv) - exp) < tol
    right = np.cross(up, dirv)
    right /= np.linalg.norm(right)
    assert np.dot(d[cy * w + (w - 1)], right) > 0 > np.dot(d[cy * w], right)
    assert np.dot(d[(h - 1) * w + cx], up) > np.dot(d[cx], up)


def test_viewport_rays_catch_a_planted_offset():
    """A generator without the half-pixel offset fails the centre-pixel pin."""
    vp = W.look_at_viewport((700.0, 400.0, -100.0), (256.0, 256.0, 256.0), 33, 21)
    w, h = 33, 21
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w, h)
    # shifting by half a pixel is what forgetting the offset does
    shifted = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, w * 2, h * 2)
    dirv = np.asarray(vp.direction, np.float64)
    assert _angle(rays[10 * w + 16, 3:], dirv) < 4e-4
    assert _angle(shifted[(2 * 10) * (2 * w) + 2 * 16, 3:], dirv) > 1e-3


# --- boundary hits ----------------------------------------------------------------

def test_boundary_flag_marks_edge_hits_only():
    """A ray through the shared edge of two voxel cells is a boundary hit; a
    ray onto the middle of a face is not."""
    t = Octree.new(4, 1)
    t.insert((1, 1, 1), OctreeEntry.visual(0x808080FF))
    o = oracle.OracleTree(t.export_model())
    face = ray((1.5, 1.5, -3.0), (0.0, 0.0, 1.0))
    edge = ray((1.0, 1.5, -3.0), (0.0, 0.0, 1.0))  # along the plane x = 1
    corner = ray((1.0, 1.0, -3.0), (0.0, 0.0, 1.0))  # onto the corner (1, 1, 1)
    side = ray((-1.0, 1.5, 1.0), normalized((1.0, 0.0, 0.0)))  # along the plane z = 1
    r = o.cast(np.stack([face, edge, corner, side]))
    assert r["hit"].tolist() == [1, 1, 1, 1]
    assert r["boundary"].tolist() == [0, 1, 1, 1]


def test_boundary_hits_on_the_workloads():
    """Boundary hits counted on the plain workload (every ray) and on a caves
    frame; the kernel's traversal source equals the oracle on each one."""
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                cfg.height)
    _, c = oracle.OracleTree(tree.export_model()).cast(rays, counters=True)
    assert c.boundary_hits / c.rays < 1e-6
    cfg = W.CAVES
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg, seed=1)[0]
    rays = oracle.viewport_rays(vp.origin, vp.direction, vp.frustum, vp.fov, cfg.width,
                                cfg.height)
    r = oracle.OracleTree(tree.export_model()).cast(rays, threads=8)
    b = np.nonzero(r["boundary"])[0]
    assert 0 < len(b) < 1e-4 * len(rays)
    # at coordinates of 16 and more a boundary hit lies exactly on two cell
    # planes, and the reference's normal there has two non-zero components
    far = b[(r["impact"][b] >= 16).all(axis=1)]
    assert len(far) > 0
    assert ((r["normal"][far] != 0).sum(axis=1) >= 2).all()
    got = hostsim.cast(tree.flatten().arrays(), rays[b])
    for f in ("hit", "pix", "voxel", "distance", "impact", "normal", "flags"):
        assert np.array_equal(got[f].view(np.uint32), r[f][b].view(np.uint32)), f
