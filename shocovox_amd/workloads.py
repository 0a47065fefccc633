"""Benchmark workloads: seeded scenes and pinhole viewports.

plain   128^3 octree, brick dim 8, heightfield built by insert (scene kind 1);
        one 256x256 viewport, 60 deg vertical fov, eye on a sphere of radius
        1.5*128 at azimuth 45 deg / elevation 30 deg around the centre.
caves   512^3 octree, brick dim 16, porous noise terrain with worm caves and
        seeded brick forms (scene kind 2); 1920x1080, 8 orbit poses
        (azimuth 2*pi*k/8 + seeded jitter, horizontal radius 1.2*512, height
        0.8*512) looking at the centre.
spheres 2048^3 octree, brick dim 32, seeded spheres over a heightfield with
        50/50 Parted/Solid bricks (scene kind 3); 3840x2160 plus incoherent
        rays (origins uniform in the bounds, directions uniform on S^2).

Viewports follow the reference's Viewport (bevy/types.rs:55-71): the glass
sits `fov` = 1 in front of the eye and spans frustum = (2 tan(30deg) * aspect,
2 tan(30deg), depth). All values are f32 so the GPU and the checker receive
identical inputs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .octree import Octree, Viewport

F = np.float32


@dataclass(frozen=True)
class Config:
    name: str
    size: int
    brick_dim: int
    scene_kind: int
    width: int
    height: int
    poses: int


PLAIN = Config("plain", 128, 8, 1, 256, 256, 1)
CAVES = Config("caves", 512, 16, 2, 1920, 1080, 8)
SPHERES = Config("spheres", 2048, 32, 3, 3840, 2160, 1)
CONFIGS = {c.name: c for c in (PLAIN, CAVES, SPHERES)}


def _normalized(v: np.ndarray) -> np.ndarray:
    v = v.astype(F)
    ln = F(np.sqrt(F(F(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
    return np.array([v[0] / ln, v[1] / ln, v[2] / ln], dtype=F)


def look_at_viewport(eye, target, width: int, height: int, vfov_deg: float = 60.0) -> Viewport:
    eye = np.asarray(eye, dtype=F)
    d = _normalized(np.asarray(target, dtype=F) - eye)
    gh = F(2.0 * math.tan(math.radians(vfov_deg) / 2.0))
    gw = F(gh * F(width) / F(height))
    return Viewport(tuple(float(v) for v in eye), tuple(float(v) for v in d),
                    (float(gw), float(gh), float(1e9)), 1.0)


def viewports(cfg: Config, seed: int = 0) -> list:
    c = cfg.size / 2.0
    centre = (c, c, c)
    if cfg.poses == 1:
        r = 1.5 * cfg.size
        az, el = math.radians(45.0), math.radians(30.0)
        eye = (c + r * math.cos(el) * math.cos(az), c + r * math.sin(el),
               c + r * math.cos(el) * math.sin(az))
        return [look_at_viewport(eye, centre, cfg.width, cfg.height)]
    rng = np.random.default_rng(seed + 1000)
    out = []
    for k in range(cfg.poses):
        az = 2.0 * math.pi * k / cfg.poses + float(rng.uniform(-0.2, 0.2))
        eye = (c + 1.2 * cfg.size * math.cos(az), 0.8 * cfg.size,
               c + 1.2 * cfg.size * math.sin(az))
        out.append(look_at_viewport(eye, centre, cfg.width, cfg.height))
    return out


def build_scene(cfg: Config, seed: int = 1) -> Octree:
    return Octree.scene(cfg.scene_kind, cfg.size, cfg.brick_dim, seed)


def incoherent_rays(size: int, n: int, seed: int = 3) -> np.ndarray:
    """Origins uniform in [0,size)^3, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.0, size, size=(n, 3)).astype(F)
    z = rng.uniform(-1.0, 1.0, size=n)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n)
    s = np.sqrt(1.0 - z * z)
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1).astype(F)
    ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(F)
    d = (d / ln[:, None]).astype(F)
    return np.concatenate([o, d], axis=1).astype(F)


# --- statistics of a built tree (reported beside benchmark results) --------

_OCT_OFF = np.array([[o & 1, (o >> 2) & 1, (o >> 1) & 1] for o in range(8)], np.int64)


def model_nodes(m: dict):
    """(keys, mins, sizes) of the nodes reachable from the root of an
    Octree.export_model() dump, level by level: an Internal node with
    Children reaches each valid child key at its octant's offset (octant =
    x + 2z + 4y, OCTANT_OFFSET_REGION_LUT of lut.rs:156-197)."""
    n = int(m["pool_len"])
    reserved = m["reserved"][:n].astype(bool)
    kids = m["children"][: 8 * n].reshape(n, 8).astype(np.int64)
    keys = [np.array([0], np.int64)]
    mins = [np.zeros((1, 3), np.int64)]
    sizes = [np.array([int(m["octree_size"])], np.int64)]
    k, mn, sz = keys[0], mins[0], sizes[0]
    while len(k):
        internal = (m["content_kind"][k] == 1) & (m["children_kind"][k] == 1)
        k, mn, sz = k[internal], mn[internal], sz[internal]
        ck = kids[k]  # (len, 8)
        valid = (ck < n) & reserved[np.minimum(ck, n - 1)]
        half = sz // 2
        cmin = mn[:, None, :] + _OCT_OFF[None, :, :] * half[:, None, None]
        k = ck[valid]
        mn = cmin[valid]
        sz = np.broadcast_to(half[:, None], ck.shape)[valid]
        keys.append(k)
        mins.append(mn)
        sizes.append(sz)
    return np.concatenate(keys), np.concatenate(mins), np.concatenate(sizes)


def empty_values(m: dict, values: np.ndarray) -> np.ndarray:
    """pix_points_to_empty for an array of palette index values (node.rs:405-427):
    no colour (index 0xFFFF) or a transparent one, and no data or data 0."""
    v = np.asarray(values, np.uint32)
    ci = (v & 0xFFFF).astype(np.int64)
    di = (v >> 16).astype(np.int64)
    colors, data = m["colors"], m["data"]
    cnone = (ci == 0xFFFF) | (ci >= int(m["color_count"]))
    dnone = (di == 0xFFFF) | (di >= int(m["data_count"]))
    c_empty = cnone | ((colors[np.minimum(ci, len(colors) - 1)] & 0xFF) == 0)
    d_empty = dnone | (data[np.minimum(di, len(data) - 1)] == 0)
    return c_empty & d_empty


def tree_statistics(tree) -> dict:
    """Filled voxels and brick forms of a tree: `occupancy` is the fraction of
    the size^3 voxels that are not empty; the `*_slots` fractions split the
    (size/dim)^3 leaf brick slots into Solid, Parted and Empty (a slot no
    node covers is Empty; merged leaves count for every slot they cover)."""
    m = tree.export_model()
    size, dim = int(m["octree_size"]), int(m["brick_dim"])
    keys, _, sizes = model_nodes(m)
    kind = m["content_kind"][keys]
    bk = m["brick_kind"].reshape(-1, 8)
    bv = m["brick_value"].reshape(-1, 8)
    n3 = dim ** 3
    vox = m["voxels"]
    voxels = 0
    slots = {"solid": 0, "parted": 0}
    for is_leaf, cover in ((kind == 2, 2), (kind == 3, 1)):
        ks, ss = keys[is_leaf], sizes[is_leaf]
        octs = 8 if cover == 2 else 1
        for o in range(octs):
            form = bk[ks, o]
            val = bv[ks, o]
            span = ss // cover  # edge of the region one brick covers
            unit3 = (span // dim) ** 3  # voxels per brick value
            solid = (form == 1) & ~empty_values(m, val)
            voxels += int((span[solid] ** 3).sum())
            slots["solid"] += int(((span[solid] // dim) ** 3).sum())
            parted = form == 2
            for i in np.nonzero(parted)[0]:
                b = vox[int(val[i]) * n3:(int(val[i]) + 1) * n3]
                voxels += int((~empty_values(m, b)).sum()) * int(unit3[i])
            slots["parted"] += int(((span[parted] // dim) ** 3).sum())
    total = (size // dim) ** 3
    return {
        "voxels": voxels, "occupancy": round(voxels / size ** 3, 4),
        "brick_slots": total,
        "solid_slots": round(slots["solid"] / total, 4),
        "parted_slots": round(slots["parted"] / total, 4),
        "empty_slots": round(1.0 - (slots["solid"] + slots["parted"]) / total, 4),
        "nodes": {"internal": int((kind == 1).sum()), "leaf": int((kind == 2).sum()),
                  "uniform_leaf": int((kind == 3).sum())},
    }


# --- pools of the batched and served workloads ------------------------------

BATCH_RESOLUTIONS = ((128, 128), (512, 512), (1280, 720), (1920, 1080))


def log_uniform_sizes(n: int, max_size: int, rng) -> list:
    """n tree sizes log-uniform over the powers of two 64..max_size: each size
    gets n // k trees (k sizes; the remainder go to distinct sizes drawn at
    random), in a seeded random order - a stratified draw of the log-uniform
    law, so that a pool of 32 or 64 trees holds every size about equally
    often."""
    exps = list(range(6, int(math.log2(max_size)) + 1))
    sizes = [2 ** e for e in exps] * (n // len(exps))
    rest = n - len(sizes)
    sizes += [2 ** int(e) for e in rng.choice(exps, size=rest, replace=False)]
    rng.shuffle(sizes)
    return sizes


def batched_pool_specs(seed: int, n_trees: int = 64, max_size: int = 1024):
    """The batched workload: (trees, views). trees = [(scene kind, size, brick
    dim, seed)] with sizes log-uniform over 64..max_size, brick dims 4/8/16
    and the content of the plain workload (the heightfield, built bottom-up,
    kind 4) or of the caves (kind 2) by seed; views = [(tree index, Viewport,
    width, height)], 1-4 per tree, resolutions drawn from BATCH_RESOLUTIONS,
    eyes on a seeded circle around the tree looking at its lower middle."""
    rng = np.random.default_rng(seed + 700)
    trees, views = [], []
    for i, size in enumerate(log_uniform_sizes(n_trees, max_size, rng)):
        dim = int(rng.choice([4, 8, 16]))
        kind = int(rng.choice([2, 4]))
        trees.append((kind, size, dim, seed * 1000 + 500 + i))
        c = size / 2.0
        for _ in range(int(rng.integers(1, 5))):
            w, h = BATCH_RESOLUTIONS[int(rng.integers(0, len(BATCH_RESOLUTIONS)))]
            az = float(rng.uniform(0, 2 * math.pi))
            eye = (c + 1.3 * size * math.cos(az), 0.9 * size, c + 1.3 * size * math.sin(az))
            views.append((i, look_at_viewport(eye, (c, c / 2, c), w, h), w, h))
    return trees, views


def served_pool_specs(seed: int, n_trees: int = 32, max_size: int = 2048) -> list:
    """The served workload's pool: [(scene kind, size, brick dim, seed)] with
    sizes log-uniform over 64..max_size, brick dims 4/8/16 below 1024^3 and 16
    from 1024^3 on (32 for the spheres at 2048^3, as in the 2048^3 workload);
    content by seed: caves (2) or the plain heightfield (4) below 1024^3, any
    of caves, spheres over a heightfield (3) and the heightfield at 1024^3,
    spheres or the heightfield at 2048^3."""
    rng = np.random.default_rng(seed + 500)
    specs = []
    for i, size in enumerate(log_uniform_sizes(n_trees, max_size, rng)):
        dim = 16 if size >= 1024 else int(rng.choice([4, 8, 16]))
        if size >= 2048:
            kind = int(rng.choice([3, 4]))
            if kind == 3:
                dim = 32  # the 2048^3 spheres workload's brick dimension
        elif size == 1024:
            kind = int(rng.choice([2, 3, 4]))
        else:
            kind = int(rng.choice([2, 4]))
        specs.append((kind, size, dim, seed * 1000 + i))
    return specs
