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
