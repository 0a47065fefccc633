"""Shaded RGBA8 renders (svo_render_rgba8) against a numpy f32 restatement of
the two reference drawing paths, with the traversal done by the oracle.

convention 0: examples/cpu_render.rs:78-137 (eye rays through pixel corners,
    light normalize(0,-1,1), `as u8` channels, grey 128 misses, flipped rows);
convention 1: viewport_render.wgsl:1028-1089 (rays from the glass point, light
    (-0.5,0.5,-0.5), unorm8 channels, 0.25 grey misses).
Every pixel must match exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from shocovox_amd import DeviceTree, workloads as W

pytestmark = pytest.mark.gpu

F = np.float32


def _normalized(v: np.ndarray) -> np.ndarray:
    ln = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
    return (v / ln[..., None]).astype(F)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]], dtype=F)


def _rays(vp, w: int, h: int, conv: int) -> np.ndarray:
    """Rays of output pixels in row-major order from the top row."""
    o = np.array(vp.origin, F)
    d = np.array(vp.direction, F)
    gw, gh, fov = F(vp.frustum[0]), F(vp.frustum[1]), F(vp.fov)
    up = np.array([0, 1, 0], F)
    right = _normalized(_cross(up, d))
    rows, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    xs = xs.reshape(-1).astype(F)
    rows = rows.reshape(-1)
    if conv == 0:
        pw, ph = F(gw / F(w)), F(gh / F(h))
        bl = ((o + d * fov) - up * F(gh / F(2))) - right * F(gw / F(2))
        ys = (h - 1 - rows).astype(F)
        glass = (bl + (right[None, :] * xs[:, None]) * pw) + (up[None, :] * ys[:, None]) * ph
        origin = np.broadcast_to(o, glass.shape)
    else:
        w8, h8 = F((w + 7) // 8 * 8), F((h + 7) // 8 * 8)
        bl = ((o + d * fov) - right * F(gw / F(2))) - up * F(gh / F(2))
        fx = (xs / w8).astype(F)
        fy = (F(1) - rows.astype(F) / h8).astype(F)
        glass = (bl + (right * gw)[None, :] * fx[:, None]) + (up * gh)[None, :] * fy[:, None]
        origin = glass
    dirs = _normalized((glass - o).astype(F))
    return np.concatenate([origin, dirs], axis=1).astype(F)


def _shade(hits: np.ndarray, conv: int) -> np.ndarray:
    n = hits["normal"].astype(F)
    c = hits["albedo"].astype(np.uint32)
    rgb = np.stack([(c >> 24) & 0xFF, (c >> 16) & 0xFF, (c >> 8) & 0xFF], axis=1).astype(F)
    if conv == 0:
        light = _normalized(np.array([0, -1, 1], F))
    else:
        light = np.array([-0.5, 0.5, -0.5], F)
    dot = (n[:, 0] * light[0] + n[:, 1] * light[1]) + n[:, 2] * light[2]
    with np.errstate(invalid="ignore"):
        if conv == 0:
            s = (F(1) - (dot / F(2) + F(0.5))).astype(F)
            v = rgb * s[:, None]
            ch = np.where(v > 0, np.minimum(np.trunc(np.nan_to_num(v, nan=0.0)), 255), 0)
            miss = np.array([128, 128, 128])
        else:
            k = (dot / F(2) + F(0.5)).astype(F)
            v = np.clip((rgb / F(255)) * k[:, None], F(0), F(1))
            ch = np.floor(v * F(255) + F(0.5))
            miss = np.array([64, 64, 64])
    ch = ch.astype(np.uint32)
    hit = hits["hit"] == 1
    ch[~hit] = miss
    return (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | np.uint32(0xFF000000)).astype(np.uint32)


@pytest.mark.parametrize("conv", [0, 1])
@pytest.mark.parametrize("size", [(256, 256), (203, 117)])
def test_rgba8_matches_restatement(gpu_device, conv, size):
    import torch
    w, h = size
    cfg = W.PLAIN
    tree = W.build_scene(cfg, seed=1)
    vp = W.viewports(cfg)[0]
    print(f"rgba8 convention {conv} {w}x{h}", flush=True)
    d = DeviceTree(tree)
    img = d.render_rgba8(vp, w, h, conv)
    torch.cuda.synchronize()
    got = img.cpu().numpy().view(np.uint32).reshape(-1)
    rays = _rays(vp, w, h, conv)
    hits = oracle.OracleTree(tree.export_model()).cast(rays)
    exp = _shade(hits, conv)
    bad = np.nonzero(got != exp)[0]
    assert len(bad) == 0, (f"{len(bad)} of {w * h} pixels differ; first {bad[:5]}: "
                           f"{got[bad[:5]]} vs {exp[bad[:5]]}")
    assert (hits["hit"] == 1).sum() > w * h // 10
