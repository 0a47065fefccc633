"""ORACLE - test infrastructure only, never part of the product.

A pure-Python restatement of the reference's MagicaVoxel import,
Octree::load_vox_file (src/convert/magicavoxel.rs:259-395), on top of the
reader it uses, the dot_vox crate 5.1.1 (Cargo.toml:20; not vendored in the
reference). The crate's behaviour restated here is the published .vox
format (ephtracy/voxel-model, MagicaVoxel-file-format-vox.txt and
-vox-extension.txt) as dot_vox exposes it:
  * models are SIZE / XYZI chunk pairs under MAIN, in file order;
  * a voxel's colour index is stored in memory one below the file's value
    (file indices run 1-255), saturating at 0, and indexes the RGBA chunk's
    colours directly;
  * scene nodes (nTRN / nGRP / nSHP) are kept in file order and addressed by
    that position.
Only tests/ import this module, as the checker of the product's loader.

Pinning: parse_rotation_matrix by the reference's own known-answer test
(magicavoxel.rs:398-414) and by the format document's bit description over
every valid byte; the placement arithmetic by hand-derived positions of
synthetic files (single models, translated pairs, rotated models compared
with numpy's rotations of the dense grid); the asset files by their scene
structure (navigate.vox is the union of navigate_x/y/z.vox, moved).
"""
from __future__ import annotations

import math
import struct

I32_MIN, I32_MAX = -(1 << 31), (1 << 31) - 1


class VoxError(ValueError):
    pass


def _dict(b: bytes, p: int):
    (n,), p = struct.unpack_from("<i", b, p), p + 4
    d = {}
    for _ in range(n):
        (kl,), p = struct.unpack_from("<i", b, p), p + 4
        k, p = b[p:p + kl].decode(), p + kl
        (vl,), p = struct.unpack_from("<i", b, p), p + 4
        v, p = b[p:p + vl].decode(), p + vl
        d[k] = v
    return d, p


def parse(data: bytes) -> dict:
    """dot_vox::load: {'models': [(size, [(x, y, z, i)])], 'palette': [(r, g, b, a)] or None,
    'scenes': [node]} with voxel indices already lowered by one."""
    if data[:4] != b"VOX " or len(data) < 20 or data[8:12] != b"MAIN":
        raise VoxError("not a .vox file")
    n, m = struct.unpack_from("<ii", data, 12)
    p, end = 20 + n, 20 + n + m
    if end > len(data):
        raise VoxError("truncated")
    models, scenes, palette, size = [], [], None, None
    while p < end:
        cid = data[p:p + 4]
        n, m = struct.unpack_from("<ii", data, p + 4)
        c = data[p + 12:p + 12 + n]
        if cid == b"SIZE":
            size = struct.unpack_from("<iii", c)
        elif cid == b"XYZI":
            (k,) = struct.unpack_from("<i", c)
            vox = [(c[4 + 4 * j], c[5 + 4 * j], c[6 + 4 * j], max(0, c[7 + 4 * j] - 1))
                   for j in range(k)]
            if size is not None:
                models.append((size, vox))
        elif cid == b"RGBA":
            palette = [tuple(c[4 * j:4 * j + 4]) for j in range(len(c) // 4)]
        elif cid == b"nTRN":
            _, q = _dict(c, 4)
            child, _res, layer, nf = struct.unpack_from("<iiii", c, q)
            q += 16
            frames = []
            for _ in range(nf):
                f, q = _dict(c, q)
                frames.append(f)
            scenes.append(("T", child, frames))
        elif cid == b"nGRP":
            _, q = _dict(c, 4)
            (k,) = struct.unpack_from("<i", c, q)
            scenes.append(("G", list(struct.unpack_from("<%di" % k, c, q + 4))))
        elif cid == b"nSHP":
            _, q = _dict(c, 4)
            (k,) = struct.unpack_from("<i", c, q)
            q += 4
            ms = []
            for _ in range(k):
                (mid,) = struct.unpack_from("<i", c, q)
                a, q = _dict(c, q + 4)
                ms.append((mid, a))
            scenes.append(("S", ms))
        p += 12 + n + m
    return {"models": models, "palette": palette, "scenes": scenes}


def parse_rotation_matrix(b: int):
    """magicavoxel.rs:63-91: row r holds its one non-zero entry in the column
    given by bits 2r..2r+1 (the third row's column is the remaining one),
    with sign bit 4+r. Returned as rows."""
    c0, c1 = b & 3, (b >> 2) & 3
    c2 = ~(c0 ^ c1) & 3
    if c0 == 3 or c1 == 3 or c0 == c1:
        raise VoxError("invalid rotation byte")
    m = [[0, 0, 0] for _ in range(3)]
    for r, col in enumerate((c0, c1, c2)):
        m[r][col] = -1 if b & (0x10 << r) else 1
    return m


def _mul(a, b):
    return [[sum(a[i][k] * b[k][j] for k in range(3)) for j in range(3)] for i in range(3)]


def _apply(m, v):
    return tuple(sum(m[r][k] * v[k] for k in range(3)) for r in range(3))


def _trunc_half(v):
    return tuple(int(x / 2) for x in v)  # i32 division truncates toward zero


_IDENT = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]


def _parse_int(s: str) -> int:
    t = s[1:] if s[:1] in "+-" else s
    if not t or not t.isdigit() or not t.isascii():
        raise VoxError("not an integer: %r" % s)
    return int(s)


def walk(vox: dict, frame: int = 0):
    """iterate_vox_tree (magicavoxel.rs:104-202): (model index, translation,
    rotation) in visiting order."""
    sc = vox["scenes"]
    if not sc or sc[0][0] != "T":
        raise VoxError("the scene root must be a transform")
    stack = [[sc[0][1], (0, 0, 0), _IDENT, 0]]
    out = []
    while stack:
        if len(stack) > len(sc) + 1:
            raise VoxError("scene graph has a cycle")
        node, tr, rot, idx = stack[-1]
        if not 0 <= node < len(sc):
            raise VoxError("scene node out of range")
        nd = sc[node]
        if nd[0] == "T":
            frames = nd[2]
            fr = frames[frame if frame < len(frames) else 0]
            t2 = tr
            if "_t" in fr:
                parts = [_parse_int(x) for x in fr["_t"].split(" ")]
                t2 = tuple(tr[k] + parts[k] for k in range(3))
            rot2 = _mul(rot, parse_rotation_matrix(_parse_int(fr["_r"]))) if "_r" in fr else _IDENT
            if idx == 0:
                stack[-1][3] += 1
                stack.append([nd[1], t2, rot2, 0])
            else:
                stack.pop()
        elif nd[0] == "G":
            ch = nd[1]
            if idx < len(ch):
                stack[-1][3] += 1
                stack.append([ch[idx], tr, rot, 0])
            else:
                stack.pop()
        else:
            for mid, attrs in nd[1]:
                if _parse_int(attrs.get("_f", "0")) == frame:
                    if not 0 <= mid < len(vox["models"]):
                        raise VoxError("model index out of range")
                    out.append((mid, tr, rot))
            stack.pop()
            if stack:
                stack[-1][3] += 1
    return out


def _rz_to_ly(v):
    """convert_coordinate Rzup <-> Lyup (spatial/math/mod.rs:196-199)."""
    return (v[0], v[2], v[1])


def load(data: bytes, brick_dim: int):
    """Octree::load_vox_file: (tree size, [(x, y, z, rgba u32)] in insert
    order). rgba is packed as Albedo::from(u32): r<<24 | g<<16 | b<<8 | a."""
    vox = parse(data)
    if vox["palette"] is None:
        raise VoxError("no RGBA chunk")
    visits = walk(vox)
    if not visits:
        raise VoxError("no model in frame 0")
    lo, hi = [I32_MAX] * 3, [I32_MIN] * 3
    for mid, p, rot in visits:
        h = _trunc_half(_apply(rot, vox["models"][mid][0]))
        for k in range(3):
            lo[k] = min(lo[k], p[k] - h[k], p[k] + h[k])
            hi[k] = max(hi[k], p[k] + h[k], p[k] - h[k])
    lo_ly, hi_ly = _rz_to_ly(lo), _rz_to_ly(hi)
    ext = max(hi_ly[k] - lo_ly[k] for k in range(3))
    f = struct.unpack("<f", struct.pack("<f", float(ext)))[0]
    e = math.ceil(math.log2(f)) if f > 0 else 0
    size = 1 << max(0, e)
    lo_rz = _rz_to_ly(lo_ly)
    inserts = []
    for mid, p, rot in visits:
        (sx, sy, sz), voxels = vox["models"][mid]
        h = _trunc_half(_apply(rot, (sx, sy, sz)))
        base = tuple(p[k] - h[k] - lo_rz[k] + (-1 if h[k] < 0 else 0) for k in range(3))
        for x, y, z, i in voxels:
            v = _apply(rot, (x, y, z))
            q = _rz_to_ly(tuple(base[k] + v[k] for k in range(3)))
            r, g, b, a = vox["palette"][i]
            inserts.append((q[0], q[1], q[2], (r << 24) | (g << 16) | (b << 8) | a))
    return size, inserts


def final_voxels(inserts) -> dict:
    """Position -> colour after every insert (the last insert of a position wins)."""
    out = {}
    for x, y, z, c in inserts:
        out[(x, y, z)] = c
    return out
