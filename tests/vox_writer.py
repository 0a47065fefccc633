"""Writes small MagicaVoxel files for the import tests, in the published
chunk format (ephtracy/voxel-model, MagicaVoxel-file-format-vox.txt and
-vox-extension.txt)."""
from __future__ import annotations

import struct


def _chunk(cid: bytes, content: bytes, children: bytes = b"") -> bytes:
    return cid + struct.pack("<ii", len(content), len(children)) + content + children


def _string(s: str) -> bytes:
    b = s.encode()
    return struct.pack("<i", len(b)) + b


def _dict(d: dict) -> bytes:
    return struct.pack("<i", len(d)) + b"".join(_string(k) + _string(v) for k, v in d.items())


def transform(child: int, frame: dict | None = None, frames: list | None = None) -> tuple:
    return ("T", child, frames if frames is not None else [frame or {}])


def group(children: list) -> tuple:
    return ("G", list(children))


def shape(models: list) -> tuple:
    """models: [model index] or [(model index, attributes)]."""
    return ("S", [m if isinstance(m, tuple) else (m, {}) for m in models])


def write(models: list, scenes: list, palette: list | None = None, version: int = 200,
          extra_chunks: bytes = b"") -> bytes:
    """models: [((sx, sy, sz), [(x, y, z, file colour index 1..255)])];
    scenes: nodes in file order (node id = position); palette: 256 (r, g, b, a)
    or None for no RGBA chunk."""
    body = b""
    for (sx, sy, sz), vox in models:
        body += _chunk(b"SIZE", struct.pack("<iii", sx, sy, sz))
        body += _chunk(b"XYZI", struct.pack("<i", len(vox)) + b"".join(bytes(v) for v in vox))
    for nid, nd in enumerate(scenes):
        if nd[0] == "T":
            c = struct.pack("<i", nid) + _dict({}) + struct.pack("<iiii", nd[1], -1, 0, len(nd[2]))
            c += b"".join(_dict(f) for f in nd[2])
            body += _chunk(b"nTRN", c)
        elif nd[0] == "G":
            c = struct.pack("<i", nid) + _dict({}) + struct.pack("<i", len(nd[1]))
            c += b"".join(struct.pack("<i", k) for k in nd[1])
            body += _chunk(b"nGRP", c)
        else:
            c = struct.pack("<i", nid) + _dict({}) + struct.pack("<i", len(nd[1]))
            c += b"".join(struct.pack("<i", m) + _dict(a) for m, a in nd[1])
            body += _chunk(b"nSHP", c)
    body += extra_chunks
    if palette is not None:
        body += _chunk(b"RGBA", b"".join(bytes(c) for c in palette))
    return b"VOX " + struct.pack("<i", version) + _chunk(b"MAIN", b"", body)


def palette_ramp() -> list:
    """256 distinct opaque colours."""
    return [((7 * i) & 255, (13 * i + 5) & 255, i, 255) for i in range(256)]


def placed(models: list, placements: list) -> tuple:
    """A scene root -> group -> one transform + shape per placement.
    placements: [(model index, frame attribute dict)]."""
    scenes = [transform(1), group([2 + 2 * k for k in range(len(placements))])]
    for m, frame in placements:
        scenes.append(transform(len(scenes) + 1, frame))
        scenes.append(shape([m]))
    return scenes
