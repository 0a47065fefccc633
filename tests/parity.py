"""Comparison of GPU results with the oracle, field by field.

The bar: hit flag, palette index, albedo, voxel coordinate and flags equal
exactly; impact point, normal and distance equal bit for bit as well (the
kernel follows the reference's f32 operation sequence), with distance also
reported against the 1e-4 voxel-unit tolerance. One exception: a NaN equals
any NaN. IEEE 754 leaves the sign and payload of a NaN produced by an invalid
operation (inf - inf, 0 * inf) to the platform: x86 SSE, where the oracle and
the reference's CPU path run, produces 0xFFC00000, AMD GPUs 0x7FC00000 (as
ARM CPUs do), so only rays whose coordinates are NaN or infinite to begin with
can differ there, and only in those bits."""
from __future__ import annotations

import numpy as np

DIST_TOL = 1e-4  # voxel units

EXACT_FIELDS = ("hit", "pix", "albedo", "voxel", "flags")
FLOAT_FIELDS = ("distance", "impact", "normal")


def gpu_to_numpy(bufs: dict) -> dict:
    out = {}
    for k, t in bufs.items():
        a = t.detach().cpu().numpy()
        if a.dtype == np.int32:
            a = a.view(np.uint32)
        out[k] = a
    return out


def compare(expected: np.ndarray, got: dict) -> dict:
    """Mismatch counts per field (rays that differ anywhere in the field)."""
    n = len(expected)
    res = {"rays": n}
    for f in EXACT_FIELDS + FLOAT_FIELDS:
        if f not in got:
            continue
        e = np.ascontiguousarray(expected[f]).reshape(n, -1)
        g = np.ascontiguousarray(got[f]).reshape(n, -1)
        if e.dtype.kind == "f":
            g = np.ascontiguousarray(g, dtype=np.float32)
            both_nan = np.isnan(e) & np.isnan(g)
            differ = (e.view(np.uint32) != g.view(np.uint32)) & ~both_nan
        else:
            differ = e != g
        res[f] = int(differ.any(axis=1).sum())
    if "distance" in got:
        hit = expected["hit"] == 1
        # infinite distances (rays from beyond f32 range) compare bitwise above
        with np.errstate(invalid="ignore"):
            d = np.abs(expected["distance"][hit].astype(np.float64) -
                       got["distance"][hit].astype(np.float64))
        res["distance_over_tol"] = int((d > DIST_TOL).sum())
    return res


def assert_parity(expected: np.ndarray, got: dict, label: str = "") -> dict:
    res = compare(expected, got)
    bad = {k: v for k, v in res.items() if k != "rays" and v}
    assert not bad, f"{label}: mismatching rays per field {bad} of {res['rays']}"
    return res


_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def result_digest(h) -> int:
    """Digest of a row-major render result, as the serving loop computes it
    on the device: the wrapping sum over pixels i of a splitmix64 chain over
    (i, hit, albedo, voxel x, y, z, distance bits)."""
    n = len(h["hit"])
    a = _mix64(np.arange(n, dtype=np.uint64) + _G)
    vox = np.asarray(h["voxel"]).reshape(n, 3)
    dist = np.ascontiguousarray(h["distance"], dtype=np.float32).view(np.uint32)
    for f in (h["hit"], h["albedo"], vox[:, 0], vox[:, 1], vox[:, 2], dist):
        a = _mix64(a ^ np.asarray(f).astype(np.uint32).astype(np.uint64))
    return int(a.sum(dtype=np.uint64))


def canonical_layout(a: dict):
    """The device layout `a` (FlatTree.arrays()) as a structure independent of
    record and slot numbering: from the root record, each record's kind,
    occupancy, MIP brick and children (Internal) or bricks per octant (Leaf,
    UniformLeaf), where a Parted brick is represented by its pix words and
    its occupancy-sidecar words. Two layouts of the same tree must give equal
    structures, whatever order their records and slots were allocated in."""
    nodes, pix, occ = a["nodes"], a["brick_pix"], a["brick_occ"]
    n3 = a["brick_dim"] ** 3
    words = (n3 + 31) // 32

    def brick(form, word):
        if form == 2:
            s = int(word)
            return (2, pix[s * n3:(s + 1) * n3].tobytes(), occ[s * words:(s + 1) * words].tobytes())
        return (int(form), int(word) if form == 1 else 0)

    def rec(i):
        r = nodes[i]
        kind = int(r["meta"]) & 3
        out = [kind, int(r["occ"]), brick(int(r["mip"]), r["mip_word"])]
        for o in range(8):
            c = int(r["child"][o])
            if kind == 1:
                out.append(None if c == 0xFFFFFFFF else rec(c))
            elif kind in (2, 3):
                out.append(brick((int(r["meta"]) >> (8 + 2 * o)) & 3, c))
        return tuple(out)

    return rec(0)
