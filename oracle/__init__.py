"""ORACLE - test infrastructure only.

ctypes front end of libsvo_oracle.so, the CPU restatement of the reference's
get_by_ray / get (see svo_oracle.cpp). Imported only by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, always as the
checker or the timed CPU baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsvo_oracle.so"

HIT_DTYPE = np.dtype([("hit", "<u4"), ("pix", "<u4"), ("albedo", "<u4"), ("flags", "<u4"),
                      ("voxel", "<u4", (3,)), ("distance", "<f4"), ("impact", "<f4", (3,)),
                      ("normal", "<f4", (3,)), ("boundary", "<u4")])


class _Dump(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("octree_size", "brick_dim", "pool_len", "parted_count",
                                          "color_count", "data_count")] + \
               [(n, C.c_void_p) for n in ("reserved", "content_kind", "internal_bits",
                                          "children_kind", "children", "children_bitmap",
                                          "brick_kind", "brick_value", "voxels", "colors",
                                          "data")] + \
               [(n, C.c_uint32) for n in ("mip_parted_count", "mips_enabled")] + \
               [(n, C.c_void_p) for n in ("mip_kind", "mip_value", "mip_voxels")]


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "stack_iterations", "child_reads",
                                          "brick_reads", "voxel_tests", "data_palette_reads",
                                          "hits", "boundary_hits")]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        L.ora_tree_load.restype = P
        L.ora_tree_load.argtypes = [C.POINTER(_Dump)]
        L.ora_tree_free.argtypes = [P]
        L.ora_get.restype = C.c_uint32
        L.ora_get.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        L.ora_get_many.argtypes = [P, P, C.c_uint64, P]
        L.ora_cast.argtypes = [P, P, C.c_uint64, P, C.c_uint32, C.POINTER(Counters)]
        L.ora_cast_lod.argtypes = [P, P, C.c_uint64, P, C.c_uint32, C.POINTER(Counters),
                                   C.c_float]
        L.ora_viewport_rays.argtypes = [P, C.c_uint32, C.c_uint32, P]
        L.ora_luts.argtypes = [P, P, P, P]
        L.ora_intersect_ray.restype = C.c_int
        L.ora_intersect_ray.argtypes = [P, P, P]
        L.ora_cube_impact_normal.argtypes = [P, P, P]
        L.ora_hash_region.restype = C.c_uint32
        L.ora_hash_region.argtypes = [P, C.c_float]
        L.ora_flat_projection.restype = C.c_uint64
        L.ora_flat_projection.argtypes = [C.c_uint64] * 4
        L.ora_position_in_bitmap_64bits.restype = C.c_uint64
        L.ora_position_in_bitmap_64bits.argtypes = [C.c_uint64] * 4
        L.ora_set_occupancy.restype = C.c_uint64
        L.ora_set_occupancy.argtypes = [C.c_uint64] * 5 + [C.c_int, C.c_uint64]
        L.ora_child_bounds.argtypes = [P, C.c_uint32, P]
        L.ora_nodestack_run.restype = C.c_uint32
        L.ora_nodestack_run.argtypes = [C.c_uint32, P, C.c_uint32, P]
        L.ora_dda_step.argtypes = [P, P, P, P]
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleTree:
    """The reference data model loaded from an Octree.export_model() dump."""

    def __init__(self, model: dict):
        self._keep = {k: np.ascontiguousarray(v) for k, v in model.items()
                      if isinstance(v, np.ndarray)}
        d = _Dump(model["octree_size"], model["brick_dim"], model["pool_len"],
                  model["parted_count"], model["color_count"], model["data_count"],
                  *(self._keep[k].ctypes.data for k in (
                      "reserved", "content_kind", "internal_bits", "children_kind", "children",
                      "children_bitmap", "brick_kind", "brick_value", "voxels", "colors",
                      "data")),
                  model.get("mip_parted_count", 0), model.get("mips_enabled", 0),
                  *((self._keep[k].ctypes.data if k in self._keep else None)
                    for k in ("mip_kind", "mip_value", "mip_voxels")))
        self._h = lib().ora_tree_load(C.byref(d))
        self._keep = None  # the oracle copied what it needs

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ora_tree_free(h)
            self._h = None

    def get(self, pos) -> int:
        """Palette index value the reference's get() resolves (0xFFFFFFFF = Empty)."""
        return int(lib().ora_get(self._h, int(pos[0]), int(pos[1]), int(pos[2])))

    def get_many(self, xyz: np.ndarray) -> np.ndarray:
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32)
        out = np.zeros(len(xyz), np.uint32)
        lib().ora_get_many(self._h, xyz.ctypes.data, len(xyz), out.ctypes.data)
        return out

    def cast(self, rays: np.ndarray, threads: int = 1, counters: bool = False,
             viewing_distance: float | None = None):
        """get_by_ray for each row (n x 6 f32: origin, direction), or
        get_by_ray_at_lod when a viewing distance is given."""
        rays = _f32(rays).reshape(-1, 6)
        out = np.zeros(len(rays), HIT_DTYPE)
        c = Counters()
        if viewing_distance is None:
            lib().ora_cast(self._h, rays.ctypes.data, len(rays), out.ctypes.data, threads,
                           C.byref(c) if counters else None)
        else:
            lib().ora_cast_lod(self._h, rays.ctypes.data, len(rays), out.ctypes.data, threads,
                               C.byref(c) if counters else None, float(viewing_distance))
        return (out, c) if counters else out


def viewport_rays(origin, direction, frustum, fov, w: int, h: int) -> np.ndarray:
    vp = _f32([*origin, *direction, *frustum, fov])
    rays = np.zeros((h * w, 6), np.float32)
    lib().ora_viewport_rays(vp.ctypes.data, w, h, rays.ctypes.data)
    return rays


def luts() -> dict:
    ray = np.zeros((64, 8), np.uint64)
    step = np.zeros((3, 3, 3), np.uint32)
    masks = np.zeros(8, np.uint64)
    bidx = np.zeros((4, 4, 4), np.uint64)
    lib().ora_luts(ray.ctypes.data, step.ctypes.data, masks.ctypes.data, bidx.ctypes.data)
    return {"ray_to_node_occupancy": ray, "octant_step_result": step,
            "bitmap_mask_for_octant": masks, "bitmap_index": bidx}


def intersect_ray(cube_min, size, origin, direction):
    """Cube::intersect_ray: None, ('inside', None) or ('hit', t)."""
    c = _f32([*cube_min, size])
    r = _f32([*origin, *direction])
    t = np.zeros(1, np.float32)
    k = lib().ora_intersect_ray(c.ctypes.data, r.ctypes.data, t.ctypes.data)
    if k == 0:
        return None
    if k == 1:
        return ("inside", None)
    return ("hit", float(t[0]))


def cube_impact_normal(cube_min, size, p) -> np.ndarray:
    c = _f32([*cube_min, size])
    pp = _f32(p)
    out = np.zeros(3, np.float32)
    lib().ora_cube_impact_normal(c.ctypes.data, pp.ctypes.data, out.ctypes.data)
    return out


def hash_region(offset, half) -> int:
    o = _f32(offset)
    return int(lib().ora_hash_region(o.ctypes.data, half))


def flat_projection(x, y, z, size) -> int:
    return int(lib().ora_flat_projection(x, y, z, size))


def position_in_bitmap_64bits(x, y, z, brick_size) -> int:
    return int(lib().ora_position_in_bitmap_64bits(x, y, z, brick_size))


def set_occupancy(pos, size, brick_dim, occupied, bitmap) -> int:
    return int(lib().ora_set_occupancy(pos[0], pos[1], pos[2], size, brick_dim, int(occupied),
                                       bitmap))


def child_bounds(cube_min, size, octant):
    c = _f32([*cube_min, size])
    out = np.zeros(4, np.float32)
    lib().ora_child_bounds(c.ctypes.data, octant, out.ctypes.data)
    return tuple(float(v) for v in out[:3]), float(out[3])


def nodestack_run(size: int, ops) -> list:
    o = np.ascontiguousarray(ops, dtype=np.int32)
    res = np.zeros(len(o), np.int32)
    n = lib().ora_nodestack_run(size, o.ctypes.data, len(o), res.ctypes.data)
    return [None if v == np.iinfo(np.int32).min else int(v) for v in res[:n]]


def dda_step(origin, direction, point, cube_min, size):
    r = _f32([*origin, *direction])
    p = _f32(point).copy()
    c = _f32([*cube_min, size])
    s = np.zeros(3, np.float32)
    lib().ora_dda_step(r.ctypes.data, p.ctypes.data, c.ctypes.data, s.ctypes.data)
    return p, s
