"""ctypes front end of libhostsim.so: the product's traversal header compiled
for the host (tests only; see hostsim.cpp)."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libhostsim.so"
HIT_DTYPE = np.dtype([("hit", "<u4"), ("pix", "<u4"), ("albedo", "<u4"), ("flags", "<u4"),
                      ("voxel", "<u4", (3,)), ("distance", "<f4"), ("impact", "<f4", (3,)),
                      ("normal", "<f4", (3,))])


class _Tree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("brick_pix", C.c_void_p), ("brick_occ", C.c_void_p),
                ("colors", C.c_void_p), ("ray_lut", C.c_void_p), ("size", C.c_float),
                ("dim", C.c_uint32), ("occ_words", C.c_uint32), ("color_count", C.c_uint32),
                ("max_steps", C.c_uint32), ("mips_enabled", C.c_uint32),
                ("viewing_distance", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB))
        L.hs_cast.argtypes = [C.POINTER(_Tree), C.c_void_p, C.c_uint64, C.c_void_p, C.c_int,
                              C.c_int]
        L.hs_viewport_rays.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        _lib = L
    return _lib


def ray_lut() -> np.ndarray:
    from shocovox_amd import _native as N
    out = np.zeros(512, np.uint64)
    N.check(N.lib().svo_ray_occupancy_lut(out.ctypes.data))
    return out


def cast(flat_arrays: dict, rays: np.ndarray, resumable: bool = False,
         max_steps: int = 0, viewing_distance: float | None = None,
         general: bool = False, regrouped: bool | str = False) -> np.ndarray:
    """The product traversal on the host: cast_ray, with resumable=True its
    ray_begin / ray_step form, with a viewing distance cast_ray_lod, with
    regrouped=True the while-while form cast_ray_ww (regrouped="nocrawl":
    that form without the crawl shortcut). As on the device, power-of-two
    trees take the P2 form unless general=True."""
    a = flat_arrays
    keep = [np.ascontiguousarray(a["nodes"]), np.ascontiguousarray(a["brick_pix"]),
            np.ascontiguousarray(a["brick_occ"]), np.ascontiguousarray(a["colors"]), ray_lut()]
    dim = a["brick_dim"]
    t = _Tree(*(k.ctypes.data for k in keep), float(a["octree_size"]), dim,
              (dim ** 3 + 31) // 32, len(a["colors"]), max_steps,
              int(a.get("mips_enabled", 0)),
              float(viewing_distance) if viewing_distance is not None else 0.0)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    out = np.zeros(len(rays), HIT_DTYPE)
    mode = 2 if viewing_distance is not None else (
        4 if regrouped == "nocrawl" else (3 if regrouped else int(resumable)))
    lib().hs_cast(C.byref(t), rays.ctypes.data, len(rays), out.ctypes.data, mode, int(general))
    return out


def viewport_rays(vp, w: int, h: int) -> np.ndarray:
    v = np.array([*vp.origin, *vp.direction, *vp.frustum, vp.fov], dtype=np.float32)
    rays = np.zeros((w * h, 6), np.float32)
    lib().hs_viewport_rays(v.ctypes.data, w, h, rays.ctypes.data)
    return rays
