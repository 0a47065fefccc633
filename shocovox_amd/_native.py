"""ctypes binding of libsvo.so (declarations in include/svo.h).

The library is built in-tree by shocovox_amd.build; importing this module
never falls back to any other implementation: a missing or stale library
raises immediately.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsvo.so"

SVO_OK = 0
STATUS_NAMES = {
    0: "OK", 1: "InvalidSize", 2: "InvalidBrickDimension", 3: "InvalidStructure",
    4: "InvalidPosition", 5: "InvalidArgument", 6: "OutOfDeviceMemory", 7: "HipError",
    8: "NoDevice",
}


class SvoError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Entry(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("albedo", C.c_uint32), ("data", C.c_uint32)]


class ModelSizes(C.Structure):
    _fields_ = [("octree_size", C.c_uint32), ("brick_dim", C.c_uint32), ("pool_len", C.c_uint32),
                ("parted_count", C.c_uint32), ("color_count", C.c_uint32),
                ("data_count", C.c_uint32), ("mip_parted_count", C.c_uint32),
                ("mips_enabled", C.c_uint32)]


class ModelDump(C.Structure):
    _fields_ = [("reserved", C.c_void_p), ("content_kind", C.c_void_p),
                ("internal_bits", C.c_void_p), ("children_kind", C.c_void_p),
                ("children", C.c_void_p), ("children_bitmap", C.c_void_p),
                ("brick_kind", C.c_void_p), ("brick_value", C.c_void_p), ("voxels", C.c_void_p),
                ("colors", C.c_void_p), ("data", C.c_void_p), ("mip_kind", C.c_void_p),
                ("mip_value", C.c_void_p), ("mip_voxels", C.c_void_p)]


class TreeDesc(C.Structure):
    _fields_ = [("octree_size", C.c_uint32), ("brick_dim", C.c_uint32),
                ("node_count", C.c_uint32), ("brick_count", C.c_uint32),
                ("color_count", C.c_uint32), ("top_level_nodes", C.c_uint32),
                ("nodes", C.c_void_p), ("brick_pix", C.c_void_p), ("brick_occ", C.c_void_p),
                ("colors", C.c_void_p), ("mips_enabled", C.c_uint32)]


class Hits(C.Structure):
    _fields_ = [("hit", C.c_void_p), ("pix", C.c_void_p), ("albedo", C.c_void_p),
                ("voxel", C.c_void_p), ("distance", C.c_void_p), ("impact", C.c_void_p),
                ("normal", C.c_void_p), ("flags", C.c_void_p)]


class Viewport(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("direction", C.c_float * 3),
                ("frustum", C.c_float * 3), ("fov", C.c_float)]


class Job(C.Structure):
    _fields_ = [("tree", C.c_void_p), ("kind", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("viewport", Viewport), ("rays", C.c_void_p),
                ("out", Hits), ("viewing_distance", C.c_float)]


class ServeConfig(C.Structure):
    _fields_ = [("qps", C.c_double), ("duration_s", C.c_double), ("seed", C.c_uint64),
                ("mode", C.c_uint32), ("max_inflight", C.c_uint32),
                ("max_batch_rays", C.c_uint64), ("max_batch_jobs", C.c_uint32),
                ("max_queue", C.c_uint32), ("edit_fraction", C.c_double),
                ("max_edit_voxels", C.c_uint32), ("zipf_s", C.c_double),
                ("n_resolutions", C.c_uint32), ("resolutions", C.POINTER(C.c_uint32)),
                ("deadline_ms", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("kind", C.c_uint32), ("tree", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("viewport", Viewport),
                ("edit_voxels", C.c_uint32), ("edit_clear", C.c_uint32),
                ("edit_seed", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = ([(k, C.c_uint64) for k in ("requests", "renders", "edits", "rejected", "rays",
                                            "launches")]
                + [(k, C.c_double) for k in ("elapsed_s", "mrays_per_s", "p50_ms", "p99_ms",
                                             "mean_ms", "max_ms", "render_p50_ms",
                                             "render_p99_ms", "edit_p50_ms", "edit_p99_ms",
                                             "mean_jobs_per_launch")]
                + [("edit_records", C.c_uint64), ("edit_bricks", C.c_uint64)]
                + [(k, C.c_double) for k in ("edit_wait_p99_ms", "edit_apply_p99_ms",
                                             "edit_upload_p99_ms")]
                + [("timed_out", C.c_uint64), ("single_launches", C.c_uint64)])


_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_SIGS = {
    "svo_last_error": (C.c_char_p, []),
    "svo_version": (C.c_char_p, []),
    "svo_octree_new": (C.c_int, [_U32, _U32, C.POINTER(_P)]),
    "svo_octree_free": (None, [_P]),
    "svo_octree_get_size": (_U32, [_P]),
    "svo_octree_get_brick_dim": (_U32, [_P]),
    "svo_octree_set_auto_simplify": (C.c_int, [_P, C.c_int]),
    "svo_octree_insert": (C.c_int, [_P, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_octree_insert_at_lod": (C.c_int, [_P, _U32, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_octree_update": (C.c_int, [_P, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_octree_clear": (C.c_int, [_P, _U32, _U32, _U32]),
    "svo_octree_clear_at_lod": (C.c_int, [_P, _U32, _U32, _U32, _U32]),
    "svo_octree_get": (C.c_int, [_P, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_octree_switch_mip_maps": (C.c_int, [_P, C.c_int]),
    "svo_octree_mip_maps_enabled": (C.c_int, [_P]),
    "svo_octree_set_mip_method": (C.c_int, [_P, _U32, _U32, C.c_float]),
    "svo_octree_get_mip_method": (C.c_int, [_P, _U32, C.POINTER(_U32), C.POINTER(C.c_float)]),
    "svo_octree_set_mip_color_threshold": (C.c_int, [_P, _U32, C.c_float]),
    "svo_octree_get_mip_color_threshold": (C.c_int, [_P, _U32, C.POINTER(C.c_float)]),
    "svo_octree_reset_mip_strategy": (C.c_int, [_P]),
    "svo_octree_recalculate_mips": (C.c_int, [_P]),
    "svo_octree_sample_root_mip": (C.c_int, [_P, _U32, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_octree_insert_many": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "svo_octree_clear_many": (C.c_int, [_P, _P, C.c_size_t]),
    "svo_octree_get_many": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "svo_octree_export": (C.c_int, [_P, C.POINTER(ModelSizes), C.POINTER(ModelDump)]),
    "svo_scene_build": (C.c_int, [_U32, _U32, _U32, C.c_uint64, C.POINTER(_P)]),
    "svo_octree_to_bytes": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "svo_octree_from_bytes": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "svo_octree_save": (C.c_int, [_P, C.c_char_p]),
    "svo_octree_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "svo_octree_clone": (C.c_int, [_P, C.POINTER(_P)]),
    "svo_octree_from_vox_bytes": (C.c_int, [_P, C.c_size_t, C.c_uint32, C.POINTER(_P)]),
    "svo_octree_load_vox": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_P)]),
    "svo_tree_load": (C.c_int, [C.c_char_p, C.POINTER(_P), C.POINTER(_P)]),
    "svo_octree_flatten": (C.c_int, [_P, C.POINTER(_P)]),
    "svo_flat_desc": (C.c_int, [_P, C.POINTER(TreeDesc)]),
    "svo_flat_sync": (C.c_int, [_P, _P]),
    "svo_flat_free": (None, [_P]),
    "svo_flat_get": (C.c_int, [_P, _P, _U32, _U32, _U32, C.POINTER(Entry)]),
    "svo_tree_create": (C.c_int, [C.POINTER(TreeDesc), C.POINTER(_P)]),
    "svo_tree_from_octree": (C.c_int, [_P, C.POINTER(_P)]),
    "svo_tree_update": (C.c_int, [_P, _P, _P]),
    "svo_tree_update_stats": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "svo_tree_destroy": (None, [_P]),
    "svo_tree_device_bytes": (C.c_uint64, [_P]),
    "svo_tree_set_step_limit": (C.c_int, [_P, _U32]),
    "svo_cast": (C.c_int, [_P, _P, _U32, C.POINTER(Hits), _P]),
    "svo_render": (C.c_int, [_P, C.POINTER(Viewport), _U32, _U32, C.POINTER(Hits), _P]),
    "svo_cast_at_lod": (C.c_int, [_P, _P, _U32, C.c_float, C.POINTER(Hits), _P]),
    "svo_render_at_lod": (C.c_int, [_P, C.POINTER(Viewport), _U32, _U32, C.c_float,
                                    C.POINTER(Hits), _P]),
    "svo_render_rgba8": (C.c_int, [_P, C.POINTER(Viewport), _U32, _U32, _U32, _P, _P]),
    "svo_ray_occupancy_lut": (C.c_int, [_P]),
    "svo_batch_create": (C.c_int, [_U32, C.POINTER(_P)]),
    "svo_batch_destroy": (None, [_P]),
    "svo_batch_run": (C.c_int, [_P, _P, _U32, _P]),
    "svo_batch_run_static": (C.c_int, [_P, _U32, _P]),
    "svo_batch_tiles": (C.c_uint64, [_P]),
    "svo_batch_variant_count": (C.c_int, []),
    "svo_batch_variant_name": (C.c_char_p, [C.c_int]),
    "svo_select_batch_variant": (C.c_int, [C.c_int]),
    "svo_current_batch_variant": (C.c_int, []),
    "svo_batch_variant_scratch": (C.c_int, [C.c_int, C.POINTER(C.c_uint32)]),
    "svo_serve_plan": (C.c_int, [C.POINTER(ServeConfig), _P, _U32, _P, _U64,
                                 C.POINTER(_U64)]),
    "svo_serve_apply_edit": (C.c_int, [_P, C.POINTER(ServeRequest)]),
    "svo_serve_run": (C.c_int, [_P, _P, _U32, C.POINTER(ServeConfig), C.POINTER(ServeStats), _P,
                                _P, _U64]),
    "svo_device_memory": (C.c_int, [C.POINTER(_U64), C.POINTER(_U64)]),
    "svo_tree_admit": (C.c_int, [_P, _U64, _U64, C.POINTER(_P)]),
    "svo_render_variant_count": (C.c_int, []),
    "svo_render_variant_name": (C.c_char_p, [C.c_int]),
    "svo_select_render_variant": (C.c_int, [C.c_int]),
    "svo_current_render_variant": (C.c_int, []),
    "svo_render_variant_scratch": (C.c_int, [C.c_int, C.POINTER(C.c_uint32)]),
    "svo_set_timing": (C.c_int, [C.c_int]),
    "svo_cast_set_order_threshold": (C.c_int, [_U32]),
    "svo_collect_kernel_ms": (_U32, [C.POINTER(C.c_float), _U32]),
}

_lib = None


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib() -> C.CDLL:
    """Loads libsvo.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m shocovox_amd.build`")
    # torch bundles its own HIP runtime with the same SONAME (libamdhip64.so.7).
    # Loading torch first makes libsvo bind to that copy, so device pointers
    # from torch tensors and libsvo's kernels share one runtime per process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    handle = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return handle


def check(status: int) -> None:
    if status != SVO_OK:
        msg = lib().svo_last_error()
        raise SvoError(status, msg.decode() if msg else "")
