"""Python mirror of the reference's octree API over libsvo.

`Octree` mirrors shocovox-rs `Octree<u32>` (src/octree/mod.rs, update/*.rs):
new / insert / insert_at_lod / update / clear / clear_at_lod / get /
get_size, with `OctreeEntry` values. `DeviceTree` is the device-resident form
that answers `get_by_ray` for batches of rays on the GPU
(src/raytracing/raytracing_on_cpu.rs:316) and renders viewports
(the bevy `OctreeGPUView`, src/raytracing/bevy/types.rs:55-134).
Device buffers are torch tensors; the traversal itself runs in libsvo's HIP
kernels, and nothing here computes on the CPU in its place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

EMPTY, VISUAL, INFORMATIVE, COMPLEX = 0, 1, 2, 3
EMPTY_MARKER = 0xFFFFFFFF


def albedo(r: int, g: int, b: int, a: int) -> int:
    """Packs an Albedo like Albedo::from(u32) unpacks it (detail.rs:92-105)."""
    return ((r & 0xFF) << 24) | ((g & 0xFF) << 16) | ((b & 0xFF) << 8) | (a & 0xFF)


def albedo_rgba(v: int) -> tuple[int, int, int, int]:
    return ((v >> 24) & 0xFF, (v >> 16) & 0xFF, (v >> 8) & 0xFF, v & 0xFF)


@dataclass(frozen=True)
class OctreeEntry:
    """OctreeEntry (src/octree/types.rs:23-36)."""

    kind: int = EMPTY
    albedo: int = 0
    data: int = 0

    @staticmethod
    def visual(albedo_u32: int) -> "OctreeEntry":
        return OctreeEntry(VISUAL, albedo_u32 & 0xFFFFFFFF, 0)

    @staticmethod
    def informative(data: int) -> "OctreeEntry":
        return OctreeEntry(INFORMATIVE, 0, data & 0xFFFFFFFF)

    @staticmethod
    def complex(albedo_u32: int, data: int) -> "OctreeEntry":
        return OctreeEntry(COMPLEX, albedo_u32 & 0xFFFFFFFF, data & 0xFFFFFFFF)

    def is_none(self) -> bool:
        """OctreeEntry::is_none (src/octree/mod.rs:97-104)."""
        if self.kind == EMPTY:
            return True
        if self.kind == VISUAL:
            return (self.albedo & 0xFF) == 0
        if self.kind == INFORMATIVE:
            return self.data == 0
        return (self.albedo & 0xFF) == 0 and self.data == 0

    def is_some(self) -> bool:
        return not self.is_none()

    def to_c(self) -> N.Entry:
        return N.Entry(self.kind, self.albedo, self.data)

    @staticmethod
    def from_c(e: N.Entry) -> "OctreeEntry":
        return OctreeEntry(int(e.kind), int(e.albedo), int(e.data))


def voxel_data(d: int) -> OctreeEntry:
    """The reference's `voxel_data!` macro (src/octree/mod.rs:62-70)."""
    return OctreeEntry.informative(d)


class OctreeError(N.SvoError):
    pass


def _release(fn: str, handle: int) -> None:
    """Frees a library handle; quiet during interpreter shutdown, when module
    globals may already be gone."""
    try:
        getattr(N.lib(), fn)(handle)
    except Exception:  # noqa: BLE001 - nothing to report to at shutdown
        pass


def _check(status: int) -> None:
    if status != N.SVO_OK:
        msg = N.lib().svo_last_error()
        raise OctreeError(status, msg.decode() if msg else "")


@dataclass(frozen=True)
class MIPResamplingMethods:
    """MIPResamplingMethods (octree/types.rs:104-140): BoxFilter, PointFilter,
    PointFilterBD, Posterize(thr), PosterizeBD(thr)."""
    kind: int
    thr: float = 0.0

    @staticmethod
    def Posterize(thr: float) -> "MIPResamplingMethods":  # noqa: N802 (reference name)
        return MIPResamplingMethods(3, float(thr))

    @staticmethod
    def PosterizeBD(thr: float) -> "MIPResamplingMethods":  # noqa: N802
        return MIPResamplingMethods(4, float(thr))


MIPResamplingMethods.BoxFilter = MIPResamplingMethods(0)
MIPResamplingMethods.PointFilter = MIPResamplingMethods(1)
MIPResamplingMethods.PointFilterBD = MIPResamplingMethods(2)


class StrategyUpdater:
    """The chainable MIP strategy editor Octree::albedo_mip_map_resampling_strategy
    returns (octree/mod.rs:379-381, mipmap.rs:706-937)."""

    def __init__(self, tree: "Octree"):
        self._t = tree

    def _h(self):
        return self._t.handle

    def switch_albedo_mip_maps(self, enabled: bool) -> "StrategyUpdater":
        _check(N.lib().svo_octree_switch_mip_maps(self._h(), int(bool(enabled))))
        return self

    def is_enabled(self) -> bool:
        return bool(N.lib().svo_octree_mip_maps_enabled(self._h()))

    def set_method_at(self, mip_level: int, method: MIPResamplingMethods) -> "StrategyUpdater":
        _check(N.lib().svo_octree_set_mip_method(self._h(), int(mip_level), method.kind,
                                                 float(method.thr)))
        return self

    def set_method(self, levels) -> "StrategyUpdater":
        for level, method in (levels.items() if isinstance(levels, dict) else levels):
            self.set_method_at(level, method)
        return self

    def get_method_at(self, mip_level: int) -> MIPResamplingMethods:
        m, t = C.c_uint32(), C.c_float()
        _check(N.lib().svo_octree_get_mip_method(self._h(), int(mip_level), C.byref(m),
                                                 C.byref(t)))
        return MIPResamplingMethods(int(m.value), float(t.value) if m.value >= 3 else 0.0)

    def set_color_similarity_thr_at(self, mip_level: int, thr: float) -> "StrategyUpdater":
        _check(N.lib().svo_octree_set_mip_color_threshold(self._h(), int(mip_level), float(thr)))
        return self

    def set_color_similarity_thr(self, levels) -> "StrategyUpdater":
        for level, thr in (levels.items() if isinstance(levels, dict) else levels):
            self.set_color_similarity_thr_at(level, thr)
        return self

    def get_new_color_similarity_at(self, mip_level: int) -> float:
        t = C.c_float()
        _check(N.lib().svo_octree_get_mip_color_threshold(self._h(), int(mip_level), C.byref(t)))
        return float(t.value)

    def reset(self) -> "StrategyUpdater":
        _check(N.lib().svo_octree_reset_mip_strategy(self._h()))
        return self

    def recalculate_mips(self) -> "StrategyUpdater":
        _check(N.lib().svo_octree_recalculate_mips(self._h()))
        return self

    def sample_root_mip(self, octant: int, position) -> "OctreeEntry":
        e = N.Entry()
        _check(N.lib().svo_octree_sample_root_mip(self._h(), int(octant), int(position[0]),
                                                  int(position[1]), int(position[2]),
                                                  C.byref(e)))
        return OctreeEntry.from_c(e)


class Octree:
    """Host octree, Octree<u32> of the reference."""

    def __init__(self, size: int, brick_dim: int, _handle: int | None = None):
        lib = N.lib()
        if _handle is None:
            h = C.c_void_p()
            _check(lib.svo_octree_new(size, brick_dim, C.byref(h)))
            _handle = h.value
        self._h = _handle

    @classmethod
    def new(cls, size: int, brick_dim: int) -> "Octree":
        return cls(size, brick_dim)

    @classmethod
    def from_bytes(cls, data: bytes) -> "Octree":
        """Octree::from_bytes (octree/mod.rs:143-147): the reference's bencode."""
        h = C.c_void_p()
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
        _check(N.lib().svo_octree_from_bytes(buf, len(data), C.byref(h)))
        return cls(0, 0, _handle=h.value)

    @classmethod
    def load(cls, path: str) -> "Octree":
        """Octree::load (octree/mod.rs:161-168)."""
        h = C.c_void_p()
        _check(N.lib().svo_octree_load(str(path).encode(), C.byref(h)))
        return cls(0, 0, _handle=h.value)

    def clone(self) -> "Octree":
        """An independent copy (the reference's Octree derives Clone)."""
        h = C.c_void_p()
        _check(N.lib().svo_octree_clone(self._h, C.byref(h)))
        return Octree(0, 0, _handle=h.value)

    @classmethod
    def load_vox_file(cls, filename: str, brick_dimension: int) -> "Octree":
        """Octree::load_vox_file (convert/magicavoxel.rs:259-278): a MagicaVoxel
        scene, frame 0, in a tree of edge 2^ceil(log2(extent))."""
        h = C.c_void_p()
        _check(N.lib().svo_octree_load_vox(str(filename).encode(), int(brick_dimension),
                                           C.byref(h)))
        return cls(0, 0, _handle=h.value)

    @classmethod
    def from_vox_bytes(cls, data: bytes, brick_dimension: int) -> "Octree":
        """load_vox_file on a file image already in memory."""
        h = C.c_void_p()
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
        _check(N.lib().svo_octree_from_vox_bytes(buf, len(data), int(brick_dimension),
                                                 C.byref(h)))
        return cls(0, 0, _handle=h.value)

    def to_bytes(self) -> bytes:
        """Octree::to_bytes (octree/mod.rs:136-140)."""
        n = C.c_size_t()
        _check(N.lib().svo_octree_to_bytes(self._h, None, 0, C.byref(n)))
        buf = (C.c_uint8 * max(1, n.value))()
        _check(N.lib().svo_octree_to_bytes(self._h, buf, n.value, C.byref(n)))
        return C.string_at(buf, n.value)

    def save(self, path: str) -> None:
        """Octree::save (octree/mod.rs:151-157)."""
        _check(N.lib().svo_octree_save(self._h, str(path).encode()))

    @classmethod
    def scene(cls, kind: int, size: int, brick_dim: int, seed: int) -> "Octree":
        """Builds one of the seeded benchmark scenes (see include/svo.h)."""
        h = C.c_void_p()
        _check(N.lib().svo_scene_build(kind, size, brick_dim, seed, C.byref(h)))
        return cls(0, 0, _handle=h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _release("svo_octree_free", h)

    @property
    def handle(self) -> int:
        return self._h

    def get_size(self) -> int:
        return int(N.lib().svo_octree_get_size(self._h))

    @property
    def brick_dim(self) -> int:
        return int(N.lib().svo_octree_get_brick_dim(self._h))

    def set_auto_simplify(self, enabled: bool) -> None:
        _check(N.lib().svo_octree_set_auto_simplify(self._h, int(enabled)))

    def insert(self, pos, entry: OctreeEntry) -> None:
        e = entry.to_c()
        _check(N.lib().svo_octree_insert(self._h, pos[0], pos[1], pos[2], C.byref(e)))

    def insert_at_lod(self, pos, size: int, entry: OctreeEntry) -> None:
        e = entry.to_c()
        _check(N.lib().svo_octree_insert_at_lod(self._h, pos[0], pos[1], pos[2], size,
                                                 C.byref(e)))

    def update(self, pos, entry: OctreeEntry) -> None:
        e = entry.to_c()
        _check(N.lib().svo_octree_update(self._h, pos[0], pos[1], pos[2], C.byref(e)))

    def clear(self, pos) -> None:
        _check(N.lib().svo_octree_clear(self._h, pos[0], pos[1], pos[2]))

    def clear_at_lod(self, pos, size: int) -> None:
        _check(N.lib().svo_octree_clear_at_lod(self._h, pos[0], pos[1], pos[2], size))

    def get(self, pos) -> OctreeEntry:
        e = N.Entry()
        _check(N.lib().svo_octree_get(self._h, pos[0], pos[1], pos[2], C.byref(e)))
        return OctreeEntry.from_c(e)

    def albedo_mip_map_resampling_strategy(self) -> StrategyUpdater:
        """Octree::albedo_mip_map_resampling_strategy (octree/mod.rs:379-381)."""
        return StrategyUpdater(self)

    def insert_many(self, xyz: np.ndarray, entries: np.ndarray) -> None:
        """insert() for each row of xyz (n,3 u32) with entries (n,3 u32: kind, albedo, data)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32)
        entries = np.ascontiguousarray(entries, dtype=np.uint32)
        _check(N.lib().svo_octree_insert_many(self._h, xyz.ctypes.data, entries.ctypes.data,
                                              len(xyz)))

    def clear_many(self, xyz: np.ndarray) -> None:
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32)
        _check(N.lib().svo_octree_clear_many(self._h, xyz.ctypes.data, len(xyz)))

    def get_many(self, xyz: np.ndarray) -> np.ndarray:
        """get() for each row; returns (n,3) u32 rows of (kind, albedo, data)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32)
        out = np.zeros((len(xyz), 3), dtype=np.uint32)
        _check(N.lib().svo_octree_get_many(self._h, xyz.ctypes.data, len(xyz), out.ctypes.data))
        return out

    def export_model(self) -> dict:
        """Raw dump of the node model (see svo_octree_export)."""
        lib = N.lib()
        s = N.ModelSizes()
        _check(lib.svo_octree_export(self._h, C.byref(s), None))
        dim = s.brick_dim
        n = s.pool_len
        out = {
            "octree_size": s.octree_size, "brick_dim": dim,
            "reserved": np.zeros(n, np.uint8), "content_kind": np.zeros(n, np.uint8),
            "internal_bits": np.zeros(n, np.uint64), "children_kind": np.zeros(n, np.uint8),
            "children": np.zeros(n * 8, np.uint32), "children_bitmap": np.zeros(n, np.uint64),
            "brick_kind": np.zeros(n * 8, np.uint8), "brick_value": np.zeros(n * 8, np.uint32),
            "voxels": np.zeros(max(1, s.parted_count * dim ** 3), np.uint32),
            "colors": np.zeros(max(1, s.color_count), np.uint32),
            "data": np.zeros(max(1, s.data_count), np.uint32),
            "mip_kind": np.zeros(max(1, n), np.uint8), "mip_value": np.zeros(max(1, n), np.uint32),
            "mip_voxels": np.zeros(max(1, s.mip_parted_count * dim ** 3), np.uint32),
        }
        d = N.ModelDump(*(out[k].ctypes.data for k in (
            "reserved", "content_kind", "internal_bits", "children_kind", "children",
            "children_bitmap", "brick_kind", "brick_value", "voxels", "colors", "data",
            "mip_kind", "mip_value", "mip_voxels")))
        _check(lib.svo_octree_export(self._h, C.byref(s), C.byref(d)))
        out["pool_len"] = n
        out["parted_count"] = s.parted_count
        out["color_count"] = s.color_count
        out["data_count"] = s.data_count
        out["mip_parted_count"] = s.mip_parted_count
        out["mips_enabled"] = s.mips_enabled
        return out

    def flatten(self) -> "FlatTree":
        h = C.c_void_p()
        _check(N.lib().svo_octree_flatten(self._h, C.byref(h)))
        return FlatTree(h.value, self)


NODE_DTYPE = np.dtype([("occ", "<u8"), ("meta", "<u4"), ("mip", "<u4"),
                       ("child", "<u4", (8,)), ("mip_word", "<u4"), ("reserved", "<u4", (3,))])


class FlatTree:
    """Flattened device layout of an Octree, still on the host."""

    def __init__(self, handle: int, owner: Octree):
        self._h = handle
        self._owner = owner

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _release("svo_flat_free", h)

    def desc(self) -> N.TreeDesc:
        d = N.TreeDesc()
        _check(N.lib().svo_flat_desc(self._h, C.byref(d)))
        return d

    def arrays(self) -> dict:
        d = self.desc()
        n3 = d.brick_dim ** 3
        words = (n3 + 31) // 32

        def view(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dtype).copy()

        return {
            "octree_size": d.octree_size, "brick_dim": d.brick_dim,
            "top_level_nodes": d.top_level_nodes,
            "nodes": view(d.nodes, d.node_count, NODE_DTYPE),
            "brick_pix": view(d.brick_pix, d.brick_count * n3, np.uint32),
            "brick_occ": view(d.brick_occ, d.brick_count * words, np.uint32),
            "colors": view(d.colors, d.color_count, np.uint32),
            "mips_enabled": int(d.mips_enabled),
        }

    def sync(self, octree: Octree) -> None:
        """Applies the octree's edits since the last sync in place."""
        _check(N.lib().svo_flat_sync(self._h, octree.handle))

    def get(self, pos) -> OctreeEntry:
        e = N.Entry()
        _check(N.lib().svo_flat_get(self._h, self._owner.handle, pos[0], pos[1], pos[2],
                                    C.byref(e)))
        return OctreeEntry.from_c(e)


@dataclass
class Viewport:
    """Viewport (src/raytracing/bevy/types.rs:55-71)."""

    origin: tuple
    direction: tuple
    frustum: tuple  # glass width, glass height, depth
    fov: float  # distance of the glass from the eye

    def to_c(self) -> N.Viewport:
        v = N.Viewport()
        for i in range(3):
            v.origin[i] = float(self.origin[i])
            v.direction[i] = float(self.direction[i])
            v.frustum[i] = float(self.frustum[i])
        v.fov = float(self.fov)
        return v


HIT_FIELDS = ("hit", "pix", "albedo", "voxel", "distance", "impact", "normal", "flags")


def _torch():
    import torch  # device memory and streams only

    return torch


def alloc_hits(n: int, device="cuda", fields=HIT_FIELDS) -> dict:
    """Device output buffers for n rays."""
    torch = _torch()
    shapes = {"hit": (n,), "pix": (n,), "albedo": (n,), "voxel": (n, 3), "distance": (n,),
              "impact": (n, 3), "normal": (n, 3), "flags": (n,)}
    dtypes = {"hit": torch.int32, "pix": torch.int32, "albedo": torch.int32,
              "voxel": torch.int32, "distance": torch.float32, "impact": torch.float32,
              "normal": torch.float32, "flags": torch.int32}
    return {k: torch.empty(shapes[k], dtype=dtypes[k], device=device) for k in fields}


_FIELD_WIDTH = {"hit": 1, "pix": 1, "albedo": 1, "voxel": 3, "distance": 1, "impact": 3,
                "normal": 3, "flags": 1}


def _check_device_tensor(t, name: str, dtype, elems: int, device: bool = True) -> None:
    """Raises ValueError unless `t` is a contiguous device tensor of `dtype`
    holding at least `elems` elements: the kernels write through its raw
    address and know nothing of its size. (device=False leaves out the
    device test, for callers that make it in a second pass.)"""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if t.numel() < elems:
        raise ValueError(f"{name}: {t.numel()} elements, at least {elems} needed")
    if device and t.device.type != "cuda":
        raise ValueError(f"{name}: tensor is on {t.device}, not a HIP device")


def _hits_struct(bufs: dict, n: int) -> N.Hits:
    """Output pointers for n rays, after checking every buffer: int32 for hit,
    pix, albedo, voxel and flags, float32 for distance, impact and normal,
    contiguous, on a HIP device, with n (x3 for voxel, impact and normal)
    elements at least."""
    torch = _torch()
    unknown = set(bufs) - set(HIT_FIELDS)
    if unknown:
        raise ValueError(f"unknown output fields {sorted(unknown)}")
    h = N.Hits()
    for k in HIT_FIELDS:
        t = bufs.get(k)
        if t is not None:
            dt = torch.float32 if k in ("distance", "impact", "normal") else torch.int32
            _check_device_tensor(t, k, dt, n * _FIELD_WIDTH[k], device=False)
    for k in HIT_FIELDS:
        t = bufs.get(k)
        if t is not None and t.device.type != "cuda":
            raise ValueError(f"{k}: tensor is on {t.device}, not a HIP device")
        setattr(h, k, t.data_ptr() if t is not None else None)
    return h


def _check_rays(rays, n: int):
    """A contiguous (>= n) x 6 float32 device tensor of origins and directions."""
    torch = _torch()
    _check_device_tensor(rays, "rays", torch.float32, 6 * n, device=False)
    if rays.dim() != 2 or rays.shape[1] != 6:
        raise ValueError(f"rays: shape {tuple(rays.shape)}, expected (n, 6)")
    if rays.device.type != "cuda":
        raise ValueError(f"rays: tensor is on {rays.device}, not a HIP device")
    return rays


def _stream_ptr(stream) -> int | None:
    if stream is None:
        torch = _torch()
        return torch.cuda.current_stream().cuda_stream or None
    return getattr(stream, "cuda_stream", stream) or None


class DeviceTree:
    """A tree resident in HBM; answers get_by_ray for batches of rays."""

    @classmethod
    def load(cls, path: str, keep_host: bool = False):
        """A saved octree file straight into device memory (svo_tree_load).
        With keep_host, returns (DeviceTree, Octree) so that edits can follow."""
        h, t = C.c_void_p(), C.c_void_p()
        _check(N.lib().svo_tree_load(str(path).encode(), C.byref(t) if keep_host else None,
                                     C.byref(h)))
        d = cls.__new__(cls)
        d._h = h.value
        return (d, Octree(0, 0, _handle=t.value)) if keep_host else d

    def __init__(self, octree: Octree | None = None, desc: N.TreeDesc | None = None):
        lib = N.lib()
        h = C.c_void_p()
        if octree is not None:
            _check(lib.svo_tree_from_octree(octree.handle, C.byref(h)))
        elif desc is not None:
            _check(lib.svo_tree_create(C.byref(desc), C.byref(h)))
        else:
            raise ValueError("need an octree or a tree description")
        self._h = h.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _release("svo_tree_destroy", h)

    def set_step_limit(self, steps: int) -> None:
        """Iteration bound per ray (0: default); rays beyond it end as misses
        with flag bit 0 set (svo_tree_set_step_limit)."""
        _check(N.lib().svo_tree_set_step_limit(self._h, int(steps)))

    @property
    def device_bytes(self) -> int:
        return int(N.lib().svo_tree_device_bytes(self._h))

    def update(self, octree: Octree, stream=None) -> tuple:
        """Re-syncs the device copy with the host octree (OctreeGPUView::reload):
        uploads only the records and bricks that changed; returns their counts."""
        _check(N.lib().svo_tree_update(self._h, octree.handle, _stream_ptr(stream)))
        r, s = C.c_uint32(), C.c_uint32()
        _check(N.lib().svo_tree_update_stats(self._h, C.byref(r), C.byref(s)))
        return int(r.value), int(s.value)

    def cast_into(self, rays, n: int, bufs: dict, stream=None) -> None:
        _check_rays(rays, n)
        hs = _hits_struct(bufs, n)
        _check(N.lib().svo_cast(self._h, rays.data_ptr(), n, C.byref(hs), _stream_ptr(stream)))

    def cast(self, rays, fields=HIT_FIELDS, stream=None) -> dict:
        """get_by_ray for every row of `rays` (n x 6 f32 device tensor)."""
        torch = _torch()
        rays = rays.to(dtype=torch.float32).contiguous()
        n = rays.shape[0]
        bufs = alloc_hits(n, rays.device, fields)
        self.cast_into(rays, n, bufs, stream)
        return bufs

    def render_into(self, vp: Viewport, w: int, h: int, bufs: dict, stream=None) -> None:
        hs = _hits_struct(bufs, w * h)
        v = vp.to_c()
        _check(N.lib().svo_render(self._h, C.byref(v), w, h, C.byref(hs), _stream_ptr(stream)))

    def render_rgba8(self, vp: Viewport, w: int, h: int, convention: int = 0, out=None,
                     stream=None):
        """Shaded image as an (h, w) int32 tensor of RGBA8 pixels (r in the low
        byte), row 0 at the top. convention 0: the reference CPU example's
        camera and shading; 1: the WGSL viewport kernel's."""
        return _render_rgba8(self, vp, w, h, convention, out, stream)

    def render(self, vp: Viewport, w: int, h: int, fields=HIT_FIELDS, device="cuda",
               stream=None) -> dict:
        bufs = alloc_hits(w * h, device, fields)
        self.render_into(vp, w, h, bufs, stream)
        return bufs

    def cast_at_lod(self, rays, viewing_distance: float, fields=HIT_FIELDS,
                    stream=None) -> dict:
        """get_by_ray_at_lod (raytracing_on_cpu.rs:325-565) for every row of
        `rays` (n x 6 f32 device tensor)."""
        torch = _torch()
        rays = rays.to(dtype=torch.float32).contiguous()
        n = rays.shape[0]
        bufs = alloc_hits(n, rays.device, fields)
        hs = _hits_struct(bufs, n)
        _check(N.lib().svo_cast_at_lod(self._h, rays.data_ptr(), n, float(viewing_distance),
                                       C.byref(hs), _stream_ptr(stream)))
        return bufs

    def render_at_lod(self, vp: Viewport, w: int, h: int, viewing_distance: float,
                      fields=HIT_FIELDS, device="cuda", stream=None) -> dict:
        """get_by_ray_at_lod for the primary rays of a viewport."""
        bufs = alloc_hits(w * h, device, fields)
        hs = _hits_struct(bufs, w * h)
        v = vp.to_c()
        _check(N.lib().svo_render_at_lod(self._h, C.byref(v), w, h, float(viewing_distance),
                                         C.byref(hs), _stream_ptr(stream)))
        return bufs


def _render_rgba8(tree: "DeviceTree", vp: "Viewport", w: int, h: int, convention: int,
                  out=None, stream=None):
    torch = _torch()
    if out is None:
        out = torch.empty((h, w), dtype=torch.int32, device="cuda")
    _check_device_tensor(out, "out", torch.int32, w * h)
    v = vp.to_c()
    _check(N.lib().svo_render_rgba8(tree._h, C.byref(v), w, h, convention, out.data_ptr(),
                                    _stream_ptr(stream)))
    return out


class Batch:
    """Batched mode: many (tree, viewport | rays) jobs in one persistent launch
    captured in a hipGraph (svo_batch_*)."""

    def __init__(self, max_jobs: int):
        h = C.c_void_p()
        _check(N.lib().svo_batch_create(max_jobs, C.byref(h)))
        self._h = h.value
        self.max_jobs = max_jobs
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            _release("svo_batch_destroy", h)

    @staticmethod
    def render_job(tree: "DeviceTree", vp: "Viewport", w: int, h: int, bufs: dict,
                   viewing_distance: float = 0.0) -> N.Job:
        """A viewport render; with viewing_distance > 0, get_by_ray_at_lod.
        The job holds references to the tree and the output tensors."""
        j = N.Job(tree._h, 0, w, h, vp.to_c(), None, _hits_struct(bufs, w * h),
                  float(viewing_distance))
        j._refs = (tree, dict(bufs))
        return j

    @staticmethod
    def cast_job(tree: "DeviceTree", rays, bufs: dict, viewing_distance: float = 0.0) -> N.Job:
        """A ray-array cast; with viewing_distance > 0, get_by_ray_at_lod.
        The job holds references to the tree, the rays and the outputs."""
        n = int(rays.shape[0])
        _check_rays(rays, n)
        j = N.Job(tree._h, 1, n, 1, N.Viewport(), rays.data_ptr(), _hits_struct(bufs, n),
                  float(viewing_distance))
        j._refs = (tree, rays, dict(bufs))
        return j

    def run(self, jobs: list, stream=None) -> None:
        """Launches the jobs asynchronously on `stream`. The jobs (and through
        them their trees and tensors) stay referenced until the next run, so
        the memory a launch in flight writes is not released under it."""
        if len(jobs) > self.max_jobs:
            raise ValueError(f"{len(jobs)} jobs, the batch was created for {self.max_jobs}")
        arr = (N.Job * max(1, len(jobs)))(*jobs)
        _check(N.lib().svo_batch_run(self._h, arr, len(jobs), _stream_ptr(stream)))
        self._keep = list(jobs)

    @property
    def tiles(self) -> int:
        return int(N.lib().svo_batch_tiles(self._h))


def batch_variants() -> list:
    """Names of the batched-launch kernel variants, in index order."""
    L = N.lib()
    return [L.svo_batch_variant_name(i).decode() for i in range(L.svo_batch_variant_count())]


def select_batch_variant(index: int) -> None:
    """Kernel variant of batch contexts created from now on."""
    _check(N.lib().svo_select_batch_variant(index))


def current_batch_variant() -> int:
    return int(N.lib().svo_current_batch_variant())


def render_variants() -> list:
    """Names of the render kernel's launch variants, in index order."""
    L = N.lib()
    return [L.svo_render_variant_name(i).decode() for i in range(L.svo_render_variant_count())]


def select_render_variant(index: int) -> None:
    _check(N.lib().svo_select_render_variant(index))


def current_render_variant() -> int:
    return int(N.lib().svo_current_render_variant())


def render_variant_scratch(index: int) -> int:
    """Scratch bytes per lane of a render variant's build (non-zero: spills)."""
    v = C.c_uint32(0)
    _check(N.lib().svo_render_variant_scratch(index, C.byref(v)))
    return int(v.value)


def batch_variant_scratch(index: int) -> int:
    """Scratch bytes per lane of a batched variant's build (non-zero: spills)."""
    v = C.c_uint32(0)
    _check(N.lib().svo_batch_variant_scratch(index, C.byref(v)))
    return int(v.value)


def set_timing(enabled: bool) -> None:
    N.lib().svo_set_timing(int(enabled))


def set_cast_order_threshold(n: int) -> None:
    """svo_cast orders its rays (origin cell, direction octant) from n rays up
    (default 2^20; 0 never). Process-wide."""
    _check(N.lib().svo_cast_set_order_threshold(int(n)))


def collect_kernel_ms(cap: int = 4096) -> list:
    buf = (C.c_float * cap)()
    n = N.lib().svo_collect_kernel_ms(buf, cap)
    return [buf[i] for i in range(n)]
