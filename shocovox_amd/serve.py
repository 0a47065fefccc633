"""Served mode: a real-time request loop over device-resident trees.

Requests arrive as a seeded Poisson process; each picks a tree and a
resolution by Zipf; a share are sparse voxel edits applied in place. The
loop itself is native (svo_serve_run): continuous batching into short
batched launches (mode "batched") or one plain launch per request (mode
"plain"), with per-tree arrival order between edits and renders.
Trees are admitted by free device memory (admit()).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .octree import DeviceTree, Octree, Viewport, _check

# resolutions of the served workload, most popular first
SERVED_RESOLUTIONS = ((256, 256), (1280, 720), (1920, 1080), (3840, 2160))

# batched: continuous batching (a launch holding one whole request runs as a
# plain render); plain: one launch per request; batched-kernel: continuous
# batching with the persistent batched kernel for every launch
MODES = {"batched": 0, "plain": 1, "batched-kernel": 2}


@dataclass
class ServeConfig:
    qps: float
    duration_s: float
    seed: int = 1
    mode: str = "batched"
    max_inflight: int = 2
    max_batch_rays: int = 1 << 23
    max_batch_jobs: int = 1024
    max_queue: int = 0
    edit_fraction: float = 0.05
    max_edit_voxels: int = 1000
    zipf_s: float = 1.1
    resolutions: tuple = field(default=SERVED_RESOLUTIONS)
    deadline_ms: float = 0.0

    def to_c(self) -> tuple:
        res = (C.c_uint32 * (2 * len(self.resolutions)))(
            *[int(v) for wh in self.resolutions for v in wh])
        c = N.ServeConfig(float(self.qps), float(self.duration_s), int(self.seed),
                          MODES[self.mode], int(self.max_inflight), int(self.max_batch_rays),
                          int(self.max_batch_jobs), int(self.max_queue),
                          float(self.edit_fraction), int(self.max_edit_voxels),
                          float(self.zipf_s), len(self.resolutions), res,
                          float(self.deadline_ms))
        return c, res  # keep the resolution array alive with the struct


@dataclass
class Request:
    arrival_s: float
    kind: int  # 0 render, 1 edit
    tree: int
    width: int
    height: int
    viewport: Viewport
    edit_voxels: int
    edit_clear: int
    edit_seed: int

    @staticmethod
    def from_c(r: N.ServeRequest) -> "Request":
        v = r.viewport
        vp = Viewport(tuple(v.origin), tuple(v.direction), tuple(v.frustum), float(v.fov))
        return Request(float(r.arrival_s), int(r.kind), int(r.tree), int(r.width),
                       int(r.height), vp, int(r.edit_voxels), int(r.edit_clear),
                       int(r.edit_seed))

    def to_c(self) -> N.ServeRequest:
        return N.ServeRequest(self.arrival_s, self.kind, self.tree, self.width, self.height,
                              self.viewport.to_c(), self.edit_voxels, self.edit_clear,
                              self.edit_seed)


def plan(cfg: ServeConfig, tree_sizes) -> list:
    """The request sequence svo_serve_run follows for these trees."""
    c, _res = cfg.to_c()
    sizes = (C.c_uint32 * len(tree_sizes))(*[int(s) for s in tree_sizes])
    count = C.c_uint64()
    L = N.lib()
    _check(L.svo_serve_plan(C.byref(c), sizes, len(tree_sizes), None, 0, C.byref(count)))
    arr = (N.ServeRequest * max(1, count.value))()
    _check(L.svo_serve_plan(C.byref(c), sizes, len(tree_sizes), arr, count.value,
                            C.byref(count)))
    return [Request.from_c(arr[i]) for i in range(count.value)]


def apply_edit(tree: Octree, req: Request) -> None:
    """The host edit the loop applies for an edit request."""
    r = req.to_c()
    _check(N.lib().svo_serve_apply_edit(tree.handle, C.byref(r)))


def run(trees: list, dtrees: list, cfg: ServeConfig, digests: bool = False) -> dict:
    """Runs a serving session; returns the stats, per-request latencies (ms;
    -1 when rejected, -2 when timed out) and, if asked, per-request result
    digests."""
    if len(trees) != len(dtrees) or not trees:
        raise ValueError("need one device tree per host tree")
    c, _res = cfg.to_c()
    n = len(plan(cfg, [t.get_size() for t in trees]))
    th = (C.c_void_p * len(trees))(*[t.handle for t in trees])
    dh = (C.c_void_p * len(dtrees))(*[d._h for d in dtrees])
    stats = N.ServeStats()
    lat = np.zeros(max(1, n), np.float32)
    dig = np.zeros(max(1, n), np.uint64) if digests else None
    _check(N.lib().svo_serve_run(th, dh, len(trees), C.byref(c), C.byref(stats),
                                 lat.ctypes.data, dig.ctypes.data if digests else None, n))
    out = {k: getattr(stats, k) for k, _ in N.ServeStats._fields_}
    out["latencies_ms"] = lat[:n]
    if digests:
        out["digests"] = dig[:n]
    return out


def device_memory() -> tuple:
    """(free, total) bytes of device memory on the current device."""
    f, t = C.c_uint64(), C.c_uint64()
    _check(N.lib().svo_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def admit(tree: Octree, budget_bytes: int = 0, reserve_bytes: int = 0) -> DeviceTree:
    """Uploads the tree if its device bytes plus the reserve fit in the free
    device memory (or in budget_bytes when given); raises OctreeError with
    status SVO_OUT_OF_DEVICE_MEMORY otherwise, before touching the device."""
    h = C.c_void_p()
    _check(N.lib().svo_tree_admit(tree.handle, int(budget_bytes), int(reserve_bytes),
                                  C.byref(h)))
    d = DeviceTree.__new__(DeviceTree)
    d._h = h.value
    return d
