"""shocovox_amd: MI355X-native ray traversal of shocovox sparse voxel octrees.

The traversal hot path (Octree::get_by_ray) runs as hand-written HIP
kernels for gfx950 inside libsvo.so, behind the C ABI of include/svo.h.
"""
from .octree import (  # noqa: F401
    COMPLEX, EMPTY, EMPTY_MARKER, INFORMATIVE, VISUAL, Batch, DeviceTree, FlatTree,
    MIPResamplingMethods, Octree, StrategyUpdater,
    OctreeEntry, OctreeError, Viewport, albedo, albedo_rgba, alloc_hits, batch_variant_scratch,
    batch_variants, collect_kernel_ms, current_batch_variant, current_render_variant,
    render_variant_scratch, render_variants, select_batch_variant, select_render_variant,
    set_cast_order_threshold, set_timing, voxel_data,
)

__all__ = [
    "COMPLEX", "EMPTY", "EMPTY_MARKER", "INFORMATIVE", "VISUAL", "Batch", "DeviceTree", "FlatTree",
    "MIPResamplingMethods", "Octree", "StrategyUpdater", "OctreeEntry", "OctreeError", "Viewport", "albedo", "albedo_rgba", "alloc_hits",
    "batch_variant_scratch", "batch_variants", "collect_kernel_ms", "current_batch_variant", "current_render_variant",
    "render_variant_scratch", "render_variants", "select_batch_variant", "select_render_variant",
    "set_cast_order_threshold", "set_timing",
    "voxel_data",
]
