"""Argument checks that guard the kernels and the file reader, run on the CPU.

* svo_render / svo_render_at_lod / svo_render_rgba8 refuse images of 2^32
  or more pixels (the render kernels index pixels in 32 bits) before they
  touch a device;
* the Python binding refuses output and ray buffers of the wrong dtype,
  layout, length or device, since the kernels write through raw addresses;
* the bencode reader skips arbitrarily nested values without recursion: a
  file of a million nested lists is read (when balanced) or refused (when
  cut short) without exhausting the host stack.
"""
from __future__ import annotations

import ctypes as C

import pytest
import torch

from shocovox_amd import Octree, OctreeEntry, OctreeError, _native as N
from shocovox_amd.octree import _check_rays, _hits_struct


def _last_error() -> str:
    return N.lib().svo_last_error().decode()


@pytest.mark.parametrize("w,h", [(65536, 65536), (1 << 20, 1 << 12), (0xFFFFFFFF, 2)])
def test_render_entry_points_refuse_images_of_2_32_pixels(w, h):
    L = N.lib()
    vp = N.Viewport()
    hits = N.Hits()
    assert L.svo_render(None, C.byref(vp), w, h, C.byref(hits), None) == 5
    assert "2^32" in _last_error()
    assert L.svo_render_at_lod(None, C.byref(vp), w, h, 1.0, C.byref(hits), None) == 5
    assert "2^32" in _last_error()
    assert L.svo_render_rgba8(None, C.byref(vp), w, h, 0, None, None) == 5
    assert "2^32" in _last_error()


def test_render_entry_points_accept_the_largest_image_below_2_32():
    # 65535 x 65537 = 2^32 - 1 pixels passes the size check and stops at the
    # missing tree instead
    L = N.lib()
    vp = N.Viewport()
    hits = N.Hits()
    assert L.svo_render(None, C.byref(vp), 65535, 65537, C.byref(hits), None) == 5
    assert _last_error() == "NULL argument"


def _bufs(n, **over):
    b = {"hit": torch.zeros(n, dtype=torch.int32), "voxel": torch.zeros((n, 3), dtype=torch.int32),
         "distance": torch.zeros(n, dtype=torch.float32)}
    b.update(over)
    return b


def test_output_buffers_are_checked():
    n = 100
    with pytest.raises(ValueError, match="hit: dtype"):
        _hits_struct(_bufs(n, hit=torch.zeros(n, dtype=torch.int64)), n)
    with pytest.raises(ValueError, match="distance: dtype"):
        _hits_struct(_bufs(n, distance=torch.zeros(n, dtype=torch.int32)), n)
    with pytest.raises(ValueError, match="voxel: 297 elements"):
        _hits_struct(_bufs(n, voxel=torch.zeros((99, 3), dtype=torch.int32)), n)
    with pytest.raises(ValueError, match="voxel: tensor is not contiguous"):
        _hits_struct(_bufs(n, voxel=torch.zeros((3, n), dtype=torch.int32).t()), n)
    with pytest.raises(ValueError, match="unknown output fields"):
        _hits_struct(_bufs(n, colour=torch.zeros(n, dtype=torch.int32)), n)
    with pytest.raises(ValueError, match="hit: expected a torch tensor"):
        _hits_struct(_bufs(n, hit=[0] * n), n)
    # right dtype, layout and length, but host memory: refused as well
    with pytest.raises(ValueError, match="not a HIP device"):
        _hits_struct(_bufs(n), n)


def test_ray_buffers_are_checked():
    with pytest.raises(ValueError, match="rays: dtype"):
        _check_rays(torch.zeros((10, 6), dtype=torch.float64), 10)
    with pytest.raises(ValueError, match="rays: 54 elements"):
        _check_rays(torch.zeros((9, 6), dtype=torch.float32), 10)
    with pytest.raises(ValueError, match="shape"):
        _check_rays(torch.zeros((20, 3), dtype=torch.float32), 10)
    with pytest.raises(ValueError, match="not a HIP device"):
        _check_rays(torch.zeros((10, 6), dtype=torch.float32), 10)


def _tree_bytes_with_solid_brick() -> tuple[bytes, int]:
    t = Octree.new(8, 2)
    t.insert_at_lod((0, 0, 0), 2, OctreeEntry.visual(0x11223344))
    b = t.to_bytes()
    at = b.index(b"3:#b#")
    end = b.index(b"e", b.index(b"i", at)) + 1  # after the Solid value
    return b, end


def test_deeply_nested_value_in_a_brick_is_skipped():
    b, end = _tree_bytes_with_solid_brick()
    depth = 1_000_000
    crafted = b[:end] + b"l" * depth + b"e" * depth + b[end:]
    t = Octree.from_bytes(crafted)
    assert t.get((0, 0, 0)) == OctreeEntry.visual(0x11223344)


def test_unterminated_deep_nesting_is_an_error():
    b, end = _tree_bytes_with_solid_brick()
    crafted = b[:end] + b"l" * 1_000_000
    with pytest.raises(OctreeError, match="InvalidStructure"):
        Octree.from_bytes(crafted)
