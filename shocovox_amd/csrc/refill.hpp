// Lane refill for the traversal kernels.
//
// A wave walks a range of ray ids [next, end_id). Every lane runs one ray
// with ray_step (one stack-loop iteration per call); whenever at least
// REFILL lanes (or all 64) are idle, those lanes take the next ids of the
// range in lane order and start new rays. Lanes stay busy until the range
// runs out, instead of idling until the slowest ray of a fixed group ends.
// The hand-out position is wave-uniform (it advances by the ballot count of
// idle lanes), so the loop needs no atomics; each pass either hands out ids
// or steps a ray that ends within the tree's iteration bound, so every wave
// leaves.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace svo {

// Pixels of a pinhole camera, in tile order: id = local_tile * 64 + k with
// k = (x & 7) + 8 * (y & 7), local_tile counted from first_tile. Outputs go
// to y * width + x, or to id when tile_major is set.
struct RenderSource {
  Camera cam;
  uint32_t width, height, tiles_x, first_tile, tile_major;
  __device__ bool ray(uint32_t id, V3f* o, V3f* d, uint32_t* out_index) const {
    const uint32_t tile = first_tile + (id >> 6), k = id & 63u;
    const uint32_t x = (tile % tiles_x) * 8u + (k & 7u);
    const uint32_t y = (tile / tiles_x) * 8u + (k >> 3);
    if (x >= width || y >= height) return false;
    *o = cam.origin;
    *d = camera_ray_dir(cam, x, y);
    *out_index = tile_major ? id : y * width + x;
    return true;
  }
  // the origin is the same for every ray, so lanes keep none
  __device__ V3f origin(V3f) const { return cam.origin; }
};

// Rows {origin, direction} of a ray array; id = row.
struct ArraySource {
  const float* rays;
  uint32_t count;
  __device__ bool ray(uint32_t id, V3f* o, V3f* d, uint32_t* out_index) const {
    if (id >= count) return false;
    const float* r = rays + 6ull * id;
    *o = V3f{r[0], r[1], r[2]};
    *d = V3f{r[3], r[4], r[5]};
    *out_index = id;
    return true;
  }
  __device__ V3f origin(V3f o) const { return o; }
};

template <int REFILL, class Source>
__device__ __attribute__((always_inline)) inline void refill_loop(const TreeView& t,
                                                                 const HitPtrs& out,
                                                                 uint32_t next, uint32_t end_id,
                                                                 const Source& src) {
  static_assert(REFILL >= 1 && REFILL <= 64, "refill threshold in 1..64");
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1ull;
  RayState st;
  bool active = false;
  uint32_t idx = 0;
  V3f o{0.0f, 0.0f, 0.0f};
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = static_cast<uint32_t>(__popcll(idle));
    if (next < end_id && (nidle >= static_cast<uint32_t>(REFILL) || nidle == 64u)) {
      if (!active) {
        const uint32_t id = next + static_cast<uint32_t>(__popcll(idle & below));
        V3f d;
        if (id < end_id && src.ray(id, &o, &d, &idx)) {
          if (ray_begin(t, o, d, st))
            active = true;
          else
            store_result(t, out, idx, miss_result(0u), src.origin(o));
        }
      }
      next += nidle;
    }
    if (__ballot(active) == 0ull) {
      if (next >= end_id) break;
      continue;
    }
    if (active) {
      CastResult r;
      if (ray_step(t, t.ray_lut, st, r)) {
        store_result(t, out, idx, r, src.origin(o));
        active = false;
      }
    }
  }
}

}  // namespace svo
