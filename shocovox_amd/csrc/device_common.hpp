// Device-side helpers shared by the single-launch and batched kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/svo.h"
#include "traverse.hpp"

namespace svo {

struct HitPtrs {
  uint32_t* hit;
  uint32_t* pix;
  uint32_t* albedo;
  uint32_t* voxel;
  float* distance;
  float* impact;
  float* normal;
  uint32_t* flags;
};

inline HitPtrs to_ptrs(const svo_hits& h) {
  return HitPtrs{h.hit, h.pix, h.albedo, h.voxel, h.distance, h.impact, h.normal, h.flags};
}

// Result slot i of the output arrays; 64-bit, since 3 * i overflows 32 bits
// from i = 2^32 / 3 on.
__device__ inline void store_result(const TreeView& t, const HitPtrs& o, uint64_t i,
                                    const CastResult& r, V3f org) {
  if (o.hit) o.hit[i] = r.hit;
  if (o.pix) o.pix[i] = r.hit ? r.pix : EMPTY_MARKER;
  if (o.albedo) {
    uint32_t a = 0u;
    if (r.hit) {
      const uint32_t c = pix_color_index(r.pix);
      if (c < NONE16 && c < t.color_count) a = t.colors[c];
    }
    o.albedo[i] = a;
  }
  if (o.voxel) {
    uint32_t vx = EMPTY_MARKER, vy = EMPTY_MARKER, vz = EMPTY_MARKER;
    if (r.hit) {
      vx = hit_voxel_axis(r.impact.x, r.cell.min.x, r.cell.size);
      vy = hit_voxel_axis(r.impact.y, r.cell.min.y, r.cell.size);
      vz = hit_voxel_axis(r.impact.z, r.cell.min.z, r.cell.size);
    }
    o.voxel[3 * i + 0] = vx;
    o.voxel[3 * i + 1] = vy;
    o.voxel[3 * i + 2] = vz;
  }
  if (o.distance) o.distance[i] = r.hit ? length(r.impact - org) : __builtin_inff();
  if (o.impact) {
    o.impact[3 * i + 0] = r.hit ? r.impact.x : 0.0f;
    o.impact[3 * i + 1] = r.hit ? r.impact.y : 0.0f;
    o.impact[3 * i + 2] = r.hit ? r.impact.z : 0.0f;
  }
  if (o.normal) {
    const V3f nrm = hit_normal(r);
    o.normal[3 * i + 0] = nrm.x;
    o.normal[3 * i + 1] = nrm.y;
    o.normal[3 * i + 2] = nrm.z;
  }
  if (o.flags) o.flags[i] = r.flags;
}

}  // namespace svo
