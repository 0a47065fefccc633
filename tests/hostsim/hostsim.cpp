// Test-only build of the product's device traversal (traverse.hpp) for the
// host CPU, so the kernel's logic can be compared with the oracle on
// machines without a GPU. It is not part of libsvo and no product code path
// loads it.
#include <cstdint>

#include "../../shocovox_amd/csrc/traverse.hpp"

extern "C" {

struct hs_tree {
  const svo::NodeRec* nodes;
  const uint32_t* brick_pix;
  const uint32_t* brick_occ;
  const uint32_t* colors;
  const uint64_t* ray_lut;
  float size;
  uint32_t dim;
  uint32_t occ_words;
  uint32_t color_count;
  uint32_t max_steps;  // 0: the default bound
  uint32_t mips_enabled;
  float viewing_distance;  // mode 2
};

struct hs_hit {
  uint32_t hit, pix, albedo, flags;
  uint32_t voxel[3];
  float distance;
  float impact[3];
  float normal[3];
};

// mode 0: cast_ray; mode 1: the resumable form (ray_begin, then ray_step
// until the ray finishes), as the refill kernels run it; mode 2: cast_ray_lod
// with t->viewing_distance; mode 3: cast_ray_ww; mode 4: cast_ray_ww without
// the crawl shortcut. Modes 0, 2 and 3 use the P2 form, as the kernels
// do, when exact_pow2_sizes holds and `general` is 0, else the general form.
void hs_cast(const hs_tree* t, const float* rays, uint64_t n, hs_hit* out, int mode,
             int general) {
  const bool p2 = !general && svo::exact_pow2_sizes(static_cast<uint64_t>(t->size), t->dim);
  svo::TreeView v{t->nodes, t->brick_pix, t->brick_occ, t->colors, t->ray_lut, t->size, t->dim,
                  t->occ_words, t->color_count, svo::pow2_rcp(static_cast<float>(t->dim)), 0u,
                  t->max_steps ? t->max_steps : svo::MAX_TRAVERSAL_STEPS, t->mips_enabled};
  for (uint64_t i = 0; i < n; ++i) {
    const float* r = rays + 6 * i;
    const svo::V3f o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
    svo::CastResult c;
    if (mode == 0) {
      c = p2 ? svo::cast_ray<true>(v, o, d) : svo::cast_ray<false>(v, o, d);
    } else if (mode == 3) {
      c = p2 ? svo::cast_ray_ww<true>(v, v.ray_lut, o, d, nullptr, 0u)
             : svo::cast_ray_ww<false>(v, v.ray_lut, o, d, nullptr, 0u);
    } else if (mode == 4) {
      c = p2 ? svo::cast_ray_ww<true, uint64_t, false>(v, v.ray_lut, o, d, nullptr, 0u)
             : svo::cast_ray_ww<false, uint64_t, false>(v, v.ray_lut, o, d, nullptr, 0u);
    } else if (mode == 2) {
      c = p2 ? svo::cast_ray_lod<true>(v, v.ray_lut, o, d, t->viewing_distance)
             : svo::cast_ray_lod<false>(v, v.ray_lut, o, d, t->viewing_distance);
    } else {
      svo::RayState st;
      if (svo::ray_begin(v, o, d, st)) {
        while (!svo::ray_step(v, v.ray_lut, st, c)) {
        }
      } else {
        c = svo::miss_result(0u);
      }
    }
    hs_hit& h = out[i];
    h.hit = c.hit;
    h.flags = c.flags;
    h.pix = c.hit ? c.pix : 0xFFFFFFFFu;
    const uint32_t ci = c.hit ? (c.pix & 0xFFFFu) : 0xFFFFu;
    h.albedo = (c.hit && ci < 0xFFFFu && ci < t->color_count) ? t->colors[ci] : 0u;
    if (c.hit) {
      h.voxel[0] = svo::hit_voxel_axis(c.impact.x, c.cell.min.x, c.cell.size);
      h.voxel[1] = svo::hit_voxel_axis(c.impact.y, c.cell.min.y, c.cell.size);
      h.voxel[2] = svo::hit_voxel_axis(c.impact.z, c.cell.min.z, c.cell.size);
      h.distance = svo::length(c.impact - o);
    } else {
      h.voxel[0] = h.voxel[1] = h.voxel[2] = 0xFFFFFFFFu;
      h.distance = __builtin_inff();
    }
    h.impact[0] = c.hit ? c.impact.x : 0.0f;
    h.impact[1] = c.hit ? c.impact.y : 0.0f;
    h.impact[2] = c.hit ? c.impact.z : 0.0f;
    const svo::V3f nrm = svo::hit_normal(c);
    h.normal[0] = nrm.x;
    h.normal[1] = nrm.y;
    h.normal[2] = nrm.z;
  }
}

void hs_viewport_rays(const float* vp, uint32_t w, uint32_t h, float* rays) {
  using svo::V3f;
  const V3f origin{vp[0], vp[1], vp[2]};
  const svo::Camera cam =
      svo::make_camera(origin, V3f{vp[3], vp[4], vp[5]}, vp[6], vp[7], vp[9], w, h);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const V3f d = svo::camera_ray_dir(cam, x, y);
      float* r = rays + 6 * (static_cast<uint64_t>(y) * w + x);
      r[0] = origin.x;
      r[1] = origin.y;
      r[2] = origin.z;
      r[3] = d.x;
      r[4] = d.y;
      r[5] = d.z;
    }
}

}  // extern "C"
