// gfx950 kernels of the traversal path and device residency of a tree.
//
// cast_kernel: one ray per lane from a caller-provided ray array.
// render_kernel: primary rays generated on the device; each 64-lane wave
//   owns an 8x8 pixel tile so neighbouring lanes follow nearby paths through
//   the tree and share node and brick cache lines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "device_common.hpp"
#include "flatten.hpp"
#include "refill.hpp"

namespace svo {

namespace {

// `order` (optional): lane i traverses ray order[i] (svo_cast's traversal
// order for large ray sets, order.hip); each ray writes its own outputs.
__global__ __launch_bounds__(256) void cast_kernel(TreeView t, const float* __restrict__ rays,
                                                   uint32_t n, const uint32_t* __restrict__ order,
                                                   HitPtrs out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = order ? order[k] : k;
  const float* r = rays + 6ull * i;
  const V3f o{r[0], r[1], r[2]};
  const V3f d{r[3], r[4], r[5]};
  const CastResult res = cast_ray_ww<true>(t, t.ray_lut, o, d, static_cast<const NodeRec*>(nullptr),
                                           0u);
  store_result(t, out, i, res, o);
}

// get_by_ray_at_lod for a ray array (MIP bricks probed per the viewing
// distance when the tree's MIPs are enabled).
__global__ __launch_bounds__(256) void cast_lod_kernel(TreeView t, const float* __restrict__ rays,
                                                       uint32_t n, float viewing_distance,
                                                       const uint32_t* __restrict__ order,
                                                       HitPtrs out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = order ? order[k] : k;  // as in cast_kernel
  const float* r = rays + 6ull * i;
  const V3f o{r[0], r[1], r[2]};
  const V3f d{r[3], r[4], r[5]};
  const CastResult res = cast_ray_lod<true>(t, t.ray_lut, o, d, viewing_distance);
  store_result(t, out, i, res, o);
}

// get_by_ray_at_lod for the primary rays of a viewport: four 64-lane waves
// per workgroup, each over one 8x8 pixel tile.
__global__ __launch_bounds__(256) void render_lod_kernel(TreeView t, Camera cam, uint32_t w,
                                                         uint32_t h, float viewing_distance,
                                                         HitPtrs out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t x = (blockIdx.x * 2u + (wave & 1u)) * 8u + (lane % 8u);
  const uint32_t y = (blockIdx.y * 2u + (wave >> 1)) * 8u + (lane / 8u);
  if (x >= w || y >= h) return;
  const V3f d = camera_ray_dir(cam, x, y);
  const CastResult res = cast_ray_lod<true>(t, t.ray_lut, cam.origin, d, viewing_distance);
  store_result(t, out, y * w + x, res, cam.origin);
}

// Primary-ray render. A workgroup is four 64-lane waves arranged 2x2 (G1:
// a single wave); each wave covers a TW x TH pixel tile (TW * TH = 64). A
// workgroup's slot is released when its last wave ends, so with four waves
// the fastest three wait for the slowest. MIN_WAVES asks the
// compiler for at least that many waves per SIMD (it caps VGPRs to fit);
// LDS_LUT stages the 4 KiB RAY_TO_NODE_OCCUPANCY_BITMASK_LUT in LDS.
// NODES > 0 stages the first NODES node records (the top levels of the
// breadth-first numbering) in LDS. XCD remaps the workgroup order so that
// each of the 8 XCDs (which receive workgroups round-robin in dispatch
// order) renders one contiguous band of the image, keeping that band's nodes
// and bricks in its own L2. WW runs the while-while arrangement of the
// traversal (cast_ray_ww), which probes the due bricks of a wave together.
template <int MIN_WAVES, bool LDS_LUT, int TW, int TH, int NODES = 0, bool XCD = false,
          bool WW = false, bool G1 = false>
__global__ __launch_bounds__(G1 ? 64 : 256) __attribute__((amdgpu_waves_per_eu(MIN_WAVES)))
void render_kernel(TreeView t, Camera cam, uint32_t w, uint32_t h, HitPtrs out) {
  static_assert(TW * TH == 64, "a wave covers 64 pixels");
  constexpr uint32_t kThreads = G1 ? 64u : 256u;
  __shared__ uint64_t s_lut[LDS_LUT ? 512 : 1];
  __shared__ NodeRec s_nodes[NODES > 0 ? NODES : 1];
  const uint64_t* lut = t.ray_lut;
  if constexpr (LDS_LUT) {
    for (uint32_t i = threadIdx.x; i < 512; i += kThreads) s_lut[i] = t.ray_lut[i];
    lut = s_lut;
  }
  uint32_t ntop = 0;
  if constexpr (NODES > 0) {
    ntop = min(static_cast<uint32_t>(NODES), t.node_count);
    const uint4* src = reinterpret_cast<const uint4*>(t.nodes);
    uint4* dst = reinterpret_cast<uint4*>(s_nodes);
    for (uint32_t i = threadIdx.x; i < ntop * 4u; i += kThreads) dst[i] = src[i];
  }
  if constexpr (LDS_LUT || NODES > 0) __syncthreads();
  uint32_t bx = blockIdx.x, by = blockIdx.y;
  if constexpr (XCD) {
    const uint32_t groups = gridDim.x * gridDim.y;
    const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t xcd = b % 8u, k = b / 8u, q = groups / 8u, rem = groups % 8u;
    const uint32_t logical = xcd * q + min(xcd, rem) + k;
    bx = logical % gridDim.x;
    by = logical / gridDim.x;
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t gx = G1 ? 1u : 2u;
  const uint32_t x = (bx * gx + (wave & 1u)) * TW + (lane % TW);
  const uint32_t y = (by * gx + (wave >> 1)) * TH + (lane / TW);
  if (x >= w || y >= h) return;
  const V3f d = camera_ray_dir(cam, x, y);
  CastResult res;
  if constexpr (WW)
    res = cast_ray_ww<true>(t, lut, cam.origin, d, s_nodes, ntop);
  else
    res = cast_ray<true>(t, lut, cam.origin, d, s_nodes, ntop);
  store_result(t, out, y * w + x, res, cam.origin);
}

// Primary-ray render with lane refill (refill.hpp). Each 64-lane wave owns
// TPW consecutive 8x8 tiles of the image; finished lanes take the next
// pixels of that range while the others keep stepping.
template <int MIN_WAVES, int TPW, int REFILL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIN_WAVES)))
void render_refill_kernel(TreeView t, Camera cam, uint32_t w, uint32_t h, HitPtrs out) {
  const uint32_t tiles_x = (w + 7u) / 8u;
  const uint32_t total_tiles = tiles_x * ((h + 7u) / 8u);
  const uint32_t first_tile = (blockIdx.x * 4u + (threadIdx.x >> 6)) * TPW;
  if (first_tile >= total_tiles) return;
  const RenderSource src{cam, w, h, tiles_x, 0u, 0u};
  refill_loop<REFILL>(t, out, first_tile * 64u, min(total_tiles, first_tile + TPW) * 64u, src);
}

// Shaded render to RGBA8 (r in the low byte), output row 0 at the top.
// CONV 0 follows the reference's CPU example (examples/cpu_render.rs:78-137):
//   ray from the eye through the glass point bottom_left + right * x *
//   pixel_w + up * y * pixel_h (pixel corners), pixel (x, y) written to row
//   h - 1 - y; a hit is shaded albedo * (1 - (normal . L / 2 + 0.5)) with
//   L = normalize(0, -1, 1), each channel (c as f32 * s) as u8; a miss is
//   grey 128.
// CONV 1 follows the WGSL update() kernel (viewport_render.wgsl:1028-1089):
//   the ray starts at the glass point bottom_left + right * frustum.x *
//   (x / W8) + up * frustum.y * (1 - row / H8), W8 and H8 being the image
//   size rounded up to the 8x8 workgroup grid, and points away from the eye;
//   a hit is albedo * (normal . (-0.5, 0.5, -0.5) / 2 + 0.5), a miss grey
//   0.25; channels are stored as unorm8 floor(clamp(v, 0, 1) * 255 + 0.5).
struct ShadeParams {
  V3f origin, bottom_left, right, up, light;
  float pixel_w, pixel_h, glass_w, glass_h;
  uint32_t w8, h8;
};

__device__ inline uint32_t sat_u8(float v) {
  // Rust `as u8`: NaN -> 0, saturating, truncating toward zero
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return static_cast<uint32_t>(v);
}

__device__ inline uint32_t unorm8(float v) {
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return static_cast<uint32_t>(floorf(v * 255.0f + 0.5f));
}

template <int CONV>
__global__ __launch_bounds__(256) void shade_kernel(TreeView t, ShadeParams sp, uint32_t w,
                                                    uint32_t h, uint32_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t x = (blockIdx.x * 2u + (wave & 1u)) * 8u + (lane & 7u);
  const uint32_t row = (blockIdx.y * 2u + (wave >> 1)) * 8u + (lane >> 3);
  if (x >= w || row >= h) return;
  V3f o, d;
  if (CONV == 0) {
    const uint32_t y = h - 1u - row;
    const V3f glass = (sp.bottom_left + (sp.right * static_cast<float>(x)) * sp.pixel_w) +
                      (sp.up * static_cast<float>(y)) * sp.pixel_h;
    o = sp.origin;
    d = normalized(glass - sp.origin);
  } else {
    const V3f glass =
        (sp.bottom_left +
         (sp.right * sp.glass_w) * (static_cast<float>(x) / static_cast<float>(sp.w8))) +
        (sp.up * sp.glass_h) * (1.0f - static_cast<float>(row) / static_cast<float>(sp.h8));
    o = glass;
    d = normalized(glass - sp.origin);
  }
  const CastResult r =
      cast_ray_ww<true>(t, t.ray_lut, o, d, static_cast<const NodeRec*>(nullptr), 0u);
  uint32_t rgba;
  if (r.hit) {
    const uint32_t ci = pix_color_index(r.pix);
    const uint32_t c = (ci < NONE16 && ci < t.color_count) ? t.colors[ci] : 0u;
    const uint32_t cr = c >> 24, cg = (c >> 16) & 0xFFu, cb = (c >> 8) & 0xFFu;
    const V3f nrm = hit_normal(r);
    const float dot = (nrm.x * sp.light.x + nrm.y * sp.light.y) + nrm.z * sp.light.z;
    if (CONV == 0) {
      const float strength = 1.0f - (dot / 2.0f + 0.5f);
      rgba = sat_u8(static_cast<float>(cr) * strength) |
             (sat_u8(static_cast<float>(cg) * strength) << 8) |
             (sat_u8(static_cast<float>(cb) * strength) << 16) | 0xFF000000u;
    } else {
      const float k = dot / 2.0f + 0.5f;
      rgba = unorm8((static_cast<float>(cr) / 255.0f) * k) |
             (unorm8((static_cast<float>(cg) / 255.0f) * k) << 8) |
             (unorm8((static_cast<float>(cb) / 255.0f) * k) << 16) | 0xFF000000u;
    }
  } else {
    rgba = CONV == 0 ? 0xFF808080u : (unorm8(0.25f) * 0x010101u) | 0xFF000000u;
  }
  out[row * w + x] = rgba;
}

using RenderFn = void (*)(TreeView, Camera, uint32_t, uint32_t, HitPtrs);

struct RenderVariant {
  const char* name;
  RenderFn fn;
  uint32_t tile_w, tile_h;
  uint32_t tiles_per_wave;  // refill kernels: 8x8 tiles per wave (0: one tile per wave)
  uint32_t single_wave_groups;  // 1: one-wave workgroups (G1), 0: 2x2 waves
};

const RenderVariant kRenderVariants[] = {
    {"w1-glut-8x8", render_kernel<1, false, 8, 8>, 8, 8},
    {"w5-glut-8x8", render_kernel<5, false, 8, 8>, 8, 8},
    {"w6-glut-8x8", render_kernel<6, false, 8, 8>, 8, 8},
    {"w8-glut-8x8", render_kernel<8, false, 8, 8>, 8, 8},
    {"w1-slut-8x8", render_kernel<1, true, 8, 8>, 8, 8},
    {"w6-slut-8x8", render_kernel<6, true, 8, 8>, 8, 8},
    {"w1-glut-16x4", render_kernel<1, false, 16, 4>, 16, 4},
    {"w1-glut-4x16", render_kernel<1, false, 4, 16>, 4, 16},
    {"w6-slut-16x4", render_kernel<6, true, 16, 4>, 16, 4},
    {"w5-glut-8x8-xcd", render_kernel<5, false, 8, 8, 0, true>, 8, 8},
    {"w1-glut-8x8-xcd", render_kernel<1, false, 8, 8, 0, true>, 8, 8},
    {"w5-glut-8x8-n73", render_kernel<5, false, 8, 8, 73>, 8, 8},
    {"w1-glut-8x8-n73", render_kernel<1, false, 8, 8, 73>, 8, 8},
    {"w1-glut-8x8-n585", render_kernel<1, false, 8, 8, 585>, 8, 8},
    {"w5-glut-8x8-n73-xcd", render_kernel<5, false, 8, 8, 73, true>, 8, 8},
    {"w5-slut-8x8-n73", render_kernel<5, true, 8, 8, 73>, 8, 8},
    {"w1-slut-8x8-n73", render_kernel<1, true, 8, 8, 73>, 8, 8},
    {"w6-glut-8x8-n73", render_kernel<6, false, 8, 8, 73>, 8, 8},
    {"w6-slut-8x8-n73", render_kernel<6, true, 8, 8, 73>, 8, 8},
    {"w6-glut-8x8-xcd", render_kernel<6, false, 8, 8, 0, true>, 8, 8},
    {"w1-glut-8x8-ww", render_kernel<1, false, 8, 8, 0, false, true>, 8, 8},
    {"w5-glut-8x8-ww", render_kernel<5, false, 8, 8, 0, false, true>, 8, 8},
    {"w6-slut-8x8-ww", render_kernel<6, true, 8, 8, 0, false, true>, 8, 8},
    {"w5-slut-8x8-n73-ww", render_kernel<5, true, 8, 8, 73, false, true>, 8, 8},
    {"w6-glut-8x8-ww-g1", render_kernel<6, false, 8, 8, 0, false, true, true>, 8, 8, 0, 1},
    {"w6-slut-8x8-ww-g1", render_kernel<6, true, 8, 8, 0, false, true, true>, 8, 8, 0, 1},
    {"w5-glut-8x8-ww-g1", render_kernel<5, false, 8, 8, 0, false, true, true>, 8, 8, 0, 1},
    {"w6-glut-8x8-g1", render_kernel<6, false, 8, 8, 0, false, false, true>, 8, 8, 0, 1},
    // lane refill measured 3.5x slower than one tile per wave on coherent
    // 1080p primary rays (lanes drift apart and lose the shared node lines);
    // kept as a variant for divergent workloads
    {"r1-t4-f16", render_refill_kernel<1, 4, 16>, 8, 8, 4},
};
constexpr int kRenderVariantCount = sizeof(kRenderVariants) / sizeof(kRenderVariants[0]);
int initial_render_variant() {
  for (int i = 0; i < kRenderVariantCount; ++i)
    if (std::strcmp(kRenderVariants[i].name, "w6-slut-8x8-ww") == 0) return i;
  return 0;
}
int g_render_variant = initial_render_variant();

// Scatter of patched records: src[i] -> dst[idx[i]] (16-byte quarters per lane)
__global__ __launch_bounds__(256) void scatter_records_kernel(NodeRec* __restrict__ dst,
                                                              const NodeRec* __restrict__ src,
                                                              const uint32_t* __restrict__ idx,
                                                              uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4u * n) return;
  const uint32_t r = i >> 2, q = i & 3u;
  const uint4* s4 = reinterpret_cast<const uint4*>(src + r);
  uint4* d4 = reinterpret_cast<uint4*>(dst + idx[r]);
  d4[q] = s4[q];
}

// Fill of patched brick slots: fills[3k] = slot, fills[3k+1] = pix value,
// fills[3k+2] = occupancy value.
__global__ __launch_bounds__(256) void fill_slots_kernel(uint32_t* __restrict__ pix,
                                                         uint32_t* __restrict__ occ,
                                                         const uint32_t* __restrict__ fills,
                                                         uint32_t n, uint32_t n3, uint32_t words) {
  const uint32_t per = n3 + words;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(per) * n) return;
  const uint32_t f = static_cast<uint32_t>(i / per), w = static_cast<uint32_t>(i % per);
  const uint64_t slot = fills[3 * f];
  if (w < n3)
    pix[slot * n3 + w] = fills[3 * f + 1];
  else
    occ[slot * words + (w - n3)] = fills[3 * f + 2];
}

// Word writes into patched brick slots: {slot, word, value} triples, the
// first n_pix into pix words, the next n_occ into occupancy words.
__global__ __launch_bounds__(256) void scatter_words_kernel(uint32_t* __restrict__ pix,
                                                            uint32_t* __restrict__ occ,
                                                            const uint32_t* __restrict__ writes,
                                                            uint32_t n_pix, uint32_t n_occ,
                                                            uint32_t n3, uint32_t words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pix + n_occ) return;
  const uint64_t slot = writes[3 * i];
  const uint32_t word = writes[3 * i + 1], value = writes[3 * i + 2];
  if (i < n_pix)
    pix[slot * n3 + word] = value;
  else
    occ[slot * words + word] = value;
}

// Scatter of patched bricks: words [n3 + occ_words] per slot, packed in src.
__global__ __launch_bounds__(256) void scatter_bricks_kernel(
    uint32_t* __restrict__ pix, uint32_t* __restrict__ occ, const uint32_t* __restrict__ src,
    const uint32_t* __restrict__ slots, uint32_t n, uint32_t n3, uint32_t words) {
  const uint32_t per = n3 + words;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(per) * n) return;
  const uint32_t b = static_cast<uint32_t>(i / per), w = static_cast<uint32_t>(i % per);
  const uint64_t slot = slots[b];
  if (w < n3)
    pix[slot * n3 + w] = src[i];
  else
    occ[slot * words + (w - n3)] = src[i];
}

// --- timing ---------------------------------------------------------------

struct EventPair {
  hipEvent_t start, stop;
};
bool g_timing = false;
std::mutex g_timing_mu;
std::vector<EventPair> g_events;

struct LaunchTimer {
  EventPair ev{nullptr, nullptr};
  hipStream_t s;
  bool on;
  explicit LaunchTimer(hipStream_t st) : s(st), on(g_timing) {
    if (on && hipEventCreate(&ev.start) == hipSuccess && hipEventCreate(&ev.stop) == hipSuccess)
      (void)hipEventRecord(ev.start, s);
    else
      on = false;
  }
  void done() {
    if (!on) return;
    (void)hipEventRecord(ev.stop, s);
    std::lock_guard<std::mutex> g(g_timing_mu);
    g_events.push_back(ev);
  }
};

int hip_fail(hipError_t e, const char* what, std::string* err) {
  *err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SVO_OUT_OF_DEVICE_MEMORY : SVO_HIP_ERROR;
}

// The render kernels index pixels y * w + x in 32 bits.
bool image_fits(uint32_t w, uint32_t h, std::string* err) {
  if (static_cast<uint64_t>(w) * h < (uint64_t(1) << 32)) return true;
  *err = "image of 2^32 or more pixels (w * h must be < 2^32)";
  return false;
}

// Checks that every index the kernels follow stays inside the arrays.
bool validate_desc(const svo_tree_desc& d, std::string* err) {
  if (d.node_count == 0 || d.nodes == nullptr) {
    *err = "tree has no nodes";
    return false;
  }
  // exact_pow2_sizes is what lets every kernel run the traversal's P2 form
  if (!exact_pow2_sizes(d.octree_size, d.brick_dim) || d.octree_size < 2 * d.brick_dim) {
    *err = "octree size and brick dimension must be powers of two, size >= 2 * dim, size <= 2^24";
    return false;
  }
  if (d.brick_count && (d.brick_pix == nullptr || d.brick_occ == nullptr)) {
    *err = "brick arrays missing";
    return false;
  }
  if (static_cast<uint64_t>(d.brick_count) * occ_words_per_brick(d.brick_dim) * 4u >
      OCC_RANGE_MAX) {
    *err = "occupancy sidecar beyond the 4 GiB the traversal's 32-bit offsets address";
    return false;
  }
  for (uint32_t i = 0; i < d.node_count; ++i) {
    const svo_node_rec& r = d.nodes[i];
    const uint32_t kind = r.meta & 3u;
    if (r.mip > FORM_PARTED || (r.mip == FORM_PARTED && r.mip_word >= d.brick_count)) {
      *err = "MIP brick reference out of range";
      return false;
    }
    if (kind == KIND_INTERNAL) {
      for (uint32_t o = 0; o < 8; ++o)
        if (r.child[o] != INVALID_INDEX && r.child[o] >= d.node_count) {
          *err = "child index out of range";
          return false;
        }
    } else if (kind == KIND_LEAF || kind == KIND_UNIFORM_LEAF) {
      const uint32_t n = kind == KIND_LEAF ? 8u : 1u;
      for (uint32_t o = 0; o < n; ++o) {
        const uint32_t form = brick_form(r.meta, o);
        if (form == 3u || (form == FORM_PARTED && r.child[o] >= d.brick_count)) {
          *err = "brick reference out of range";
          return false;
        }
      }
    }
  }
  return true;
}

}  // namespace

void timing_enable(bool on) { g_timing = on; }

int render_variant_count() { return kRenderVariantCount; }
const char* render_variant_name(int i) {
  return (i >= 0 && i < kRenderVariantCount) ? kRenderVariants[i].name : nullptr;
}
bool render_variant_select(int i) {
  if (i < 0 || i >= kRenderVariantCount) return false;
  g_render_variant = i;
  return true;
}
int render_variant_current() { return g_render_variant; }
int render_variant_scratch(int i, uint32_t* bytes, std::string* err) {
  if (i < 0 || i >= kRenderVariantCount) {
    *err = "no such variant";
    return SVO_INVALID_ARGUMENT;
  }
  hipFuncAttributes a{};
  const hipError_t e =
      hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kRenderVariants[i].fn));
  if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes", err);
  *bytes = static_cast<uint32_t>(a.localSizeBytes);
  return SVO_OK;
}

uint32_t timing_collect(float* out_ms, uint32_t cap) {
  std::lock_guard<std::mutex> g(g_timing_mu);
  uint32_t n = 0;
  for (const EventPair& e : g_events) {
    float ms = 0.0f;
    if (hipEventSynchronize(e.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, e.start, e.stop) == hipSuccess && n < cap)
      out_ms[n++] = ms;
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  g_events.clear();
  return n;
}

namespace {
// Capacity an upload allocates for `need` elements: a quarter of headroom so
// that edits which add nodes or bricks are patched in place.
uint64_t grown_capacity(uint64_t need) { return need + need / 4 + 1; }
// Slots of an occupancy allocation: `want`, or fewer (but at least `need`)
// so that the allocation stays within the traversal's 32-bit read offsets.
uint64_t occ_slots(uint64_t want, uint64_t need, uint64_t slot_bytes) {
  return std::max(need, std::min(want, OCC_RANGE_MAX / slot_bytes));
}
// Palettes grow with almost every edit that brings a new colour, and an entry
// costs 4 bytes, so they get a fixed allowance on top.
uint64_t grown_color_capacity(uint64_t need) { return grown_capacity(need) + 4096; }
// A tree that outgrew its arrays is being edited: it grows by half again.
uint64_t regrown_capacity(uint64_t need) { return need + need / 2 + 64; }
}  // namespace

uint64_t device_tree_bytes_for(const svo_tree_desc& d) {
  const uint64_t n3 = static_cast<uint64_t>(d.brick_dim) * d.brick_dim * d.brick_dim;
  const uint64_t words = occ_words_per_brick(d.brick_dim);
  const uint64_t bricks = grown_capacity(d.brick_count ? d.brick_count : 1);
  return grown_capacity(d.node_count) * sizeof(NodeRec) + bricks * (n3 + words) * sizeof(uint32_t) +
         grown_color_capacity(d.color_count ? d.color_count : 1) * sizeof(uint32_t) +
         64 * 8 * sizeof(uint64_t);
}

int device_tree_upload(DeviceTree* t, const svo_tree_desc& d, void* stream, std::string* err) {
  if (!validate_desc(d, err)) return SVO_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(t->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice", err);
  const uint64_t n3 = static_cast<uint64_t>(d.brick_dim) * d.brick_dim * d.brick_dim;
  const uint64_t words = occ_words_per_brick(d.brick_dim);
  auto grow = [&](void** p, uint64_t* cap, uint64_t need, uint64_t elem) -> int {
    if (need <= *cap && *p) return SVO_OK;
    uint64_t nc = p == reinterpret_cast<void**>(&t->colors) ? grown_color_capacity(need)
                                                             : grown_capacity(need);
    void* np = nullptr;
    hipError_t ge = hipMalloc(&np, nc * elem);
    if (ge != hipSuccess) return hip_fail(ge, "hipMalloc", err);
    if (*p) {
      // keep the old buffer alive until queued work on it has finished
      (void)hipStreamSynchronize(s);
      (void)hipFree(*p);
    }
    *p = np;
    *cap = nc;
    return SVO_OK;
  };
  int st = grow(reinterpret_cast<void**>(&t->nodes), &t->node_capacity, d.node_count,
                sizeof(NodeRec));
  if (st != SVO_OK) return st;
  // The brick arrays are compared in bytes, not slots: the description may
  // carry another brick dimension than the resident tree (a slot of 16^3
  // values is 64 times one of 4^3), and both arrays must hold its bricks.
  const uint64_t slots = d.brick_count ? d.brick_count : 1;
  const uint64_t pix_need = slots * n3 * sizeof(uint32_t);
  const uint64_t occ_need = slots * words * sizeof(uint32_t);
  auto grow_bytes = [&](uint32_t** p, uint64_t* cap_bytes, uint64_t need,
                        uint64_t slot_bytes) -> int {
    if (need <= *cap_bytes && *p) return SVO_OK;
    const uint64_t nb = (p == &t->brick_occ ? occ_slots(grown_capacity(slots), slots, slot_bytes)
                                            : grown_capacity(slots)) *
                        slot_bytes;
    void* np = nullptr;
    hipError_t ge = hipMalloc(&np, nb);
    if (ge != hipSuccess) return hip_fail(ge, "hipMalloc(bricks)", err);
    if (*p) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(*p);
    }
    *p = static_cast<uint32_t*>(np);
    *cap_bytes = nb;
    return SVO_OK;
  };
  st = grow_bytes(&t->brick_pix, &t->pix_bytes, pix_need, n3 * sizeof(uint32_t));
  if (st != SVO_OK) return st;
  st = grow_bytes(&t->brick_occ, &t->occ_bytes, occ_need, words * sizeof(uint32_t));
  if (st != SVO_OK) return st;
  t->brick_capacity = std::min(t->pix_bytes / (n3 * sizeof(uint32_t)),
                               t->occ_bytes / (words * sizeof(uint32_t)));
  st = grow(reinterpret_cast<void**>(&t->colors), &t->color_capacity,
            d.color_count ? d.color_count : 1, sizeof(uint32_t));
  if (st != SVO_OK) return st;

  e = hipMemcpyAsync(t->nodes, d.nodes, static_cast<size_t>(d.node_count) * sizeof(NodeRec),
                     hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(nodes)", err);
  if (d.brick_count) {
    e = hipMemcpyAsync(t->brick_pix, d.brick_pix, d.brick_count * n3 * sizeof(uint32_t),
                       hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(bricks)", err);
    e = hipMemcpyAsync(t->brick_occ, d.brick_occ, d.brick_count * words * sizeof(uint32_t),
                       hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(occupancy)", err);
  }
  if (d.color_count) {
    e = hipMemcpyAsync(t->colors, d.colors, d.color_count * sizeof(uint32_t),
                       hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(colors)", err);
  }
  // Host buffers of the description may be released after this call.
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize", err);

  t->node_count = d.node_count;
  t->brick_count = d.brick_count;
  t->color_count = d.color_count;
  t->top_level_nodes = d.top_level_nodes;
  t->bytes = t->node_capacity * sizeof(NodeRec) + t->pix_bytes + t->occ_bytes +
             t->color_capacity * sizeof(uint32_t) + 64 * 8 * sizeof(uint64_t);
  t->view.nodes = t->nodes;
  t->view.brick_pix = t->brick_pix;
  t->view.brick_occ = t->brick_occ;
  t->view.colors = t->colors;
  t->view.ray_lut = t->ray_lut;
  t->view.size = static_cast<float>(d.octree_size);
  t->view.dim = d.brick_dim;
  t->view.occ_words = static_cast<uint32_t>(words);
  t->view.color_count = d.color_count;
  t->view.inv_dim = pow2_rcp(static_cast<float>(d.brick_dim));
  t->view.node_count = d.node_count;
  t->view.max_steps = t->step_limit ? t->step_limit : MAX_TRAVERSAL_STEPS;
  t->view.mips_enabled = d.mips_enabled;
  t->view.occ_range = static_cast<uint32_t>(t->occ_bytes);
  return SVO_OK;
}

int device_tree_create(const svo_tree_desc& d, DeviceTree** out, std::string* err) {
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    *err = "no HIP device available";
    return SVO_NO_DEVICE;
  }
  DeviceTree* t = new DeviceTree();
  e = hipGetDevice(&t->device);
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "hipGetDevice", err);
  }
  e = hipMalloc(reinterpret_cast<void**>(&t->ray_lut), 64 * 8 * sizeof(uint64_t));
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "hipMalloc(lut)", err);
  }
  uint64_t lut[64 * 8];
  generate_ray_occupancy_lut(lut);
  e = hipMemcpy(t->ray_lut, lut, sizeof(lut), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    device_tree_destroy(t);
    return hip_fail(e, "hipMemcpy(lut)", err);
  }
  const int st = device_tree_upload(t, d, nullptr, err);
  if (st != SVO_OK) {
    device_tree_destroy(t);
    return st;
  }
  *out = t;
  return SVO_OK;
}

namespace {

// Moves one device array into a larger allocation: the live prefix is copied
// device to device on `s`; the old buffer is retired (kernels already queued
// may still read it). Retired memory beyond the tree's own size is returned
// to the device first, which waits for the device once.
int regrow(DeviceTree* t, void** p, uint64_t* cap, uint64_t new_cap, uint64_t live_bytes,
           uint64_t elem, hipStream_t s, std::string* err) {
  if (t->retired_bytes > std::max<uint64_t>(t->bytes, uint64_t(64) << 20)) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize", err);
    for (void* q : t->retired) (void)hipFree(q);
    t->retired.clear();
    t->retired_bytes = 0;
  }
  void* np = nullptr;
  hipError_t e = hipMalloc(&np, new_cap * elem);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(grow)", err);
  if (*p && live_bytes) {
    e = hipMemcpyAsync(np, *p, live_bytes, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
      (void)hipFree(np);
      return hip_fail(e, "hipMemcpyAsync(grow)", err);
    }
  }
  if (*p) {
    t->retired.push_back(*p);
    t->retired_bytes += *cap * elem;
  }
  *p = np;
  *cap = new_cap;
  return SVO_OK;
}

// Grows the arrays of a resident tree to hold `d`, keeping their contents.
int grow_resident(DeviceTree* t, const svo_tree_desc& d, hipStream_t s, std::string* err) {
  const uint64_t n3 = static_cast<uint64_t>(d.brick_dim) * d.brick_dim * d.brick_dim;
  const uint64_t words = occ_words_per_brick(d.brick_dim);
  int st = SVO_OK;
  if (d.node_count > t->node_capacity)
    st = regrow(t, reinterpret_cast<void**>(&t->nodes), &t->node_capacity,
                regrown_capacity(d.node_count), t->node_count * sizeof(NodeRec), sizeof(NodeRec), s,
                err);
  if (st == SVO_OK && d.brick_count > t->brick_capacity) {
    // (same brick dimension as the resident tree: device_tree_patch sends a
    // change of dimension to device_tree_upload)
    const uint64_t nc = regrown_capacity(d.brick_count);
    uint64_t pix_cap = t->pix_bytes / (n3 * 4), occ_cap = t->occ_bytes / (words * 4);
    st = regrow(t, reinterpret_cast<void**>(&t->brick_pix), &pix_cap, nc,
                t->brick_count * n3 * 4, n3 * 4, s, err);
    if (st == SVO_OK) t->pix_bytes = pix_cap * n3 * 4;
    if (st == SVO_OK)
      st = regrow(t, reinterpret_cast<void**>(&t->brick_occ), &occ_cap,
                  occ_slots(nc, d.brick_count, words * 4), t->brick_count * words * 4, words * 4, s,
                  err);
    if (st == SVO_OK) t->occ_bytes = occ_cap * words * 4;
    t->brick_capacity = std::min(t->pix_bytes / (n3 * 4), t->occ_bytes / (words * 4));
  }
  if (st == SVO_OK && d.color_count > t->color_capacity)
    st = regrow(t, reinterpret_cast<void**>(&t->colors), &t->color_capacity,
                grown_color_capacity(d.color_count), t->color_count * 4ull, 4, s, err);
  if (st != SVO_OK) return st;
  t->bytes = t->node_capacity * sizeof(NodeRec) + t->pix_bytes + t->occ_bytes +
             t->color_capacity * sizeof(uint32_t) + 64 * 8 * sizeof(uint64_t);
  t->view.nodes = t->nodes;
  t->view.brick_pix = t->brick_pix;
  t->view.brick_occ = t->brick_occ;
  t->view.colors = t->colors;
  t->view.occ_range = static_cast<uint32_t>(t->occ_bytes);
  return SVO_OK;
}

}  // namespace

namespace {

// Makes staging set `ps` hold `bytes`; the set is not in use (its last patch
// has finished).
int reserve_set(DeviceTree* t, DeviceTree::PatchStaging& ps, size_t bytes, std::string* err) {
  hipError_t e;
  ps.tree = t;
  if (bytes <= ps.dev_bytes && bytes <= ps.host_bytes) return SVO_OK;
  const size_t cap = std::max<size_t>(bytes + bytes / 2, size_t(8) << 20);
  if (bytes > ps.dev_bytes) {
    // a smaller buffer is retired, not freed: freeing waits for the whole
    // device; growth is geometric, so few buffers are ever retired
    void* grown = nullptr;
    e = hipMalloc(&grown, cap);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(staging)", err);
    if (ps.dev) {
      t->retired.push_back(ps.dev);
      t->retired_bytes += ps.dev_bytes;
    }
    ps.dev = grown;
    ps.dev_bytes = cap;
  }
  if (bytes > ps.host_bytes) {
    if (ps.host) (void)hipHostFree(ps.host);  // the set's last copy has finished
    ps.host = nullptr;
    ps.host_bytes = 0;
    e = hipHostMalloc(&ps.host, cap, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(staging)", err);
    ps.host_bytes = cap;
  }
  return SVO_OK;
}

// Waits (on the host) until staging set `ps` is no longer read by a patch.
int drain_set(DeviceTree* t, DeviceTree::PatchStaging& ps, std::string*) {
  std::unique_lock<std::mutex> g(t->staging_mu);
  t->staging_cv.wait(g, [&] { return !ps.pending; });
  return SVO_OK;
}

// Host function enqueued after a patch's copies: its staging set is free.
void staging_released(void* arg) {
  DeviceTree::PatchStaging* ps = static_cast<DeviceTree::PatchStaging*>(arg);
  DeviceTree* t = ps->tree;
  {
    std::lock_guard<std::mutex> g(t->staging_mu);
    ps->pending = false;
  }
  t->staging_cv.notify_all();
}

}  // namespace

int device_tree_reserve_staging(DeviceTree* t, size_t bytes, std::string* err) {
  hipError_t e = hipSetDevice(t->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice", err);
  for (DeviceTree::PatchStaging& ps : t->staging) {
    if (bytes > ps.dev_bytes || bytes > ps.host_bytes) {
      const int st = drain_set(t, ps, err);
      if (st != SVO_OK) return st;
    }
    const int st = reserve_set(t, ps, bytes, err);
    if (st != SVO_OK) return st;
  }
  return SVO_OK;
}

int device_record_event(void* event, void* stream, std::string* err) {
  const hipError_t e =
      hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream));
  return e == hipSuccess ? SVO_OK : hip_fail(e, "hipEventRecord", err);
}

int device_tree_patch(DeviceTree* t, const svo_tree_desc& d, const std::vector<uint32_t>& records,
                      const BrickPatch& bp, uint32_t colors_from, void* stream, void* start_event,
                      void* const* end_events, uint32_t n_end, std::string* err) {
  if (d.brick_dim != t->view.dim) return 1;  // caller re-uploads everything
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(t->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice", err);
  const uint32_t n3 = d.brick_dim * d.brick_dim * d.brick_dim;
  const uint32_t words = occ_words_per_brick(d.brick_dim);
  if (static_cast<uint64_t>(d.brick_count) * words * 4u > OCC_RANGE_MAX) {
    *err = "occupancy sidecar beyond the 4 GiB the traversal's 32-bit offsets address";
    return SVO_INVALID_ARGUMENT;
  }
  for (uint32_t r : records)
    if (r >= d.node_count) {
      *err = "patched record out of range";
      return SVO_INVALID_ARGUMENT;
    }
  auto slots_ok = [&](const uint32_t* v, size_t n, size_t stride, uint32_t word_limit) {
    for (size_t i = 0; i < n; ++i)
      if (v[i * stride] >= d.brick_count || (word_limit && v[i * stride + 1] >= word_limit))
        return false;
    return true;
  };
  if (!slots_ok(bp.whole, bp.n_whole, 1, 0) || !slots_ok(bp.fills, bp.n_fills, 3, 0) ||
      !slots_ok(bp.pix_writes, bp.n_pix, 3, n3) || !slots_ok(bp.occ_writes, bp.n_occ, 3, words)) {
    *err = "patched brick slot or word out of range";
    return SVO_INVALID_ARGUMENT;
  }
  const uint64_t per_slot = static_cast<uint64_t>(n3) + words;
  if (bp.n_pix + bp.n_occ > 0x7FFFFFFFull || (bp.n_fills * per_slot + 255) / 256 > 0x7FFFFFFFull ||
      (bp.n_whole * per_slot + 255) / 256 > 0x7FFFFFFFull) {
    *err = "patch too large";
    return SVO_INVALID_ARGUMENT;
  }
  const uint32_t new_colors = d.color_count > colors_from ? d.color_count - colors_from : 0u;
  // Staging layout (16-byte aligned sections), identical on host and device:
  // [records][record indices][whole slots' words][whole slot indices]
  // [fills][pix writes][occupancy writes][new colours].
  // The host side is pinned memory, so the whole patch is one asynchronous
  // DMA on the caller's stream: a copy from pageable memory, or a hipFree,
  // would wait for other streams' work and stall behind busy launches.
  auto align16 = [](size_t v) { return (v + 15) & ~size_t(15); };
  const size_t o_ridx = align16(records.size() * sizeof(NodeRec));
  const size_t o_bw = align16(o_ridx + records.size() * 4);
  const size_t o_slots = align16(o_bw + bp.n_whole * static_cast<size_t>(n3 + words) * 4);
  const size_t o_fills = align16(o_slots + bp.n_whole * 4);
  const size_t o_writes = align16(o_fills + bp.n_fills * 12);
  const size_t o_colors = align16(o_writes + (bp.n_pix + bp.n_occ) * 12);
  const size_t bytes = align16(o_colors + new_colors * 4u);
  // the set the patch two calls back used: its copy and scatter must have
  // finished before the host refills it
  DeviceTree::PatchStaging& ps = t->staging[t->next_staging];
  int sst = drain_set(t, ps, err);
  if (sst == SVO_OK) sst = reserve_set(t, ps, bytes, err);
  if (sst != SVO_OK) return sst;
  char* h = static_cast<char*>(ps.host);
  for (size_t i = 0; i < records.size(); ++i)
    std::memcpy(h + i * sizeof(NodeRec), &d.nodes[records[i]], sizeof(NodeRec));
  if (!records.empty()) std::memcpy(h + o_ridx, records.data(), records.size() * 4);
  for (size_t i = 0; i < bp.n_whole; ++i) {
    uint32_t* dst = reinterpret_cast<uint32_t*>(h + o_bw) + i * (n3 + words);
    std::memcpy(dst, d.brick_pix + static_cast<size_t>(bp.whole[i]) * n3, n3 * sizeof(uint32_t));
    std::memcpy(dst + n3, d.brick_occ + static_cast<size_t>(bp.whole[i]) * words,
                words * sizeof(uint32_t));
  }
  if (bp.n_whole) std::memcpy(h + o_slots, bp.whole, bp.n_whole * 4);
  if (bp.n_fills) std::memcpy(h + o_fills, bp.fills, bp.n_fills * 12);
  if (bp.n_pix) std::memcpy(h + o_writes, bp.pix_writes, bp.n_pix * 12);
  if (bp.n_occ) std::memcpy(h + o_writes + bp.n_pix * 12, bp.occ_writes, bp.n_occ * 12);
  if (new_colors) std::memcpy(h + o_colors, d.colors + colors_from, new_colors * 4u);
  // host work done: the device part starts here
  if (start_event) {
    e = hipEventRecord(static_cast<hipEvent_t>(start_event), s);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord(patch start)", err);
  }
  if (d.node_count > t->node_capacity || d.brick_count > t->brick_capacity ||
      d.color_count > t->color_capacity) {
    const int st = grow_resident(t, d, s, err);
    if (st != SVO_OK) return st;
  }
  char* base = static_cast<char*>(ps.dev);
  if (bytes > 0) {
    e = hipMemcpyAsync(base, h, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(patch)", err);
  }
  if (!records.empty()) {
    const uint32_t n = static_cast<uint32_t>(records.size());
    hipLaunchKernelGGL(scatter_records_kernel, dim3((4 * n + 255) / 256), dim3(256), 0, s,
                       t->nodes, reinterpret_cast<const NodeRec*>(base),
                       reinterpret_cast<const uint32_t*>(base + o_ridx), n);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "scatter_records_kernel", err);
  }
  if (bp.n_fills) {
    const uint64_t total = static_cast<uint64_t>(bp.n_fills) * (n3 + words);
    hipLaunchKernelGGL(fill_slots_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)),
                       dim3(256), 0, s, t->brick_pix, t->brick_occ,
                       reinterpret_cast<const uint32_t*>(base + o_fills),
                       static_cast<uint32_t>(bp.n_fills), n3, words);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "fill_slots_kernel", err);
  }
  if (bp.n_whole) {
    const uint64_t total = static_cast<uint64_t>(bp.n_whole) * (n3 + words);
    hipLaunchKernelGGL(scatter_bricks_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)),
                       dim3(256), 0, s, t->brick_pix, t->brick_occ,
                       reinterpret_cast<const uint32_t*>(base + o_bw),
                       reinterpret_cast<const uint32_t*>(base + o_slots),
                       static_cast<uint32_t>(bp.n_whole), n3, words);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "scatter_bricks_kernel", err);
  }
  if (bp.n_pix + bp.n_occ) {
    const uint32_t n = static_cast<uint32_t>(bp.n_pix + bp.n_occ);
    hipLaunchKernelGGL(scatter_words_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       t->brick_pix, t->brick_occ,
                       reinterpret_cast<const uint32_t*>(base + o_writes),
                       static_cast<uint32_t>(bp.n_pix), static_cast<uint32_t>(bp.n_occ), n3,
                       words);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "scatter_words_kernel", err);
  }
  if (new_colors) {
    e = hipMemcpyAsync(t->colors + colors_from, base + o_colors, new_colors * 4u,
                       hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(colors)", err);
  }
  for (uint32_t i = 0; i < n_end; ++i) {
    e = hipEventRecord(static_cast<hipEvent_t>(end_events[i]), s);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord(patch end)", err);
  }
  // no wait here: the set stays pending until the stream has run the copies
  // and then staging_released
  {
    std::lock_guard<std::mutex> g(t->staging_mu);
    ps.pending = true;
  }
  e = hipLaunchHostFunc(s, staging_released, &ps);
  if (e != hipSuccess) {
    // the copies are queued but nothing will report their end: wait here
    (void)hipStreamSynchronize(s);
    {
      std::lock_guard<std::mutex> g(t->staging_mu);
      ps.pending = false;
    }
    return hip_fail(e, "hipLaunchHostFunc(staging)", err);
  }
  t->next_staging ^= 1u;
  t->node_count = d.node_count;
  t->brick_count = d.brick_count;
  t->color_count = d.color_count;
  t->view.color_count = d.color_count;
  t->view.node_count = d.node_count;
  t->view.mips_enabled = d.mips_enabled;
  return SVO_OK;
}

void device_tree_destroy(DeviceTree* t) {
  if (!t) return;
  if (t->device >= 0) (void)hipSetDevice(t->device);
  (void)hipDeviceSynchronize();
  {
    // the host functions of the last patches may still be on their way
    // (bounded: a stream in error may never call them)
    std::unique_lock<std::mutex> g(t->staging_mu);
    (void)t->staging_cv.wait_for(g, std::chrono::seconds(5), [&] {
      return !t->staging[0].pending && !t->staging[1].pending;
    });
  }
  for (DeviceTree::PatchStaging& ps : t->staging) {
    if (ps.host) (void)hipHostFree(ps.host);
    if (ps.dev) (void)hipFree(ps.dev);
  }
  for (void* p : t->retired) (void)hipFree(p);
  if (t->nodes) (void)hipFree(t->nodes);
  if (t->brick_pix) (void)hipFree(t->brick_pix);
  if (t->brick_occ) (void)hipFree(t->brick_occ);
  if (t->colors) (void)hipFree(t->colors);
  if (t->ray_lut) (void)hipFree(t->ray_lut);
  delete t;
}

int device_cast(const DeviceTree& t, const float* rays, uint32_t n, const svo_hits& out,
                void* stream, std::string* err) {
  if (n == 0) return SVO_OK;
  if (rays == nullptr) {
    *err = "rays is NULL";
    return SVO_INVALID_ARGUMENT;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  LaunchTimer timer(s);
  uint32_t* order = nullptr;
  const uint32_t threshold = cast_order_threshold();
  if (threshold != 0 && n >= threshold) {
    // without device memory for the sort the rays keep their own order
    const int st = order_rays(rays, n, t.view.size, &order, s, err);
    if (st != SVO_OK && st != SVO_OUT_OF_DEVICE_MEMORY) return st;
  }
  // n < 2^32, so the grid (at most 2^24 workgroups) is computed without
  // wrapping in 64 bits
  const uint32_t block = 256;
  const uint32_t grid = static_cast<uint32_t>((static_cast<uint64_t>(n) + block - 1) / block);
  hipLaunchKernelGGL(cast_kernel, dim3(grid), dim3(block), 0, s, t.view, rays, n, order,
                     to_ptrs(out));
  hipError_t e = hipGetLastError();
  if (order) (void)hipFreeAsync(order, s);
  if (e != hipSuccess) return hip_fail(e, "cast_kernel launch", err);
  timer.done();
  return SVO_OK;
}

namespace {
std::atomic<uint32_t> g_cast_order_threshold{1u << 20};
}  // namespace

void set_cast_order_threshold(uint32_t n) { g_cast_order_threshold.store(n); }
uint32_t cast_order_threshold() { return g_cast_order_threshold.load(); }

int device_cast_lod(const DeviceTree& t, const float* rays, uint32_t n, float viewing_distance,
                    const svo_hits& out, void* stream, std::string* err) {
  if (n == 0) return SVO_OK;
  if (rays == nullptr) {
    *err = "rays is NULL";
    return SVO_INVALID_ARGUMENT;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  LaunchTimer timer(s);
  uint32_t* order = nullptr;
  const uint32_t threshold = cast_order_threshold();
  if (threshold != 0 && n >= threshold) {
    const int st = order_rays(rays, n, t.view.size, &order, s, err);
    if (st != SVO_OK && st != SVO_OUT_OF_DEVICE_MEMORY) return st;
  }
  const uint32_t grid = static_cast<uint32_t>((static_cast<uint64_t>(n) + 255u) / 256u);
  hipLaunchKernelGGL(cast_lod_kernel, dim3(grid), dim3(256), 0, s, t.view, rays, n,
                     viewing_distance, order, to_ptrs(out));
  hipError_t e = hipGetLastError();
  if (order) (void)hipFreeAsync(order, s);
  if (e != hipSuccess) return hip_fail(e, "cast_lod_kernel launch", err);
  timer.done();
  return SVO_OK;
}

int device_render_lod(const DeviceTree& t, const Camera& cam, uint32_t w, uint32_t h,
                      float viewing_distance, const svo_hits& out, void* stream,
                      std::string* err) {
  if (w == 0 || h == 0) return SVO_OK;
  if (!image_fits(w, h, err)) return SVO_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  LaunchTimer timer(s);
  const dim3 grid((w + 15u) / 16u, (h + 15u) / 16u);
  hipLaunchKernelGGL(render_lod_kernel, grid, dim3(256), 0, s, t.view, cam, w, h,
                     viewing_distance, to_ptrs(out));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render_lod_kernel launch", err);
  timer.done();
  return SVO_OK;
}

int device_render(const DeviceTree& t, const Camera& cam, uint32_t w, uint32_t h,
                  const svo_hits& out, void* stream, std::string* err) {
  if (w == 0 || h == 0) return SVO_OK;
  if (!image_fits(w, h, err)) return SVO_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  LaunchTimer timer(s);
  const RenderVariant& v = kRenderVariants[g_render_variant];
  const uint32_t gx = v.single_wave_groups ? 1u : 2u;
  dim3 grid((w + gx * v.tile_w - 1) / (gx * v.tile_w), (h + gx * v.tile_h - 1) / (gx * v.tile_h));
  if (v.tiles_per_wave) {
    // four waves per workgroup, each over tiles_per_wave consecutive 8x8 tiles
    const uint64_t tiles = static_cast<uint64_t>((w + 7) / 8) * ((h + 7) / 8);
    if (tiles * 64 >= (1ull << 32)) {
      *err = "image too large for the refill kernels";
      return SVO_INVALID_ARGUMENT;
    }
    const uint64_t waves = (tiles + v.tiles_per_wave - 1) / v.tiles_per_wave;
    grid = dim3(static_cast<uint32_t>((waves + 3) / 4));
  }
  hipLaunchKernelGGL(v.fn, grid, dim3(v.single_wave_groups ? 64 : 256), 0, s, t.view, cam, w, h, to_ptrs(out));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render_kernel launch", err);
  timer.done();
  return SVO_OK;
}

int device_render_rgba8(const DeviceTree& t, V3f origin, V3f dir, float glass_w, float glass_h,
                        float fov, uint32_t w, uint32_t h, uint32_t convention, uint32_t* out,
                        void* stream, std::string* err) {
  if (convention > 1) {
    *err = "convention must be 0 (CPU example) or 1 (WGSL viewport)";
    return SVO_INVALID_ARGUMENT;
  }
  if (w == 0 || h == 0) return SVO_OK;
  if (!image_fits(w, h, err)) return SVO_INVALID_ARGUMENT;
  if (!out) {
    *err = "output is NULL";
    return SVO_INVALID_ARGUMENT;
  }
  const V3f up{0.0f, 1.0f, 0.0f};
  ShadeParams sp;
  sp.origin = origin;
  sp.up = up;
  sp.right = normalized(cross(up, dir));
  sp.glass_w = glass_w;
  sp.glass_h = glass_h;
  sp.pixel_w = glass_w / static_cast<float>(w);
  sp.pixel_h = glass_h / static_cast<float>(h);
  sp.w8 = (w + 7u) / 8u * 8u;
  sp.h8 = (h + 7u) / 8u * 8u;
  if (convention == 0) {
    // viewport_bottom_left of cpu_render.rs:92-94, light of :97
    sp.bottom_left = ((origin + dir * fov) - up * (glass_h / 2.0f)) - sp.right * (glass_w / 2.0f);
    sp.light = normalized(V3f{0.0f, -1.0f, 1.0f});
  } else {
    // bottom left of viewport_render.wgsl:1033-1043, light of :1060
    sp.bottom_left = ((origin + dir * fov) - sp.right * (glass_w / 2.0f)) - up * (glass_h / 2.0f);
    sp.light = V3f{-0.5f, 0.5f, -0.5f};
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  LaunchTimer timer(s);
  const dim3 grid((w + 15) / 16, (h + 15) / 16);
  if (convention == 0)
    hipLaunchKernelGGL(shade_kernel<0>, grid, dim3(256), 0, s, t.view, sp, w, h, out);
  else
    hipLaunchKernelGGL(shade_kernel<1>, grid, dim3(256), 0, s, t.view, sp, w, h, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "shade_kernel launch", err);
  timer.done();
  return SVO_OK;
}

}  // namespace svo
