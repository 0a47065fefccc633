// Per-ray octree traversal over the device layout (layout.hpp).
//
// The control flow and every floating-point operation follow
// get_by_ray_at_lod with MIP maps disabled (raytracing_on_cpu.rs:325-565),
// probe_brick (:256-312) and traverse_brick (:156-252), so that one ray takes
// the same path through the tree, bit for bit, as the reference CPU code:
//   * a 4-entry ring stack whose push overwrites the oldest entry and whose
//     pop past the retained entries restarts from the root after a 0.1 nudge
//     along the ray (:19-82, :548-562);
//   * the bitmap position is clamped to [1e-5, 4-1e-5] once per stack-loop
//     iteration and then advanced by step*4/size without a clamp (:425-436,
//     :510) - both are kept;
//   * brick voxels are tested through the occupancy sidecar, which holds
//     !pix_points_to_empty() for every voxel (node.rs:405-427).
// Two guards exist that the reference lacks: an iteration bound (the
// reference loop has no termination proof) and a bitmap-index range check
// (the reference would index its LUT out of bounds and panic). Either sets a
// flag on the ray; on well-formed inputs both stay clear.
//
// What differs from the reference is only how the same operations are laid
// out for 64-lane waves: the restart loop and the stack loop are one loop;
// POP and ADVANCE lanes share their first DDA step; the ring stack is a
// shift register; a miss carries only its flags; cast_ray_ww regroups a
// wave's brick probes; and trees whose sizes are powers of two (all device
// trees) use the exact power-of-two forms of the size arithmetic (P2).
#pragma once

#include "layout.hpp"
#include "svo_math.hpp"

namespace svo {

constexpr uint32_t FLAG_ITERATION_GUARD = 1u;
constexpr uint32_t FLAG_BITMAP_RANGE = 2u;
constexpr uint32_t MAX_TRAVERSAL_STEPS = 1u << 22;
// largest occupancy allocation (bytes) of a device tree: its reads take
// 32-bit byte offsets
constexpr uint64_t OCC_RANGE_MAX = 0xFFFFFFFFull;

struct TreeView {
  const NodeRec* nodes;
  const uint32_t* brick_pix;
  const uint32_t* brick_occ;
  const uint32_t* colors;
  const uint64_t* ray_lut;  // RAY_TO_NODE_OCCUPANCY_BITMASK_LUT, [pos*8 + dir]
  float size;
  uint32_t dim;
  uint32_t occ_words;
  uint32_t color_count;
  float inv_dim;  // pow2_rcp(dim): exact since the brick dimension is a power of two
  uint32_t node_count;
  uint32_t max_steps;  // iteration bound of one ray (MAX_TRAVERSAL_STEPS unless lowered)
  uint32_t mips_enabled;  // MIPMapStrategy::enabled: LOD queries probe node MIPs
  // bytes of the device's brick_occ allocation (below 2^32): the range of
  // its 32-bit offset reads (occ_word); 0 on the host, where it is unused
  uint32_t occ_range;
};

// Occupancy word flat >> 5 of the brick whose words begin at byte offset
// brick_off of brick_occ, for a probe whose cell index may lie outside the
// brick (`in` false: the caller then ignores the word). On the device the
// read is a buffer load whose range is the whole occupancy allocation, so
// such an index reads some word of the allocation, or 0 past its end, and
// needs no clamp; on the host it reads the brick's first word instead.
SVO_HD inline uint32_t occ_word(const TreeView& t, uint32_t brick_off, uint32_t flat, bool in) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)in;
  // 0x00020000: 32-bit data format, the word-read descriptor of gfx9 parts
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t*>(t.brick_occ), 0, static_cast<int>(t.occ_range), 0x00020000);
  // (the word index made opaque: the offset is then one shift-add of the
  // brick's word offset, where the compiler would otherwise form
  // (flat >> 3) & ~3 and add it separately)
  uint32_t w = flat >> 5;
  asm volatile("" : "+v"(w));
  return __builtin_amdgcn_raw_buffer_load_b32(rsrc, static_cast<int>(brick_off + (w << 2)), 0, 0);
#else
  return t.brick_occ[(brick_off >> 2) + (in ? (flat >> 5) : 0u)];
#endif
}

// Every bounds size a traversal meets is the tree size halved k times (PUSH)
// or doubled back (POP), never below the brick dimension, and a brick's
// voxel size is a bounds size over the brick dimension. When the tree size
// is a power of two no larger than 2^24 and the brick dimension a power of
// two, all of these are powers of two in [1, 2^24] and all bounds
// coordinates are integers below 2^24. Then x / s equals x * (1/s) exactly
// and the POP's a - fmodf(a, b) equals floor(a / b) * b exactly (see
// pow2_rcp and sub_fmod), so the traversal can be compiled with P2 = true,
// without the general division and fmodf fallbacks; it is the same
// arithmetic, bit for bit. (Octree::new checks both powers through f32
// logarithms, which can admit a non-power at large sizes; such trees take
// the P2 = false form; the device upload refuses them.)
SVO_HD inline bool exact_pow2_sizes(uint64_t size, uint64_t dim) {
  return size != 0 && (size & (size - 1)) == 0 && size <= (1ull << 24) && dim != 0 &&
         (dim & (dim - 1)) == 0 && dim <= size;
}

// 1 / s for a bounds size s (a power of two in [1, 2^24] when P2): the
// exponent is negated by one integer subtraction
template <bool P2>
SVO_HD inline float size_rcp(float s) {
  if constexpr (P2) {
    return __builtin_bit_cast(float, 0x7F000000u - __builtin_bit_cast(uint32_t, s));
  } else {
    return pow2_rcp(s);
  }
}
template <bool P2>
SVO_HD inline float div_size(float x, float s, float rcp) {
  if constexpr (P2) {
    return x * rcp;
  } else {
    return div_by(x, s, rcp);
  }
}
template <bool P2>
SVO_HD inline V3f div3_size(V3f v, float s, float rcp) {
  if constexpr (P2) {
    return V3f{v.x * rcp, v.y * rcp, v.z * rcp};
  } else {
    return div3_by(v, s, rcp);
  }
}
// (p - min) * 4 / size, the position of p in the 4x4x4 bitmap of the node
// with bounds b (raytracing_on_cpu.rs:425-427). With P2, 4 and 1 / size are
// powers of two, so one scaling by their product rounds the same real
// number once, as the two steps do.
template <bool P2>
SVO_HD inline V3f bitmap_pos(V3f p, Cube b, float rcp) {
  if constexpr (P2) {
    const float k = rcp * 4.0f;
    return V3f{(p.x - b.min.x) * k, (p.y - b.min.y) * k, (p.z - b.min.z) * k};
  } else {
    return div3_by((p - b.min) * 4.0f, b.size, rcp);
  }
}
// st * 4 / s, the move of a node's bitmap position for a DDA step st (per
// axis +0.0, +-1 or NaN). With P2, 4 / s is a power of two and both
// products are exact, so one multiplication gives the same value.
template <bool P2>
SVO_HD inline V3f bitmap_move(V3f st, float s, float rcp) {
  if constexpr (P2) {
    const float k = rcp * 4.0f;
    return V3f{st.x * k, st.y * k, st.z * k};
  } else {
    return div3_by(st * 4.0f, s, rcp);
  }
}
// The POP's octant of the node within its parent (raytracing_on_cpu.rs:
// 450-461): hash_region(center - parent min, size). With P2 every bounds
// coordinate is a multiple of its size below 2^24 and min - parent min is
// exactly 0 or the size per axis, so the test center - parent min >= size
// is min != parent min.
template <bool P2>
SVO_HD inline uint32_t pop_octant(Cube b, V3f pmin) {
  if constexpr (P2) {
    return static_cast<uint32_t>(b.min.x != pmin.x) + static_cast<uint32_t>(b.min.z != pmin.z) * 2u +
           static_cast<uint32_t>(b.min.y != pmin.y) * 4u;
  } else {
    return hash_region(b.min + v3u(b.size / 2.0f) - pmin, b.size);
  }
}
// a - fmodf(a, b) for a bounds coordinate a >= +0 and b twice a bounds size
template <bool P2>
SVO_HD inline float sub_fmod_size(float a, float b, float rcp_b) {
  if constexpr (P2) {
    return floorf(a * rcp_b) * b;
  } else {
    return sub_fmod(a, b, rcp_b);
  }
}

// A ray's result. The normal of a hit is a function of the hit cell and the
// impact point (hit_normal); it is derived only where a caller asks for it.
// For a miss only `hit` (0) and `flags` are defined; consumers write the
// miss values themselves (see miss_result for what a miss reports).
struct CastResult {
  uint32_t hit;
  uint32_t pix;
  uint32_t flags;
  V3f impact;
  Cube cell;  // the voxel cell (Parted) or brick (Solid) that was hit
};

// get_by_ray's normal (raytracing_on_cpu.rs:296-311): cube_impact_normal of
// the hit cell at the impact point; zero for a miss.
SVO_HD inline V3f hit_normal(const CastResult& r) {
  return r.hit ? cube_impact_normal(r.cell, r.impact) : V3f{0.0f, 0.0f, 0.0f};
}

// Rust f32::clamp(min, max): NaN passes through
SVO_HD inline float rclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// BITMAP_INDEX_LUT[floor(x) as usize][floor(y)][floor(z)] = x + 4y + 16z
// The reference converts with `as usize`; a saturating u32 conversion gives
// the same value for 0..3 and a value above 3 (flagged) for everything else,
// without the 64-bit conversion sequence. The conversion truncates, which
// saturates to the same value as floor would: equal for x >= 0, and 0 for
// negative x and NaN either way, so the floor is left out.
SVO_HD inline uint32_t bitmap_index(V3f bp, uint32_t* flags) {
  uint32_t x = f2u32(bp.x);
  uint32_t y = f2u32(bp.y);
  uint32_t z = f2u32(bp.z);
  if (x > 3 || y > 3 || z > 3) {
    *flags |= FLAG_BITMAP_RANGE;
    x = x > 3 ? 3 : x;
    y = y > 3 ? 3 : y;
    z = z > 3 ? 3 : z;
  }
  return x + 4u * y + 16u * z;
}

// A ray that hit nothing: the only defined fields are hit and flags.
// (Materialising the miss values here would make the compiler set them in
// every loop iteration, ahead of the loop's exits.)
SVO_HD inline CastResult miss_flags(uint32_t flags) {
  CastResult r;
  r.hit = 0;
  r.flags = flags;
  return r;
}

// A miss with the values get_by_ray's callers see: no palette value, impact
// and cell zero.
SVO_HD inline CastResult miss_result(uint32_t flags) {
  CastResult r;
  r.hit = 0;
  r.pix = EMPTY_MARKER;
  r.flags = flags;
  r.impact = V3f{0.0f, 0.0f, 0.0f};
  r.cell = Cube{V3f{0.0f, 0.0f, 0.0f}, 0.0f};
  return r;
}

// bitmap_index(rclamp(bp, 1e-5, 4 - 1e-5)) without forming the clamped
// position: per axis, a saturating conversion (truncation, NaN -> 0) and a
// clamp of the integer to 0..3 give the same cell as clamping first (below
// 1e-5 and above 4 - 1e-5 the integer is <= 0 and >= 3; NaN passes the float
// clamp and converts to 0 either way), and no range test is needed.
// Truncation and floor differ only below 0, where both clamp to 0.
SVO_HD inline uint32_t clamp03(int32_t v) { return static_cast<uint32_t>(v < 0 ? 0 : (v > 3 ? 3 : v)); }
SVO_HD inline uint32_t bitmap_cell(V3f bp) {
  return clamp03(f2i32(bp.x)) + 4u * clamp03(f2i32(bp.y)) + 16u * clamp03(f2i32(bp.z));
}

// probe_brick (raytracing_on_cpu.rs:256-312) with traverse_brick (:156-252)
template <bool P2>
SVO_HD __attribute__((always_inline)) inline bool probe_brick(const TreeView& t, uint32_t form,
                                                              uint32_t word, Cube bb, V3f d,
                                                              V3f sig, V3f scale, V3f* p,
                                                              CastResult* r) {
  if (form == FORM_EMPTY) return false;
  if (form == FORM_SOLID) {
    r->pix = word;
    r->impact = *p;
    r->cell = bb;
    return true;
  }
  const int dim = static_cast<int>(t.dim);
  const float fdim = static_cast<float>(t.dim);
  const float rcp_size = size_rcp<P2>(bb.size);
  V3f pib;
  if constexpr (P2) {
    // (p - min) * dim / size: dim and 1 / size are powers of two, so scaling
    // by their product rounds the same real number once, as the two steps do
    const float k = fdim * rcp_size;
    pib = (*p - bb.min) * k;
  } else {
    pib = div3_size<P2>((*p - bb.min) * fdim, bb.size, rcp_size);
  }
  auto clampi = [dim](int v) {
    const int a = v > 0 ? v : 0;
    return a < dim - 1 ? a : dim - 1;
  };
  int ix = clampi(f2i32(pib.x));
  int iy = clampi(f2i32(pib.y));
  int iz = clampi(f2i32(pib.z));
  const float unit = div_size<P2>(bb.size, fdim, t.inv_dim);
  // A step moves the cell by signum(direction) on the axes that stepped and
  // leaves the others (their step component is +0.0, and cell corners are
  // never -0.0): the per-axis moves are fixed for the whole probe.
  const V3f umove{sig.x * unit, sig.y * unit, sig.z * unit};
  const uint32_t shift = static_cast<uint32_t>(__builtin_ctz(t.dim));
  // ix + iy dim + iz dim^2 (P2: the factors as shifts)
  auto flat_of = [&](int x, int y, int z) -> uint32_t {
    if constexpr (P2) {
      return static_cast<uint32_t>(x) | (static_cast<uint32_t>(y) << shift) |
             (static_cast<uint32_t>(z) << (2u * shift));
    } else {
      return static_cast<uint32_t>(x + y * dim + z * dim * dim);
    }
  };
  // The loop has one exit: the cell is outside the brick, or occupied, or
  // (without P2) the step bound is reached (then `in` still holds and
  // `occupied` not). The occupancy word is read for every cell, from index 0
  // when the cell is outside, so that the test needs no branch.
  bool in, occupied;
  if constexpr (P2) {
    // Cell coordinates mirrored on the axes whose signum is -1: a = ix
    // (signum +1) or dim-1-ix = ix ^ (dim-1) (signum -1), so that every move
    // is +1 and an axis leaves the brick at a = dim; the cell's flat index is
    // the mirrored one xor the mirror mask. With dim a power of two, all
    // three are below dim exactly when their bitwise or is.
    const uint32_t dm1 = t.dim - 1u;
    const uint32_t nx = sig.x < 0.0f ? dm1 : 0u;
    const uint32_t ny = sig.y < 0.0f ? dm1 : 0u;
    const uint32_t nz = sig.z < 0.0f ? dm1 : 0u;
    const uint32_t mirror = nx | (ny << shift) | (nz << (2u * shift));
    uint32_t ax = static_cast<uint32_t>(ix) ^ nx;
    uint32_t ay = static_cast<uint32_t>(iy) ^ ny;
    uint32_t az = static_cast<uint32_t>(iz) ^ nz;
    // occupancy words per brick: dim^3 / 32 rounded up, a power of two
    const uint32_t occ_off = (word << static_cast<uint32_t>(__builtin_ctz(t.occ_words))) * 4u;
    // cmin holds the cell's far corner (the corner the direction points to):
    // min + (i + 1) unit on an axis whose signum is +1, min + i unit on the
    // others. Every term is a multiple of unit (a power of two) no larger
    // than the tree size (at most 2^24), so each is exact in f32: the fused
    // form gives the same value as the reference's min + i unit, then + unit.
    V3f cmin{fmaf(static_cast<float>(static_cast<uint32_t>(ix) + (sig.x < 0.0f ? 0u : 1u)), unit,
                  bb.min.x),
             fmaf(static_cast<float>(static_cast<uint32_t>(iy) + (sig.y < 0.0f ? 0u : 1u)), unit,
                  bb.min.y),
             fmaf(static_cast<float>(static_cast<uint32_t>(iz) + (sig.z < 0.0f ? 0u : 1u)), unit,
                  bb.min.z)};
    for (;;) {
      in = (ax | ay | az) < t.dim;
      const uint32_t flat = (ax | (ay << shift) | (az << (2u * shift))) ^ mirror;
      const uint32_t bits = occ_word(t, occ_off, flat, in);
      occupied = in & (bit_of(bits, flat) != 0u);
      if (!in | occupied) break;
      bool sx, sy, sz;
      const float m = dda_step_far(d, p, cmin, scale, &sx, &sy, &sz);
      cmin = V3f{sx ? cmin.x + umove.x : cmin.x, sy ? cmin.y + umove.y : cmin.y,
                 sz ? cmin.z + umove.z : cmin.z};
      ax += static_cast<uint32_t>(sx);
      ay += static_cast<uint32_t>(sy);
      az += static_cast<uint32_t>(sz);
      // A step of length m (not NaN) moves at least one axis, each axis only
      // in its direction's sign (signum is +-1 there: a NaN direction makes
      // every distance NaN), so the cell leaves the brick within 3 dim
      // steps. A NaN step moves nothing (p becomes NaN and stays so), so the
      // loop would repeat that state: it ends here.
      if (!(m == m)) break;
    }
    ix = static_cast<int>(ax ^ nx);
    iy = static_cast<int>(ay ^ ny);
    iz = static_cast<int>(az ^ nz);
  } else {
    V3f cmin = bb.min + V3f{static_cast<float>(ix), static_cast<float>(iy),
                            static_cast<float>(iz)} *
                            unit;
    const uint32_t* occ = t.brick_occ + static_cast<uint64_t>(word) * t.occ_words;
    const int mx = step_int(sig.x), my = step_int(sig.y), mz = step_int(sig.z);
    for (int left = 4 * dim + 16;;) {
      in = ix >= 0 && ix < dim && iy >= 0 && iy < dim && iz >= 0 && iz < dim;
      const uint32_t flat = in ? flat_of(ix, iy, iz) : 0u;
      const uint32_t bits = occ[flat >> 5];
      occupied = in & (bit_of(bits, flat) != 0u);
      if (!in | occupied) break;
      bool sx, sy, sz;
      dda_step_axes(d, sig, p, cmin, unit, scale, &sx, &sy, &sz);
      cmin = V3f{sx ? cmin.x + umove.x : cmin.x, sy ? cmin.y + umove.y : cmin.y,
                 sz ? cmin.z + umove.z : cmin.z};
      ix += sx ? mx : 0;
      iy += sy ? my : 0;
      iz += sz ? mz : 0;
      if (--left == 0) break;
    }
  }
  if (occupied) {
    const uint64_t n3 = static_cast<uint64_t>(t.dim) * t.dim * t.dim;
    r->pix = t.brick_pix[static_cast<uint64_t>(word) * n3 + flat_of(ix, iy, iz)];
    const V3f scaled = V3f{static_cast<float>(ix), static_cast<float>(iy),
                           static_cast<float>(iz)} *
                       bb.size;
    const Cube hb{bb.min + div3_size<P2>(scaled, fdim, t.inv_dim),
                  div_size<P2>(bb.size, fdim, t.inv_dim)};
    r->impact = *p;
    r->cell = hb;
    return true;
  }
  if (in) r->flags |= FLAG_ITERATION_GUARD;
  return false;
}

// The rest of a ray whose POP of the root empties the stack because the
// root has no occupied bitmap cell ahead of the ray (occupancy & LUT == 0).
// The reference then restarts from the root at p + 0.1 d, fails the same
// test, and so on until p leaves the tree: the LUT box of a cell holds the
// boxes of every cell further along the ray (each coordinate moves towards
// the box's far corner or stays), so no later position passes the test and
// the ray ends as a miss. Only the nudges and the iteration count remain,
// replayed here exactly (the restart's interior test, then one counted
// iteration per restart). Not for MIP-enabled LOD rays (their MIP test runs
// on every iteration) nor for a leaf root (its brick probe does).
SVO_HD inline CastResult crawl_out_of_root(const TreeView& t, V3f p, V3f d, uint32_t steps,
                                           uint32_t flags) {
  for (;;) {
    p = p + d * 0.1f;
    if (!((p.x < t.size) & (p.y < t.size) & (p.z < t.size) & (p.x > 0.0f) & (p.y > 0.0f) &
          (p.z > 0.0f)))
      return miss_flags(flags);
    if (++steps > t.max_steps) return miss_flags(flags | FLAG_ITERATION_GUARD);
  }
}

// The reference's NodeStack<u32, 4> (raytracing_on_cpu.rs:19-82): a ring of
// four whose push overwrites the oldest entry once full and whose count
// says how many entries are retained. Kept here as a shift register, s0 the
// top: a push shifts down (dropping the oldest), a pop shifts up, so the top
// is read without selecting among the four registers. Entries beyond
// `count` are never read, which makes the two forms equivalent.
struct RingStack {
  uint32_t s0, s1, s2, s3;
  uint32_t count;
  SVO_HD void push(uint32_t v) {
    s3 = s2;
    s2 = s1;
    s1 = s0;
    s0 = v;
    count = count < 4u ? count + 1u : 4u;
  }
  SVO_HD void pop() {
    s0 = s1;
    s1 = s2;
    s2 = s3;
    count -= 1u;
  }
};

// get_by_ray_at_lod (raytracing_on_cpu.rs:325-565). Without LOD it is
// get_by_ray (viewing distance f32::MAX, under which the MIP test never
// passes). `lut` is the RAY_TO_NODE_OCCUPANCY_BITMASK_LUT, wherever the
// caller staged it. `top` optionally holds a copy of the first `ntop` node
// records (the top levels, as records are numbered breadth-first), e.g.
// staged in LDS; those records are read from it instead of t.nodes. P2
// requires exact_pow2_sizes(t.size, t.dim).
//
// With LOD and the tree's MIPs enabled, every stack-loop iteration first
// tests the node's level against the distance travelled (:368-386): the
// level starts at log2(size / dim), falls by one per PUSH, rises by one per
// POP and is not reset by a restart from the root; when the level is below
// |origin - round(p / (2 level)) * 2 level| / viewing_distance, the node's
// MIP brick is probed over the node's bounds like a uniform leaf's brick.
template <bool LOD, bool P2, class LutT>
SVO_HD __attribute__((always_inline)) inline CastResult cast_ray_impl(
    const TreeView& t, const LutT* __restrict__ lut, V3f o, V3f d, const NodeRec* top,
    uint32_t ntop, float viewing_distance) {
  // Only the flags live across the loops; the probe fills pix, impact and
  // cell on the hit path, and a miss returns miss_flags(flags). (Zeroing
  // the hit fields up front would make the compiler carry them around every
  // loop.)
  CastResult r;
  r.flags = 0;

  const V3f scale = dda_scale_factors(d);
  const uint32_t dir_idx = hash_direction(d);
  const Cube root{v3u(0.0f), t.size};
  const float lo = 0.00001f;
  const float hi = 4.0f - 0.00001f;

  RingStack stk{0u, 0u, 0u, 0u, 1u};  // the root pushed
  Cube bounds = root;
  V3f p;
  uint32_t target;
  {
    float tt = 0.0f;
    const int ir = intersect_ray(root, o, d, &tt);
    if (ir != 0) {
      p = o + d * (ir == 2 ? tt : 0.0f);
      target = hash_region(p - root.min, root.size / 2.0f);
    } else {
      p = o;
      target = OOB_OCTANT;
    }
  }
  if (target == OOB_OCTANT) return miss_flags(0u);
  uint32_t steps = 0;
  // log2(size / dim) of two powers of two, exactly
  float mip_level =
      LOD ? static_cast<float>(__builtin_ctz(static_cast<uint32_t>(t.size) / t.dim)) : 0.0f;
  // One loop runs the reference's two: an iteration of its stack loop, and
  // when the stack has emptied, first its restart from the root (which ends
  // the ray when the nudged point leaves the tree).
  for (;;) {
    // signum(d) per axis, formed again in every iteration from d's sign bits
    // (one instruction each) rather than kept in three registers: the empty
    // asm makes d opaque here, so the compiler cannot hoist these (or other
    // values derived from d) out of the loop. For a NaN component the
    // reference's signum is NaN and this one +-1; then every scale factor is
    // NaN (each involves all three components), every DDA distance is NaN
    // and no step selects a signum, so the difference is never observed.
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(d.x), "+v"(d.y), "+v"(d.z));
#endif
    V3f sig{sign_one(d.x), sign_one(d.y), sign_one(d.z)};
    if (++steps > t.max_steps) {
      r.flags |= FLAG_ITERATION_GUARD;
      break;
    }
    const uint32_t key = stk.s0;
    const NodeRec& n = (key < ntop ? top : t.nodes)[key];
    const uint64_t occ = n.occ;
    const uint32_t meta = n.meta;
    // the child word of the current target, read together with the header
    // (same record, one round trip); with target OOB_OCTANT it is child[0]
    // and unused
    const uint32_t tword = n.child[target & 7u];
    const uint32_t kind = node_kind(meta);
    bool backtrack = kind == KIND_UNIFORM_LEAF;
    if (LOD && t.mips_enabled) {
      const float grid = mip_level * 2.0f;
      const V3f snapped{roundf(p.x / grid) * grid, roundf(p.y / grid) * grid,
                        roundf(p.z / grid) * grid};
      if (mip_level < length(o - snapped) / viewing_distance &&
          probe_brick<P2>(t, n.mip, n.mip_word, bounds, d, sig, scale, &p, &r))
        goto hit_found;
    }
    if (target != OOB_OCTANT && (kind == KIND_UNIFORM_LEAF || kind == KIND_LEAF)) {
      const bool uniform = kind == KIND_UNIFORM_LEAF;
      const uint32_t slot = uniform ? 0u : target;
      if (probe_brick<P2>(t, brick_form(meta, slot), slot == target ? tword : n.child[0],
                          uniform ? bounds : child_bounds_for(bounds, target), d, sig, scale,
                          &p, &r))
        goto hit_found;
    }
    // signum(d) formed again past the probe, so that the values formed
    // before it are dead inside the probe's DDA loop (whose registers then
    // hold the per-axis cell moves rather than recomputing them each step)
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(d.x), "+v"(d.y), "+v"(d.z));
#endif
    sig = V3f{sign_one(d.x), sign_one(d.y), sign_one(d.z)};
    const float rcp_size = size_rcp<P2>(bounds.size);
    // the position in the node's 4x4x4 bitmap; the clamp to [lo, hi]
    // (:425-436) is applied to the float position only where ADVANCE
    // continues from it, the cell index comes from bitmap_cell
    V3f bp = bitmap_pos<P2>(p, bounds, rcp_size);
    uint32_t flat = bitmap_cell(bp);
    const bool pop = backtrack || target == OOB_OCTANT || occ == 0ull ||
                     (occ & lut[flat * 8u + dir_idx]) == 0ull;
    Cube tb = child_bounds_for(bounds, target);
    // the child word is read only when the occupancy test passes (pure
    // reads: the same decision as testing the child first)
    const bool push = !pop && kind == KIND_INTERNAL &&
                      (occ & octant_bitmap_mask(target)) != 0ull && tword != INVALID_INDEX;
    // PUSH, POP and ADVANCE leave the next bounds and target here, and the
    // stack moves once after them, by selects: the loop has a single back
    // edge, so its state stays in the same registers on every path.
    Cube next_bounds = bounds;
    uint32_t next_target = target;
    if (push) {
      // PUSH (:484-492)
      next_bounds = tb;
      next_target = hash_region(p - tb.min, tb.size / 2.0f);
    } else {
      // POP (:445-475) and ADVANCE (:497-544) both begin with one DDA step,
      // out of the node (POP, unless the stack empties) or out of the target
      // child (ADVANCE); it is issued once for the lanes of both. (The POP's
      // other values depend on the bounds only, so taking the step before
      // them changes nothing.)
      if (!pop && ++steps > t.max_steps) {
        r.flags |= FLAG_ITERATION_GUARD;
        break;
      }
      // (left unset when the POP empties the stack: then it is not read)
      V3f st;
      if (!pop || stk.count > 1u) st = dda_step(d, sig, &p, pop ? bounds : tb, scale);
      if (pop) {
        if (stk.count == 1u && key == 0u && !backtrack && target != OOB_OCTANT &&
            kind != KIND_LEAF && !(LOD && t.mips_enabled))
          return crawl_out_of_root(t, p, d, steps, r.flags);
        if (stk.count > 1u) {
          // the parent's bounds and the sibling the step leads to
          const float s2 = bounds.size * 2.0f;
          const float rcp_s2 = size_rcp<P2>(s2);
          const V3f pmin{sub_fmod_size<P2>(bounds.min.x, s2, rcp_s2),
                         sub_fmod_size<P2>(bounds.min.y, s2, rcp_s2),
                         sub_fmod_size<P2>(bounds.min.z, s2, rcp_s2)};
          const uint32_t from = pop_octant<P2>(bounds, pmin);
          next_target = step_octant(from, st);
          next_bounds = Cube{pmin, s2};
        } else {
          // the POP empties the stack: the reference's restart from the
          // root (:551-562), which ends the ray once the nudged point has
          // left the tree (taken here, not at the next iteration's start:
          // nothing happens in between)
          p = p + d * 0.1f;
          next_target = ((p.x < t.size) & (p.y < t.size) & (p.z < t.size) & (p.x > 0.0f) &
                         (p.y > 0.0f) & (p.z > 0.0f))
                            ? hash_region(p, t.size / 2.0f)
                            : OOB_OCTANT;
          if (next_target == OOB_OCTANT) break;
          next_bounds = root;
        }
      } else {
        // the rest of ADVANCE
        bool guard = false;
        next_target = step_octant(target, st);
        bp = V3f{rclamp(bp.x, lo, hi), rclamp(bp.y, lo, hi), rclamp(bp.z, lo, hi)};
        while (next_target != OOB_OCTANT) {
          tb = child_bounds_for(bounds, next_target);
          bp = bp + bitmap_move<P2>(st, bounds.size, rcp_size);
          flat = bitmap_index(bp, &r.flags);
          // stop at a reachable occupied child or a non-empty leaf brick; the
          // tests run cheapest first and the child word is read last
          if (kind == KIND_LEAF && brick_form(meta, next_target) != FORM_EMPTY) break;
          if (kind == KIND_INTERNAL && (occ & octant_bitmap_mask(next_target)) != 0ull &&
              (lut[flat * 8u + dir_idx] & occ) != 0ull && n.child[next_target] != INVALID_INDEX)
            break;
          if (++steps > t.max_steps) {
            guard = true;
            break;
          }
          st = dda_step(d, sig, &p, tb, scale);
          next_target = step_octant(next_target, st);
        }
        if (guard) {
          r.flags |= FLAG_ITERATION_GUARD;
          break;
        }
      }
    }
    // PUSH shifts the stack down (its oldest entry drops out once four are
    // held), POP shifts it up, and a POP that empties it is followed by the
    // restart's push of the root; ADVANCE leaves it
    const uint32_t s0 = stk.s0, s1 = stk.s1, s2 = stk.s2, s3 = stk.s3;
    const bool emptied = pop && stk.count == 1u;
    stk.s0 = push ? tword : (pop ? (emptied ? 0u : s1) : s0);
    stk.s1 = push ? s0 : (pop ? s2 : s1);
    stk.s2 = push ? s1 : (pop ? s3 : s2);
    stk.s3 = push ? s2 : s3;
    stk.count = push ? (stk.count < 4u ? stk.count + 1u : 4u)
                     : (pop ? (emptied ? 1u : stk.count - 1u) : stk.count);
    if (LOD) mip_level += push ? -1.0f : (pop ? 1.0f : 0.0f);
    bounds = next_bounds;
    target = next_target;
  }
  return miss_flags(r.flags);
hit_found:
  r.hit = 1;
  return r;
}

// The same traversal as cast_ray_impl<false, P2> with the work of a wave
// regrouped (the "while-while" arrangement of GPU ray traversal): each lane
// runs stack-loop iterations until its next iteration starts with a brick
// probe, or its ray ends; the wave then probes the due bricks of all such
// lanes together, and the rounds repeat. Every lane performs the same
// operations in the same order as in cast_ray_impl (an iteration interrupted
// for its probe resumes after it, re-reading its node record and without
// counting a second step), so results are identical bit for bit; only the
// interleaving across lanes differs, which keeps more lanes active in the
// brick DDA loop. (Ending a round's first phase early, once a quorum of
// the wave's lanes wait for their probes, measured slower.)
//
// CRAWL = false leaves out the crawl_out_of_root shortcut (the restarts are
// then run one by one, with the same result; the host tests run both forms).
template <bool P2, class LutT, bool CRAWL = true>
SVO_HD __attribute__((always_inline)) inline CastResult cast_ray_ww(
    const TreeView& t, const LutT* __restrict__ lut, V3f o, V3f d, const NodeRec* top,
    uint32_t ntop) {
  // (here, unlike in cast_ray_impl, zeroing the hit fields up front gives
  // the smaller register allocation, and measured 4% faster)
  CastResult r = miss_result(0u);

  const V3f scale = dda_scale_factors(d);
  const uint32_t dir_idx = hash_direction(d);
  const Cube root{v3u(0.0f), t.size};
  const float lo = 0.00001f;
  const float hi = 4.0f - 0.00001f;

  RingStack stk{0u, 0u, 0u, 0u, 1u};  // the root pushed
  Cube bounds = root;
  V3f p;
  uint32_t target;
  {
    float tt = 0.0f;
    const int ir = intersect_ray(root, o, d, &tt);
    if (ir != 0) {
      p = o + d * (ir == 2 ? tt : 0.0f);
      target = hash_region(p - root.min, root.size / 2.0f);
    } else {
      p = o;
      target = OOB_OCTANT;
    }
  }
  bool finished = target == OOB_OCTANT;
  bool probed = false;  // the current iteration's probe has run
  bool crawl = false;   // the ray ends in crawl_out_of_root, after the loop
  uint32_t steps = 0;
  for (;;) {
    // phase 1: stack-loop iterations up to the next due probe
    bool due = false;
    while (!finished) {
      // signum(d), formed in each iteration as in cast_ray_impl
#if defined(__HIP_DEVICE_COMPILE__)
      asm volatile("" : "+v"(d.x), "+v"(d.y), "+v"(d.z));
#endif
      const V3f sig{sign_one(d.x), sign_one(d.y), sign_one(d.z)};
      if (stk.count == 0u) {
        // restart from the root (:551-562)
        p = p + d * 0.1f;
        target = ((p.x < t.size) & (p.y < t.size) & (p.z < t.size) & (p.x > 0.0f) & (p.y > 0.0f) &
                   (p.z > 0.0f))
                     ? hash_region(p, t.size / 2.0f)
                     : OOB_OCTANT;
        if (target == OOB_OCTANT) {
          finished = true;
          break;
        }
        bounds = root;
        stk.push(0u);
      }
      if (!probed && ++steps > t.max_steps) {
        r.flags |= FLAG_ITERATION_GUARD;
        finished = true;
        break;
      }
      const uint32_t key = stk.s0;
      const NodeRec& n = (key < ntop ? top : t.nodes)[key];
      const uint64_t occ = n.occ;
      const uint32_t meta = n.meta;
      const uint32_t tword = n.child[target & 7u];
      const uint32_t kind = node_kind(meta);
      if (!probed && target != OOB_OCTANT && (kind == KIND_UNIFORM_LEAF || kind == KIND_LEAF)) {
        due = true;
        break;
      }
      probed = false;
      const bool backtrack = kind == KIND_UNIFORM_LEAF;
      const float rcp_size = size_rcp<P2>(bounds.size);
      V3f bp = bitmap_pos<P2>(p, bounds, rcp_size);
      uint32_t flat = bitmap_cell(bp);  // cell of the clamped position
      if (backtrack || target == OOB_OCTANT || occ == 0ull ||
          (occ & lut[flat * 8u + dir_idx]) == 0ull) {
        // POP (:445-475); a root with nothing ahead of the ray ends the
        // ray as crawl_out_of_root does, which runs after the loop (where
        // the loop's state is no longer live)
        if (CRAWL && stk.count == 1u && key == 0u && !backtrack && target != OOB_OCTANT &&
            kind != KIND_LEAF) {
          crawl = true;
          finished = true;
          break;
        }
        stk.pop();
        if (stk.count != 0u) {
          const float s2 = bounds.size * 2.0f;
          const float rcp_s2 = size_rcp<P2>(s2);
          const V3f pmin{sub_fmod_size<P2>(bounds.min.x, s2, rcp_s2),
                         sub_fmod_size<P2>(bounds.min.y, s2, rcp_s2),
                         sub_fmod_size<P2>(bounds.min.z, s2, rcp_s2)};
          const uint32_t from = pop_octant<P2>(bounds, pmin);
          const V3f st = dda_step(d, sig, &p, bounds, scale);
          target = step_octant(from, st);
          bounds.size *= 2.0f;
          bounds.min = pmin;
        }
        continue;
      }
      Cube tb = child_bounds_for(bounds, target);
      uint32_t tck = INVALID_INDEX;
      if (kind == KIND_INTERNAL && (occ & octant_bitmap_mask(target)) != 0ull) tck = tword;
      if (tck != INVALID_INDEX) {
        // PUSH (:484-492)
        bounds = tb;
        target = hash_region(p - tb.min, tb.size / 2.0f);
        stk.push(tck);
      } else {
        // ADVANCE (:497-544), from the clamped position
        bp = V3f{rclamp(bp.x, lo, hi), rclamp(bp.y, lo, hi), rclamp(bp.z, lo, hi)};
        for (;;) {
          if (++steps > t.max_steps) {
            r.flags |= FLAG_ITERATION_GUARD;
            finished = true;
            break;
          }
          const V3f st = dda_step(d, sig, &p, tb, scale);
          target = step_octant(target, st);
          if (target == OOB_OCTANT) break;
          tb = child_bounds_for(bounds, target);
          bp = bp + bitmap_move<P2>(st, bounds.size, rcp_size);
          flat = bitmap_index(bp, &r.flags);
          if (kind == KIND_LEAF && brick_form(meta, target) != FORM_EMPTY) break;
          if (kind == KIND_INTERNAL && (occ & octant_bitmap_mask(target)) != 0ull &&
              (lut[flat * 8u + dir_idx] & occ) != 0ull && n.child[target] != INVALID_INDEX)
            break;
        }
      }
    }
    if (finished) break;
    if (!due) continue;
    // phase 2: the due probe of the iteration, lanes of the wave together
    {
#if defined(__HIP_DEVICE_COMPILE__)
      asm volatile("" : "+v"(d.x), "+v"(d.y), "+v"(d.z));
#endif
      const V3f sig{sign_one(d.x), sign_one(d.y), sign_one(d.z)};
      const NodeRec& n = (stk.s0 < ntop ? top : t.nodes)[stk.s0];
      const uint32_t meta = n.meta;
      const bool uniform = node_kind(meta) == KIND_UNIFORM_LEAF;
      const uint32_t slot = uniform ? 0u : target;
      if (probe_brick<P2>(t, brick_form(meta, slot), n.child[slot],
                          uniform ? bounds : child_bounds_for(bounds, target), d, sig, scale,
                          &p, &r)) {
        r.hit = 1;
        return r;
      }
      probed = true;
    }
  }
  if (crawl) r.flags = crawl_out_of_root(t, p, d, steps, r.flags).flags;
  return r;
}

template <bool P2 = false, class LutT>
SVO_HD __attribute__((always_inline)) inline CastResult cast_ray(const TreeView& t,
                                                                const LutT* __restrict__ lut,
                                                                V3f o, V3f d,
                                                                const NodeRec* top = nullptr,
                                                                uint32_t ntop = 0u) {
  return cast_ray_impl<false, P2>(t, lut, o, d, top, ntop, 0.0f);
}

template <bool P2 = false, class LutT>
SVO_HD __attribute__((always_inline)) inline CastResult cast_ray_lod(
    const TreeView& t, const LutT* __restrict__ lut, V3f o, V3f d, float viewing_distance) {
  return cast_ray_impl<true, P2>(t, lut, o, d, nullptr, 0u, viewing_distance);
}

template <bool P2 = false>
SVO_HD __attribute__((always_inline)) inline CastResult cast_ray(const TreeView& t, V3f o,
                                                                V3f d) {
  return cast_ray<P2>(t, t.ray_lut, o, d);
}

// The same traversal as cast_ray in resumable form: ray_begin sets a ray up,
// each ray_step runs one iteration of the stack loop (and the restart from
// the root when the stack empties). A kernel can then hand a finished lane a
// new ray while the other lanes of its wave keep stepping. The operation
// sequence per ray is identical to cast_ray's. Only the traversal state and
// the flags live across steps; the result is written when the ray finishes.
struct RayState {
  V3f d, scale, p;
  Cube bounds;
  uint32_t head, count, target, steps, flags;
  // the ring stack's four slots as separate scalars: a dynamically indexed
  // array inside the state would be kept in scratch memory
  uint32_t stk0, stk1, stk2, stk3;
  SVO_HD uint32_t stk_get(uint32_t i) const {
    return i == 0u ? stk0 : (i == 1u ? stk1 : (i == 2u ? stk2 : stk3));
  }
  SVO_HD void stk_set(uint32_t i, uint32_t v) {
    stk0 = i == 0u ? v : stk0;
    stk1 = i == 1u ? v : stk1;
    stk2 = i == 2u ? v : stk2;
    stk3 = i == 3u ? v : stk3;
  }
};

// false when the ray misses the tree outright (its result is miss_result(0))
SVO_HD __attribute__((always_inline)) inline bool ray_begin(const TreeView& t, V3f o, V3f d,
                                                           RayState& s) {
  s.d = d;
  s.scale = dda_scale_factors(d);
  const Cube root{v3u(0.0f), t.size};
  float tt = 0.0f;
  const int ir = intersect_ray(root, o, d, &tt);
  if (ir == 0) return false;
  s.p = o + d * (ir == 2 ? tt : 0.0f);
  s.target = hash_region(s.p - root.min, root.size / 2.0f);
  if (s.target == OOB_OCTANT) return false;
  s.bounds = root;
  s.stk0 = 0u;
  s.stk1 = 0u;
  s.stk2 = 0u;
  s.stk3 = 0u;
  s.head = 1u;
  s.count = 1u;
  s.steps = 0u;
  s.flags = 0u;
  return true;
}

// true when the ray is finished; r then holds its result
template <class LutT>
SVO_HD __attribute__((always_inline)) inline bool ray_step(const TreeView& t,
                                                          const LutT* __restrict__ lut,
                                                          RayState& s, CastResult& r) {
  const float lo = 0.00001f;
  const float hi = 4.0f - 0.00001f;
  if (++s.steps > t.max_steps) {
    r = miss_result(s.flags | FLAG_ITERATION_GUARD);
    return true;
  }
  // cheap functions of the direction, recomputed rather than kept per lane
  const V3f sig{rsignum(s.d.x), rsignum(s.d.y), rsignum(s.d.z)};
  const uint32_t dir_idx = hash_direction(s.d);
  const NodeRec& n = t.nodes[s.stk_get(s.head)];
  const uint64_t occ = n.occ;
  const uint32_t meta = n.meta;
  const uint32_t kind = node_kind(meta);
  bool backtrack = kind == KIND_UNIFORM_LEAF;
  if (s.target != OOB_OCTANT && kind != KIND_INTERNAL && kind != KIND_NOTHING) {
    const bool uniform = kind == KIND_UNIFORM_LEAF;
    const uint32_t slot = uniform ? 0u : s.target;
    CastResult hr = miss_result(0u);
    if (probe_brick<false>(t, brick_form(meta, slot), n.child[slot],
                           uniform ? s.bounds : child_bounds_for(s.bounds, s.target), s.d, sig,
                           s.scale, &s.p, &hr)) {
      r = hr;
      r.hit = 1;
      r.flags = s.flags | hr.flags;
      return true;
    }
    s.flags |= hr.flags;
  }
  const float rcp_size = pow2_rcp(s.bounds.size);
  const V3f rel = (s.p - s.bounds.min) * 4.0f;
  V3f bp = div3_by(rel, s.bounds.size, rcp_size);
  bp = V3f{rclamp(bp.x, lo, hi), rclamp(bp.y, lo, hi), rclamp(bp.z, lo, hi)};
  uint32_t flat = bitmap_index(bp, &s.flags);
  if (backtrack || s.target == OOB_OCTANT || occ == 0ull ||
      (occ & lut[flat * 8u + dir_idx]) == 0ull) {
    // POP
    s.head = (s.head + 3u) & 3u;
    s.count -= 1u;
    if (s.count != 0u) {
      const V3f center = s.bounds.min + v3u(s.bounds.size / 2.0f);
      const float s2 = s.bounds.size * 2.0f;
      const float rcp_s2 = pow2_rcp(s2);
      const V3f pmin{sub_fmod(s.bounds.min.x, s2, rcp_s2), sub_fmod(s.bounds.min.y, s2, rcp_s2),
                     sub_fmod(s.bounds.min.z, s2, rcp_s2)};
      const uint32_t from = hash_region(center - pmin, s.bounds.size);
      const V3f st = dda_step(s.d, sig, &s.p, s.bounds, s.scale);
      s.target = step_octant(from, st);
      s.bounds.size *= 2.0f;
      s.bounds.min = pmin;
    }
  } else {
    Cube tb = child_bounds_for(s.bounds, s.target);
    uint32_t tck = INVALID_INDEX;
    if (kind == KIND_INTERNAL && (occ & octant_bitmap_mask(s.target)) != 0ull)
      tck = n.child[s.target];
    if (tck != INVALID_INDEX) {
      // PUSH
      s.bounds = tb;
      s.target = hash_region(s.p - tb.min, tb.size / 2.0f);
      s.head = (s.head + 1u) & 3u;
      s.count = s.count < 4u ? s.count + 1u : 4u;
      s.stk_set(s.head, tck);
    } else {
      // ADVANCE
      for (;;) {
        if (++s.steps > t.max_steps) {
          r = miss_result(s.flags | FLAG_ITERATION_GUARD);
          return true;
        }
        const V3f st = dda_step(s.d, sig, &s.p, tb, s.scale);
        s.target = step_octant(s.target, st);
        if (s.target == OOB_OCTANT) break;
        tb = child_bounds_for(s.bounds, s.target);
        const V3f sd = st * 4.0f;
        bp = bp + div3_by(sd, s.bounds.size, rcp_size);
        flat = bitmap_index(bp, &s.flags);
        if (kind == KIND_LEAF && brick_form(meta, s.target) != FORM_EMPTY) break;
        if (kind == KIND_INTERNAL && (occ & octant_bitmap_mask(s.target)) != 0ull &&
            (lut[flat * 8u + dir_idx] & occ) != 0ull && n.child[s.target] != INVALID_INDEX)
          break;
      }
    }
  }
  if (s.count == 0u) {
    // restart from the root
    s.p = s.p + s.d * 0.1f;
    s.target = (s.p.x < t.size && s.p.y < t.size && s.p.z < t.size && s.p.x > 0.0f &&
                s.p.y > 0.0f && s.p.z > 0.0f)
                   ? hash_region(s.p, t.size / 2.0f)
                   : OOB_OCTANT;
    if (s.target == OOB_OCTANT) {
      r = miss_result(s.flags);
      return true;
    }
    s.bounds = Cube{v3u(0.0f), t.size};
    s.head = (s.head + 1u) & 3u;
    s.count = 1u;
    s.stk_set(s.head, 0u);
  }
  return false;
}

// Voxel coordinate of a hit: floor(impact) clamped into the hit cell.
SVO_HD inline uint32_t hit_voxel_axis(float p, float cmin, float csize) {
  float f = floorf(p);
  const float hi = cmin + csize - 1.0f;
  f = f < cmin ? cmin : (f > hi ? hi : f);
  return f2u32(f);
}

// Ray of pixel (x, y) of a pinhole viewport (y = 0 is the bottom row), as the
// reference's CPU harness builds it (examples/cpu_render.rs:104-114) with the
// ray through the pixel centre.
struct Camera {
  V3f origin;
  V3f bottom_left;
  V3f right;
  V3f up;
  float pixel_w;
  float pixel_h;
};

// Camera of a Viewport{origin, direction, frustum, fov} (bevy/types.rs:55-71)
// as the reference's CPU harness derives it (examples/cpu_render.rs:78-94):
// up = +y, right = normalize(up x direction), glass `fov` in front of the eye.
SVO_HD inline Camera make_camera(V3f origin, V3f dir, float glass_w, float glass_h, float fov,
                                 uint32_t w, uint32_t h) {
  const V3f up{0.0f, 1.0f, 0.0f};
  Camera cam;
  cam.origin = origin;
  cam.up = up;
  cam.right = normalized(cross(up, dir));
  cam.pixel_w = glass_w / static_cast<float>(w);
  cam.pixel_h = glass_h / static_cast<float>(h);
  cam.bottom_left = ((origin + dir * fov) - up * (glass_h / 2.0f)) - cam.right * (glass_w / 2.0f);
  return cam;
}

SVO_HD inline V3f camera_ray_dir(const Camera& c, uint32_t x, uint32_t y) {
  const float fx = static_cast<float>(x) + 0.5f;
  const float fy = static_cast<float>(y) + 0.5f;
  const V3f glass = (c.bottom_left + (c.right * fx) * c.pixel_w) + (c.up * fy) * c.pixel_h;
  return normalized(glass - c.origin);
}

}  // namespace svo
