// ORACLE - test infrastructure only, never part of the product.
//
// A CPU restatement, in plain C++, of the reference's ray query
// Octree::get_by_ray / get_by_ray_at_lod (shocovox-rs 0.11.1) and of the
// reference point query Octree::get, operating on the reference's own data
// model (ObjectPool<NodeContent>, Vec<NodeChildren>, palettes) as dumped by
// svo_octree_export. It shares no code with the product: only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and only
// to check or to time against the GPU path.
//
// Built with g++ -O3 -ffp-contract=off (no fast-math), so every f32 operation
// is a single IEEE-754 binary32 operation in source order, as Rust evaluates
// it. Each function cites the reference lines it restates.
//
// Pinning (what checks each part independently of this file):
//   * LUTs: regenerated here from the reference's generator functions and
//     compared bit-for-bit with the literal tables (lut.rs:154-896), which the
//     tests hold as a golden fixture;
//   * intersect_ray / cube_impact_normal / hash_region / flat projection /
//     set_occupancy_in_bitmap_64bits / NodeStack: the reference's own
//     known-answer tests (spatial/*/tests.rs, raytracing/tests.rs:817-911);
//   * get_by_ray_at_lod: the 21 ray known-answer tests of
//     raytracing/tests.rs:144-813 on trees built by the product's insert,
//     plus invariants (hit voxel is non-empty under get(); a miss never has
//     a non-empty voxel on a brute-force f64 march away from boundaries);
//   * get: the reference's insert/clear known-answer tests.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace ora {

// ---------------------------------------------------------------------------
// V3c<T> (spatial/math/vector.rs)

template <class T>
struct V3c {
  T x, y, z;
};
using V3f = V3c<float>;

static inline V3f add(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3f sub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3f mul(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3f div(V3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3f unit(float s) { return {s, s, s}; }

// f32::min / f32::max: llvm.minnum / maxnum as lowered on x86-64
static inline float fmin_r(float self, float other) {
  if (std::isnan(self)) return other;
  return other < self ? other : self;
}
static inline float fmax_r(float self, float other) {
  if (std::isnan(self)) return other;
  return other > self ? other : self;
}
// f32::signum
static inline float signum_f(float v) {
  if (std::isnan(v)) return v;
  return std::signbit(v) ? -1.0f : 1.0f;
}
// i32::signum
static inline int32_t signum_i(int32_t v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }
// f32 `as i32` (saturating, NaN -> 0)
static inline int32_t as_i32(float v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(v);
}
// f32 `as usize` (saturating, NaN -> 0)
static inline uint64_t as_usize(float v) {
  if (std::isnan(v) || v <= 0.0f) return 0;
  if (v >= 18446744073709551616.0f) return UINT64_MAX;
  return static_cast<uint64_t>(v);
}
// f32::clamp
static inline float clamp_f(float self, float lo, float hi) {
  float x = self;
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// V3c<f32>::length (vector.rs:75-77)
static inline float length(V3f v) { return std::sqrt((v.x * v.x) + (v.y * v.y) + (v.z * v.z)); }
// V3c<f32>::normalized (vector.rs:79-81)
static inline V3f normalized(V3f v) { return div(v, length(v)); }
// V3c::cross (vector.rs:186-192)
static inline V3f cross(V3f a, V3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// From<V3c<f32>> for V3c<i32> (vector.rs:354-364)
static inline V3c<int32_t> to_i32_round(V3f v) {
  return {as_i32(std::round(v.x)), as_i32(std::round(v.y)), as_i32(std::round(v.z))};
}

// ---------------------------------------------------------------------------
// LUTs (spatial/lut.rs), regenerated from the generator functions

const uint8_t OOB_OCTANT = 8;
const float FLOAT_ERROR_TOLERANCE = 0.00001f;  // raytracing/mod.rs:5
const size_t BITMAP_DIMENSION = 4;             // math/mod.rs:213

const V3f OCTANT_OFFSET_REGION_LUT[8] = {  // lut.rs:156-197
    {0.f, 0.f, 0.f}, {1.f, 0.f, 0.f}, {0.f, 0.f, 1.f}, {1.f, 0.f, 1.f},
    {0.f, 1.f, 0.f}, {1.f, 1.f, 0.f}, {0.f, 1.f, 1.f}, {1.f, 1.f, 1.f}};

// hash_region (math/mod.rs:185-193)
static inline uint8_t hash_region(V3f offset, float size_half) {
  return static_cast<uint8_t>(static_cast<uint8_t>(offset.x >= size_half) +
                              static_cast<uint8_t>(offset.z >= size_half) * 2 +
                              static_cast<uint8_t>(offset.y >= size_half) * 4);
}
// hash_direction (math/mod.rs:196-200)
static inline uint8_t hash_direction(V3f direction) {
  return hash_region(add(unit(1.f), direction), 1.f);
}
// flat_projection (math/mod.rs:209-211)
static inline size_t flat_projection(size_t x, size_t y, size_t z, size_t size) {
  return x + (y * size) + (z * size * size);
}
// position_in_bitmap_64bits (math/mod.rs:256-280)
static inline size_t position_in_bitmap_64bits(V3c<size_t> idx, size_t brick_size) {
  return flat_projection(idx.x * BITMAP_DIMENSION / brick_size, idx.y * BITMAP_DIMENSION / brick_size,
                         idx.z * BITMAP_DIMENSION / brick_size, BITMAP_DIMENSION);
}
// set_occupancy_in_bitmap_64bits (math/mod.rs:288-336)
static void set_occupancy_in_bitmap_64bits(V3c<size_t> position, size_t size, size_t brick_dim,
                                           bool occupied, uint64_t* bitmap) {
  if (brick_dim == 1) {
    *bitmap = occupied ? UINT64_MAX : 0;
    return;
  }
  const size_t update_count = as_usize(
      std::ceil(static_cast<float>(size) * static_cast<float>(BITMAP_DIMENSION) / static_cast<float>(brick_dim)));
  const V3f sp = {std::floor(static_cast<float>(position.x * BITMAP_DIMENSION) / static_cast<float>(brick_dim)),
                  std::floor(static_cast<float>(position.y * BITMAP_DIMENSION) / static_cast<float>(brick_dim)),
                  std::floor(static_cast<float>(position.z * BITMAP_DIMENSION) / static_cast<float>(brick_dim))};
  const V3c<size_t> start = {as_usize(std::round(sp.x)), as_usize(std::round(sp.y)), as_usize(std::round(sp.z))};
  for (size_t x = start.x; x < std::min(start.x + update_count, BITMAP_DIMENSION); ++x)
    for (size_t y = start.y; y < std::min(start.y + update_count, BITMAP_DIMENSION); ++y)
      for (size_t z = start.z; z < std::min(start.z + update_count, BITMAP_DIMENSION); ++z) {
        const uint64_t pos_mask = 1ull << position_in_bitmap_64bits({x, y, z}, BITMAP_DIMENSION);
        if (occupied)
          *bitmap |= pos_mask;
        else
          *bitmap &= ~pos_mask;
      }
}

struct Luts {
  uint64_t ray_to_node_occupancy[64][8];  // RAY_TO_NODE_OCCUPANCY_BITMASK_LUT
  uint32_t octant_step_result[3][3][3];   // OCTANT_STEP_RESULT_LUT
  uint64_t bitmap_mask_for_octant[8];     // BITMAP_MASK_FOR_OCTANT_LUT
  size_t bitmap_index[4][4][4];           // BITMAP_INDEX_LUT
};

// generate_lut_64_bits (lut.rs:32-85)
static void generate_lut_64_bits(uint64_t lut[64][8]) {
  std::memset(lut, 0, sizeof(uint64_t) * 64 * 8);
  for (int32_t x = 0; x < 4; ++x)
    for (int32_t y = 0; y < 4; ++y)
      for (int32_t z = 0; z < 4; ++z) {
        const size_t bitmask_position = position_in_bitmap_64bits(
            {static_cast<size_t>(x), static_cast<size_t>(y), static_cast<size_t>(z)}, 4);
        for (int32_t dx = -1; dx <= 1; ++dx)
          for (int32_t dy = -1; dy <= 1; ++dy)
            for (int32_t dz = -1; dz <= 1; ++dz) {
              if (0 == dx || 0 == dy || 0 == dz) continue;
              const uint8_t direction_position = hash_direction(
                  {static_cast<float>(dx), static_cast<float>(dy), static_cast<float>(dz)});
              auto clamp3 = [](int32_t v) { return v < 0 ? 0 : (v > 3 ? 3 : v); };
              const int32_t moved_x = clamp3(x + dx * 4), moved_y = clamp3(y + dy * 4),
                            moved_z = clamp3(z + dz * 4);
              const int32_t min_x = std::min(moved_x, x), max_x = std::max(moved_x, x);
              const int32_t min_y = std::min(moved_y, y), max_y = std::max(moved_y, y);
              const int32_t min_z = std::min(moved_z, z), max_z = std::max(moved_z, z);
              uint64_t result_bitmask = 0;
              for (int32_t bx = min_x; bx <= max_x; ++bx)
                for (int32_t by = min_y; by <= max_y; ++by)
                  for (int32_t bz = min_z; bz <= max_z; ++bz)
                    set_occupancy_in_bitmap_64bits({static_cast<size_t>(bx), static_cast<size_t>(by),
                                                    static_cast<size_t>(bz)},
                                                   1, 4, true, &result_bitmask);
              lut[bitmask_position][direction_position] = result_bitmask;
            }
      }
}

// generate_octant_step_result_lut (lut.rs:88-140)
static void generate_octant_step_result_lut(uint32_t lut[3][3][3]) {
  auto octant_after_step = [](V3c<int32_t> step_vector, size_t octant) -> uint8_t {
    const float SPACE_SIZE = 12.f;
    const V3f off = OCTANT_OFFSET_REGION_LUT[octant];
    const float cx = SPACE_SIZE / 4.f + off.x * (SPACE_SIZE / 2.f);
    const float cy = SPACE_SIZE / 4.f + off.y * (SPACE_SIZE / 2.f);
    const float cz = SPACE_SIZE / 4.f + off.z * (SPACE_SIZE / 2.f);
    const float sx = static_cast<float>(signum_i(step_vector.x));
    const float sy = static_cast<float>(signum_i(step_vector.y));
    const float sz = static_cast<float>(signum_i(step_vector.z));
    const V3f after = {cx + sx * (SPACE_SIZE / 2.f), cy + sy * (SPACE_SIZE / 2.f), cz + sz * (SPACE_SIZE / 2.f)};
    if (after.x < 0.f || after.x > SPACE_SIZE || after.y < 0.f || after.y > SPACE_SIZE || after.z < 0.f ||
        after.z > SPACE_SIZE)
      return OOB_OCTANT;
    return hash_region(after, SPACE_SIZE / 2.f);
  };
  std::memset(lut, 0, sizeof(uint32_t) * 27);
  for (size_t octant = 0; octant < 8; ++octant) {
    const uint32_t pos = static_cast<uint32_t>(4 * octant);
    for (int32_t z = -1; z <= 1; ++z)
      for (int32_t y = -1; y <= 1; ++y)
        for (int32_t x = -1; x <= 1; ++x) {
          const uint8_t r = octant_after_step({x, y, z}, octant);
          lut[x + 1][y + 1][z + 1] |= (static_cast<uint32_t>(r) & 0x0F) << pos;
        }
  }
}

// BITMAP_MASK_FOR_OCTANT_LUT (lut.rs:199-208): the 2x2x2 bitmap cells of an
// octant, as convert_8bit_bitmap_to_64bit (lut.rs:7-29) produces them
static void generate_octant_masks(uint64_t out[8]) {
  for (size_t octant = 0; octant < 8; ++octant) {
    const V3f mp = OCTANT_OFFSET_REGION_LUT[octant];
    const size_t mx = static_cast<size_t>(mp.x) * 2, my = static_cast<size_t>(mp.y) * 2,
                 mz = static_cast<size_t>(mp.z) * 2;
    uint64_t bitmap = 0;
    for (size_t x = mx; x < mx + 2; ++x)
      for (size_t y = my; y < my + 2; ++y)
        for (size_t z = mz; z < mz + 2; ++z) set_occupancy_in_bitmap_64bits({x, y, z}, 1, 4, true, &bitmap);
    out[octant] = bitmap;
  }
}

// generate_bitmap_flat_index_lut (lut.rs:143-152)
static void generate_bitmap_flat_index_lut(size_t lut[4][4][4]) {
  for (size_t x = 0; x < 4; ++x)
    for (size_t y = 0; y < 4; ++y)
      for (size_t z = 0; z < 4; ++z) lut[x][y][z] = position_in_bitmap_64bits({x, y, z}, 4);
}

static const Luts& luts() {
  static const Luts L = [] {
    Luts l;
    generate_lut_64_bits(l.ray_to_node_occupancy);
    generate_octant_step_result_lut(l.octant_step_result);
    generate_octant_masks(l.bitmap_mask_for_octant);
    generate_bitmap_flat_index_lut(l.bitmap_index);
    return l;
  }();
  return L;
}

// ---------------------------------------------------------------------------
// Cube, Ray (spatial/mod.rs:31-53, spatial/raytracing/mod.rs)

struct Cube {
  V3f min_position;
  float size;
  static Cube root_bounds(float size) { return {unit(0.f), size}; }
  // child_bounds_for (spatial/mod.rs:45-52)
  Cube child_bounds_for(uint8_t octant) const {
    const float child_size = size / 2.f;
    return {add(min_position, mul(OCTANT_OFFSET_REGION_LUT[octant], child_size)), child_size};
  }
};

struct Ray {
  V3f origin, direction;
  V3f point_at(float d) const { return add(origin, mul(direction, d)); }
};

// Cube::intersect_ray (spatial/raytracing/mod.rs:32-61); 0 None,
// 1 Some(impact_distance: None), 2 Some(Some(t))
static int intersect_ray(const Cube& c, const Ray& ray, float* out) {
  const V3f max_position = add(c.min_position, unit(c.size));
  const float t1 = (c.min_position.x - ray.origin.x) / ray.direction.x;
  const float t2 = (max_position.x - ray.origin.x) / ray.direction.x;
  const float t3 = (c.min_position.y - ray.origin.y) / ray.direction.y;
  const float t4 = (max_position.y - ray.origin.y) / ray.direction.y;
  const float t5 = (c.min_position.z - ray.origin.z) / ray.direction.z;
  const float t6 = (max_position.z - ray.origin.z) / ray.direction.z;
  const float tmin = fmax_r(fmax_r(fmin_r(t1, t2), fmin_r(t3, t4)), fmin_r(t5, t6));
  const float tmax = fmin_r(fmin_r(fmax_r(t1, t2), fmax_r(t3, t4)), fmax_r(t5, t6));
  if (tmax < 0.f || tmin > tmax) return 0;
  if (tmin < 0.0f) return 1;
  *out = tmin;
  return 2;
}

// step_octant (spatial/raytracing/mod.rs:68-80)
static uint8_t step_octant(uint8_t octant, V3f step) {
  const uint32_t pos = 4u * octant;
  const uint32_t v =
      luts().octant_step_result[signum_i(as_i32(step.x)) + 1][signum_i(as_i32(step.y)) + 1][signum_i(as_i32(step.z)) + 1];
  return static_cast<uint8_t>((v & (0x0Fu << pos)) >> pos);
}

// cube_impact_normal (spatial/raytracing/mod.rs:106-134)
static V3f cube_impact_normal(const Cube& cube, V3f impact_point) {
  const V3f mid_to_impact = sub(add(cube.min_position, unit(cube.size / 2.f)), impact_point);
  const float max_component =
      fmax_r(fmax_r(std::fabs(mid_to_impact.x), std::fabs(mid_to_impact.y)), std::fabs(mid_to_impact.z));
  const V3f impact_normal = {std::fabs(mid_to_impact.x) == max_component ? -mid_to_impact.x : 0.f,
                             std::fabs(mid_to_impact.y) == max_component ? -mid_to_impact.y : 0.f,
                             std::fabs(mid_to_impact.z) == max_component ? -mid_to_impact.z : 0.f};
  return normalized(impact_normal);
}

// ---------------------------------------------------------------------------
// NodeStack<T, SIZE> (raytracing_on_cpu.rs:19-82)

template <class T, size_t SIZE = 4>
struct NodeStack {
  T data[SIZE] = {};
  size_t head_index = 0;
  uint8_t count = 0;
  bool is_empty() const { return 0 == count; }
  void push(T d) {
    head_index = (head_index + 1) % SIZE;
    count = static_cast<uint8_t>(std::min<size_t>(count + 1, SIZE));
    data[head_index] = d;
  }
  bool pop(T* out) {
    if (0 == count) return false;
    count -= 1;
    const T result = data[head_index];
    if (head_index == 0)
      head_index = SIZE - 1;
    else
      head_index -= 1;
    if (out) *out = result;
    return true;
  }
  T* last() { return 0 == count ? nullptr : &data[head_index]; }
};

// ---------------------------------------------------------------------------
// Data model (octree/types.rs), filled from the dump

enum class BrickData : uint8_t { Empty = 0, Solid = 1, Parted = 2 };
enum class NodeContent : uint8_t { Nothing = 0, Internal = 1, Leaf = 2, UniformLeaf = 3 };
enum class NodeChildren : uint8_t { NoChildren = 0, Children = 1, OccupancyBitmap = 2 };

struct Brick {
  BrickData kind;
  uint32_t solid;
  const uint32_t* parted;  // dim^3 values
};

struct Node {
  bool reserved;
  NodeContent content;
  uint64_t internal_bits;
  NodeChildren children_kind;
  uint32_t children[8];
  uint64_t bitmap;
  Brick bricks[8];  // Leaf: 8, UniformLeaf: bricks[0]
};

struct Counters {
  uint64_t stack_iterations = 0, child_reads = 0, brick_reads = 0, voxel_tests = 0,
           data_palette_reads = 0, hits = 0, boundary_hits = 0;
};

struct Octree {
  uint32_t octree_size = 0, brick_dim = 0;
  std::vector<Node> nodes;
  std::vector<uint32_t> voxels;
  std::vector<uint32_t> colors;  // Albedo packed r<<24|g<<16|b<<8|a
  std::vector<uint32_t> data;    // user data (u32)
  // node_mips and MIPMapStrategy::enabled (types.rs:141-186)
  std::vector<Brick> mips;
  std::vector<uint32_t> mip_voxels;
  bool mips_enabled = false;

  // ObjectPool::key_is_valid (object_pool.rs:238-240)
  bool key_is_valid(size_t key) const { return key < nodes.size() && nodes[key].reserved; }
  // NodeConnection::child (node.rs:49-54); empty_marker::<usize>() otherwise
  size_t child(size_t key, uint8_t octant) const {
    const Node& n = nodes[key];
    return n.children_kind == NodeChildren::Children ? static_cast<size_t>(n.children[octant]) : SIZE_MAX;
  }
  // stored_occupied_bits (detail.rs:524-544)
  uint64_t stored_occupied_bits(size_t key) const {
    const Node& n = nodes[key];
    switch (n.content) {
      case NodeContent::Leaf:
      case NodeContent::UniformLeaf:
        return n.children_kind == NodeChildren::OccupancyBitmap ? n.bitmap : 0;
      case NodeContent::Nothing:
        return 0;
      case NodeContent::Internal:
        return n.internal_bits;
    }
    return 0;
  }
  // pix_points_to_empty (node.rs:405-427); Albedo::is_transparent (detail.rs:78-80);
  // VoxelData for Zero types (detail.rs:40-44)
  bool pix_points_to_empty(uint32_t index, Counters* c = nullptr) const {
    const size_t ci = index & 0x0000FFFFu;
    const size_t di = (index & 0xFFFF0000u) >> 16;
    const bool color_none = !(ci < 0xFFFFu);
    const bool data_none = di == 0xFFFFu;
    if (c && !data_none) c->data_palette_reads++;
    return (color_none || (colors[ci] & 0xFFu) == 0) && (data_none || data[di] == 0);
  }
};

// ---------------------------------------------------------------------------
// Ray query (raytracing_on_cpu.rs)

struct Hit {
  uint32_t pix;
  V3f impact;
  V3f normal;
  Cube cell;
};

struct Query {
  const Octree& t;
  const Ray& ray;
  Counters* c;
  uint32_t flags = 0;  // bit 0: iteration guard, bit 1: BITMAP_INDEX_LUT index out of bounds
  float viewing_distance = 3.40282347e+38f;  // f32::MAX: get_by_ray (:316-318)
};

const uint64_t MAX_STEPS = 1ull << 22;

// get_dda_scale_factors (:99-112); powf(2.) evaluated as x*x
static V3f get_dda_scale_factors(const Ray& ray) {
  const float zx = ray.direction.z / ray.direction.x, yx = ray.direction.y / ray.direction.x;
  const float xy = ray.direction.x / ray.direction.y, zy = ray.direction.z / ray.direction.y;
  const float xz = ray.direction.x / ray.direction.z, yz = ray.direction.y / ray.direction.z;
  return {std::sqrt(1.f + zx * zx + yx * yx), std::sqrt(xy * xy + 1.f + zy * zy),
          std::sqrt((xz * xz + 1.f) + yz * yz)};
}

// dda_step_to_next_sibling (:124-152)
static V3f dda_step_to_next_sibling(const Ray& ray, V3f* ray_current_point, const Cube& current_bounds,
                                    V3f ray_scale_factors) {
  const V3f signum_vec = {signum_f(ray.direction.x), signum_f(ray.direction.y), signum_f(ray.direction.z)};
  const V3f diff_from_min = sub(*ray_current_point, current_bounds.min_position);
  const V3f steps_needed = {current_bounds.size * fmax_r(signum_vec.x, 0.f) - signum_vec.x * diff_from_min.x,
                            current_bounds.size * fmax_r(signum_vec.y, 0.f) - signum_vec.y * diff_from_min.y,
                            current_bounds.size * fmax_r(signum_vec.z, 0.f) - signum_vec.z * diff_from_min.z};
  const float d_x = std::fabs(steps_needed.x * ray_scale_factors.x);
  const float d_y = std::fabs(steps_needed.y * ray_scale_factors.y);
  const float d_z = std::fabs(steps_needed.z * ray_scale_factors.z);
  const float min_step = fmin_r(fmin_r(d_x, d_y), d_z);
  *ray_current_point = add(*ray_current_point, mul(ray.direction, min_step));
  return {min_step == d_x ? signum_vec.x : 0.f, min_step == d_y ? signum_vec.y : 0.f,
          min_step == d_z ? signum_vec.z : 0.f};
}

// traverse_brick (:156-252); returns the hit index and flat index
static bool traverse_brick(Query& q, V3f* ray_current_point, const uint32_t* brick, const Cube& brick_bounds,
                           size_t brick_dim, V3f ray_scale_factors, V3c<size_t>* hit_index, size_t* hit_flat) {
  const V3f position_in_brick =
      div(mul(sub(*ray_current_point, brick_bounds.min_position), static_cast<float>(brick_dim)), brick_bounds.size);
  const int32_t hi = static_cast<int32_t>(brick_dim - 1);
  auto clamp_i = [hi](int32_t v) { return v < 0 ? 0 : (v > hi ? hi : v); };
  V3c<int32_t> current_index = {clamp_i(as_i32(position_in_brick.x)), clamp_i(as_i32(position_in_brick.y)),
                                clamp_i(as_i32(position_in_brick.z))};
  const int32_t flat_delta_x = static_cast<int32_t>(flat_projection(1, 0, 0, brick_dim));
  const int32_t flat_delta_y = static_cast<int32_t>(flat_projection(0, 1, 0, brick_dim));
  const int32_t flat_delta_z = static_cast<int32_t>(flat_projection(0, 0, 1, brick_dim));
  int32_t current_flat_index = static_cast<int32_t>(
      flat_projection(static_cast<size_t>(current_index.x), static_cast<size_t>(current_index.y),
                      static_cast<size_t>(current_index.z), brick_dim));
  const float brick_unit = brick_bounds.size / static_cast<float>(brick_dim);
  Cube current_bounds = {
      add(brick_bounds.min_position,
          mul({static_cast<float>(current_index.x), static_cast<float>(current_index.y),
               static_cast<float>(current_index.z)},
              brick_unit)),
      brick_unit};
  V3f step = unit(0.f);
  const int32_t bd = static_cast<int32_t>(brick_dim);
  for (uint64_t guard = 0;; ++guard) {
    if (guard > 4 * brick_dim + 16) {
      q.flags |= 1u;
      return false;
    }
    if (current_index.x < 0 || current_index.x >= bd || current_index.y < 0 || current_index.y >= bd ||
        current_index.z < 0 || current_index.z >= bd)
      return false;
    current_flat_index += as_i32(step.x) * flat_delta_x + as_i32(step.y) * flat_delta_y + as_i32(step.z) * flat_delta_z;
    if (q.c) q.c->voxel_tests++;
    if (!q.t.pix_points_to_empty(brick[current_flat_index], q.c)) {
      *hit_index = {static_cast<size_t>(current_index.x), static_cast<size_t>(current_index.y),
                    static_cast<size_t>(current_index.z)};
      *hit_flat = static_cast<size_t>(current_flat_index);
      return true;
    }
    step = dda_step_to_next_sibling(q.ray, ray_current_point, current_bounds, ray_scale_factors);
    current_bounds.min_position = add(current_bounds.min_position, mul(step, brick_unit));
    const V3c<int32_t> si = to_i32_round(step);
    current_index = {current_index.x + si.x, current_index.y + si.y, current_index.z + si.z};
  }
}

// probe_brick (:256-312)
static bool probe_brick(Query& q, V3f* ray_current_point, const Brick& brick, const Cube& brick_bounds,
                        V3f ray_scale_factors, Hit* hit) {
  switch (brick.kind) {
    case BrickData::Empty:
      return false;
    case BrickData::Solid:
      if (q.c) q.c->brick_reads++;
      hit->pix = brick.solid;
      hit->impact = *ray_current_point;
      hit->normal = cube_impact_normal(brick_bounds, *ray_current_point);
      hit->cell = brick_bounds;
      return true;
    case BrickData::Parted: {
      if (q.c) q.c->brick_reads++;
      V3c<size_t> idx;
      size_t flat;
      if (traverse_brick(q, ray_current_point, brick.parted, brick_bounds, q.t.brick_dim, ray_scale_factors, &idx,
                         &flat)) {
        const Cube hit_bounds = {
            add(brick_bounds.min_position,
                div(mul({static_cast<float>(idx.x), static_cast<float>(idx.y), static_cast<float>(idx.z)},
                        brick_bounds.size),
                    static_cast<float>(q.t.brick_dim))),
            brick_bounds.size / static_cast<float>(q.t.brick_dim)};
        hit->pix = brick.parted[flat];
        hit->impact = *ray_current_point;
        hit->normal = cube_impact_normal(hit_bounds, *ray_current_point);
        hit->cell = hit_bounds;
        return true;
      }
      return false;
    }
  }
  return false;
}

static inline size_t bitmap_index(Query& q, V3f p) {
  size_t x = as_usize(std::floor(p.x)), y = as_usize(std::floor(p.y)), z = as_usize(std::floor(p.z));
  if (x > 3 || y > 3 || z > 3) {  // the reference panics here (array index out of bounds)
    q.flags |= 2u;
    x = std::min<size_t>(x, 3);
    y = std::min<size_t>(y, 3);
    z = std::min<size_t>(z, 3);
  }
  return luts().bitmap_index[x][y][z];
}

// get_by_ray_at_lod (:325-565). The MIP branch (:368-386) runs only when the
// tree's MIP strategy is enabled: at every stack-loop iteration the node's MIP
// brick is probed over the node's bounds once the node's level falls below
// the ray's travelled distance (to the point snapped to the level's grid)
// divided by viewing_distance. mip_level starts at log2(size / dim), drops
// by one per PUSH and rises by one per POP; it is not reset when the ray
// restarts from the root.
static bool get_by_ray(Query& q, Hit* hit) {
  const Octree& t = q.t;
  const Ray& ray = q.ray;
  const Luts& L = luts();
  const V3f ray_scale_factors = get_dda_scale_factors(ray);
  const size_t direction_lut_index = hash_direction(ray.direction);

  NodeStack<uint32_t> node_stack;
  Cube current_bounds = Cube::root_bounds(static_cast<float>(t.octree_size));
  V3f ray_current_point;
  uint8_t target_octant;
  {
    float d = 0.f;
    const int r = intersect_ray(current_bounds, ray, &d);
    if (r != 0) {
      ray_current_point = ray.point_at(r == 2 ? d : 0.f);
      target_octant =
          hash_region(sub(ray_current_point, current_bounds.min_position), current_bounds.size / 2.f);
    } else {
      ray_current_point = ray.origin;
      target_octant = OOB_OCTANT;
    }
  }
  size_t current_node_key;
  uint64_t steps = 0;
  float mip_level = std::log2(static_cast<float>(t.octree_size) / static_cast<float>(t.brick_dim));
  while (target_octant != OOB_OCTANT) {
    current_node_key = 0;  // ROOT_NODE_KEY (detail.rs:160)
    current_bounds = Cube::root_bounds(static_cast<float>(t.octree_size));
    node_stack.push(0);
    while (!node_stack.is_empty()) {
      if (++steps > MAX_STEPS) {
        q.flags |= 1u;
        return false;
      }
      if (q.c) q.c->stack_iterations++;
      const uint64_t current_node_occupied_bits = t.stored_occupied_bits(*node_stack.last());
      const Node& node = t.nodes[current_node_key];
      bool do_backtrack_after_leaf_miss = node.content == NodeContent::UniformLeaf;
      if (t.mips_enabled) {
        const float grid = mip_level * 2.f;
        const V3f snapped = {std::round(ray_current_point.x / grid) * grid,
                             std::round(ray_current_point.y / grid) * grid,
                             std::round(ray_current_point.z / grid) * grid};
        if (mip_level < length(sub(ray.origin, snapped)) / q.viewing_distance) {
          static const Brick kEmptyMip{};
          const Brick& mip = current_node_key < t.mips.size() ? t.mips[current_node_key] : kEmptyMip;
          if (probe_brick(q, &ray_current_point, mip, current_bounds, ray_scale_factors, hit)) return true;
        }
      }
      if (target_octant != OOB_OCTANT) {
        if (node.content == NodeContent::UniformLeaf) {
          if (probe_brick(q, &ray_current_point, node.bricks[0], current_bounds, ray_scale_factors, hit)) return true;
          do_backtrack_after_leaf_miss = true;
        } else if (node.content == NodeContent::Leaf) {
          if (probe_brick(q, &ray_current_point, node.bricks[target_octant],
                          current_bounds.child_bounds_for(target_octant), ray_scale_factors, hit))
            return true;
        }
      }

      V3f bitmap_pos_in_node = div(mul(sub(ray_current_point, current_bounds.min_position),
                                       static_cast<float>(BITMAP_DIMENSION)),
                                   current_bounds.size);
      bitmap_pos_in_node = {clamp_f(bitmap_pos_in_node.x, FLOAT_ERROR_TOLERANCE, 4.f - FLOAT_ERROR_TOLERANCE),
                            clamp_f(bitmap_pos_in_node.y, FLOAT_ERROR_TOLERANCE, 4.f - FLOAT_ERROR_TOLERANCE),
                            clamp_f(bitmap_pos_in_node.z, FLOAT_ERROR_TOLERANCE, 4.f - FLOAT_ERROR_TOLERANCE)};
      size_t flat_pos_in_bitmap = bitmap_index(q, bitmap_pos_in_node);

      if (do_backtrack_after_leaf_miss || target_octant == OOB_OCTANT || 0 == current_node_occupied_bits ||
          0 == (current_node_occupied_bits & L.ray_to_node_occupancy[flat_pos_in_bitmap][direction_lut_index])) {
        // POP
        node_stack.pop(nullptr);
        mip_level += 1.f;
        if (uint32_t* parent = node_stack.last()) {
          current_node_key = *parent;
          const V3f current_bound_center = add(current_bounds.min_position, unit(current_bounds.size / 2.f));
          const float m = current_bounds.size * 2.f;
          const V3f modulo = {std::fmod(current_bounds.min_position.x, m), std::fmod(current_bounds.min_position.y, m),
                              std::fmod(current_bounds.min_position.z, m)};
          const V3f parent_bound_min_position = sub(current_bounds.min_position, modulo);
          const uint8_t from = hash_region(sub(current_bound_center, parent_bound_min_position), current_bounds.size);
          const V3f step = dda_step_to_next_sibling(ray, &ray_current_point, current_bounds, ray_scale_factors);
          target_octant = step_octant(from, step);
          current_bounds.size *= 2.f;
          current_bounds.min_position = parent_bound_min_position;
        }
        continue;
      }

      Cube target_bounds = current_bounds.child_bounds_for(target_octant);
      if (q.c) q.c->child_reads++;
      uint32_t target_child_key = static_cast<uint32_t>(t.child(current_node_key, target_octant));
      if (t.key_is_valid(target_child_key) &&
          0 != (current_node_occupied_bits & L.bitmap_mask_for_octant[target_octant])) {
        // PUSH
        current_node_key = target_child_key;
        current_bounds = target_bounds;
        target_octant = hash_region(sub(ray_current_point, target_bounds.min_position), target_bounds.size / 2.f);
        node_stack.push(target_child_key);
        mip_level -= 1.f;
      } else {
        // ADVANCE
        for (;;) {
          if (++steps > MAX_STEPS) {
            q.flags |= 1u;
            return false;
          }
          const V3f step_vec = dda_step_to_next_sibling(ray, &ray_current_point, target_bounds, ray_scale_factors);
          target_octant = step_octant(target_octant, step_vec);
          if (OOB_OCTANT != target_octant) {
            target_bounds = current_bounds.child_bounds_for(target_octant);
            if (q.c) q.c->child_reads++;
            target_child_key = static_cast<uint32_t>(t.child(current_node_key, target_octant));
            bitmap_pos_in_node = add(bitmap_pos_in_node, div(mul(step_vec, 4.f), current_bounds.size));
            flat_pos_in_bitmap = bitmap_index(q, bitmap_pos_in_node);
          }
          bool leaf_stop = false;
          if (target_octant != OOB_OCTANT && node.content == NodeContent::Leaf)
            leaf_stop = node.bricks[target_octant].kind != BrickData::Empty;
          if (target_octant == OOB_OCTANT ||
              (t.key_is_valid(target_child_key) &&
               0 != (current_node_occupied_bits & L.bitmap_mask_for_octant[target_octant]) &&
               0 != (L.ray_to_node_occupancy[flat_pos_in_bitmap][direction_lut_index] & current_node_occupied_bits)) ||
              leaf_stop)
            break;
        }
      }
    }
    // restart from the root (:551-562)
    ray_current_point = add(ray_current_point, mul(ray.direction, 0.1f));
    const float s = static_cast<float>(t.octree_size);
    target_octant = (ray_current_point.x < s && ray_current_point.y < s && ray_current_point.z < s &&
                     ray_current_point.x > 0.f && ray_current_point.y > 0.f && ray_current_point.z > 0.f)
                        ? hash_region(ray_current_point, s / 2.f)
                        : OOB_OCTANT;
  }
  return false;
}

// bound_contains (detail.rs:20-27)
static bool bound_contains(const Cube& b, V3f p) {
  return p.x >= b.min_position.x && p.x < b.min_position.x + b.size && p.y >= b.min_position.y &&
         p.y < b.min_position.y + b.size && p.z >= b.min_position.z && p.z < b.min_position.z + b.size;
}

// matrix_index_for (math/mod.rs:218-252)
static V3c<size_t> matrix_index_for(const Cube& bounds, V3f position, uint32_t dim) {
  const V3f m = div(mul(sub(position, bounds.min_position), static_cast<float>(dim)), bounds.size);
  return {as_usize(std::round(std::floor(m.x))), as_usize(std::round(std::floor(m.y))),
          as_usize(std::round(std::floor(m.z)))};
}

// Octree::get / get_internal (octree/mod.rs:209-371); returns the palette
// index value whose pix_get_ref() the reference returns, EMPTY for Empty.
static uint32_t get(const Octree& t, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t EMPTY = 0xFFFFFFFFu;
  const V3f position = {static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)};
  Cube current_bounds = Cube::root_bounds(static_cast<float>(t.octree_size));
  if (!bound_contains(current_bounds, position)) return EMPTY;
  size_t key = 0;
  for (;;) {
    const Node& n = t.nodes[key];
    switch (n.content) {
      case NodeContent::Nothing:
        return EMPTY;
      case NodeContent::Leaf: {
        const uint8_t oct = hash_region(sub(position, current_bounds.min_position), current_bounds.size / 2.f);
        const Brick& b = n.bricks[oct];
        if (b.kind == BrickData::Empty) return EMPTY;
        if (b.kind == BrickData::Solid) return b.solid;
        current_bounds = current_bounds.child_bounds_for(oct);
        const V3c<size_t> mi = matrix_index_for(current_bounds, position, t.brick_dim);
        const uint32_t v = b.parted[flat_projection(mi.x, mi.y, mi.z, t.brick_dim)];
        return t.pix_points_to_empty(v) ? EMPTY : v;
      }
      case NodeContent::UniformLeaf: {
        const Brick& b = n.bricks[0];
        if (b.kind == BrickData::Empty) return EMPTY;
        if (b.kind == BrickData::Parted) {
          const V3c<size_t> mi = matrix_index_for(current_bounds, position, t.brick_dim);
          const uint32_t v = b.parted[flat_projection(mi.x, mi.y, mi.z, t.brick_dim)];
          return t.pix_points_to_empty(v) ? EMPTY : v;
        }
        return t.pix_points_to_empty(b.solid) ? EMPTY : b.solid;
      }
      case NodeContent::Internal: {
        const uint8_t oct = hash_region(sub(position, current_bounds.min_position), current_bounds.size / 2.f);
        const size_t child = t.child(key, oct);
        if (!t.key_is_valid(child)) return EMPTY;
        key = child;
        current_bounds = current_bounds.child_bounds_for(oct);
        break;
      }
    }
  }
}

}  // namespace ora

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes)

extern "C" {

struct ora_dump {
  uint32_t octree_size, brick_dim, pool_len, parted_count, color_count, data_count;
  const uint8_t* reserved;
  const uint8_t* content_kind;
  const uint64_t* internal_bits;
  const uint8_t* children_kind;
  const uint32_t* children;
  const uint64_t* children_bitmap;
  const uint8_t* brick_kind;
  const uint32_t* brick_value;
  const uint32_t* voxels;
  const uint32_t* colors;
  const uint32_t* data;
  uint32_t mip_parted_count, mips_enabled;
  const uint8_t* mip_kind;     // [pool_len]
  const uint32_t* mip_value;   // [pool_len]: Solid value or Parted index
  const uint32_t* mip_voxels;  // [mip_parted_count * dim^3]
};

struct ora_hit {
  uint32_t hit, pix, albedo, flags;
  uint32_t voxel[3];
  float distance;
  float impact[3];
  float normal[3];
  uint32_t boundary;  // 1: a boundary hit (see boundary_hit)
};

struct ora_counters {
  uint64_t rays, stack_iterations, child_reads, brick_reads, voxel_tests, data_palette_reads, hits,
      boundary_hits;
};

void* ora_tree_load(const ora_dump* d) {
  using namespace ora;
  Octree* t = new Octree();
  t->octree_size = d->octree_size;
  t->brick_dim = d->brick_dim;
  const size_t n3 = static_cast<size_t>(d->brick_dim) * d->brick_dim * d->brick_dim;
  t->voxels.assign(d->voxels, d->voxels + static_cast<size_t>(d->parted_count) * n3);
  t->colors.assign(d->colors, d->colors + d->color_count);
  t->data.assign(d->data, d->data + d->data_count);
  t->nodes.resize(d->pool_len);
  for (uint32_t k = 0; k < d->pool_len; ++k) {
    Node& n = t->nodes[k];
    n.reserved = d->reserved[k] != 0;
    n.content = static_cast<NodeContent>(d->content_kind[k]);
    n.internal_bits = d->internal_bits[k];
    n.children_kind = static_cast<NodeChildren>(d->children_kind[k]);
    for (int o = 0; o < 8; ++o) n.children[o] = d->children[8 * k + o];
    n.bitmap = d->children_bitmap[k];
    for (int o = 0; o < 8; ++o) {
      Brick& b = n.bricks[o];
      b.kind = static_cast<BrickData>(d->brick_kind[8 * k + o]);
      b.solid = d->brick_value[8 * k + o];
      b.parted = b.kind == BrickData::Parted ? t->voxels.data() + static_cast<size_t>(d->brick_value[8 * k + o]) * n3
                                             : nullptr;
    }
  }
  t->mips_enabled = d->mips_enabled != 0;
  if (d->mip_kind) {
    t->mip_voxels.assign(d->mip_voxels, d->mip_voxels + static_cast<size_t>(d->mip_parted_count) * n3);
    t->mips.resize(d->pool_len);
    for (uint32_t k = 0; k < d->pool_len; ++k) {
      Brick& b = t->mips[k];
      b.kind = static_cast<BrickData>(d->mip_kind[k]);
      b.solid = d->mip_value[k];
      b.parted = b.kind == BrickData::Parted ? t->mip_voxels.data() + static_cast<size_t>(d->mip_value[k]) * n3
                                             : nullptr;
    }
  }
  return t;
}

void ora_tree_free(void* t) { delete static_cast<ora::Octree*>(t); }

uint32_t ora_get(const void* tp, uint32_t x, uint32_t y, uint32_t z) {
  return ora::get(*static_cast<const ora::Octree*>(tp), x, y, z);
}

void ora_get_many(const void* tp, const uint32_t* xyz, uint64_t n, uint32_t* out) {
  const ora::Octree& t = *static_cast<const ora::Octree*>(tp);
  for (uint64_t i = 0; i < n; ++i) out[i] = ora::get(t, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

// A boundary hit: the reference's f32 impact point lies within 1e-6 voxel
// units of a second face of the hit cell (the voxel cell of a Parted brick,
// the brick of a Solid one), so that the ray meets an edge or a corner of the
// cell there and the cell it is taken to enter is decided by an exact tie of
// f32 values. Per axis the distance to the nearer face of the cell is taken
// in f64; the impact lies on (or next to) its entry face, the smallest of the
// three; the second smallest must exceed 1e-6 for the hit cell to be
// unambiguous. (For coordinates of 16 and more f32 values are 2^-19 apart or
// more, so there the second face is hit exactly.)
static uint32_t boundary_hit(const ora::Hit& h) {
  const double eps = 1e-6;
  const double p[3] = {h.impact.x, h.impact.y, h.impact.z};
  const double m[3] = {h.cell.min_position.x, h.cell.min_position.y, h.cell.min_position.z};
  const double u = h.cell.size;
  double e[3];
  for (int i = 0; i < 3; ++i) e[i] = std::min(std::fabs(p[i] - m[i]), std::fabs(m[i] + u - p[i]));
  std::sort(e, e + 3);
  return e[1] <= eps ? 1u : 0u;
}

static void cast_one(const ora::Octree& t, const float* r, ora_hit* out, ora::Counters* c,
                     float viewing_distance) {
  using namespace ora;
  const Ray ray = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
  Query q{t, ray, c};
  q.viewing_distance = viewing_distance;
  Hit h{};
  const bool hit = get_by_ray(q, &h);
  out->hit = hit ? 1u : 0u;
  out->flags = q.flags;
  out->boundary = hit ? boundary_hit(h) : 0u;
  if (c && out->boundary) c->boundary_hits++;
  if (!hit) {
    out->pix = 0xFFFFFFFFu;
    out->albedo = 0;
    out->voxel[0] = out->voxel[1] = out->voxel[2] = 0xFFFFFFFFu;
    out->distance = INFINITY;
    for (int i = 0; i < 3; ++i) out->impact[i] = out->normal[i] = 0.f;
    return;
  }
  if (c) c->hits++;
  out->pix = h.pix;
  const uint32_t ci = h.pix & 0xFFFFu;
  out->albedo = (ci < 0xFFFFu && ci < t.colors.size()) ? t.colors[ci] : 0u;
  const float p[3] = {h.impact.x, h.impact.y, h.impact.z};
  const float m[3] = {h.cell.min_position.x, h.cell.min_position.y, h.cell.min_position.z};
  for (int i = 0; i < 3; ++i) {
    float f = std::floor(p[i]);
    const float hi = m[i] + h.cell.size - 1.f;
    f = f < m[i] ? m[i] : (f > hi ? hi : f);
    out->voxel[i] = static_cast<uint32_t>(as_usize(f));
  }
  out->distance = length(sub(h.impact, ray.origin));
  out->impact[0] = h.impact.x;
  out->impact[1] = h.impact.y;
  out->impact[2] = h.impact.z;
  out->normal[0] = h.normal.x;
  out->normal[1] = h.normal.y;
  out->normal[2] = h.normal.z;
}

// get_by_ray for n rays {ox,oy,oz,dx,dy,dz}; `threads` > 1 splits the rays
// into contiguous blocks over std::threads.
// get_by_ray_at_lod with the given viewing distance for n rays (f32::MAX is
// get_by_ray); `threads` as for ora_cast.
void ora_cast_lod(const void* tp, const float* rays, uint64_t n, ora_hit* out, uint32_t threads,
                  ora_counters* counters, float viewing_distance) {
  const ora::Octree& t = *static_cast<const ora::Octree*>(tp);
  if (threads <= 1) {
    ora::Counters c;
    for (uint64_t i = 0; i < n; ++i)
      cast_one(t, rays + 6 * i, out + i, counters ? &c : nullptr, viewing_distance);
    if (counters) {
      counters->rays += n;
      counters->stack_iterations += c.stack_iterations;
      counters->child_reads += c.child_reads;
      counters->brick_reads += c.brick_reads;
      counters->voxel_tests += c.voxel_tests;
      counters->data_palette_reads += c.data_palette_reads;
      counters->hits += c.hits;
      counters->boundary_hits += c.boundary_hits;
    }
    return;
  }
  std::vector<std::thread> pool;
  const uint64_t chunk = (n + threads - 1) / threads;
  for (uint32_t k = 0; k < threads; ++k) {
    const uint64_t b = k * chunk, e = std::min<uint64_t>(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&, b, e]() {
      for (uint64_t i = b; i < e; ++i) cast_one(t, rays + 6 * i, out + i, nullptr, viewing_distance);
    });
  }
  for (auto& th : pool) th.join();
}

void ora_cast(const void* tp, const float* rays, uint64_t n, ora_hit* out, uint32_t threads,
              ora_counters* counters) {
  ora_cast_lod(tp, rays, n, out, threads, counters, 3.40282347e+38f);
}

// Primary rays of a w x h pinhole viewport {origin[3], direction[3],
// frustum[3], fov}: the pixel-centre form of the reference harness's ray
// construction (examples/cpu_render.rs:78-114), row 0 at the bottom.
void ora_viewport_rays(const float* vp, uint32_t w, uint32_t h, float* rays) {
  using namespace ora;
  const V3f origin = {vp[0], vp[1], vp[2]};
  const V3f direction = {vp[3], vp[4], vp[5]};
  const float width = vp[6], height = vp[7], fov = vp[9];
  const V3f up = {0.f, 1.f, 0.f};
  const V3f right = normalized(cross(up, direction));
  const float pixel_width = width / static_cast<float>(w);
  const float pixel_height = height / static_cast<float>(h);
  const V3f bottom_left =
      sub(sub(add(origin, mul(direction, fov)), mul(up, height / 2.f)), mul(right, width / 2.f));
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const V3f glass = add(add(bottom_left, mul(mul(right, static_cast<float>(x) + 0.5f), pixel_width)),
                            mul(mul(up, static_cast<float>(y) + 0.5f), pixel_height));
      const V3f d = normalized(sub(glass, origin));
      float* r = rays + 6 * (static_cast<uint64_t>(y) * w + x);
      r[0] = origin.x;
      r[1] = origin.y;
      r[2] = origin.z;
      r[3] = d.x;
      r[4] = d.y;
      r[5] = d.z;
    }
}

// --- hooks for the reference's known-answer tests ---------------------------

void ora_luts(uint64_t* ray_lut /*64*8*/, uint32_t* step_lut /*27*/, uint64_t* octant_masks /*8*/,
              uint64_t* bitmap_index /*64*/) {
  const ora::Luts& L = ora::luts();
  std::memcpy(ray_lut, L.ray_to_node_occupancy, sizeof(L.ray_to_node_occupancy));
  std::memcpy(step_lut, L.octant_step_result, sizeof(L.octant_step_result));
  std::memcpy(octant_masks, L.bitmap_mask_for_octant, sizeof(L.bitmap_mask_for_octant));
  for (int i = 0; i < 64; ++i) bitmap_index[i] = (&L.bitmap_index[0][0][0])[i];
}

int ora_intersect_ray(const float* cube /*min xyz, size*/, const float* ray, float* t) {
  const ora::Cube c = {{cube[0], cube[1], cube[2]}, cube[3]};
  const ora::Ray r = {{ray[0], ray[1], ray[2]}, {ray[3], ray[4], ray[5]}};
  return ora::intersect_ray(c, r, t);
}

void ora_cube_impact_normal(const float* cube, const float* p, float* out) {
  const ora::Cube c = {{cube[0], cube[1], cube[2]}, cube[3]};
  const ora::V3f n = ora::cube_impact_normal(c, {p[0], p[1], p[2]});
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
}

uint32_t ora_hash_region(const float* off, float half) { return ora::hash_region({off[0], off[1], off[2]}, half); }

uint64_t ora_flat_projection(uint64_t x, uint64_t y, uint64_t z, uint64_t size) {
  return ora::flat_projection(x, y, z, size);
}

uint64_t ora_position_in_bitmap_64bits(uint64_t x, uint64_t y, uint64_t z, uint64_t brick_size) {
  return ora::position_in_bitmap_64bits({x, y, z}, brick_size);
}

uint64_t ora_set_occupancy(uint64_t x, uint64_t y, uint64_t z, uint64_t size, uint64_t brick_dim, int occupied,
                           uint64_t bitmap) {
  ora::set_occupancy_in_bitmap_64bits({x, y, z}, size, brick_dim, occupied != 0, &bitmap);
  return bitmap;
}

void ora_child_bounds(const float* cube, uint32_t octant, float* out) {
  const ora::Cube c = {{cube[0], cube[1], cube[2]}, cube[3]};
  const ora::Cube r = c.child_bounds_for(static_cast<uint8_t>(octant));
  out[0] = r.min_position.x;
  out[1] = r.min_position.y;
  out[2] = r.min_position.z;
  out[3] = r.size;
}

// NodeStack<i32, SIZE> driven by an op list: op >= 0 push(op), -1 pop,
// -2 last; results: pop/last -> value or INT32_MIN for None.
uint32_t ora_nodestack_run(uint32_t size, const int32_t* ops, uint32_t n, int32_t* results) {
  uint32_t nr = 0;
  auto run = [&](auto& st) {
    for (uint32_t i = 0; i < n; ++i) {
      if (ops[i] >= 0) {
        st.push(ops[i]);
      } else if (ops[i] == -1) {
        int32_t v;
        results[nr++] = st.pop(&v) ? v : INT32_MIN;
      } else {
        const int32_t* v = st.last();
        results[nr++] = v ? *v : INT32_MIN;
      }
    }
  };
  if (size == 3) {
    ora::NodeStack<int32_t, 3> st;
    run(st);
  } else {
    ora::NodeStack<int32_t, 4> st;
    run(st);
  }
  return nr;
}

void ora_dda_step(const float* ray, float* point, const float* cube, float* step_out) {
  const ora::Ray r = {{ray[0], ray[1], ray[2]}, {ray[3], ray[4], ray[5]}};
  const ora::Cube c = {{cube[0], cube[1], cube[2]}, cube[3]};
  ora::V3f p = {point[0], point[1], point[2]};
  const ora::V3f s = ora::dda_step_to_next_sibling(r, &p, c, ora::get_dda_scale_factors(r));
  point[0] = p.x;
  point[1] = p.y;
  point[2] = p.z;
  step_out[0] = s.x;
  step_out[1] = s.y;
  step_out[2] = s.z;
}

}  // extern "C"
