// Seeded synthetic scenes used as benchmark workloads.
//
// kind 1 (heightfield, the CPU-runnable case): a 2D value-noise heightfield
//   h(x,z) = S/8 + (3S/4) * fbm2(x,z) (4 octaves, period S/4, gain 0.5,
//   lacunarity 2, smoothstep bilinear, normalised to [0,1]); voxel solid iff
//   y <= floor(h); colour = palette16[band(y)]. Built with Octree::insert in
//   x -> y -> z order, so the tree is the one the reference's insert builds.
// kind 2 (caves): porous 3D value-noise terrain, voxel solid iff
//   fbm3(p, period 48) >= 0.40 + 0.22 y/S (denser at the bottom), minus worm
//   caves (seeded random walks carving spheres of radius 2-6); colour =
//   palette16[(band(y) + tint) mod 16] with a 4-level tint from period-8
//   noise. Built bottom-up; every leaf brick slot draws its form from the
//   seed: 40% Empty, 10% Solid (colour of the brick's height band), the rest
//   Parted from the content (Parted bricks that come out homogeneous are
//   simplified to Empty / Solid as the reference's brick simplify does).
// kind 3 (spheres): seeded spheres (radius 2-64) over a noise heightfield,
//   built bottom-up; non-empty leaf bricks are Parted or Solid 50/50.
// kind 4: the kind 1 heightfield built bottom-up (same voxels; usable at any
//   size, where inserting voxel by voxel would take too long).
// Bottom-up builds produce the reference's node model directly: Leaf nodes
// with OccupancyBitmap occupancy, Internal nodes whose occupancy bits are the
// union of their children's, and the reference's simplification rules
// applied on the way up.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <climits>
#include <cstdint>
#include <thread>
#include <vector>

#include "octree.hpp"

namespace svo {

namespace {

inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352dU;
  h ^= h >> 15;
  h *= 0x846ca68bU;
  h ^= h >> 16;
  return h;
}

inline uint32_t hash4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  uint32_t h = mix32(a * 0x9E3779B1u + 0x85EBCA77u);
  h = mix32(h ^ (b * 0xC2B2AE3Du));
  h = mix32(h ^ (c * 0x27D4EB2Fu));
  h = mix32(h ^ (d * 0x165667B1u));
  return h;
}

inline float unit_from(uint32_t h) { return static_cast<float>(h >> 8) * (1.0f / 16777216.0f); }

inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); }

// 2D value noise of one octave: lattice values at integer multiples of `period`
float value2(float x, float z, float period, uint32_t octave, uint32_t seed) {
  const float fx = x / period, fz = z / period;
  const float x0f = floorf(fx), z0f = floorf(fz);
  const int32_t x0 = static_cast<int32_t>(x0f), z0 = static_cast<int32_t>(z0f);
  const float tx = smooth(fx - x0f), tz = smooth(fz - z0f);
  auto v = [&](int32_t i, int32_t k) {
    return unit_from(hash4(static_cast<uint32_t>(i), static_cast<uint32_t>(k), octave, seed));
  };
  const float a = v(x0, z0) + (v(x0 + 1, z0) - v(x0, z0)) * tx;
  const float b = v(x0, z0 + 1) + (v(x0 + 1, z0 + 1) - v(x0, z0 + 1)) * tx;
  return a + (b - a) * tz;
}

float fbm2(float x, float z, float period, uint32_t seed) {
  float sum = 0.0f, amp = 1.0f, norm = 0.0f;
  for (uint32_t o = 0; o < 4; ++o) {
    sum += amp * value2(x, z, period, o, seed);
    norm += amp;
    amp *= 0.5f;
    period *= 0.5f;
  }
  return sum / norm;
}

float value3(float x, float y, float z, float period, uint32_t octave, uint32_t seed) {
  const float fx = x / period, fy = y / period, fz = z / period;
  const float x0f = floorf(fx), y0f = floorf(fy), z0f = floorf(fz);
  const int32_t x0 = static_cast<int32_t>(x0f), y0 = static_cast<int32_t>(y0f),
                z0 = static_cast<int32_t>(z0f);
  const float tx = smooth(fx - x0f), ty = smooth(fy - y0f), tz = smooth(fz - z0f);
  auto v = [&](int32_t i, int32_t j, int32_t k) {
    return unit_from(hash4(static_cast<uint32_t>(i) * 73856093u ^ static_cast<uint32_t>(j),
                           static_cast<uint32_t>(k), octave + 16u, seed));
  };
  auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
  const float c00 = lerp(v(x0, y0, z0), v(x0 + 1, y0, z0), tx);
  const float c10 = lerp(v(x0, y0 + 1, z0), v(x0 + 1, y0 + 1, z0), tx);
  const float c01 = lerp(v(x0, y0, z0 + 1), v(x0 + 1, y0, z0 + 1), tx);
  const float c11 = lerp(v(x0, y0 + 1, z0 + 1), v(x0 + 1, y0 + 1, z0 + 1), tx);
  return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
}

// value3 over one brick with the lattice values it needs computed once:
// fill() hashes the lattice points the brick [b, b + d)^3 reaches at one
// octave, eval() is value3 at a point of the brick, the same operations in
// the same order with the lattice values taken from the table.
struct Value3Brick {
  float period = 1.0f;
  int32_t i0 = 0, j0 = 0, k0 = 0;
  int32_t ni = 0, nj = 0, nk = 0;
  std::vector<float> v;
  void fill(float per, uint32_t octave, uint32_t seed, uint32_t bx, uint32_t by, uint32_t bz,
            uint32_t d) {
    period = per;
    auto lo = [&](uint32_t b) { return static_cast<int32_t>(floorf(static_cast<float>(b) / per)); };
    auto hi = [&](uint32_t b) {
      return static_cast<int32_t>(floorf(static_cast<float>(b + d - 1) / per)) + 1;
    };
    i0 = lo(bx);
    j0 = lo(by);
    k0 = lo(bz);
    ni = hi(bx) - i0 + 1;
    nj = hi(by) - j0 + 1;
    nk = hi(bz) - k0 + 1;
    v.resize(static_cast<size_t>(ni) * nj * nk);
    for (int32_t k = 0; k < nk; ++k)
      for (int32_t j = 0; j < nj; ++j)
        for (int32_t i = 0; i < ni; ++i)
          v[i + ni * (j + nj * k)] =
              unit_from(hash4(static_cast<uint32_t>(i0 + i) * 73856093u ^
                                  static_cast<uint32_t>(j0 + j),
                              static_cast<uint32_t>(k0 + k), octave + 16u, seed));
  }
  float eval(float x, float y, float z) const {
    const float fx = x / period, fy = y / period, fz = z / period;
    const float x0f = floorf(fx), y0f = floorf(fy), z0f = floorf(fz);
    const int32_t x0 = static_cast<int32_t>(x0f) - i0, y0 = static_cast<int32_t>(y0f) - j0,
                  z0 = static_cast<int32_t>(z0f) - k0;
    const float tx = smooth(fx - x0f), ty = smooth(fy - y0f), tz = smooth(fz - z0f);
    auto at = [&](int32_t i, int32_t j, int32_t k) { return v[i + ni * (j + nj * k)]; };
    auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
    const float c00 = lerp(at(x0, y0, z0), at(x0 + 1, y0, z0), tx);
    const float c10 = lerp(at(x0, y0 + 1, z0), at(x0 + 1, y0 + 1, z0), tx);
    const float c01 = lerp(at(x0, y0, z0 + 1), at(x0 + 1, y0, z0 + 1), tx);
    const float c11 = lerp(at(x0, y0 + 1, z0 + 1), at(x0 + 1, y0 + 1, z0 + 1), tx);
    return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
  }
};

// fbm3 over one brick (the four octaves' lattice tables)
struct Fbm3Brick {
  Value3Brick oct[4];
  void fill(float period, uint32_t seed, uint32_t bx, uint32_t by, uint32_t bz, uint32_t d) {
    for (uint32_t o = 0; o < 4; ++o) {
      oct[o].fill(period, o, seed, bx, by, bz, d);
      period *= 0.5f;
    }
  }
  // fbm3's sum, in fbm3's order
  float eval(float x, float y, float z) const {
    float sum = 0.0f, amp = 1.0f, norm = 0.0f;
    for (uint32_t o = 0; o < 4; ++o) {
      sum += amp * oct[o].eval(x, y, z);
      norm += amp;
      amp *= 0.5f;
    }
    return sum / norm;
  }
};

float fbm3(float x, float y, float z, float period, uint32_t seed) {
  float sum = 0.0f, amp = 1.0f, norm = 0.0f;
  for (uint32_t o = 0; o < 4; ++o) {
    sum += amp * value3(x, y, z, period, o, seed);
    norm += amp;
    amp *= 0.5f;
    period *= 0.5f;
  }
  return sum / norm;
}

const uint32_t PALETTE16[16] = {
    0x3B2F2FFFu, 0x5C4033FFu, 0x6B4F3AFFu, 0x7A5C43FFu, 0x8B6B4AFFu, 0x4F7942FFu,
    0x3F704DFFu, 0x2E6B3AFFu, 0x5E8C31FFu, 0x7BA05BFFu, 0x9AB973FFu, 0xA9A9A9FFu,
    0x8C8C8CFFu, 0xBEBEBEFFu, 0xDCDCDCFFu, 0xFFFAFAFFu,
};

inline uint32_t band_color(uint32_t y, uint32_t size) {
  uint32_t b = static_cast<uint32_t>((static_cast<uint64_t>(y) * 16u) / size);
  return PALETTE16[b > 15u ? 15u : b];
}

template <class F>
void parallel_for(size_t n, F&& f) {
  unsigned hw = std::thread::hardware_concurrency();
  unsigned nt = std::max(1u, std::min(hw ? hw : 1u, 16u));
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto& th : pool) th.join();
}

struct Sphere {
  float x, y, z, r;
};

// Content of one leaf brick, filled by a scene's voxel function.
struct LeafBrickPlan {
  BrickKind kind = BrickKind::Empty;
  uint32_t solid = 0;
  std::vector<uint32_t> voxels;
};

// Bottom-up assembly of a tree from per-leaf brick plans.
class BottomUpBuilder {
 public:
  BottomUpBuilder(Octree& t) : t_(t), dim_(t.brick_dim()), size_(t.size()) {}

  // `leaf(lx, ly, lz, octant, plan)` fills the brick of octant `octant` of the
  // leaf node at leaf-grid coordinate (lx, ly, lz).
  template <class LeafFn>
  void build(LeafFn&& leaf) {
    leaf_size_ = 2 * dim_;
    const uint32_t n = size_ / leaf_size_;
    const size_t count = static_cast<size_t>(n) * n * n;
    leaves_.assign(count, NodeContent{});
    leaf_bits_.assign(count, 0);
    parallel_for(count, [&](size_t i) {
      const uint32_t lx = static_cast<uint32_t>(i % n), ly = static_cast<uint32_t>((i / n) % n),
                     lz = static_cast<uint32_t>(i / (static_cast<size_t>(n) * n));
      NodeContent c;
      c.kind = NodeKind::Leaf;
      for (uint32_t o = 0; o < 8; ++o) {
        LeafBrickPlan plan;
        leaf(lx, ly, lz, o, plan);
        Brick b;
        if (plan.kind == BrickKind::Solid) {
          b = Brick::make_solid(plan.solid);
        } else if (plan.kind == BrickKind::Parted) {
          b = Brick::make_parted(std::move(plan.voxels));
        }
        c.bricks[o] = std::move(b);
      }
      leaves_[i] = std::move(c);
    });
    // Brick simplification and leaf occupancy (read-only palette lookups; the
    // node pool is not touched until build_node).
    parallel_for(count, [&](size_t i) {
      NodeContent& c = leaves_[i];
      uint64_t bits = 0;
      for (uint32_t o = 0; o < 8; ++o) {
        Brick& b = c.bricks[o];
        simplify_brick(b);
        bits |= brick_bits_in_leaf(b, o);
      }
      leaf_bits_[i] = bits;
    });
    const uint32_t root = build_node(0, 0, 0, size_, n);
    (void)root;
  }

 private:
  void simplify_brick(Brick& b) const {
    if (b.kind != BrickKind::Parted) {
      if (b.kind == BrickKind::Solid && t_.pix_points_to_empty(b.solid)) b = Brick::empty();
      return;
    }
    const uint32_t v0 = b.parted[0];
    for (uint32_t v : b.parted)
      if (v != v0) return;
    b = t_.pix_points_to_empty(v0) ? Brick::empty() : Brick::make_solid(v0);
  }

  // Occupancy bits of a brick inside its leaf node's 4x4x4 bitmap: each
  // bitmap cell covers (dim/2)^3 voxels of one brick.
  uint64_t brick_bits_in_leaf(const Brick& b, uint32_t octant) const {
    const uint64_t mask = octant_bitmap_mask(octant);
    if (b.kind == BrickKind::Empty) return 0;
    if (b.kind == BrickKind::Solid) return mask;
    uint64_t bits = 0;
    const uint32_t ox = octant & 1u, oy = (octant >> 2) & 1u, oz = (octant >> 1) & 1u;
    const uint32_t d = dim_;
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t y = 0; y < d; ++y)
        for (uint32_t x = 0; x < d; ++x) {
          if (t_.pix_points_to_empty(b.parted[x + y * d + z * d * d])) continue;
          const uint64_t p[3] = {x + ox * d, y + oy * d, z + oz * d};
          set_occupancy_in_bitmap_64bits(p, 1, 2 * d, true, &bits);
        }
    return bits;
  }

  // Returns the pool key of the node covering [origin, origin+size), or
  // EMPTY_MARKER if the region holds nothing.
  uint32_t build_node(uint32_t x0, uint32_t y0, uint32_t z0, uint32_t size, uint32_t n) {
    if (size == leaf_size_) {
      const size_t i = (x0 / leaf_size_) + static_cast<size_t>(y0 / leaf_size_) * n +
                       static_cast<size_t>(z0 / leaf_size_) * n * n;
      const uint64_t bits = leaf_bits_[i];
      NodeContent& c = leaves_[i];
      bool all_empty = true;
      for (const Brick& b : c.bricks)
        if (b.kind != BrickKind::Empty) all_empty = false;
      if (all_empty) return EMPTY_MARKER;
      const uint32_t key = push(std::move(c), NodeChildren::bitmap_of(bits));
      t_.simplify(key);
      return key;
    }
    const uint32_t half = size / 2;
    uint32_t kids[8];
    bool any = false;
    for (uint32_t o = 0; o < 8; ++o) {
      const uint32_t cx = x0 + (o & 1u) * half, cy = y0 + ((o >> 2) & 1u) * half,
                     cz = z0 + ((o >> 1) & 1u) * half;
      kids[o] = build_node(cx, cy, cz, half, n);
      any = any || kids[o] != EMPTY_MARKER;
    }
    if (!any) return EMPTY_MARKER;
    uint64_t bits = 0;
    for (uint32_t o = 0; o < 8; ++o) {
      if (kids[o] == EMPTY_MARKER) continue;
      const uint64_t cb = t_.stored_occupied_bits(kids[o]);
      bits |= fold_child_bits(cb, o);
    }
    NodeChildren nc;
    nc.kind = ChildKind::Children;
    for (uint32_t o = 0; o < 8; ++o) nc.c[o] = kids[o];
    const uint32_t key = size == size_ ? set_root(bits, nc) : push(NodeContent::internal(bits), nc);
    t_.simplify(key);
    return key;
  }

  // The 4x4x4 bitmap of a child seen from its parent: each parent cell of the
  // child's octant covers a 2x2x2 block of child cells.
  static uint64_t fold_child_bits(uint64_t child, uint32_t octant) {
    if (child == 0) return 0;
    const uint32_t ox = (octant & 1u) * 2, oy = ((octant >> 2) & 1u) * 2,
                   oz = ((octant >> 1) & 1u) * 2;
    uint64_t out = 0;
    for (uint32_t pz = 0; pz < 2; ++pz)
      for (uint32_t py = 0; py < 2; ++py)
        for (uint32_t px = 0; px < 2; ++px) {
          const uint64_t block = 0x0000000000330033ull << (px * 2 + py * 8 + pz * 32);
          if (child & block) out |= 1ull << ((ox + px) + 4 * (oy + py) + 16 * (oz + pz));
        }
    return out;
  }

  uint32_t push(NodeContent c, NodeChildren nc) {
    const uint32_t key = t_.nodes_mut().push(std::move(c));
    auto& kids = t_.node_children_mut();
    if (kids.size() < static_cast<size_t>(key) + 1) kids.resize(key + 1);
    kids[key] = nc;
    return key;
  }

  uint32_t set_root(uint64_t bits, NodeChildren nc) {
    t_.nodes_mut().get(Octree::ROOT_NODE_KEY) = NodeContent::internal(bits);
    t_.node_children_mut()[Octree::ROOT_NODE_KEY] = nc;
    return Octree::ROOT_NODE_KEY;
  }

  Octree& t_;
  uint32_t dim_, size_, leaf_size_ = 0;
  std::vector<NodeContent> leaves_;
  std::vector<uint64_t> leaf_bits_;
};

Status build_heightfield(Octree& t, uint64_t seed) {
  const uint32_t s = t.size();
  const float period = static_cast<float>(s) / 4.0f;
  const uint32_t sd = static_cast<uint32_t>(seed ^ (seed >> 32));
  std::vector<int32_t> height(static_cast<size_t>(s) * s);
  for (uint32_t x = 0; x < s; ++x)
    for (uint32_t z = 0; z < s; ++z) {
      const float h = static_cast<float>(s) / 8.0f +
                      (3.0f * static_cast<float>(s) / 4.0f) *
                          fbm2(static_cast<float>(x), static_cast<float>(z), period, sd);
      height[x + static_cast<size_t>(z) * s] = static_cast<int32_t>(floorf(h));
    }
  Entry e;
  e.kind = EntryKind::Visual;
  for (uint32_t x = 0; x < s; ++x)
    for (uint32_t y = 0; y < s; ++y)
      for (uint32_t z = 0; z < s; ++z) {
        if (static_cast<int32_t>(y) > height[x + static_cast<size_t>(z) * s]) continue;
        e.albedo = band_color(y, s);
        const Status st = t.insert(x, y, z, e);
        if (st != Status::Ok) return st;
      }
  return Status::Ok;
}

std::vector<Sphere> worms(uint32_t size, uint32_t seed) {
  std::vector<Sphere> out;
  const uint32_t n_worms = std::max(8u, size / 8u);
  const uint32_t steps = 200;
  for (uint32_t w = 0; w < n_worms; ++w) {
    float x = unit_from(hash4(w, 1, 7, seed)) * static_cast<float>(size);
    float y = unit_from(hash4(w, 2, 7, seed)) * static_cast<float>(size) * 0.5f;
    float z = unit_from(hash4(w, 3, 7, seed)) * static_cast<float>(size);
    float yaw = unit_from(hash4(w, 4, 7, seed)) * 6.2831853f;
    float pitch = 0.0f;
    for (uint32_t s = 0; s < steps; ++s) {
      const float r = 2.0f + 4.0f * unit_from(hash4(w, s, 8, seed));
      out.push_back(Sphere{x, y, z, r});
      yaw += (unit_from(hash4(w, s, 9, seed)) - 0.5f) * 0.8f;
      pitch = std::max(-0.6f, std::min(0.6f, pitch + (unit_from(hash4(w, s, 10, seed)) - 0.5f) *
                                                         0.4f));
      x += 3.0f * cosf(yaw) * cosf(pitch);
      z += 3.0f * sinf(yaw) * cosf(pitch);
      y += 3.0f * sinf(pitch);
    }
  }
  return out;
}

// Spheres bucketed per leaf node, for voxel tests limited to nearby spheres.
struct SphereGrid {
  uint32_t n = 0, cell = 1;
  std::vector<std::vector<uint32_t>> buckets;
  const std::vector<Sphere>* spheres = nullptr;
  void build(const std::vector<Sphere>& s, uint32_t size, uint32_t cell_size) {
    spheres = &s;
    cell = cell_size;
    n = size / cell_size;
    buckets.assign(static_cast<size_t>(n) * n * n, {});
    for (uint32_t i = 0; i < s.size(); ++i) {
      const Sphere& q = s[i];
      auto rng = [&](float c, int32_t* lo, int32_t* hi) {
        *lo = std::max<int32_t>(0, static_cast<int32_t>(floorf((c - q.r) / cell)));
        *hi = std::min<int32_t>(static_cast<int32_t>(n) - 1,
                                static_cast<int32_t>(floorf((c + q.r) / cell)));
      };
      int32_t x0, x1, y0, y1, z0, z1;
      rng(q.x, &x0, &x1);
      rng(q.y, &y0, &y1);
      rng(q.z, &z0, &z1);
      for (int32_t z = z0; z <= z1; ++z)
        for (int32_t y = y0; y <= y1; ++y)
          for (int32_t x = x0; x <= x1; ++x)
            buckets[x + static_cast<size_t>(y) * n + static_cast<size_t>(z) * n * n].push_back(i);
    }
  }
  const std::vector<uint32_t>& at(uint32_t lx, uint32_t ly, uint32_t lz) const {
    return buckets[lx + static_cast<size_t>(ly) * n + static_cast<size_t>(lz) * n * n];
  }
};

Status build_caves(Octree& t, uint64_t seed) {
  const uint32_t s = t.size(), d = t.brick_dim();
  const uint32_t sd = static_cast<uint32_t>(seed ^ (seed >> 32));
  const float fs = static_cast<float>(s);
  const std::vector<Sphere> caves = worms(s, sd);
  SphereGrid grid;
  grid.build(caves, s, 2 * d);
  std::vector<uint32_t> pix(16);
  for (uint32_t i = 0; i < 16; ++i) pix[i] = t.add_to_palette(Entry{EntryKind::Visual, PALETTE16[i], 0});
  auto band = [&](uint32_t y) {
    const uint32_t b = static_cast<uint32_t>((static_cast<uint64_t>(y) * 16u) / s);
    return b > 15u ? 15u : b;
  };
  BottomUpBuilder b(t);
  b.build([&](uint32_t lx, uint32_t ly, uint32_t lz, uint32_t o, LeafBrickPlan& plan) {
    const uint32_t bx = lx * 2 * d + (o & 1u) * d, by = ly * 2 * d + ((o >> 2) & 1u) * d,
                   bz = lz * 2 * d + ((o >> 1) & 1u) * d;
    const uint32_t slot = (bx / d) + (by / d) * 4096u + (bz / d) * 16777216u;
    const float u = unit_from(hash4(slot, 0x5EEDu, 99, sd));
    if (u < 0.40f) {
      plan.kind = BrickKind::Empty;
      return;
    }
    if (u < 0.50f) {
      plan.kind = BrickKind::Solid;
      plan.solid = pix[band(by + d / 2)];
      return;
    }
    plan.kind = BrickKind::Parted;
    plan.voxels.assign(static_cast<size_t>(d) * d * d, EMPTY_MARKER);
    const auto& near = grid.at(lx, ly, lz);
    // fbm3 and value3 of the brick's voxels with each lattice point hashed
    // once (the same values as evaluating them point by point)
    thread_local Fbm3Brick density;
    thread_local Value3Brick tint_noise;
    density.fill(48.0f, sd, bx, by, bz, d);
    tint_noise.fill(8.0f, 5u, sd, bx, by, bz, d);
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t y = 0; y < d; ++y)
        for (uint32_t x = 0; x < d; ++x) {
          const float px = static_cast<float>(bx + x), py = static_cast<float>(by + y),
                      pz = static_cast<float>(bz + z);
          // porous density: denser at the bottom, sparser towards the top
          const float dens = density.eval(px, py, pz);
          if (dens < 0.40f + 0.22f * (py / fs)) continue;
          bool carved = false;
          for (uint32_t si : near) {
            const Sphere& q = caves[si];
            const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
            if (dx * dx + dy * dy + dz * dz <= q.r * q.r) {
              carved = true;
              break;
            }
          }
          if (carved) continue;
          const uint32_t tint = static_cast<uint32_t>(4.0f * tint_noise.eval(px, py, pz)) & 3u;
          plan.voxels[x + y * d + z * d * d] = pix[(band(by + y) + tint) & 15u];
        }
  });
  return Status::Ok;
}

Status build_spheres(Octree& t, uint64_t seed) {
  const uint32_t s = t.size(), d = t.brick_dim();
  const uint32_t sd = static_cast<uint32_t>(seed ^ (seed >> 32));
  const float fs = static_cast<float>(s);
  std::vector<float> height(static_cast<size_t>(s) * s);
  for (uint32_t x = 0; x < s; ++x)
    for (uint32_t z = 0; z < s; ++z)
      height[x + static_cast<size_t>(z) * s] =
          fs * 0.05f + fs * 0.1f * fbm2(static_cast<float>(x), static_cast<float>(z), fs / 4.0f, sd);
  const uint32_t n_spheres = std::max(16u, static_cast<uint32_t>(20000.0 * (double(s) / 2048.0) *
                                                               (double(s) / 2048.0)));
  std::vector<Sphere> balls;
  balls.reserve(n_spheres);
  for (uint32_t i = 0; i < n_spheres; ++i) {
    const float r = 2.0f + 62.0f * unit_from(hash4(i, 5, 11, sd)) * std::min(1.0f, fs / 2048.0f);
    balls.push_back(Sphere{unit_from(hash4(i, 1, 11, sd)) * fs, unit_from(hash4(i, 2, 11, sd)) * fs,
                           unit_from(hash4(i, 3, 11, sd)) * fs, r});
  }
  SphereGrid grid;
  grid.build(balls, s, 2 * d);
  std::vector<uint32_t> pix(16);
  for (uint32_t i = 0; i < 16; ++i) pix[i] = t.add_to_palette(Entry{EntryKind::Visual, PALETTE16[i], 0});
  BottomUpBuilder b(t);
  b.build([&](uint32_t lx, uint32_t ly, uint32_t lz, uint32_t o, LeafBrickPlan& plan) {
    const uint32_t bx = lx * 2 * d + (o & 1u) * d, by = ly * 2 * d + ((o >> 2) & 1u) * d,
                   bz = lz * 2 * d + ((o >> 1) & 1u) * d;
    const auto& near = grid.at(lx, ly, lz);
    float hmax = 0.0f;
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t x = 0; x < d; ++x)
        hmax = std::max(hmax, height[(bx + x) + static_cast<size_t>(bz + z) * s]);
    if (near.empty() && static_cast<float>(by) > hmax) {
      plan.kind = BrickKind::Empty;
      return;
    }
    std::vector<uint32_t> vox(static_cast<size_t>(d) * d * d, EMPTY_MARKER);
    bool any = false;
    uint32_t first = EMPTY_MARKER;
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t y = 0; y < d; ++y)
        for (uint32_t x = 0; x < d; ++x) {
          const float px = static_cast<float>(bx + x) + 0.5f, py = static_cast<float>(by + y) + 0.5f,
                      pz = static_cast<float>(bz + z) + 0.5f;
          uint32_t v = EMPTY_MARKER;
          if (py <= height[(bx + x) + static_cast<size_t>(bz + z) * s]) {
            v = pix[(static_cast<uint64_t>(by + y) * 16u) / s > 15 ? 15 : (static_cast<uint64_t>(by + y) * 16u) / s];
          } else {
            for (uint32_t si : near) {
              const Sphere& q = balls[si];
              const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
              if (dx * dx + dy * dy + dz * dz <= q.r * q.r) {
                v = pix[si % 16];
                break;
              }
            }
          }
          if (v != EMPTY_MARKER) {
            if (!any) first = v;
            any = true;
          }
          vox[x + y * d + z * d * d] = v;
        }
    if (!any) {
      plan.kind = BrickKind::Empty;
      return;
    }
    const uint32_t slot = (bx / d) + (by / d) * 4096u + (bz / d) * 16777216u;
    if (unit_from(hash4(slot, 77, 12, sd)) < 0.5f) {
      plan.kind = BrickKind::Solid;
      plan.solid = first;
      return;
    }
    plan.kind = BrickKind::Parted;
    plan.voxels = std::move(vox);
  });
  return Status::Ok;
}

// The heightfield content of kind 1 built bottom-up (kind 4): the same
// heights and band colours, one height evaluation per column, for trees too
// large to build voxel by voxel with insert.
Status build_heightfield_bottom_up(Octree& t, uint64_t seed) {
  const uint32_t s = t.size(), d = t.brick_dim();
  const float period = static_cast<float>(s) / 4.0f;
  const uint32_t sd = static_cast<uint32_t>(seed ^ (seed >> 32));
  std::vector<int32_t> height(static_cast<size_t>(s) * s);
  parallel_for(s, [&](size_t x) {
    for (uint32_t z = 0; z < s; ++z) {
      const float h = static_cast<float>(s) / 8.0f +
                      (3.0f * static_cast<float>(s) / 4.0f) *
                          fbm2(static_cast<float>(x), static_cast<float>(z), period, sd);
      height[x + static_cast<size_t>(z) * s] = static_cast<int32_t>(floorf(h));
    }
  });
  std::vector<uint32_t> pix(16);
  for (uint32_t i = 0; i < 16; ++i)
    pix[i] = t.add_to_palette(Entry{EntryKind::Visual, PALETTE16[i], 0});
  auto band = [&](uint32_t y) {
    const uint32_t b = static_cast<uint32_t>((static_cast<uint64_t>(y) * 16u) / s);
    return b > 15u ? 15u : b;
  };
  BottomUpBuilder b(t);
  b.build([&](uint32_t lx, uint32_t ly, uint32_t lz, uint32_t o, LeafBrickPlan& plan) {
    const uint32_t bx = lx * 2 * d + (o & 1u) * d, by = ly * 2 * d + ((o >> 2) & 1u) * d,
                   bz = lz * 2 * d + ((o >> 1) & 1u) * d;
    int32_t hmin = INT32_MAX, hmax = INT32_MIN;
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t x = 0; x < d; ++x) {
        const int32_t h = height[(bx + x) + static_cast<size_t>(bz + z) * s];
        hmin = std::min(hmin, h);
        hmax = std::max(hmax, h);
      }
    if (static_cast<int32_t>(by) > hmax) {
      plan.kind = BrickKind::Empty;
      return;
    }
    if (static_cast<int32_t>(by + d - 1) <= hmin && band(by) == band(by + d - 1)) {
      // below the surface in one colour band: the Solid brick the builder's
      // simplification would make of the filled Parted one
      plan.kind = BrickKind::Solid;
      plan.solid = pix[band(by)];
      return;
    }
    plan.kind = BrickKind::Parted;
    plan.voxels.assign(static_cast<size_t>(d) * d * d, EMPTY_MARKER);
    for (uint32_t z = 0; z < d; ++z)
      for (uint32_t y = 0; y < d; ++y)
        for (uint32_t x = 0; x < d; ++x) {
          const int32_t h = height[(bx + x) + static_cast<size_t>(bz + z) * s];
          if (static_cast<int32_t>(by + y) <= h)
            plan.voxels[x + y * d + z * d * d] = pix[band(by + y)];
        }
    // (bricks that come out homogeneous are simplified by the builder)
  });
  return Status::Ok;
}

}  // namespace

Status build_scene(uint32_t kind, uint32_t size, uint32_t brick_dim, uint64_t seed, Octree* out) {
  Status st;
  *out = Octree::create(size, brick_dim, &st);
  if (st != Status::Ok) return st;
  switch (kind) {
    case 1:
      return build_heightfield(*out, seed);
    case 2:
      return build_caves(*out, seed);
    case 3:
      return build_spheres(*out, seed);
    case 4:
      return build_heightfield_bottom_up(*out, seed);
    default:
      return Status::InvalidStructure;
  }
}

}  // namespace svo
