// Host octree model: construction, point query, insert/update/clear and the
// simplification rules of shocovox-rs 0.11.1.  See octree.hpp.
#include "octree.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

namespace svo {

namespace {

constexpr float BITMAP_DIMENSION_F = 4.0f;

inline V3f to_f(const uint32_t p[3]) {
  return V3f{static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2])};
}

// V3c<f32> -> V3c<usize> / V3c<u32> conversions round (vector.rs:306-336)
inline uint64_t round_usize(float v) { return f2usize(roundf(v)); }
inline uint32_t round_u32(float v) { return f2u32(roundf(v)); }

// bound_contains (detail.rs:20-27)
inline bool bound_contains(Cube b, V3f p) {
  return p.x >= b.min.x && p.x < b.min.x + b.size && p.y >= b.min.y && p.y < b.min.y + b.size &&
         p.z >= b.min.z && p.z < b.min.z + b.size;
}

// child_octant_for (detail.rs:30-38)
inline uint32_t child_octant_for(Cube b, V3f p) { return hash_region(p - b.min, b.size / 2.0f); }

// Target bounds as the edit paths spell them: min + (offset * size) / 2
// (insert.rs:129-134, clear.rs:77-82)
inline Cube edit_target_bounds(Cube c, uint32_t octant) {
  const V3f off = octant_offset(octant);
  return Cube{c.min + (off * c.size) / 2.0f, c.size / 2.0f};
}

// matrix_index_for (math/mod.rs:218-252)
inline void matrix_index_for(Cube b, const uint32_t pos[3], uint32_t dim, uint64_t out[3]) {
  const V3f rel = ((to_f(pos) - b.min) * static_cast<float>(dim)) / b.size;
  out[0] = round_usize(floorf(rel.x));
  out[1] = round_usize(floorf(rel.y));
  out[2] = round_usize(floorf(rel.z));
}

// #[derive(PartialOrd)] on V3c compares (x, y, z) lexicographically.
inline bool lex_le_f(V3f a, V3f b) {
  if (a.x != b.x) return a.x < b.x;
  if (a.y != b.y) return a.y < b.y;
  return a.z <= b.z;
}
inline bool lex_le_u(const uint32_t a[3], const uint32_t b[3]) {
  if (a[0] != b[0]) return a[0] < b[0];
  if (a[1] != b[1]) return a[1] < b[1];
  return a[2] <= b[2];
}

inline bool is_leaf(NodeKind k) { return k == NodeKind::Leaf || k == NodeKind::UniformLeaf; }

// Octant offset as V3c<usize> (round of the LUT floats)
inline void offset_usize(uint32_t octant, uint64_t out[3]) {
  out[0] = octant & 1u;
  out[1] = (octant >> 2) & 1u;
  out[2] = (octant >> 1) & 1u;
}

// get_homogeneous_data (node.rs:300-313)
inline const uint32_t* homogeneous(const Brick& b) {
  switch (b.kind) {
    case BrickKind::Empty:
      return nullptr;
    case BrickKind::Solid:
      return &b.solid;
    case BrickKind::Parted: {
      // the values xor the first, or-ed over blocks of 64 (a reduction the
      // compiler vectorizes) with one test per block
      const uint32_t* p = b.parted.data();
      const size_t n = b.parted.size();
      const uint32_t f = p[0];
      size_t i = 0;
      for (; i + 64 <= n; i += 64) {
        uint32_t diff = 0;
        for (size_t j = 0; j < 64; ++j) diff |= p[i + j] ^ f;
        if (diff) return nullptr;
      }
      for (; i < n; ++i)
        if (p[i] != f) return nullptr;
      return &b.parted[0];
    }
  }
  return nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// spatial/math helpers (math/mod.rs:35-162)

uint64_t flat_projection(uint64_t x, uint64_t y, uint64_t z, uint64_t size) {
  return x + y * size + z * size * size;
}

uint64_t position_in_bitmap_64bits(const uint64_t idx[3], uint64_t brick_size) {
  return flat_projection(idx[0] * 4 / brick_size, idx[1] * 4 / brick_size,
                         idx[2] * 4 / brick_size, 4);
}

void set_occupancy_in_bitmap_64bits(const uint64_t pos[3], uint64_t size, uint64_t brick_dim,
                                    bool occupied, uint64_t* bitmap) {
  if (brick_dim == 1) {
    *bitmap = occupied ? ~0ull : 0ull;
    return;
  }
  const uint64_t count = f2usize(
      ceilf((static_cast<float>(size) * BITMAP_DIMENSION_F) / static_cast<float>(brick_dim)));
  uint64_t start[3];
  for (int i = 0; i < 3; ++i)
    start[i] =
        round_usize(floorf(static_cast<float>(pos[i] * 4) / static_cast<float>(brick_dim)));
  for (uint64_t x = start[0]; x < std::min<uint64_t>(start[0] + count, 4); ++x)
    for (uint64_t y = start[1]; y < std::min<uint64_t>(start[1] + count, 4); ++y)
      for (uint64_t z = start[2]; z < std::min<uint64_t>(start[2] + count, 4); ++z) {
        const uint64_t p[3] = {x, y, z};
        const uint64_t mask = 1ull << position_in_bitmap_64bits(p, 4);
        if (occupied)
          *bitmap |= mask;
        else
          *bitmap &= ~mask;
      }
}

// ---------------------------------------------------------------------------
// NodeChildren (node.rs:48-84)

uint32_t& NodeChildren::child_mut(uint32_t octant) {
  if (kind == ChildKind::NoChildren) {
    kind = ChildKind::Children;
    for (auto& v : c) v = EMPTY_MARKER;
  }
  // The reference panics on OccupancyBitmap; the edit paths never reach it.
  return c[octant];
}

void NodeChildren::clear(uint32_t octant) {
  if (kind != ChildKind::Children) return;
  c[octant] = EMPTY_MARKER;
  for (uint32_t v : c)
    if (v != EMPTY_MARKER) return;
  kind = ChildKind::NoChildren;
}

bool Entry::is_none() const {
  switch (kind) {
    case EntryKind::Empty:
      return true;
    case EntryKind::Visual:
      return (albedo & 0xFFu) == 0;
    case EntryKind::Informative:
      return data == 0;
    case EntryKind::Complex:
      return (albedo & 0xFFu) == 0 && data == 0;
  }
  return true;
}

// ---------------------------------------------------------------------------
// NodePool (object_pool.rs:140-241)

bool NodePool::is_next_available() const {
  return first_available_ + 1 < items_.size() && !reserved_[first_available_ + 1];
}

bool NodePool::check_first_available() {
  if (first_available_ < items_.size() && !reserved_[first_available_]) return true;
  if (is_next_available()) {
    first_available_ += 1;
    return true;
  }
  first_available_ = items_.size();
  return false;
}

uint32_t NodePool::push(NodeContent item) {
  size_t key;
  if (check_first_available()) {
    reserved_[first_available_] = 1;
    key = first_available_;
  } else {
    items_.emplace_back();
    reserved_.push_back(1);
    key = items_.size() - 1;
  }
  if (is_next_available()) first_available_ += 1;
  items_[key] = std::move(item);
  touch(static_cast<uint32_t>(key));
  return static_cast<uint32_t>(key);
}

bool NodePool::free(uint32_t key) {
  if (!key_is_valid(key)) return false;
  reserved_[key] = 0;
  first_available_ = std::min<size_t>(first_available_, key);
  touch(key);
  return true;
}

void NodePool::swap(uint32_t a, uint32_t b) {
  touch(a);
  touch(b);
  std::swap(items_[a], items_[b]);
  std::swap(reserved_[a], reserved_[b]);
}

// ---------------------------------------------------------------------------
// Octree construction and palette

Octree Octree::create(uint32_t size, uint32_t brick_dim, Status* status) {
  Octree t;
  *status = Status::Ok;
  // (brick_dimension as f32).log(2.0).fract() != 0.0  (octree/mod.rs:174)
  const float ld = logf(static_cast<float>(brick_dim)) / logf(2.0f);
  if (size == 0 || (ld - truncf(ld)) != 0.0f || ld != ld) {
    *status = Status::InvalidBrickDimension;
    return t;
  }
  const float ls = logf(static_cast<float>(size) / static_cast<float>(brick_dim)) / logf(2.0f);
  if (brick_dim > size || (ls - truncf(ls)) != 0.0f || ls != ls) {
    *status = Status::InvalidSize;
    return t;
  }
  if (size < brick_dim * 2) {
    *status = Status::InvalidStructure;
    return t;
  }
  t.octree_size_ = size;
  t.brick_dim_ = brick_dim;
  t.nodes_.push(NodeContent::nothing());
  t.node_children_.assign(1, NodeChildren{});
  return t;
}

uint32_t Octree::add_to_palette(const Entry& e) {
  auto color = [&](uint32_t albedo) {
    auto it = color_index_.find(albedo);
    if (it != color_index_.end()) return it->second;
    const uint32_t idx = static_cast<uint32_t>(color_index_.size());
    color_index_.emplace(albedo, idx);
    color_palette_.push_back(albedo);
    return idx;
  };
  auto data = [&](uint32_t d) {
    auto it = data_index_.find(d);
    if (it != data_index_.end()) return it->second;
    const uint32_t idx = static_cast<uint32_t>(data_index_.size());
    data_index_.emplace(d, idx);
    data_palette_.push_back(d);
    return idx;
  };
  switch (e.kind) {
    case EntryKind::Empty:
      return EMPTY_MARKER;
    case EntryKind::Visual:
      if (e.albedo == 0) return EMPTY_MARKER;
      return (color(e.albedo) & 0xFFFFu) | (NONE16 << 16);
    case EntryKind::Informative:
      if (e.data == 0) return EMPTY_MARKER;
      return NONE16 | ((data(e.data) & 0xFFFFu) << 16);
    case EntryKind::Complex: {
      if (e.albedo == 0) return add_to_palette(Entry{EntryKind::Informative, 0, e.data});
      if (e.data == 0) return add_to_palette(Entry{EntryKind::Visual, e.albedo, 0});
      const uint32_t ci = color(e.albedo);
      const uint32_t di = data(e.data);
      return (ci & 0xFFFFu) | ((di & 0xFFFFu) << 16);
    }
  }
  return EMPTY_MARKER;
}

// Octree::decode_bencode_object (convert/bytecode.rs:580-660): palettes and
// their reverse maps (built in index order, so a later duplicate wins)
void Octree::assign_model(std::vector<NodeContent> items, std::vector<uint8_t> reserved,
                          size_t first_available, std::vector<NodeChildren> children,
                          std::vector<uint32_t> colors, std::vector<uint32_t> data,
                          bool simplify) {
  nodes_.assign(std::move(items), std::move(reserved), first_available);
  node_children_ = std::move(children);
  color_palette_ = std::move(colors);
  data_palette_ = std::move(data);
  color_index_.clear();
  data_index_.clear();
  for (uint32_t i = 0; i < color_palette_.size(); ++i) color_index_[color_palette_[i]] = i;
  for (uint32_t i = 0; i < data_palette_.size(); ++i) data_index_[data_palette_[i]] = i;
  auto_simplify = simplify;
  ++edit_epoch_;
}

// pix_get_ref (node.rs:429-467)
Entry Octree::pix_get_ref(uint32_t pix) const {
  const uint32_t c = pix_color_index(pix), d = pix_data_index(pix);
  const bool cn = c >= NONE16, dn = d == NONE16;
  Entry e;
  if (cn && dn) return e;
  if (dn) {
    e.kind = EntryKind::Visual;
    e.albedo = c < color_palette_.size() ? color_palette_[c] : 0;
    return e;
  }
  if (cn) {
    e.kind = EntryKind::Informative;
    e.data = d < data_palette_.size() ? data_palette_[d] : 0;
    return e;
  }
  e.kind = EntryKind::Complex;
  e.albedo = c < color_palette_.size() ? color_palette_[c] : 0;
  e.data = d < data_palette_.size() ? data_palette_[d] : 0;
  return e;
}

// ---------------------------------------------------------------------------
// Point query (octree/mod.rs:209-371)

Entry Octree::get(uint32_t x, uint32_t y, uint32_t z) const {
  const uint32_t pos[3] = {x, y, z};
  return get_internal(ROOT_NODE_KEY, Cube{v3u(0.0f), static_cast<float>(octree_size_)}, pos);
}

// get_internal (octree/mod.rs:219-371): the point query from a given node
Entry Octree::get_internal(uint32_t key, Cube bounds, const uint32_t pos[3]) const {
  const V3f p = to_f(pos);
  if (!bound_contains(bounds, p)) return Entry{};
  const uint32_t dim = brick_dim_;
  for (;;) {
    const NodeContent& n = nodes_.get(key);
    switch (n.kind) {
      case NodeKind::Nothing:
        return Entry{};
      case NodeKind::Leaf: {
        const uint32_t oct = child_octant_for(bounds, p);
        const Brick& b = n.bricks[oct];
        if (b.kind == BrickKind::Empty) return Entry{};
        if (b.kind == BrickKind::Solid) return pix_get_ref(b.solid);
        bounds = child_bounds_for(bounds, oct);
        uint64_t mi[3];
        matrix_index_for(bounds, pos, dim, mi);
        const uint64_t f = flat_projection(mi[0], mi[1], mi[2], dim);
        if (!pix_points_to_empty(b.parted[f])) return pix_get_ref(b.parted[f]);
        return Entry{};
      }
      case NodeKind::UniformLeaf: {
        const Brick& b = n.bricks[0];
        if (b.kind == BrickKind::Empty) return Entry{};
        if (b.kind == BrickKind::Parted) {
          uint64_t mi[3];
          matrix_index_for(bounds, pos, dim, mi);
          const uint64_t f = flat_projection(mi[0], mi[1], mi[2], dim);
          if (pix_points_to_empty(b.parted[f])) return Entry{};
          return pix_get_ref(b.parted[f]);
        }
        if (pix_points_to_empty(b.solid)) return Entry{};
        return pix_get_ref(b.solid);
      }
      case NodeKind::Internal: {
        const uint32_t oct = child_octant_for(bounds, p);
        const uint32_t child = node_children_[key].child(oct);
        if (!nodes_.key_is_valid(child)) return Entry{};
        key = child;
        bounds = child_bounds_for(bounds, oct);
        break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Occupancy bookkeeping (detail.rs:524-569)

uint64_t Octree::stored_occupied_bits(uint32_t key) const {
  const NodeContent& n = nodes_.get(key);
  switch (n.kind) {
    case NodeKind::Leaf:
    case NodeKind::UniformLeaf: {
      const NodeChildren& c = node_children_[key];
      return c.kind == ChildKind::OccupancyBitmap ? c.bitmap : 0;
    }
    case NodeKind::Nothing:
      return 0;
    case NodeKind::Internal:
      return n.bits;
  }
  return 0;
}

void Octree::store_occupied_bits(uint32_t key, uint64_t bits) {
  NodeContent& n = nodes_.get(key);
  switch (n.kind) {
    case NodeKind::Internal:
      n.bits = bits;
      break;
    case NodeKind::Nothing:
      kids(key) = NodeChildren::bitmap_of(bits);
      break;
    case NodeKind::Leaf:
    case NodeKind::UniformLeaf:
      // The reference panics for Children here; the edit paths never reach it.
      if (node_children_[key].kind != ChildKind::Children)
        kids(key) = NodeChildren::bitmap_of(bits);
      break;
  }
}

// ---------------------------------------------------------------------------
// Brick helpers (node.rs:104-332)

bool Octree::brick_is_empty_throughout(const Brick& b, uint32_t octant) const {
  const uint64_t dim = brick_dim_;
  switch (b.kind) {
    case BrickKind::Empty:
      return true;
    case BrickKind::Solid:
      return pix_points_to_empty(b.solid);
    case BrickKind::Parted: {
      if (dim == 1) return pix_points_to_empty(b.parted[0]);
      uint64_t off[3];
      offset_usize(octant, off);
      if (dim == 2) return pix_points_to_empty(b.parted[flat_projection(off[0], off[1], off[2], 2)]);
      const uint64_t ext = dim / 2;
      for (uint64_t x = off[0] * ext; x < off[0] * ext + ext; ++x)
        for (uint64_t y = off[1] * ext; y < off[1] * ext + ext; ++y)
          for (uint64_t z = off[2] * ext; z < off[2] * ext + ext; ++z)
            if (!pix_points_to_empty(b.parted[flat_projection(x, y, z, dim)])) return false;
      return true;
    }
  }
  return true;
}

bool Octree::brick_is_part_empty_throughout(const Brick& b, uint32_t part_octant,
                                            uint32_t target_octant) const {
  const uint64_t dim = brick_dim_;
  switch (b.kind) {
    case BrickKind::Empty:
      return true;
    case BrickKind::Solid:
      return pix_points_to_empty(b.solid);
    case BrickKind::Parted: {
      if (dim == 1) return pix_points_to_empty(b.parted[0]);
      if (dim == 2) {
        uint64_t off[3];
        offset_usize(part_octant, off);
        return pix_points_to_empty(b.parted[flat_projection(off[0], off[1], off[2], 2)]);
      }
      const float outer = static_cast<float>(dim) / 2.0f;
      const float inner = static_cast<float>(dim) / 4.0f;
      const V3f o = octant_offset(part_octant) * outer + octant_offset(target_octant) * inner;
      const uint64_t off[3] = {round_usize(o.x), round_usize(o.y), round_usize(o.z)};
      const uint64_t n = f2usize(inner);
      for (uint64_t x = 0; x < n; ++x)
        for (uint64_t y = 0; y < n; ++y)
          for (uint64_t z = 0; z < n; ++z)
            if (!pix_points_to_empty(
                    b.parted[flat_projection(off[0] + x, off[1] + y, off[2] + z, dim)]))
              return false;
      return true;
    }
  }
  return true;
}

uint64_t Octree::brick_occupied_bits(const std::vector<uint32_t>& brick) const {
  const uint64_t dim = brick_dim_;
  uint64_t bitmap = 0;
  if (dim >= 4 && (dim & (dim - 1)) == 0) {
    // For these dimensions set_occupancy_in_bitmap_64bits(p, 1, dim) sets the
    // single bit of cell p / (dim / 4): the bitmap is one bit per block of
    // (dim/4)^3 voxels, set when any voxel of the block is non-empty.
    const uint64_t b = dim / 4;
    for (uint64_t z = 0; z < dim; ++z)
      for (uint64_t y = 0; y < dim; ++y) {
        const uint32_t* row = brick.data() + flat_projection(0, y, z, dim);
        const uint64_t base = 4 * (y / b) + 16 * (z / b);
        for (uint64_t bx = 0; bx < 4; ++bx) {
          const uint64_t bit = 1ull << (base + bx);
          if (bitmap & bit) continue;
          for (uint64_t i = 0; i < b; ++i)
            if (!pix_points_to_empty(row[bx * b + i])) {
              bitmap |= bit;
              break;
            }
        }
      }
    return bitmap;
  }
  for (uint64_t x = 0; x < dim; ++x)
    for (uint64_t y = 0; y < dim; ++y)
      for (uint64_t z = 0; z < dim; ++z)
        if (!pix_points_to_empty(brick[flat_projection(x, y, z, dim)])) {
          const uint64_t p[3] = {x, y, z};
          set_occupancy_in_bitmap_64bits(p, 1, dim, true, &bitmap);
        }
  return bitmap;
}

uint64_t Octree::brick_calculate_occupied_bits(const Brick& b) const {
  switch (b.kind) {
    case BrickKind::Empty:
      return 0;
    case BrickKind::Solid:
      return pix_points_to_empty(b.solid) ? 0 : ~0ull;
    case BrickKind::Parted:
      return brick_occupied_bits(b.parted);
  }
  return 0;
}

// BrickData::simplify (node.rs:316-331)
bool Octree::brick_simplify(Brick& b) const {
  if (b.kind == BrickKind::Parted && b.version != 0 && b.mixed_at == b.version) return false;
  const uint32_t* h = homogeneous(b);
  if (!h) {
    if (b.kind == BrickKind::Parted) b.mixed_at = b.version;
    return false;
  }
  const uint32_t v = *h;
  if (pix_points_to_empty(v))
    b = Brick::empty();
  else
    b = Brick::make_solid(v);
  return true;
}

// NodeContent::is_empty (node.rs:470-515)
bool Octree::content_is_empty(const NodeContent& n) const {
  switch (n.kind) {
    case NodeKind::UniformLeaf: {
      const Brick& b = n.bricks[0];
      if (b.kind == BrickKind::Empty) return true;
      if (b.kind == BrickKind::Solid) return pix_points_to_empty(b.solid);
      for (uint32_t v : b.parted)
        if (!pix_points_to_empty(v)) return false;
      return true;
    }
    case NodeKind::Leaf:
      for (const Brick& b : n.bricks) {
        if (b.kind == BrickKind::Solid && !pix_points_to_empty(b.solid)) return false;
        if (b.kind == BrickKind::Parted)
          for (uint32_t v : b.parted)
            if (!pix_points_to_empty(v)) return false;
      }
      return true;
    case NodeKind::Internal:
      return false;
    case NodeKind::Nothing:
      return true;
  }
  return true;
}

// NodeContent::is_all (node.rs:518-552)
bool Octree::content_is_all(const NodeContent& n, uint32_t data) const {
  auto brick_all = [&](const Brick& b) {
    if (b.kind == BrickKind::Empty) return false;
    const uint32_t* h = homogeneous(b);
    return h != nullptr && *h == data;
  };
  switch (n.kind) {
    case NodeKind::UniformLeaf:
      return brick_all(n.bricks[0]);
    case NodeKind::Leaf:
      for (const Brick& b : n.bricks)
        if (!brick_all(b)) return false;
      return true;
    default:
      return false;
  }
}

// NodeContent::compare (node.rs:554-573)
bool Octree::content_compare(const NodeContent& a, const NodeContent& b) {
  switch (a.kind) {
    case NodeKind::Nothing:
      return b.kind == NodeKind::Nothing;
    case NodeKind::Internal:
      return false;
    case NodeKind::UniformLeaf:
      return b.kind == NodeKind::UniformLeaf && a.bricks[0] == b.bricks[0];
    case NodeKind::Leaf:
      if (b.kind != NodeKind::Leaf) return false;
      for (int i = 0; i < 8; ++i)
        if (!(a.bricks[i] == b.bricks[i])) return false;
      return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Structural helpers (detail.rs:181-520)

uint32_t Octree::push_node(NodeContent c) { return nodes_.push(std::move(c)); }

bool Octree::should_bitmap_be_empty_at_bitmap_index(uint32_t node_key,
                                                    const uint64_t idx[3]) const {
  const V3f position = v3u(0.5f) + V3f{static_cast<float>(idx[0]), static_cast<float>(idx[1]),
                                       static_cast<float>(idx[2])};
  const uint32_t to = hash_region(position, BITMAP_DIMENSION_F / 2.0f);
  const uint32_t toc = hash_region(position - (octant_offset(to) * BITMAP_DIMENSION_F) / 2.0f,
                                   BITMAP_DIMENSION_F / 4.0f);
  return should_bitmap_be_empty_at_octants(node_key, to, toc);
}

bool Octree::should_bitmap_be_empty_at_octants(uint32_t node_key, uint32_t to,
                                               uint32_t toc) const {
  const NodeContent& n = nodes_.get(node_key);
  switch (n.kind) {
    case NodeKind::Nothing:
      return true;
    case NodeKind::Internal: {
      const uint32_t child = node_children_[node_key].child(to);
      return nodes_.key_is_valid(child) ? node_empty_at(child, toc) : true;
    }
    case NodeKind::UniformLeaf:
      return brick_is_part_empty_throughout(n.bricks[0], to, toc);
    case NodeKind::Leaf:
      return brick_is_empty_throughout(n.bricks[to], toc);
  }
  return true;
}

bool Octree::node_empty_at(uint32_t node_key, uint32_t octant) const {
  const NodeContent& n = nodes_.get(node_key);
  auto brick_empty = [&](const Brick& b) {
    if (b.kind == BrickKind::Empty) return true;
    if (b.kind == BrickKind::Solid) return pix_points_to_empty(b.solid);
    const uint32_t* h = homogeneous(b);
    return h ? pix_points_to_empty(*h) : false;
  };
  switch (n.kind) {
    case NodeKind::Nothing:
      return true;
    case NodeKind::Leaf:
      return brick_empty(n.bricks[octant]);
    case NodeKind::UniformLeaf:
      return brick_empty(n.bricks[0]);
    case NodeKind::Internal:
      // The reference looks up the child at `octant` for every child octant
      // (detail.rs:305-312); kept as is.
      for (uint32_t co = 0; co < 8; ++co) {
        const uint32_t child = node_children_[node_key].child(octant);
        if (nodes_.key_is_valid(child) && !node_empty_at(child, co)) return false;
      }
      return true;
  }
  return true;
}

// dilute_brick_data (update/mod.rs:563-628)
static void dilute_brick_data(const std::vector<uint32_t>& data, uint32_t dim,
                              std::vector<uint32_t> out[8]) {
  const size_t n = static_cast<size_t>(dim) * dim * dim;
  if (dim == 1) {
    for (int o = 0; o < 8; ++o) out[o] = data;
    return;
  }
  for (int o = 0; o < 8; ++o) out[o].assign(n, data[o]);
  if (dim == 2) return;
  for (uint32_t o = 0; o < 8; ++o) {
    uint64_t off[3];
    offset_usize(o, off);
    off[0] *= 2;
    off[1] *= 2;
    off[2] *= 2;
    std::vector<uint32_t> nb(n, data[flat_projection(off[0], off[1], off[2], dim)]);
    for (uint64_t x = 0; x < dim; ++x)
      for (uint64_t y = 0; y < dim; ++y)
        for (uint64_t z = 0; z < dim; ++z) {
          if (x < 2 && y < 2 && z < 2) continue;
          nb[flat_projection(x, y, z, dim)] =
              data[flat_projection(off[0] + x / 2, off[1] + y / 2, off[2] + z / 2, dim)];
        }
    out[o] = std::move(nb);
  }
}

void Octree::subdivide_leaf_to_nodes(uint32_t node_key, uint32_t target_octant) {
  // detail.rs:321-486
  const uint64_t occupied =
      node_children_[node_key].kind == ChildKind::OccupancyBitmap ? node_children_[node_key].bitmap
                                                                  : 0;
  NodeContent content = NodeContent::internal(occupied);
  std::swap(content, nodes_.get(node_key));
  uint32_t new_children[8];
  for (auto& v : new_children) v = EMPTY_MARKER;
  auto grow = [&](uint32_t key) {
    if (node_children_.size() < static_cast<size_t>(key) + 1) node_children_.resize(key + 1);
  };
  if (content.kind == NodeKind::Leaf) {
    for (uint32_t o = 0; o < 8; ++o) {
      Brick brick = std::move(content.bricks[o]);
      content.bricks[o] = Brick::empty();
      if (brick.kind == BrickKind::Empty) {
        if (o == target_octant) {
          new_children[o] = push_node(NodeContent::nothing());
          grow(new_children[o]);
        }
      } else if (brick.kind == BrickKind::Solid) {
        new_children[o] = push_node(NodeContent::uniform(Brick::make_solid(brick.solid)));
        grow(new_children[o]);
        node_children_[new_children[o]] = NodeChildren::bitmap_of(~0ull);
      } else {
        // The reference computes the bitmap from bricks[octant] after it has
        // been swapped out (always Empty there), so the child bitmap is 0.
        new_children[o] = push_node(NodeContent::uniform(Brick::make_parted(brick.parted)));
        grow(new_children[o]);
        node_children_[new_children[o]] =
            NodeChildren::bitmap_of(brick_calculate_occupied_bits(content.bricks[o]));
      }
    }
  } else if (content.kind == NodeKind::UniformLeaf) {
    Brick& brick = content.bricks[0];
    if (brick.kind == BrickKind::Empty) {
      new_children[target_octant] = push_node(NodeContent::nothing());
      grow(new_children[target_octant]);
      node_children_[new_children[target_octant]] = NodeChildren::bitmap_of(0);
    } else if (brick.kind == BrickKind::Solid) {
      for (uint32_t o = 0; o < 8; ++o) {
        new_children[o] = push_node(NodeContent::uniform(Brick::make_solid(brick.solid)));
        grow(new_children[o]);
        node_children_[new_children[o]] = NodeChildren::bitmap_of(~0ull);
      }
    } else {
      std::vector<uint32_t> kids[8];
      dilute_brick_data(brick.parted, brick_dim_, kids);
      for (uint32_t o = 0; o < 8; ++o) {
        const uint64_t bits = brick_occupied_bits(kids[o]);
        new_children[o] = push_node(NodeContent::uniform(Brick::make_parted(std::move(kids[o]))));
        grow(new_children[o]);
        node_children_[new_children[o]] = NodeChildren::bitmap_of(bits);
      }
    }
  }
  NodeChildren nc;
  nc.kind = ChildKind::Children;
  for (int o = 0; o < 8; ++o) nc.c[o] = new_children[o];
  kids(node_key) = nc;
}

Brick Octree::try_brick_from_node(uint32_t key) const {
  if (!nodes_.key_is_valid(key)) return Brick::empty();
  const NodeContent& n = nodes_.get(key);
  if (n.kind == NodeKind::UniformLeaf) return n.bricks[0];
  return Brick::empty();
}

void Octree::deallocate_children_of(uint32_t node) {
  if (!nodes_.key_is_valid(node)) return;
  if (node_children_[node].kind != ChildKind::Children) return;
  std::vector<uint32_t> to_free;
  for (uint32_t c : node_children_[node].c)
    if (nodes_.key_is_valid(c)) to_free.push_back(c);
  for (uint32_t c : to_free) {
    deallocate_children_of(c);
    nodes_.free(c);
    kids(c) = NodeChildren{};
  }
}

// ---------------------------------------------------------------------------
// Leaf update (update/mod.rs:160-549, 637-675)

uint64_t Octree::update_brick(bool overwrite_if_empty, std::vector<uint32_t>& brick,
                              Cube brick_bounds, uint32_t dim, const uint32_t pos[3],
                              uint32_t size, uint32_t data) {
  uint64_t mi[3];
  matrix_index_for(brick_bounds, pos, dim, mi);
  const uint64_t update_size = std::min<uint64_t>(dim - mi[0], size);
  for (uint64_t x = mi[0]; x < std::min<uint64_t>(mi[0] + size, dim); ++x)
    for (uint64_t y = mi[1]; y < std::min<uint64_t>(mi[1] + size, dim); ++y)
      for (uint64_t z = mi[2]; z < std::min<uint64_t>(mi[2] + size, dim); ++z) {
        uint32_t& v = brick[flat_projection(x, y, z, dim)];
        if (overwrite_if_empty) {
          v = data;
        } else {
          if (pix_color_index(data) < NONE16) v = (v & 0xFFFF0000u) | (data & 0x0000FFFFu);
          if (pix_data_index(data) != NONE16) v = (v & 0x0000FFFFu) | (data & 0xFFFF0000u);
        }
      }
  return update_size;
}

uint64_t Octree::leaf_update(bool overwrite_if_empty, uint32_t node_key, Cube node_bounds,
                             Cube target_bounds, uint32_t target_octant, const uint32_t pos[3],
                             uint32_t size, uint32_t tc, int depth) {
  if (depth > 8) return 0;  // the reference would recurse without end here
  const uint32_t dim = brick_dim_;
  const size_t n3 = static_cast<size_t>(dim) * dim * dim;
  NodeContent& node = nodes_.get(node_key);
  switch (node.kind) {
    case NodeKind::Leaf: {
      Brick& b = node.bricks[target_octant];
      if (b.kind == BrickKind::Empty) {
        std::vector<uint32_t> nb(n3, EMPTY_MARKER);
        const uint64_t us = update_brick(overwrite_if_empty, nb, target_bounds, dim, pos, size, tc);
        b = Brick::make_parted(std::move(nb));
        return us;
      }
      if (b.kind == BrickKind::Solid) {
        const uint32_t v = b.solid;
        if ((pix_points_to_empty(tc) && !pix_points_to_empty(v)) ||
            (!pix_points_to_empty(tc) && v != tc)) {
          std::vector<uint32_t> nb(n3, v);
          const uint64_t us =
              update_brick(overwrite_if_empty, nb, target_bounds, dim, pos, size, tc);
          nodes_.get(node_key).bricks[target_octant] = Brick::make_parted(std::move(nb));
          return us;
        }
        return 0;
      }
      b.touch();
      return update_brick(overwrite_if_empty, b.parted, target_bounds, dim, pos, size, tc);
    }
    case NodeKind::UniformLeaf: {
      Brick& mat = node.bricks[0];
      if (mat.kind == BrickKind::Empty) {
        if (!pix_points_to_empty(tc)) {
          NodeContent leaf;
          leaf.kind = NodeKind::Leaf;
          std::vector<uint32_t> nb(n3, add_to_palette(Entry{}));
          const uint64_t us =
              update_brick(overwrite_if_empty, nb, target_bounds, dim, pos, size, tc);
          leaf.bricks[target_octant] = Brick::make_parted(std::move(nb));
          nodes_.get(node_key) = std::move(leaf);
          return us;
        }
        // falls through to the trailing recursive call (update/mod.rs:486)
      } else if (mat.kind == BrickKind::Solid) {
        const uint32_t v = mat.solid;
        if (pix_points_to_empty(tc) && pix_points_to_empty(v)) {
          nodes_.get(node_key) = NodeContent::nothing();
          return 0;
        }
        if ((!pix_points_to_empty(tc) && v != tc) ||
            (pix_points_to_empty(tc) && !pix_points_to_empty(v))) {
          mat = Brick::make_parted(std::vector<uint32_t>(n3, v));
          return leaf_update(overwrite_if_empty, node_key, node_bounds, target_bounds,
                             target_octant, pos, size, tc, depth + 1);
        }
        return 0;
      } else {
        std::vector<uint32_t>& brick = mat.parted;
        uint64_t mi[3];
        matrix_index_for(node_bounds, pos, dim, mi);
        const uint64_t f = flat_projection(mi[0], mi[1], mi[2], dim);
        if (1 < dim && ((pix_points_to_empty(tc) && pix_points_to_empty(brick[f])) ||
                        (!pix_points_to_empty(tc) && brick[f] == tc)))
          return 0;
        if (node_bounds.size <= static_cast<float>(dim) && dim > 1) {
          mat.touch();
          return update_brick(overwrite_if_empty, brick, node_bounds, dim, pos, size, tc);
        }
        NodeContent leaf;
        leaf.kind = NodeKind::Leaf;
        uint64_t us = 0;
        if (dim == 1) {
          for (int o = 0; o < 8; ++o) leaf.bricks[o] = Brick::make_parted(brick);
          std::vector<uint32_t> nb = brick;
          us = update_brick(overwrite_if_empty, nb, target_bounds, dim, pos, size, tc);
          leaf.bricks[target_octant] = Brick::make_parted(std::move(nb));
        } else {
          std::vector<uint32_t> kids[8];
          std::vector<uint32_t> taken = std::move(brick);
          dilute_brick_data(taken, dim, kids);
          for (uint32_t o = 0; o < 8; ++o) {
            if (o == target_octant)
              us = update_brick(overwrite_if_empty, kids[o], target_bounds, dim, pos, size, tc);
            leaf.bricks[o] = Brick::make_parted(std::move(kids[o]));
          }
        }
        nodes_.get(node_key) = std::move(leaf);
        return us;
      }
      return leaf_update(overwrite_if_empty, node_key, node_bounds, target_bounds, target_octant,
                         pos, size, tc, depth + 1);
    }
    case NodeKind::Internal: {
      const uint64_t ocbits = node.bits;
      kids(node_key) = NodeChildren::bitmap_of(ocbits);
      // child(i) of an OccupancyBitmap is always invalid, so every brick is Empty
      NodeContent leaf;
      leaf.kind = NodeKind::Leaf;
      for (uint32_t o = 0; o < 8; ++o)
        leaf.bricks[o] = try_brick_from_node(node_children_[node_key].child(o));
      nodes_.get(node_key) = std::move(leaf);
      deallocate_children_of(node_key);
      return leaf_update(overwrite_if_empty, node_key, node_bounds, target_bounds, target_octant,
                         pos, size, tc, depth + 1);
    }
    case NodeKind::Nothing: {
      NodeContent leaf;
      leaf.kind = NodeKind::Leaf;
      for (uint32_t o = 0; o < 8; ++o)
        leaf.bricks[o] = try_brick_from_node(node_children_[node_key].child(o));
      nodes_.get(node_key) = std::move(leaf);
      deallocate_children_of(node_key);
      return leaf_update(overwrite_if_empty, node_key, node_bounds, target_bounds, target_octant,
                         pos, size, tc, depth + 1);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// simplify (update/mod.rs:689-1068)

bool Octree::simplify(uint32_t key) {
  if (!nodes_.key_is_valid(key)) return false;
  NodeContent& n = nodes_.get(key);
  switch (n.kind) {
    case NodeKind::Nothing:
      return true;
    case NodeKind::UniformLeaf: {
      Brick& b = n.bricks[0];
      if (b.kind == BrickKind::Empty) return true;
      if (b.kind == BrickKind::Solid) {
        if (pix_points_to_empty(b.solid)) {
          n = NodeContent::nothing();
          kids(key) = NodeChildren{};
          return true;
        }
        return false;
      }
      return brick_simplify(b);
    }
    case NodeKind::Leaf: {
      bool simplified = false;
      bool uniform_solid = true;
      bool have_value = false;
      uint32_t solid_value = 0;
      for (int o = 0; o < 8; ++o) {
        simplified |= brick_simplify(n.bricks[o]);
        if (uniform_solid) {
          if (n.bricks[o].kind == BrickKind::Solid) {
            if (have_value) {
              if (solid_value != n.bricks[o].solid) uniform_solid = false;
            } else {
              solid_value = n.bricks[o].solid;
              have_value = true;
            }
          } else {
            uniform_solid = false;
          }
        }
      }
      if (uniform_solid) {
        n = NodeContent::uniform(Brick::make_solid(solid_value));
        return true;
      }
      const uint32_t dim = brick_dim_;
      const uint64_t half = dim / 2;
      // The eight bricks merge into one brick of the leaf's size when every
      // Empty or Solid brick equals the first brick and every Parted brick
      // is uniform in each of its 2x2x2 blocks. That is tested first, and the
      // merged brick (each octant a half-size copy) is built only then.
      auto octant_off = [half](uint32_t o, uint64_t off[3]) {
        const V3f f = octant_offset(o) * static_cast<float>(half);
        off[0] = round_usize(f.x);
        off[1] = round_usize(f.y);
        off[2] = round_usize(f.z);
      };
      auto block_uniform = [&](const std::vector<uint32_t>& p, uint64_t x, uint64_t y,
                               uint64_t z) {
        const uint32_t v = p[flat_projection(x * 2, y * 2, z * 2, dim)];
        return v == p[flat_projection(x * 2 + 1, y * 2, z * 2, dim)] &&
               v == p[flat_projection(x * 2, y * 2 + 1, z * 2, dim)] &&
               v == p[flat_projection(x * 2, y * 2, z * 2 + 1, dim)] &&
               v == p[flat_projection(x * 2 + 1, y * 2 + 1, z * 2, dim)] &&
               v == p[flat_projection(x * 2, y * 2 + 1, z * 2 + 1, dim)] &&
               v == p[flat_projection(x * 2 + 1, y * 2, z * 2 + 1, dim)] &&
               v == p[flat_projection(x * 2 + 1, y * 2 + 1, z * 2 + 1, dim)];
      };
      bool uniform = true;
      for (uint32_t o = 0; o < 8 && uniform; ++o) {
        const Brick& b = n.bricks[o];
        if (b.kind != BrickKind::Parted) {
          uniform = b == n.bricks[0];
        } else {
          for (uint64_t x = 0; x < half && uniform; ++x)
            for (uint64_t y = 0; y < half && uniform; ++y)
              for (uint64_t z = 0; z < half && uniform; ++z)
                uniform = block_uniform(b.parted, x, y, z);
        }
      }
      if (uniform) {
        std::vector<uint32_t> unified(static_cast<size_t>(dim) * dim * dim, EMPTY_MARKER);
        for (uint32_t o = 0; o < 8; ++o) {
          uint64_t off[3];
          octant_off(o, off);
          const Brick& b = n.bricks[o];
          if (b.kind == BrickKind::Solid) {
            for (uint64_t x = off[0]; x < off[0] + half; ++x)
              for (uint64_t y = off[1]; y < off[1] + half; ++y)
                for (uint64_t z = off[2]; z < off[2] + half; ++z)
                  unified[flat_projection(x, y, z, dim)] = b.solid;
          } else if (b.kind == BrickKind::Parted) {
            for (uint64_t x = 0; x < half; ++x)
              for (uint64_t y = 0; y < half; ++y)
                for (uint64_t z = 0; z < half; ++z)
                  unified[flat_projection(off[0] + x, off[1] + y, off[2] + z, dim)] =
                      b.parted[flat_projection(x * 2, y * 2, z * 2, dim)];
          }
        }
        n = NodeContent::uniform(Brick::make_parted(std::move(unified)));
        simplified = true;
      }
      return simplified;
    }
    case NodeKind::Internal: {
      if (n.bits == 0 || node_children_[key].kind == ChildKind::NoChildren) {
        n = NodeContent::nothing();
        return true;
      }
      if (node_children_[key].kind != ChildKind::Children) return false;
      uint32_t child_keys[8];
      for (int o = 0; o < 8; ++o) child_keys[o] = node_children_[key].c[o];
      simplify(child_keys[0]);
      if (!nodes_.key_is_valid(child_keys[0])) {
        for (int o = 1; o < 8; ++o) simplify(child_keys[o]);
        return false;
      }
      for (int o = 1; o < 8; ++o) {
        simplify(child_keys[o]);
        if (!nodes_.key_is_valid(child_keys[o]) ||
            !content_compare(nodes_.get(child_keys[0]), nodes_.get(child_keys[o])))
          return false;
      }
      nodes_.swap(key, child_keys[0]);
      const NodeChildren promoted = node_children_[child_keys[0]];
      deallocate_children_of(key);
      kids(key) = promoted;
      return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// insert_at_lod_internal (update/insert.rs:99-388)

Status Octree::insert_at_lod_internal(bool overwrite_if_empty, uint32_t x, uint32_t y,
                                      uint32_t z, uint32_t insert_size, const Entry& e) {
  const Cube root{v3u(0.0f), static_cast<float>(octree_size_)};
  const uint32_t upos[3] = {x, y, z};
  const V3f position = to_f(upos);
  if (!bound_contains(root, position)) return Status::InvalidPosition;
  if (e.is_none()) return Status::Ok;
  ++edit_epoch_;

  struct Frame {
    uint32_t key;
    Cube bounds;
  };
  std::vector<Frame> stack{{ROOT_NODE_KEY, root}};
  uint64_t actual_update_size = 0;
  const uint32_t tc = add_to_palette(e);
  const uint32_t rpos[3] = {round_u32(position.x), round_u32(position.y), round_u32(position.z)};
  for (;;) {
    const Frame cur = stack.back();
    const uint32_t ck = cur.key;
    const uint32_t oct = child_octant_for(cur.bounds, position);
    const Cube tb = edit_target_bounds(cur.bounds, oct);
    uint32_t tck = node_children_[ck].child(oct);
    if (insert_size > 1 && tb.size <= static_cast<float>(insert_size) &&
        lex_le_f(position, tb.min)) {
      if (is_leaf(nodes_.get(ck).kind)) {
        subdivide_leaf_to_nodes(ck, oct);
        tck = node_children_[ck].child(oct);
      }
      if (nodes_.key_is_valid(tck)) {
        deallocate_children_of(tck);
        nodes_.get(tck) = NodeContent::uniform(Brick::make_solid(tc));
        kids(tck) = NodeChildren::bitmap_of(~0ull);
      } else {
        const uint32_t nk = push_node(NodeContent::uniform(Brick::make_solid(tc)));
        if (node_children_.size() < static_cast<size_t>(nk) + 1) node_children_.resize(nk + 1);
        kids(ck).child_mut(oct) = nk;
        kids(nk) = NodeChildren::bitmap_of(~0ull);
      }
      actual_update_size = f2usize(tb.size);
      break;
    }
    const Cube data_bounds = nodes_.get(ck).kind == NodeKind::UniformLeaf ? cur.bounds : tb;
    if (data_bounds.size > static_cast<float>(brick_dim_) || nodes_.key_is_valid(tck)) {
      if (nodes_.key_is_valid(tck)) {
        stack.push_back({node_children_[ck].child(oct), tb});
      } else if (is_leaf(nodes_.get(ck).kind)) {
        const NodeContent& n = nodes_.get(ck);
        bool target_match = false;
        if (n.kind == NodeKind::UniformLeaf) {
          const Brick& b = n.bricks[0];
          if (b.kind == BrickKind::Solid) {
            target_match = b.solid == tc;
          } else if (b.kind == BrickKind::Parted) {
            uint64_t mi[3];
            matrix_index_for(cur.bounds, rpos, brick_dim_, mi);
            target_match = b.parted[flat_projection(mi[0], mi[1], mi[2], brick_dim_)] == tc;
          }
        } else {
          const Brick& b = n.bricks[oct];
          if (b.kind == BrickKind::Solid) {
            target_match = b.solid == tc;
          } else if (b.kind == BrickKind::Parted) {
            uint64_t mi[3];
            matrix_index_for(tb, rpos, brick_dim_, mi);
            target_match = b.parted[flat_projection(mi[0], mi[1], mi[2], brick_dim_)] == tc;
          }
        }
        if (target_match || content_is_all(nodes_.get(ck), tc)) break;
        subdivide_leaf_to_nodes(ck, oct);
        stack.push_back({node_children_[ck].child(oct), tb});
      } else {
        if (nodes_.get(ck).kind == NodeKind::Nothing) nodes_.get(ck) = NodeContent::internal(0);
        const uint32_t nk = push_node(NodeContent::nothing());
        if (node_children_.size() < nodes_.len()) node_children_.resize(nodes_.len());
        kids(ck).child_mut(oct) = nk;
        stack.push_back({nk, tb});
      }
    } else {
      actual_update_size =
          leaf_update(overwrite_if_empty, ck, cur.bounds, tb, oct, rpos, insert_size, tc, 0);
      break;
    }
  }

  bool simplifyable = auto_simplify;
  for (auto it = stack.rbegin(); it != stack.rend(); ++it) {
    const uint32_t key = it->key;
    const Cube nb = it->bounds;
    if (!nodes_.key_is_valid(key)) continue;
    if (nodes_.get(key).kind == NodeKind::Nothing) nodes_.get(key) = NodeContent::internal(0);
    uint64_t bits = stored_occupied_bits(key);
    if (f2usize(nb.size) == actual_update_size) {
      bits = ~0ull;
    } else {
      const V3f rel = position - nb.min;
      const uint64_t p[3] = {round_usize(rel.x), round_usize(rel.y), round_usize(rel.z)};
      set_occupancy_in_bitmap_64bits(p, actual_update_size, f2usize(nb.size), true, &bits);
    }
    store_occupied_bits(key, bits);
    update_mip(key, nb, rpos);
    if (is_leaf(nodes_.get(key).kind)) {
      simplifyable = simplify(key);
      continue;
    }
    if (simplifyable) simplifyable = simplify(key);
  }
  return Status::Ok;
}

// ---------------------------------------------------------------------------
// clear_at_lod (update/clear.rs:55-348)

Status Octree::clear_at_lod(uint32_t x, uint32_t y, uint32_t z, uint32_t clear_size) {
  const Cube root{v3u(0.0f), static_cast<float>(octree_size_)};
  const uint32_t pos[3] = {x, y, z};
  const V3f position = to_f(pos);
  if (!bound_contains(root, position)) return Status::InvalidPosition;
  ++edit_epoch_;

  struct Frame {
    uint32_t key;
    Cube bounds;
  };
  std::vector<Frame> stack{{ROOT_NODE_KEY, root}};
  uint64_t actual_update_size = 0;
  const uint32_t dim = brick_dim_;
  for (;;) {
    const Frame cur = stack.back();
    const uint32_t ck = cur.key;
    const uint32_t oct = child_octant_for(cur.bounds, position);
    const Cube tb = edit_target_bounds(cur.bounds, oct);
    const uint32_t tck = node_children_[ck].child(oct);
    const uint32_t tbu[3] = {round_u32(tb.min.x), round_u32(tb.min.y), round_u32(tb.min.z)};
    if (clear_size > 1 && tb.size <= static_cast<float>(clear_size) && lex_le_u(pos, tbu) &&
        nodes_.key_is_valid(tck)) {
      deallocate_children_of(tck);
      nodes_.get(tck) = NodeContent::nothing();
      kids(tck) = NodeChildren{};
      actual_update_size = f2usize(tb.size);
      stack.push_back({node_children_[ck].child(oct), tb});
      break;
    }
    if (tb.size > static_cast<float>(std::max(clear_size, dim)) || nodes_.key_is_valid(tck)) {
      if (nodes_.key_is_valid(tck)) {
        stack.push_back({node_children_[ck].child(oct), tb});
      } else if (is_leaf(nodes_.get(ck).kind)) {
        const NodeContent& n = nodes_.get(ck);
        const Brick& b = n.kind == NodeKind::UniformLeaf ? n.bricks[0] : n.bricks[oct];
        bool target_match = true;
        if (b.kind == BrickKind::Solid) {
          target_match = pix_points_to_empty(b.solid);
        } else if (b.kind == BrickKind::Parted) {
          // index relative to the current node's bounds (clear.rs:136-143, 160-167)
          const uint32_t cm[3] = {round_u32(cur.bounds.min.x), round_u32(cur.bounds.min.y),
                                  round_u32(cur.bounds.min.z)};
          const uint64_t f = flat_projection(pos[0] - cm[0], pos[1] - cm[1], pos[2] - cm[2], dim);
          if (f >= b.parted.size()) return Status::InvalidStructure;  // reference panics
          target_match = pix_points_to_empty(b.parted[f]);
        }
        if (target_match || content_is_empty(n)) break;
        subdivide_leaf_to_nodes(ck, oct);
        stack.push_back({node_children_[ck].child(oct), tb});
      } else {
        break;
      }
    } else {
      actual_update_size =
          leaf_update(true, ck, cur.bounds, tb, oct, pos, clear_size, EMPTY_MARKER, 0);
      break;
    }
  }

  bool have_removed = false;
  Frame removed{0, Cube{}};
  {
    const Frame last = stack.back();
    stack.pop_back();
    if (f2usize(last.bounds.size) <= actual_update_size) {
      have_removed = true;
      removed = last;
    }
  }
  bool simplifyable = auto_simplify;
  for (auto it = stack.rbegin(); it != stack.rend(); ++it) {
    const uint32_t key = it->key;
    const Cube nb = it->bounds;
    if (have_removed) {
      const uint32_t co =
          hash_region((removed.bounds.min - nb.min) + v3u(removed.bounds.size / 2.0f),
                      nb.size / 2.0f);
      kids(key).clear(co);
      nodes_.free(removed.key);
      have_removed = false;
    }
    const uint64_t prev = stored_occupied_bits(key);
    uint64_t bits = node_children_[key].kind == ChildKind::NoChildren ? 0 : prev;
    if (f2usize(nb.size) == actual_update_size) {
      bits = 0;
    } else {
      uint64_t start[3];
      matrix_index_for(nb, pos, 4, start);
      const uint64_t upd = f2usize(
          ceilf((static_cast<float>(actual_update_size) * BITMAP_DIMENSION_F) / nb.size));
      for (uint64_t bx = start[0]; bx < std::min<uint64_t>(start[0] + upd, 4); ++bx)
        for (uint64_t by = start[1]; by < std::min<uint64_t>(start[1] + upd, 4); ++by)
          for (uint64_t bz = start[2]; bz < std::min<uint64_t>(start[2] + upd, 4); ++bz) {
            const uint64_t idx[3] = {bx, by, bz};
            if (should_bitmap_be_empty_at_bitmap_index(key, idx))
              set_occupancy_in_bitmap_64bits(idx, 1, 4, false, &bits);
          }
    }
    if (bits != 0 && node_children_[key].kind == ChildKind::Children) {
      nodes_.get(key) = NodeContent::internal(bits);
    } else {
      deallocate_children_of(key);
      kids(key) = NodeChildren{};
      have_removed = true;
      removed = *it;
      nodes_.get(key) = NodeContent::nothing();
    }
    if (bits == 0)
      kids(key) = NodeChildren{};
    else
      store_occupied_bits(key, bits);
    update_mip(key, nb, pos);
    if (simplifyable) simplifyable = simplify(key);
    if (prev == bits) break;
  }
  return Status::Ok;
}

}  // namespace svo

// ---------------------------------------------------------------------------
// MIP maps (mipmap.rs)

namespace svo {

namespace {

struct Rgba {
  uint32_t r, g, b, a;
};
inline Rgba unpack(uint32_t c) { return Rgba{c >> 24, (c >> 16) & 0xFFu, (c >> 8) & 0xFFu, c & 0xFFu}; }
inline uint32_t pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
  return (r << 24) | (g << 16) | (b << 8) | a;
}
// Rust's `f32 as u8`: truncation toward zero, saturating, NaN to 0
inline uint32_t f2u8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return static_cast<uint32_t>(v);
}

// Albedou32 (mipmap.rs:35-120) with the wrapping u32 arithmetic of a release
// build: differences below zero wrap, as do squares past 2^32.
struct Acc {
  uint32_t r, g, b, a;
};
inline Acc acc_of(Rgba c) { return Acc{c.r, c.g, c.b, c.a}; }
inline Acc acc_pow2(Acc v) { return Acc{v.r * v.r, v.g * v.g, v.b * v.b, v.a * v.a}; }
inline Acc acc_add(Acc x, Acc y) { return Acc{x.r + y.r, x.g + y.g, x.b + y.b, x.a + y.a}; }
inline Acc acc_sub(Acc x, Acc y) { return Acc{x.r - y.r, x.g - y.g, x.b - y.b, x.a - y.a}; }
inline uint32_t round_f2u32(float v) { return f2u32(roundf(v)); }
inline Acc acc_div(Acc x, uint32_t d) {
  const float fd = static_cast<float>(d);
  return Acc{round_f2u32(static_cast<float>(x.r) / fd), round_f2u32(static_cast<float>(x.g) / fd),
             round_f2u32(static_cast<float>(x.b) / fd), round_f2u32(static_cast<float>(x.a) / fd)};
}
inline Acc acc_sqrt(Acc x) {
  return Acc{round_f2u32(sqrtf(static_cast<float>(x.r))), round_f2u32(sqrtf(static_cast<float>(x.g))),
             round_f2u32(sqrtf(static_cast<float>(x.b))), round_f2u32(sqrtf(static_cast<float>(x.a)))};
}
inline float acc_length(Acc x) {
  return sqrtf(static_cast<float>(x.r * x.r + x.g * x.g + x.b * x.b + x.a * x.a));
}
inline uint32_t acc_to_albedo(Acc x) {
  return pack(std::min(x.r, 255u), std::min(x.g, 255u), std::min(x.b, 255u), std::min(x.a, 255u));
}

// Albedo::distance_from (detail.rs:82-89)
inline float color_distance(uint32_t x, uint32_t y) {
  const Rgba p = unpack(x), q = unpack(y);
  const float dr = static_cast<float>(p.r) - static_cast<float>(q.r);
  const float dg = static_cast<float>(p.g) - static_cast<float>(q.g);
  const float db = static_cast<float>(p.b) - static_cast<float>(q.b);
  const float da = static_cast<float>(p.a) - static_cast<float>(q.a);
  return sqrtf(dr * dr + dg * dg + db * db + da * da);
}

// MIPResaplingFunction::execute (mipmap.rs:132-269). `sample(x, y, z, &rgba)`
// reports whether the position holds a colour. Where the reference picks
// among equals through a HashMap (PointFilter's most frequent colour,
// Posterize's first similar group and largest group), whose iteration order
// changes from run to run, the colour met first in sampling order is taken.
template <class F>
bool resample(MipMethodSpec m, const uint32_t start[3], uint32_t size, F sample, uint32_t* out) {
  if (m.kind == MipMethod::BoxFilter) {
    bool any = false;
    uint32_t count = 0;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t x = start[0]; x < start[0] + size; ++x)
      for (uint32_t y = start[1]; y < start[1] + size; ++y)
        for (uint32_t z = start[2]; z < start[2] + size; ++z) {
          uint32_t c;
          if (!sample(x, y, z, &c)) continue;
          const Rgba v = unpack(c);
          const float q[4] = {static_cast<float>(v.r), static_cast<float>(v.g),
                              static_cast<float>(v.b), static_cast<float>(v.a)};
          if (!any) {
            any = true;
            count = 1;
            for (int k = 0; k < 4; ++k) s[k] = q[k] * q[k];
          } else {
            ++count;
            for (int k = 0; k < 4; ++k) s[k] += q[k] * q[k];
          }
        }
    if (!any) return false;
    const float n = static_cast<float>(count);
    uint32_t o[4];
    for (int k = 0; k < 4; ++k) o[k] = f2u8(fminf(sqrtf(s[k] / n), 255.0f));
    *out = pack(o[0], o[1], o[2], o[3]);
    return true;
  }
  if (m.kind == MipMethod::PointFilter || m.kind == MipMethod::PointFilterBD) {
    std::vector<std::pair<uint32_t, uint32_t>> counts;  // colour, occurrences
    for (uint32_t x = start[0]; x < start[0] + size; ++x)
      for (uint32_t y = start[1]; y < start[1] + size; ++y)
        for (uint32_t z = start[2]; z < start[2] + size; ++z) {
          uint32_t c;
          if (!sample(x, y, z, &c)) continue;
          bool found = false;
          for (auto& e : counts)
            if (e.first == c) {
              ++e.second;
              found = true;
              break;
            }
          if (!found) counts.emplace_back(c, 1u);
        }
    if (counts.empty()) return false;
    size_t best = 0;
    for (size_t i = 1; i < counts.size(); ++i)
      if (counts[i].second > counts[best].second) best = i;
    *out = counts[best].first;
    return true;
  }
  // Posterize / PosterizeBD: groups of similar colours as (sum of squares, count)
  std::vector<std::pair<Acc, uint32_t>> groups;
  const float thr = m.thr * 255.0f;
  for (uint32_t x = start[0]; x < start[0] + size; ++x)
    for (uint32_t y = start[1]; y < start[1] + size; ++y)
      for (uint32_t z = start[2]; z < start[2] + size; ++z) {
        uint32_t c;
        if (!sample(x, y, z, &c)) continue;
        const Acc col = acc_of(unpack(c));
        bool merged = false;
        for (auto& g : groups) {
          const Acc poster = acc_sqrt(acc_div(g.first, g.second));
          if (acc_length(acc_sub(poster, col)) < thr) {
            g.first = acc_add(g.first, acc_pow2(col));
            ++g.second;
            merged = true;
            break;
          }
        }
        if (!merged) groups.emplace_back(acc_pow2(col), 1u);
      }
  if (groups.empty()) return false;
  size_t best = 0;
  for (size_t i = 1; i < groups.size(); ++i)
    if (groups[i].second > groups[best].second) best = i;
  *out = acc_to_albedo(acc_sqrt(acc_div(groups[best].first, groups[best].second)));
  return true;
}

inline bool albedo_of(const Entry& e, uint32_t* out) {
  if (e.kind != EntryKind::Visual && e.kind != EntryKind::Complex) return false;
  *out = e.albedo;
  return true;
}

inline uint32_t log2_level(float ratio) {
  // (node size / brick dim).log2() as usize: truncation of an exact power of two
  return f2u32(log2f(ratio));
}

}  // namespace

const Brick& Octree::node_mip(uint32_t key) const {
  static const Brick kEmpty{};
  return key < node_mips_.size() ? node_mips_[key] : kEmpty;
}

Brick& Octree::mip_mut(uint32_t key) {
  if (key >= node_mips_.size()) node_mips_.resize(static_cast<size_t>(key) + 1);
  nodes_.touch(key);
  return node_mips_[key];
}

void Octree::set_mip_method(uint32_t level, MipMethodSpec m) {
  // set_method_at_internal (mipmap.rs:657-674)
  if (m.kind == MipMethod::Posterize || m.kind == MipMethod::PosterizeBD)
    m.thr = std::min(std::max(m.thr, 0.0f), 1.0f);
  mip_strategy_.methods[level] = m;
  ++edit_epoch_;
}

void Octree::set_mip_color_threshold(uint32_t level, float thr) {
  // set_color_similarity_thr_internal (mipmap.rs:618-631)
  mip_strategy_.color_thresholds[level] = std::min(std::max(thr, 0.0f), 1.0f);
  ++edit_epoch_;
}

void Octree::reset_mip_strategy() {
  mip_strategy_ = MipStrategy{};
  ++edit_epoch_;
}

void Octree::assign_mips(std::vector<Brick> mips, MipStrategy strategy) {
  node_mips_ = std::move(mips);
  if (node_mips_.empty()) node_mips_.resize(1);
  mip_strategy_ = std::move(strategy);
  for (uint32_t k = 0; k < node_mips_.size() && k < nodes_.len(); ++k) nodes_.touch(k);
  ++edit_epoch_;
}

void Octree::switch_albedo_mip_maps(bool enabled) {
  const bool before = mip_strategy_.enabled;
  mip_strategy_.enabled = enabled;
  if (enabled && before != enabled && nodes_.get(ROOT_NODE_KEY).kind != NodeKind::Nothing)
    recalculate_mips();
  ++edit_epoch_;
}

// recalculate_mips (mipmap.rs:798-862): depth first, children before parents
void Octree::recalculate_mips() {
  ++edit_epoch_;
  node_mips_.assign(nodes_.len() ? nodes_.len() : 1, Brick{});
  for (uint32_t k = 0; k < nodes_.len(); ++k) nodes_.touch(k);
  struct Item {
    uint32_t key;
    Cube bounds;
    uint32_t octant;
  };
  std::vector<Item> stack{{ROOT_NODE_KEY, Cube{v3u(0.0f), static_cast<float>(octree_size_)}, 0}};
  while (!stack.empty()) {
    Item& it = stack.back();
    if (it.octant == OOB_OCTANT) {
      recalculate_mip(it.key, it.bounds);
      stack.pop_back();
      if (!stack.empty()) stack.back().octant += 1;
      continue;
    }
    const NodeContent& n = nodes_.get(it.key);
    if (n.kind == NodeKind::Internal) {
      const uint32_t c = node_children_[it.key].child(it.octant);
      if (nodes_.key_is_valid(c) && nodes_.get(c).kind != NodeKind::Nothing) {
        const Item next{c, child_bounds_for(it.bounds, it.octant), 0};
        stack.push_back(next);
      } else {
        it.octant += 1;
      }
    } else {
      // Nothing cannot be reached (the root is checked by the caller and
      // empty children are skipped); leaves are evaluated at once
      it.octant = OOB_OCTANT;
    }
  }
}

// recalculate_mip (mipmap.rs:882-902)
void Octree::recalculate_mip(uint32_t key, Cube bounds) {
  if (!mip_strategy_.enabled) return;
  const float dim = static_cast<float>(brick_dim_);
  for (uint32_t x = 0; x < brick_dim_; ++x)
    for (uint32_t y = 0; y < brick_dim_; ++y)
      for (uint32_t z = 0; z < brick_dim_; ++z) {
        const V3f v{static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)};
        const V3f scaled = (v * bounds.size) / dim;
        const V3f p = bounds.min + V3f{roundf(scaled.x), roundf(scaled.y), roundf(scaled.z)};
        const uint32_t pos[3] = {round_u32(p.x), round_u32(p.y), round_u32(p.z)};
        update_mip(key, bounds, pos);
      }
}

// update_mip (mipmap.rs:297-589)
void Octree::update_mip(uint32_t key, Cube bounds, const uint32_t pos[3]) {
  if (!mip_strategy_.enabled) return;
  const uint32_t level = log2_level(bounds.size / static_cast<float>(brick_dim_));
  const MipMethodSpec sampler = mip_strategy_.method_at(level);
  const bool bottom = sampler.kind == MipMethod::PointFilterBD;
  const NodeContent& n = nodes_.get(key);
  const uint32_t dim = brick_dim_;
  uint32_t start[3];
  uint32_t size = 0;
  auto floor_u32 = [](float v) { return round_u32(floorf(v)); };
  switch (n.kind) {
    case NodeKind::Nothing:
      return;
    case NodeKind::UniformLeaf:
      // a uniform leaf is its own MIP
      if (node_mip(key).kind != BrickKind::Empty) mip_mut(key) = Brick{};
      return;
    case NodeKind::Leaf: {
      size = std::min(f2u32(bounds.size) / dim, dim * 2u);
      for (int k = 0; k < 3; ++k) {
        const uint32_t aligned = (pos[k] - pos[k] % size) * 2u * dim;
        start[k] = floor_u32(static_cast<float>(aligned) / bounds.size);
      }
      break;
    }
    case NodeKind::Internal:
      if (bottom) {
        size = f2u32(bounds.size) / dim;
        for (int k = 0; k < 3; ++k) start[k] = floor_u32(static_cast<float>(pos[k] - pos[k] % size));
      } else {
        size = 2;
        const V3f rel = to_f(pos) - bounds.min;
        const V3f s2 = ((rel * 2.0f) * static_cast<float>(dim)) / bounds.size;
        start[0] = floor_u32(s2.x);
        start[1] = floor_u32(s2.y);
        start[2] = floor_u32(s2.z);
        for (int k = 0; k < 3; ++k) start[k] -= start[k] % 2u;
      }
      break;
  }
  uint32_t color = 0;
  bool have = false;
  if (n.kind == NodeKind::Leaf || bottom) {
    have = resample(sampler, start, size,
                    [&](uint32_t x, uint32_t y, uint32_t z, uint32_t* c) {
                      const uint32_t q[3] = {x, y, z};
                      return albedo_of(get_internal(key, bounds, q), c);
                    },
                    &color);
  } else {
    const float fdim = static_cast<float>(dim);
    const NodeChildren& kids_of = node_children_[key];
    const V3f sf{static_cast<float>(start[0]), static_cast<float>(start[1]),
                 static_cast<float>(start[2])};
    const uint32_t first = hash_region(sf, fdim);
    if (first < 8u && kids_of.child(first) != EMPTY_MARKER) {
      have = resample(sampler, start, size,
                      [&](uint32_t x, uint32_t y, uint32_t z, uint32_t* c) {
                        const V3f pf{static_cast<float>(x), static_cast<float>(y),
                                     static_cast<float>(z)};
                        const uint32_t oct = hash_region(pf, fdim);
                        if (oct >= 8u) return false;
                        const uint32_t child = kids_of.child(oct);
                        if (child == EMPTY_MARKER) return false;
                        const V3f off = octant_offset(oct) * fdim;
                        const uint32_t q[3] = {x - round_u32(off.x), y - round_u32(off.y),
                                               z - round_u32(off.z)};
                        const Brick& m = node_mip(child);
                        if (m.kind == BrickKind::Empty) return false;
                        const uint32_t pix =
                            m.kind == BrickKind::Solid
                                ? m.solid
                                : m.parted[flat_projection(q[0], q[1], q[2], dim)];
                        return albedo_of(pix_get_ref(pix), c);
                      },
                      &color);
    }
  }
  if (!have) return;
  // colour matching against the palette (mipmap.rs:527-560)
  uint32_t entry;
  auto thr = mip_strategy_.color_thresholds.find(level);
  bool similar = false;
  if (thr != mip_strategy_.color_thresholds.end()) {
    const float limit = thr->second * 255.0f;
    for (uint32_t i = 0; i < color_palette_.size(); ++i)
      if (color_distance(color, color_palette_[i]) < limit) {
        entry = (i & 0xFFFFu) | (NONE16 << 16);  // NodeContent::pix_visual(i as u16)
        similar = true;
        break;
      }
  }
  if (!similar) entry = add_to_palette(Entry{EntryKind::Visual, color, 0});
  uint64_t mi[3];
  matrix_index_for(bounds, pos, dim, mi);
  const uint64_t f = flat_projection(mi[0], mi[1], mi[2], dim);
  Brick& m = mip_mut(key);
  if (m.kind == BrickKind::Empty) {
    m = Brick::make_parted(std::vector<uint32_t>(static_cast<size_t>(dim) * dim * dim, EMPTY_MARKER));
  } else if (m.kind == BrickKind::Solid) {
    m = Brick::make_parted(std::vector<uint32_t>(static_cast<size_t>(dim) * dim * dim, m.solid));
  }
  m.parted[f] = entry;
  m.touch();
}

Entry Octree::sample_root_mip(uint32_t octant, uint32_t x, uint32_t y, uint32_t z) const {
  const uint32_t key =
      octant == OOB_OCTANT ? ROOT_NODE_KEY : node_children_[ROOT_NODE_KEY].child(octant);
  if (!nodes_.key_is_valid(key)) return Entry{};
  const Brick& m = node_mip(key);
  if (m.kind == BrickKind::Empty) return Entry{};
  if (m.kind == BrickKind::Solid) return pix_get_ref(m.solid);
  return pix_get_ref(m.parted[flat_projection(x, y, z, brick_dim_)]);
}

}  // namespace svo
