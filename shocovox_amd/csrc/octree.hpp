// Host-side sparse voxel octree: the data model and edit operations of
// shocovox-rs 0.11.1 `Octree<u32>` (user data = u32), kept on the CPU.
//
// This is the builder that feeds the device layout (flatten.cpp) and applies
// sparse edits for the serving loop. It follows the reference's node model
// exactly, because the traversal's result depends on the tree's shape
// (Leaf vs UniformLeaf, Solid vs Parted bricks, occupancy bits):
//   NodeContent{Nothing, Internal(u64), Leaf([BrickData;8]), UniformLeaf(BrickData)}
//   NodeChildren{NoChildren, Children([u32;8]), OccupancyBitmap(u64)}
//   BrickData{Empty, Parted(Vec<u32>), Solid(u32)}
// Reference: src/octree/types.rs:38-207, src/object_pool.rs:87-241,
//            src/octree/node.rs, src/octree/detail.rs, src/octree/mod.rs,
//            src/octree/update/{mod,insert,clear}.rs
#pragma once

#include <atomic>
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "svo_math.hpp"

namespace svo {

enum class BrickKind : uint8_t { Empty = 0, Solid = 1, Parted = 2 };

// A fresh, process-wide unique content version for a Parted brick (never 0).
inline uint64_t next_brick_version() {
  static std::atomic<uint64_t> counter{0};
  return counter.fetch_add(1, std::memory_order_relaxed) + 1;
}

struct Brick {
  BrickKind kind = BrickKind::Empty;
  uint32_t solid = 0;
  std::vector<uint32_t> parted;
  // Content version of a Parted brick: two bricks with the same non-zero
  // version hold the same values (a copy keeps it; every write in place
  // takes a fresh one through touch()). 0 = unknown. The layout mirror uses
  // it to skip bricks an edit did not change; equality ignores it.
  uint64_t version = 0;
  // The version at which brick_simplify last found the values not all
  // equal (0: never): the same content needs no second scan.
  mutable uint64_t mixed_at = 0;

  static Brick empty() { return Brick{}; }
  static Brick make_solid(uint32_t v) {
    Brick b;
    b.kind = BrickKind::Solid;
    b.solid = v;
    return b;
  }
  static Brick make_parted(std::vector<uint32_t> v) {
    Brick b;
    b.kind = BrickKind::Parted;
    b.parted = std::move(v);
    b.version = next_brick_version();
    return b;
  }
  // Call after writing `parted` in place.
  void touch() { version = next_brick_version(); }
  bool operator==(const Brick& o) const {
    if (kind != o.kind) return false;
    if (kind == BrickKind::Solid) return solid == o.solid;
    if (kind == BrickKind::Parted) return parted == o.parted;
    return true;
  }
};

enum class NodeKind : uint8_t { Nothing = 0, Internal = 1, Leaf = 2, UniformLeaf = 3 };

struct NodeContent {
  NodeKind kind = NodeKind::Nothing;
  uint64_t bits = 0;   // Internal occupancy
  Brick bricks[8];     // Leaf: 8 bricks; UniformLeaf: bricks[0]

  static NodeContent nothing() { return NodeContent{}; }
  static NodeContent internal(uint64_t b) {
    NodeContent n;
    n.kind = NodeKind::Internal;
    n.bits = b;
    return n;
  }
  static NodeContent uniform(Brick b) {
    NodeContent n;
    n.kind = NodeKind::UniformLeaf;
    n.bricks[0] = std::move(b);
    return n;
  }
};

enum class ChildKind : uint8_t { NoChildren = 0, Children = 1, OccupancyBitmap = 2 };

struct NodeChildren {
  ChildKind kind = ChildKind::NoChildren;
  uint32_t c[8] = {EMPTY_MARKER, EMPTY_MARKER, EMPTY_MARKER, EMPTY_MARKER,
                   EMPTY_MARKER, EMPTY_MARKER, EMPTY_MARKER, EMPTY_MARKER};
  uint64_t bitmap = 0;

  static NodeChildren bitmap_of(uint64_t b) {
    NodeChildren n;
    n.kind = ChildKind::OccupancyBitmap;
    n.bitmap = b;
    return n;
  }
  // NodeConnection::child (node.rs:49-54)
  uint32_t child(uint32_t octant) const {
    return kind == ChildKind::Children ? c[octant] : EMPTY_MARKER;
  }
  // NodeConnection::child_mut (node.rs:56-64)
  uint32_t& child_mut(uint32_t octant);
  // NodeConnection::clear (node.rs:75-83)
  void clear(uint32_t octant);
};

// Entry of a voxel as the reference's OctreeEntry (types.rs:23-36).
// albedo is packed as Albedo::from(u32): r<<24 | g<<16 | b<<8 | a.
enum class EntryKind : uint8_t { Empty = 0, Visual = 1, Informative = 2, Complex = 3 };
struct Entry {
  EntryKind kind = EntryKind::Empty;
  uint32_t albedo = 0;
  uint32_t data = 0;
  bool is_none() const;  // OctreeEntry::is_none (octree/mod.rs:97-104)
};

enum class Status : int {
  Ok = 0,
  InvalidSize = 1,
  InvalidBrickDimension = 2,
  InvalidStructure = 3,
  InvalidPosition = 4,
};

// ObjectPool<NodeContent> (object_pool.rs:87-241)
// Every mutable access records the key as dirty (conservatively: a mutable
// reference counts as a write), so a device copy can be patched in place.
class NodePool {
 public:
  uint32_t push(NodeContent item);
  bool free(uint32_t key);
  bool key_is_valid(uint64_t key) const { return key < items_.size() && reserved_[key]; }
  NodeContent& get(uint32_t key) {
    touch(key);
    return items_[key];
  }
  const NodeContent& get(uint32_t key) const { return items_[key]; }
  void swap(uint32_t a, uint32_t b);
  uint32_t len() const { return static_cast<uint32_t>(items_.size()); }
  bool reserved(uint32_t key) const { return reserved_[key] != 0; }
  void reserve(size_t n) {
    items_.reserve(n);
    reserved_.reserve(n);
  }
  size_t first_available() const { return first_available_; }
  // Replaces the pool's contents (a deserialized pool); every key is dirty.
  void assign(std::vector<NodeContent> items, std::vector<uint8_t> reserved,
              size_t first_available) {
    items_ = std::move(items);
    reserved_ = std::move(reserved);
    first_available_ = first_available;
    dirty_flag_.clear();
    dirty_.clear();
    for (uint32_t k = 0; k < items_.size(); ++k) touch(k);
  }
  void touch(uint32_t key) {
    if (key >= dirty_flag_.size()) dirty_flag_.resize(static_cast<size_t>(key) + 1, 0);
    if (!dirty_flag_[key]) {
      dirty_flag_[key] = 1;
      dirty_.push_back(key);
    }
  }
  // Keys written since the last call, in first-touch order.
  std::vector<uint32_t> take_dirty() {
    for (uint32_t k : dirty_) dirty_flag_[k] = 0;
    std::vector<uint32_t> out;
    out.swap(dirty_);
    return out;
  }

 private:
  bool is_next_available() const;
  bool check_first_available();
  std::vector<NodeContent> items_;
  std::vector<uint8_t> reserved_;
  size_t first_available_ = 0;
  std::vector<uint8_t> dirty_flag_;
  std::vector<uint32_t> dirty_;
};

// MIPResamplingMethods (types.rs:104-140): the threshold of the Posterize
// forms is kept clamped to [0, 1].
enum class MipMethod : uint8_t {
  BoxFilter = 0,
  PointFilter = 1,
  PointFilterBD = 2,
  Posterize = 3,
  PosterizeBD = 4
};
struct MipMethodSpec {
  MipMethod kind = MipMethod::BoxFilter;
  float thr = 0.0f;
  bool operator==(const MipMethodSpec& o) const { return kind == o.kind && thr == o.thr; }
};

// MIPMapStrategy (types.rs:141-160) with its default (mipmap.rs:591-604):
// disabled; level 1 PointFilter, levels 2-4 BoxFilter (others BoxFilter too);
// colour-matching thresholds 0.1, 0.05 and 0.02 at levels 2, 3 and 4.
struct MipStrategy {
  bool enabled = false;
  std::map<uint32_t, MipMethodSpec> methods{{1, {MipMethod::PointFilter, 0.0f}},
                                            {2, {MipMethod::BoxFilter, 0.0f}},
                                            {3, {MipMethod::BoxFilter, 0.0f}},
                                            {4, {MipMethod::BoxFilter, 0.0f}}};
  std::map<uint32_t, float> color_thresholds{{2, 0.1f}, {3, 0.05f}, {4, 0.02f}};
  MipMethodSpec method_at(uint32_t level) const {
    auto it = methods.find(level);
    return it == methods.end() ? MipMethodSpec{} : it->second;
  }
};

class Octree {
 public:
  static constexpr uint32_t ROOT_NODE_KEY = 0;

  // Octree::new (octree/mod.rs:173-205). On failure `status` is set and the
  // returned tree must not be used.
  static Octree create(uint32_t size, uint32_t brick_dim, Status* status);

  Status insert(uint32_t x, uint32_t y, uint32_t z, const Entry& e) {
    return insert_at_lod_internal(true, x, y, z, 1, e);
  }
  Status insert_at_lod(uint32_t x, uint32_t y, uint32_t z, uint32_t size, const Entry& e) {
    return insert_at_lod_internal(true, x, y, z, size, e);
  }
  Status update(uint32_t x, uint32_t y, uint32_t z, const Entry& e) {
    return insert_at_lod_internal(false, x, y, z, 1, e);
  }
  Status clear(uint32_t x, uint32_t y, uint32_t z) { return clear_at_lod(x, y, z, 1); }
  Status clear_at_lod(uint32_t x, uint32_t y, uint32_t z, uint32_t size);
  Entry get(uint32_t x, uint32_t y, uint32_t z) const;

  // MIP maps (mipmap.rs). The strategy setters follow StrategyUpdater
  // (mipmap.rs:706-797); switch_albedo_mip_maps recalculates every MIP when
  // MIPs are turned on for a non-empty tree (:865-880).
  const MipStrategy& mip_strategy() const { return mip_strategy_; }
  void set_mip_method(uint32_t level, MipMethodSpec m);
  void set_mip_color_threshold(uint32_t level, float thr);
  void reset_mip_strategy();
  void switch_albedo_mip_maps(bool enabled);
  void recalculate_mips();
  // node_mips[key]; keys never written hold an Empty brick
  const Brick& node_mip(uint32_t key) const;
  size_t node_mip_count() const { return node_mips_.size(); }
  // sample_root_mip (mipmap.rs:905-937): octant 8 samples the root's MIP
  Entry sample_root_mip(uint32_t octant, uint32_t x, uint32_t y, uint32_t z) const;
  // Installs deserialized MIPs and strategy.
  void assign_mips(std::vector<Brick> mips, MipStrategy strategy);

  uint32_t size() const { return octree_size_; }
  uint32_t brick_dim() const { return brick_dim_; }
  bool auto_simplify = true;

  // Palette management (update/mod.rs:55-136)
  uint32_t add_to_palette(const Entry& e);
  // pix_points_to_empty (node.rs:405-427); inline, since brick scans call it
  // for every voxel
  bool pix_points_to_empty(uint32_t pix) const {
    if (pix == EMPTY_MARKER) return true;
    const uint32_t c = pix_color_index(pix), d = pix_data_index(pix);
    const bool color_empty =
        c >= NONE16 || (c < color_palette_.size() ? (color_palette_[c] & 0xFFu) == 0 : true);
    const bool data_empty =
        d == NONE16 || (d < data_palette_.size() ? data_palette_[d] == 0 : true);
    return color_empty && data_empty;
  }
  Entry pix_get_ref(uint32_t pix) const;

  // Read access used by the flattener, the model export and the scene builder.
  const NodePool& nodes() const { return nodes_; }
  NodePool& nodes_mut() { return nodes_; }
  const std::vector<NodeChildren>& node_children() const { return node_children_; }
  std::vector<NodeChildren>& node_children_mut() { return node_children_; }
  // Pool keys whose node or children changed since the last call.
  std::vector<uint32_t> take_dirty() { return nodes_.take_dirty(); }
  const std::vector<uint32_t>& color_palette() const { return color_palette_; }
  const std::vector<uint32_t>& data_palette() const { return data_palette_; }
  uint64_t stored_occupied_bits(uint32_t key) const;
  // Installs a deserialized model (node pool, children, palettes) in place of
  // this tree's; palette lookup maps are rebuilt as Octree::from_bencode does
  // (a later duplicate palette entry wins).
  void assign_model(std::vector<NodeContent> items, std::vector<uint8_t> reserved,
                    size_t first_available, std::vector<NodeChildren> children,
                    std::vector<uint32_t> colors, std::vector<uint32_t> data, bool simplify);
  void store_occupied_bits(uint32_t key, uint64_t bits);
  bool simplify(uint32_t key);
  // Counts every edit applied since the last call (for incremental re-upload).
  uint64_t edit_epoch() const { return edit_epoch_; }

 private:
  Status insert_at_lod_internal(bool overwrite_if_empty, uint32_t x, uint32_t y, uint32_t z,
                                uint32_t size, const Entry& e);
  uint64_t leaf_update(bool overwrite_if_empty, uint32_t node_key, Cube node_bounds,
                       Cube target_bounds, uint32_t target_octant, const uint32_t pos[3],
                       uint32_t size, uint32_t target_content, int depth);
  static uint64_t update_brick(bool overwrite_if_empty, std::vector<uint32_t>& brick,
                               Cube brick_bounds, uint32_t dim, const uint32_t pos[3],
                               uint32_t size, uint32_t data);
  void subdivide_leaf_to_nodes(uint32_t node_key, uint32_t target_octant);
  Entry get_internal(uint32_t key, Cube bounds, const uint32_t pos[3]) const;
  void update_mip(uint32_t key, Cube bounds, const uint32_t pos[3]);
  void recalculate_mip(uint32_t key, Cube bounds);
  Brick& mip_mut(uint32_t key);
  void deallocate_children_of(uint32_t node);
  Brick try_brick_from_node(uint32_t key) const;
  uint32_t push_node(NodeContent c);
  bool brick_simplify(Brick& b) const;
  bool brick_is_empty_throughout(const Brick& b, uint32_t octant) const;
  bool brick_is_part_empty_throughout(const Brick& b, uint32_t part_octant,
                                      uint32_t target_octant) const;
  uint64_t brick_occupied_bits(const std::vector<uint32_t>& brick) const;
  uint64_t brick_calculate_occupied_bits(const Brick& b) const;
  bool content_is_empty(const NodeContent& n) const;
  bool content_is_all(const NodeContent& n, uint32_t data) const;
  static bool content_compare(const NodeContent& a, const NodeContent& b);
  bool should_bitmap_be_empty_at_bitmap_index(uint32_t node_key, const uint64_t idx[3]) const;
  bool should_bitmap_be_empty_at_octants(uint32_t node_key, uint32_t target_octant,
                                         uint32_t target_octant_for_child) const;
  bool node_empty_at(uint32_t node_key, uint32_t octant) const;

  NodeChildren& kids(uint32_t key) {
    nodes_.touch(key);
    return node_children_[key];
  }

  uint32_t brick_dim_ = 1;
  uint32_t octree_size_ = 2;
  NodePool nodes_;
  std::vector<NodeChildren> node_children_;
  std::vector<uint32_t> color_palette_;
  std::vector<uint32_t> data_palette_;
  std::unordered_map<uint32_t, uint32_t> color_index_;
  std::unordered_map<uint32_t, uint32_t> data_index_;
  uint64_t edit_epoch_ = 0;
  MipStrategy mip_strategy_;
  std::vector<Brick> node_mips_{Brick{}};
};

// Helpers of src/spatial/math/mod.rs used by the builder.
uint64_t flat_projection(uint64_t x, uint64_t y, uint64_t z, uint64_t size);
uint64_t position_in_bitmap_64bits(const uint64_t idx[3], uint64_t brick_size);
void set_occupancy_in_bitmap_64bits(const uint64_t pos[3], uint64_t size, uint64_t brick_dim,
                                    bool occupied, uint64_t* bitmap);

}  // namespace svo
