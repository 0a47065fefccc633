// Flattening of the host octree into the device layout (layout.hpp).
#pragma once

#include <cstdint>
#include <vector>

#include "layout.hpp"
#include "octree.hpp"

namespace svo {

struct FlatTree {
  uint32_t octree_size = 0;
  uint32_t brick_dim = 0;
  std::vector<NodeRec> nodes;
  std::vector<uint32_t> brick_pix;  // brick_count * dim^3
  std::vector<uint32_t> brick_occ;  // brick_count * occ_words_per_brick(dim)
  std::vector<uint32_t> colors;     // packed albedo, r<<24|g<<16|b<<8|a
  std::vector<uint32_t> level_end;  // level_end[l] = first record index of level l+1
  bool mips_enabled = false;
  uint32_t brick_count() const {
    const uint64_t n = static_cast<uint64_t>(brick_dim) * brick_dim * brick_dim;
    return n ? static_cast<uint32_t>(brick_pix.size() / n) : 0;
  }
};

// Breadth-first renumbering of every node reachable from the root through
// valid child keys; bricks are copied into slots in visiting order.
// With `headroom`, the node and brick arrays reserve as much room again
// (address space only until used), so a mirror that grows them edit by edit
// does not copy the whole layout until its size has doubled.
// `versions`, when given, receives each slot's Brick::version.
FlatTree flatten(const Octree& tree, bool headroom = false,
                 std::vector<uint64_t>* versions = nullptr);

// Point query on the flattened layout; it must agree with Octree::get for
// every position (used as the layout round-trip check). Returns the palette
// index value at the position (EMPTY_MARKER where the reference's get()
// yields OctreeEntry::Empty).
uint32_t flat_get(const FlatTree& f, const Octree& palette_owner, uint32_t x, uint32_t y,
                  uint32_t z);

// RAY_TO_NODE_OCCUPANCY_BITMASK_LUT [64 positions][8 directions], built as the
// reference's generator does (lut.rs:32-85): for a bitmap cell and a
// direction octant, the cells of the box spanned from the cell to the node
// corner the direction points at.
void generate_ray_occupancy_lut(uint64_t out[64 * 8]);

}  // namespace svo

namespace svo {

// A host copy of a tree's device layout that follows edits of the host
// octree: records of dirty pool keys are rewritten in place, new nodes and
// bricks get new records / slots (or reuse released ones), and what changed
// is listed so that only that is uploaded. Every slot below synced_slots()
// holds on the device what it holds here (slots are never cleared when
// released), so a brick's change is sent as the cheapest of three forms:
// the changed words alone, a fill of the slot with one pix and one occupancy
// value followed by the words that differ from the fill, or the whole slot.
class LayoutMirror {
 public:
  // A fill of brick slot `slot`: every pix word = pix, every occupancy word = occ.
  struct SlotFill {
    uint32_t slot, pix, occ;
  };
  // A word write into brick slot `slot` (word index within the slot's pix
  // words, or within its occupancy words).
  struct WordWrite {
    uint32_t slot, word, value;
  };

  // Full build (breadth-first, as flatten()). The tree's log of changed
  // nodes is left alone; the caller drains it when this mirror is its reader.
  void build(const Octree& tree);
  // Applies the tree's pending dirty keys.
  void apply(Octree& tree);

  const FlatTree& flat() const { return flat_; }
  // Record indices / brick slots written since the last mark_synced().
  const std::vector<uint32_t>& dirty_records() const { return dirty_recs_; }
  const std::vector<uint32_t>& dirty_slots() const { return dirty_slots_; }
  // The brick part of the patch, by form (each dirty slot in exactly one):
  // slots sent whole, slot fills, and word writes (pix and occupancy) that
  // follow the fills.
  const std::vector<uint32_t>& whole_slots() const { return whole_slots_; }
  const std::vector<SlotFill>& slot_fills() const { return fills_; }
  const std::vector<WordWrite>& pix_writes() const { return pix_writes_; }
  const std::vector<WordWrite>& occ_writes() const { return occ_writes_; }
  bool full_upload() const { return full_; }
  uint32_t colors_synced() const { return colors_synced_; }
  // Sizes of the last patch (records, brick slots) that mark_synced() retired.
  uint32_t last_records() const { return last_recs_; }
  uint32_t last_slots() const { return last_slots_; }
  // Brick slots whose device contents equal the mirror's (all of them after
  // a full upload or a patch).
  uint32_t synced_slots() const { return synced_slots_; }
  // The device now holds the mirror: the patch lists are retired.
  void mark_synced();

 private:
  uint32_t record_for(uint32_t key);
  void write_record(const Octree& tree, uint32_t key, std::vector<uint32_t>* pending);
  // What the encoding of some slots adds to the patch (one per worker).
  struct EncodeOut {
    std::vector<uint32_t> dirty, whole;
    std::vector<SlotFill> fills;
    std::vector<WordWrite> pix, occ;
    std::vector<uint32_t> scratch;
    bool rewritten = false;
  };
  uint32_t write_brick(const Octree& tree, const Brick& b, uint32_t old_slot);
  void encode_slot(const Octree& tree, uint32_t slot, const uint32_t* new_pix, EncodeOut* out);
  void encode_deferred(const Octree& tree);
  void drop_partial_forms();
  void release_slot(uint32_t slot);
  void release_key(uint32_t key);

  FlatTree flat_;
  std::vector<uint32_t> key_to_rec_;
  std::vector<uint32_t> free_recs_, free_slots_;
  std::vector<uint32_t> dirty_recs_, dirty_slots_;
  std::vector<uint8_t> rec_dirty_flag_, slot_dirty_flag_;
  std::vector<uint64_t> slot_version_;  // Brick::version last written to each slot
  std::vector<uint32_t> whole_slots_;
  std::vector<SlotFill> fills_;
  std::vector<WordWrite> pix_writes_, occ_writes_;
  // slots written by this apply and the values they take, encoded at its
  // end (the last write of a slot wins)
  std::vector<std::pair<uint32_t, const uint32_t*>> deferred_;
  bool rewritten_ = false;  // a slot was written twice in one patch (sent whole)
  bool full_ = true;
  uint32_t colors_synced_ = 0;
  uint32_t synced_slots_ = 0;
  uint32_t last_recs_ = 0, last_slots_ = 0;
};

}  // namespace svo
