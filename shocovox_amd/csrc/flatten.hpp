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
  uint32_t brick_count() const {
    const uint64_t n = static_cast<uint64_t>(brick_dim) * brick_dim * brick_dim;
    return n ? static_cast<uint32_t>(brick_pix.size() / n) : 0;
  }
};

// Breadth-first renumbering of every node reachable from the root through
// valid child keys; bricks are copied into slots in visiting order.
FlatTree flatten(const Octree& tree);

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
