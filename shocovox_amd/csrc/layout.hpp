// Device layout of a resident octree (shared by host and device code).
//
// Nodes are renumbered breadth-first from the root, so the top levels of the
// tree occupy the first records and can be staged into LDS as a prefix.
// One node is one 64-byte record (half a 128-byte cache line):
//   occ       u64  stored_occupied_bits() of the reference node
//   meta      u32  bits 0-1 node kind (Nothing/Internal/Leaf/UniformLeaf),
//                  bits 8+2o..9+2o brick form of octant o (Empty/Solid/Parted);
//                  a UniformLeaf keeps its single brick in octant 0
//   child[o]  u32  Internal: child record index, INVALID when the reference
//                  child key is not a valid pool key;
//                  Leaf/UniformLeaf: Solid -> the brick's palette index value,
//                  Parted -> brick slot in the brick pool
// Bricks keep the reference's three forms: Empty costs nothing, Solid is the
// value stored in the child word, Parted owns a slot holding dim^3 palette
// index values (x + y*dim + z*dim^2) plus a dim^3-bit sidecar whose bit is 1
// where pix_points_to_empty() is false. The DDA reads only the sidecar; the
// palette value is read once, on a hit.
#pragma once

#include <cstdint>

namespace svo {

constexpr uint32_t INVALID_INDEX = 0xFFFFFFFFu;

enum : uint32_t { KIND_NOTHING = 0, KIND_INTERNAL = 1, KIND_LEAF = 2, KIND_UNIFORM_LEAF = 3 };
enum : uint32_t { FORM_EMPTY = 0, FORM_SOLID = 1, FORM_PARTED = 2 };

// A node's MIP brick (node_mips[key]) rides in `mip` (its FORM_*) and
// `mip_word` (the Solid value or the Parted brick slot).
struct alignas(64) NodeRec {
  uint64_t occ;
  uint32_t meta;
  uint32_t mip;
  uint32_t child[8];
  uint32_t mip_word;
  uint32_t reserved[3];
};
static_assert(sizeof(NodeRec) == 64, "node record must be 64 bytes");

inline constexpr uint32_t node_kind(uint32_t meta) { return meta & 3u; }
inline constexpr uint32_t brick_form(uint32_t meta, uint32_t octant) {
  return (meta >> (8u + 2u * octant)) & 3u;
}
inline constexpr uint32_t occ_words_per_brick(uint32_t dim) {
  return (dim * dim * dim + 31u) / 32u;
}

}  // namespace svo
