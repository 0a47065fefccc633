// The reference's persisted octree format: bencode of Octree<u32> as
// src/convert/bytecode.rs:30-674 and src/object_pool.rs:24-131 write it.
#pragma once

#include <cstdint>
#include <string>

#include "octree.hpp"

namespace svo {

// Parses a serialized octree (Octree::from_bytes, octree/mod.rs:143-147).
// MIP maps and the MIP strategy are read and dropped: the traversal here runs
// with MIPs disabled. Returns false with a message on malformed input.
bool octree_from_bencode(const uint8_t* data, size_t n, Octree* out, std::string* err);

// Serializes an octree as Octree::to_bytes (octree/mod.rs:136-140) does; node
// MIPs are written as empty bricks and the MIP strategy as disabled.
void octree_to_bencode(const Octree& t, std::string* out);

}  // namespace svo
