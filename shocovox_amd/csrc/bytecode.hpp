// The reference's persisted octree format: bencode of Octree<u32> as
// src/convert/bytecode.rs:30-674 and src/object_pool.rs:24-131 write it.
#pragma once

#include <cstdint>
#include <string>

#include "octree.hpp"

namespace svo {

// Parses a serialized octree (Octree::from_bytes, octree/mod.rs:143-147),
// its MIP maps and MIP strategy included. Returns false with a message on
// malformed input.
bool octree_from_bencode(const uint8_t* data, size_t n, Octree* out, std::string* err);

// Serializes an octree as Octree::to_bytes (octree/mod.rs:136-140) does; the
// strategy's levels are written in ascending order.
void octree_to_bencode(const Octree& t, std::string* out);

}  // namespace svo
