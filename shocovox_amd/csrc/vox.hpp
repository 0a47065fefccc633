// MagicaVoxel .vox import: the reference's Octree::load_vox_file
// (src/convert/magicavoxel.rs:259-395), reading the file format as the
// dot_vox 5.1.1 crate exposes it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "octree.hpp"

namespace svo {

enum class VoxStatus : int { Ok = 0, Malformed, TreeSize, BrickDim, Position };

// Builds an octree of edge 2^ceil(log2(extent)) with brick dimension
// brick_dim from a .vox file image, inserting every voxel of every model
// that frame 0 of the scene graph places. On failure *err says why.
VoxStatus octree_from_vox(const uint8_t* data, size_t n, uint32_t brick_dim, Octree* out,
                          std::string* err);

// parse_rotation_matrix (magicavoxel.rs:63-91) as rows: m[r][c]. Returns
// false for a byte that names no signed permutation.
bool vox_rotation(uint8_t b, int8_t m[3][3]);

}  // namespace svo
