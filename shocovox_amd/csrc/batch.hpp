// Batched mode: many trees, viewports and ray arrays in one persistent launch.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/svo.h"
#include "device.hpp"

namespace svo {

struct BatchContext;

struct BatchJobIn {
  const DeviceTree* tree;
  uint32_t kind;  // 0 render, 1 cast
  uint32_t width;
  uint32_t height;
  Camera cam;
  const float* rays;
  svo_hits out;
  // Render chunks: the job covers tiles [first_tile, first_tile + tile_limit)
  // of the image (tile_limit 0 = to the end). With tile_major set, outputs
  // are stored tile after tile (64 slots per tile, pixel (x&7) + 8*(y&7)
  // inside a tile) from out[0], instead of at y * width + x.
  uint32_t first_tile = 0;
  uint32_t tile_limit = 0;
  uint32_t tile_major = 0;
  // > 0: get_by_ray_at_lod with this viewing distance; 0: get_by_ray
  float viewing_distance = 0.0f;
};

int batch_create(uint32_t max_jobs, BatchContext** out, std::string* err);
// A context of the given variant instead of the selected one.
int batch_create_variant(uint32_t max_jobs, int variant, BatchContext** out, std::string* err);
// The variant a serving loop uses: the selected one when its tiles are 8x8
// (the tile-major chunks of the loop are laid out in 8x8 tiles), otherwise
// classic-w5-g5.
int batch_variant_serving();
void batch_destroy(BatchContext* b);
int batch_run(BatchContext* b, const BatchJobIn* jobs, uint32_t n, void* stream, std::string* err);
uint64_t batch_tiles(const BatchContext* b);

// Kernel variants of batched launches; a context uses the variant selected
// when it was created.
int batch_variant_count();
const char* batch_variant_name(int i);
// Private (scratch) bytes per lane of a variant's kernel (non-zero: spills).
int batch_variant_scratch(int i, uint32_t* bytes, std::string* err);
bool batch_variant_select(int i);
int batch_variant_current();
int batch_context_variant(const BatchContext* b);

}  // namespace svo
