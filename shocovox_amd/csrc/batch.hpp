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
};

int batch_create(uint32_t max_jobs, BatchContext** out, std::string* err);
void batch_destroy(BatchContext* b);
int batch_run(BatchContext* b, const BatchJobIn* jobs, uint32_t n, void* stream, std::string* err);
uint64_t batch_tiles(const BatchContext* b);

}  // namespace svo
