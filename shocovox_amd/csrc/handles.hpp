// Definitions of the opaque handles of include/svo.h, shared by the C ABI
// glue and the serving loop.
#pragma once

#include <string>

#include "../../include/svo.h"
#include "device.hpp"
#include "flatten.hpp"
#include "octree.hpp"

struct svo_octree {
  svo::Octree tree;
  // encoding of the last svo_octree_to_bytes size query (see there)
  mutable std::string bytes;
  mutable uint64_t bytes_epoch = 0;
  mutable bool bytes_valid = false;
};

struct svo_flat {
  svo::LayoutMirror mirror;
};

struct svo_tree {
  svo::DeviceTree* dev = nullptr;
  svo::LayoutMirror mirror;  // host copy of the layout, for in-place edits
  bool has_mirror = false;
};

namespace svo {
// Sets the message svo_last_error() returns on the calling thread.
void set_last_error(const std::string& msg);
// svo_tree_update that also records `start_event` (a hipEvent_t, or NULL) on
// `stream` when the host part of the update is done, before its first device
// operation, and the `n_end` events of `end_events` when its last device
// operation is done.
int tree_update(svo_tree* dtree, svo_octree* tree, void* stream, void* start_event,
                void* const* end_events = nullptr, uint32_t n_end = 0);
}  // namespace svo
