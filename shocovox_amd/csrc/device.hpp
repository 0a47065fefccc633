// Device residency and kernel launches of the traversal path.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/svo.h"
#include "layout.hpp"
#include "traverse.hpp"

namespace svo {

struct DeviceTree {
  int device = -1;
  uint32_t node_count = 0;
  uint32_t brick_count = 0;
  uint32_t color_count = 0;
  uint32_t top_level_nodes = 0;
  uint64_t node_capacity = 0;   // records allocated
  uint64_t brick_capacity = 0;  // brick slots that fit both brick arrays at the current dim
  uint64_t color_capacity = 0;
  // bytes allocated for brick_pix and brick_occ: a re-upload may change the
  // brick dimension, and with it the bytes of a slot, so the arrays are
  // sized in bytes and brick_capacity is derived from them
  uint64_t pix_bytes = 0;
  uint64_t occ_bytes = 0;
  NodeRec* nodes = nullptr;
  uint32_t* brick_pix = nullptr;
  uint32_t* brick_occ = nullptr;
  uint32_t* colors = nullptr;
  uint64_t* ray_lut = nullptr;
  // Staging of incremental patches, two sets used alternately: a patch fills
  // one (pinned host buffer, device buffer) while the copy and scatter of the
  // previous patch may still be reading the other; `done` marks the end of
  // the last patch that used a set.
  // A staging set is pending from the patch that fills it until that patch's
  // copies have run: a host function enqueued after them on the caller's
  // stream clears the flag (no event is kept on the caller's stream, which
  // may be destroyed before the set is reused).
  struct PatchStaging {
    void* host = nullptr;
    uint64_t host_bytes = 0;
    void* dev = nullptr;
    uint64_t dev_bytes = 0;
    DeviceTree* tree = nullptr;  // the callback's argument: this set's owner
    bool pending = false;        // guarded by staging_mu
  };
  PatchStaging staging[2];
  std::mutex staging_mu;
  std::condition_variable staging_cv;
  uint32_t next_staging = 0;
  std::vector<void*> retired;    // outgrown device buffers, freed with the tree
  uint64_t retired_bytes = 0;
  uint32_t step_limit = 0;       // lowered iteration bound per ray (0: MAX_TRAVERSAL_STEPS)
  uint64_t bytes = 0;
  TreeView view{};
};

// Every function returns an svo_status and fills `err` on failure.
int device_tree_create(const svo_tree_desc& desc, DeviceTree** out, std::string* err);
// Re-uploads the given description into an existing tree, growing it when
// needed; ordered on `stream`.
int device_tree_upload(DeviceTree* t, const svo_tree_desc& desc, void* stream, std::string* err);
void device_tree_destroy(DeviceTree* t);
// Device bytes a fresh device_tree_create of `desc` allocates (capacities
// include the growth headroom of uploads).
uint64_t device_tree_bytes_for(const svo_tree_desc& desc);
// Makes both patch stagings (pinned host and device) hold at least `bytes`
// (at least 8 MiB), so the first edits of a tree allocate nothing.
int device_tree_reserve_staging(DeviceTree* t, size_t bytes, std::string* err);
// The brick part of an incremental patch: slots sent whole (their words are
// read from the layout), slot fills {slot, pix value, occupancy value}, and
// word writes {slot, word, value} into the slots' pix and occupancy words,
// applied after the fills. Each listed slot appears in one form only.
struct BrickPatch {
  const uint32_t* whole = nullptr;
  size_t n_whole = 0;
  const uint32_t* fills = nullptr;  // 3 words per fill
  size_t n_fills = 0;
  const uint32_t* pix_writes = nullptr;  // 3 words per write
  size_t n_pix = 0;
  const uint32_t* occ_writes = nullptr;  // 3 words per write
  size_t n_occ = 0;
};

// Writes the given records and brick changes of a layout into an existing
// tree, asynchronously on `stream`: the call returns once the copy and
// scatter are queued (it waits on the host only for the patch two calls
// back, whose staging it reuses). Arrays whose capacity `desc` (the layout's
// full description) outgrows are reallocated larger and their live contents
// copied device to device on `stream` first, so a patch never re-sends the
// whole tree from the host. `start_event` (a hipEvent_t, or NULL) is
// recorded on `stream` before the patch's first device operation, after its
// host work; the `n_end` events of `end_events` right after its last one
// (ahead of the host function that releases the staging). Returns 1 (asking
// for a full upload) only when the brick dimension differs.
int device_tree_patch(DeviceTree* t, const svo_tree_desc& desc,
                      const std::vector<uint32_t>& records, const BrickPatch& bricks,
                      uint32_t colors_from, void* stream, void* start_event,
                      void* const* end_events, uint32_t n_end, std::string* err);

// hipEventRecord(event, stream) with the status mapped.
int device_record_event(void* event, void* stream, std::string* err);

int device_cast(const DeviceTree& t, const float* rays, uint32_t n, const svo_hits& out,
                void* stream, std::string* err);
// svo_cast's traversal order (order.hip): *order receives n ray indices sorted
// by (Morton code of the origin on a 64^3 grid of a tree of size
// tree_size, direction octant), in memory from hipMallocAsync on `stream`
// that the caller returns with hipFreeAsync on the same stream (*order stays
// NULL, and the rays in their own order, for n above 2^31 - 1).
int order_rays(const float* rays, uint32_t n, float tree_size, uint32_t** order, void* stream,
               std::string* err);
// Ray counts from which svo_cast traverses in that order (0: never).
void set_cast_order_threshold(uint32_t n);
uint32_t cast_order_threshold();
int device_render(const DeviceTree& t, const Camera& cam, uint32_t w, uint32_t h,
                  const svo_hits& out, void* stream, std::string* err);
// get_by_ray_at_lod over a ray array / a viewport's primary rays
int device_cast_lod(const DeviceTree& t, const float* rays, uint32_t n, float viewing_distance,
                    const svo_hits& out, void* stream, std::string* err);
int device_render_lod(const DeviceTree& t, const Camera& cam, uint32_t w, uint32_t h,
                      float viewing_distance, const svo_hits& out, void* stream,
                      std::string* err);
// Shaded RGBA8 render (conventions described at shade_kernel).
int device_render_rgba8(const DeviceTree& t, V3f origin, V3f dir, float glass_w, float glass_h,
                        float fov, uint32_t w, uint32_t h, uint32_t convention, uint32_t* out,
                        void* stream, std::string* err);

// Launch variants of the render kernel (occupancy hint, LUT placement, tile
// shape); the process-wide selection applies to later svo_render calls.
int render_variant_count();
const char* render_variant_name(int i);
// Private (scratch) bytes per lane of a variant's kernel: non-zero when the
// build spills registers.
int render_variant_scratch(int i, uint32_t* bytes, std::string* err);
bool render_variant_select(int i);
int render_variant_current();

// Kernel timing with HIP events recorded on the launch stream.
void timing_enable(bool on);
uint32_t timing_collect(float* out_ms, uint32_t cap);

}  // namespace svo
