// Serving loop: seeded Poisson arrivals of ray-cast requests over resident
// trees, served by continuous batching (mode 0) or by one plain launch per
// request (mode 1), with sparse voxel edits applied in place in per-tree
// arrival order.
//
// Edits run on a pool of host threads (host octree edit and layout update,
// then an asynchronous patch on the thread's own stream). The host never
// waits for the device to order them: an edit starts once every earlier
// render of its tree has been launched, its patch waits on the device for
// the launches still running them (events recorded on their streams) and
// for the tree's previous patch, and launches that render the tree after the
// edit wait on the device for its patch. Every wait refers to an event
// recorded earlier on the host, so the waits cannot form a cycle.
//
// Continuous batching here is iteration-level scheduling: every batched
// launch is short (at most max_batch_rays rays), and at each launch
// boundary the host packs the queue head - remaining chunks of requests
// already started, then new arrivals - into the next launch. Up to
// max_inflight launches sit in their own streams at once, each with its own
// batch context (job table, graph) and output pool, so the next launch is
// formed while the previous ones run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/svo.h"
#include "batch.hpp"
#include "handles.hpp"

namespace svo {

namespace {

int serve_fail(int st, const std::string& msg) {
  set_last_error(msg);
  return st;
}

// ---------------------------------------------------------------- plan ---

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// xoshiro256** (Blackman and Vigna), seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (uint64_t& v : s) v = splitmix64(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  uint32_t below(uint32_t n) {
    return static_cast<uint32_t>(((next() >> 32) * static_cast<uint64_t>(n)) >> 32);
  }
};

struct Zipf {
  std::vector<double> cdf;
  Zipf(uint32_t n, double s) {
    double acc = 0.0;
    for (uint32_t k = 1; k <= n; ++k) {
      acc += 1.0 / std::pow(static_cast<double>(k), s);
      cdf.push_back(acc);
    }
    for (double& c : cdf) c /= acc;
  }
  uint32_t draw(Rng& r) const {
    const double u = r.uniform();
    const auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    return static_cast<uint32_t>(std::min<size_t>(it - cdf.begin(), cdf.size() - 1));
  }
};

constexpr uint64_t kMaxRequests = 1ull << 24;
constexpr double kDrainSeconds = 120.0;
// smallest share of a batched launch given to each waiting request (64 tiles
// of 8x8 pixels = 4096 rays), so that many waiting requests do not cut a
// launch into slivers
constexpr uint64_t kMinShareTiles = 64;
// pause between polls of running launches and edits when nothing changed
constexpr double kPollSeconds = 20e-6;
// host threads that apply edits (each tree's edits stay in order on one at a time)
constexpr uint32_t kEditWorkers = 4;

int check_config(const svo_serve_config* cfg, uint32_t n_trees) {
  if (!cfg) return serve_fail(SVO_INVALID_ARGUMENT, "config is NULL");
  if (n_trees == 0) return serve_fail(SVO_INVALID_ARGUMENT, "no trees");
  if (!(cfg->qps > 0.0) || !(cfg->duration_s >= 0.0) || !std::isfinite(cfg->qps) ||
      !std::isfinite(cfg->duration_s))
    return serve_fail(SVO_INVALID_ARGUMENT, "qps must be > 0 and duration_s >= 0");
  if (cfg->qps * cfg->duration_s > static_cast<double>(kMaxRequests))
    return serve_fail(SVO_INVALID_ARGUMENT, "plan exceeds 2^24 requests");
  if (cfg->n_resolutions == 0 || !cfg->resolutions)
    return serve_fail(SVO_INVALID_ARGUMENT, "no resolutions");
  for (uint32_t i = 0; i < cfg->n_resolutions; ++i) {
    const uint32_t w = cfg->resolutions[2 * i], h = cfg->resolutions[2 * i + 1];
    if (w == 0 || h == 0 || w > 16384 || h > 16384)
      return serve_fail(SVO_INVALID_ARGUMENT, "resolution outside 1..16384");
  }
  if (cfg->edit_fraction < 0.0 || cfg->edit_fraction > 1.0)
    return serve_fail(SVO_INVALID_ARGUMENT, "edit_fraction outside [0, 1]");
  if (cfg->edit_fraction > 0.0 && cfg->max_edit_voxels == 0)
    return serve_fail(SVO_INVALID_ARGUMENT, "max_edit_voxels is 0");
  if (cfg->mode > 2) return serve_fail(SVO_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
  if (!(cfg->deadline_ms >= 0.0) || !std::isfinite(cfg->deadline_ms))
    return serve_fail(SVO_INVALID_ARGUMENT, "deadline_ms must be >= 0");
  if (cfg->max_inflight == 0 || cfg->max_inflight > 16)
    return serve_fail(SVO_INVALID_ARGUMENT, "max_inflight outside 1..16");
  if (cfg->mode != 1 && (cfg->max_batch_rays < 64 || cfg->max_batch_rays > (1ull << 30) ||
                         cfg->max_batch_jobs == 0 || cfg->max_batch_jobs > 4096))
    return serve_fail(SVO_INVALID_ARGUMENT, "batch budget outside 64..2^30 rays, 1..4096 jobs");
  return SVO_OK;
}

svo_viewport look_at(double cx, double r, double az, double el, uint32_t w, uint32_t h) {
  svo_viewport v{};
  const float eye[3] = {static_cast<float>(cx + r * std::cos(el) * std::cos(az)),
                        static_cast<float>(cx + r * std::sin(el)),
                        static_cast<float>(cx + r * std::cos(el) * std::sin(az))};
  const float c = static_cast<float>(cx);
  const float d[3] = {c - eye[0], c - eye[1], c - eye[2]};
  const float ln = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  for (int i = 0; i < 3; ++i) {
    v.origin[i] = eye[i];
    v.direction[i] = d[i] / ln;
  }
  const float gh = static_cast<float>(2.0 * std::tan(M_PI / 6.0));
  v.frustum[0] = gh * static_cast<float>(w) / static_cast<float>(h);
  v.frustum[1] = gh;
  v.frustum[2] = 1e9f;
  v.fov = 1.0f;
  return v;
}

std::vector<svo_serve_request> make_plan(const svo_serve_config& cfg, const uint32_t* sizes,
                                         uint32_t n_trees) {
  std::vector<svo_serve_request> out;
  Rng rng(cfg.seed);
  const Zipf ztree(n_trees, cfg.zipf_s);
  const Zipf zres(cfg.n_resolutions, cfg.zipf_s);
  double t = 0.0;
  for (;;) {
    t += -std::log1p(-rng.uniform()) / cfg.qps;
    if (t >= cfg.duration_s || out.size() >= kMaxRequests) break;
    svo_serve_request q{};
    q.arrival_s = t;
    q.tree = ztree.draw(rng);
    const double kind_u = rng.uniform();
    if (kind_u < cfg.edit_fraction) {
      q.kind = 1;
      q.edit_voxels = 1 + rng.below(cfg.max_edit_voxels);
      q.edit_clear = rng.below(2);
      q.edit_seed = rng.next();
    } else {
      q.kind = 0;
      const uint32_t r = zres.draw(rng);
      q.width = cfg.resolutions[2 * r];
      q.height = cfg.resolutions[2 * r + 1];
      const double az = 2.0 * M_PI * rng.uniform();
      const double el = (10.0 + 50.0 * rng.uniform()) * M_PI / 180.0;
      const double s = static_cast<double>(sizes[q.tree]);
      q.viewport = look_at(s / 2.0, 1.5 * s, az, el, q.width, q.height);
    }
    out.push_back(q);
  }
  return out;
}

int apply_edit(Octree& tree, const svo_serve_request& q) {
  Rng rng(q.edit_seed);
  const uint32_t s = tree.size();
  for (uint32_t k = 0; k < q.edit_voxels; ++k) {
    const uint32_t x = rng.below(s), y = rng.below(s), z = rng.below(s);
    Status st;
    if (q.edit_clear) {
      st = tree.clear(x, y, z);
    } else {
      Entry e;
      e.kind = EntryKind::Visual;
      // one of 16 fixed opaque colours, so that edits reuse palette entries
      const uint32_t c = rng.below(16);
      e.albedo = (((c * 53u + 40u) & 0xFFu) << 24) | (((c * 97u + 80u) & 0xFFu) << 16) |
                 (((c * 151u + 120u) & 0xFFu) << 8) | 0xFFu;
      st = tree.insert(x, y, z, e);
    }
    if (st != Status::Ok) return SVO_INVALID_POSITION;
  }
  return SVO_OK;
}

// -------------------------------------------------------------- digest ---

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct DigestJob {
  unsigned long long* target;
  const uint32_t* hit;
  const uint32_t* albedo;
  const uint32_t* voxel;
  const float* distance;
  uint32_t count;       // output slots of the job
  uint32_t tile_major;  // slot k is pixel (k & 7, (k >> 3) & 7) of tile first_tile + k / 64
  uint32_t first_tile;
  uint32_t tiles_x;
  uint32_t width;
  uint32_t height;
};

__global__ __launch_bounds__(256) void digest_kernel(const DigestJob* __restrict__ jobs) {
  const DigestJob& j = jobs[blockIdx.y];
  unsigned long long acc = 0;
  for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < j.count; k += gridDim.x * 256u) {
    uint32_t i = k;
    if (j.tile_major) {
      const uint32_t tile = j.first_tile + (k >> 6);
      const uint32_t x = (tile % j.tiles_x) * 8u + (k & 7u);
      const uint32_t y = (tile / j.tiles_x) * 8u + ((k >> 3) & 7u);
      if (x >= j.width || y >= j.height) continue;
      i = y * j.width + x;
    }
    uint64_t a = mix64(static_cast<uint64_t>(i) + 0x9E3779B97F4A7C15ull);
    a = mix64(a ^ j.hit[k]);
    a = mix64(a ^ j.albedo[k]);
    a = mix64(a ^ j.voxel[3ull * k + 0]);
    a = mix64(a ^ j.voxel[3ull * k + 1]);
    a = mix64(a ^ j.voxel[3ull * k + 2]);
    a = mix64(a ^ __float_as_uint(j.distance[k]));
    acc += a;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63u) == 0 && acc != 0) atomicAdd(j.target, acc);
}

__global__ __launch_bounds__(256) void zero_u64_kernel(unsigned long long* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) p[i] = 0;
}

// ---------------------------------------------------------------- loop ---

using Clock = std::chrono::steady_clock;

struct Chunk {
  uint32_t req;
  uint32_t first_tile;
  uint32_t tiles;
  uint64_t offset;  // first output slot in the pool
};

struct Slot {
  BatchContext* batch = nullptr;  // mode 0
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint64_t cap = 0;  // output slots
  uint32_t* hit = nullptr;
  uint32_t* albedo = nullptr;
  uint32_t* voxel = nullptr;
  float* distance = nullptr;
  DigestJob* h_digest = nullptr;
  DigestJob* d_digest = nullptr;
  uint32_t digest_cap = 0;
  std::vector<Chunk> chunks;
  bool busy = false;
};

struct EditTimes {
  double start = 0, applied = 0;  // host clock: edit started, host octree and layout edited
  double upload_ms = 0;           // device time of the patch (its copy and scatter)
};

// One edit handed to the workers. The patch waits on the device for `waits`
// (events recorded on the streams of launches that hold earlier renders of
// the tree) and, after an earlier edit of the tree, for that edit's patch;
// `start` and `done` bracket the patch on the worker's stream.
struct EditJob {
  uint32_t req = 0;
  bool after_edit = false;
  std::vector<hipEvent_t> waits;
  hipEvent_t start = nullptr, done = nullptr;
};

// Edits run on a pool of host threads, each with its own stream: edits of
// different trees proceed in parallel, edits of one tree never overlap (the
// loop hands out the next edit of a tree only after the previous one's host
// part has finished).
struct EditPool {
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<EditJob> todo;
  std::vector<std::pair<EditJob, EditTimes>> finished;  // host part done, patch queued
  bool stop = false;
  int status = SVO_OK;
  std::string error;
  uint64_t records = 0, bricks = 0;
};

struct Session {
  svo_octree* const* trees;
  svo_tree* const* dtrees;
  uint32_t n_trees;
  svo_serve_config cfg;
  std::vector<svo_serve_request> plan;
  int device = 0;
  Clock::time_point t0;
  hipEvent_t ref = nullptr;
  std::vector<Slot> slots;
  unsigned long long* d_digests = nullptr;
  EditPool edits;
  // per tree: recorded after its last patch; launches that render the tree
  // after an edit wait for it on the device
  std::vector<hipEvent_t> tree_ready;
  std::vector<uint8_t> has_ready;
  std::vector<hipEvent_t> events;  // every event the session created
  std::vector<hipEvent_t> free_events;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  hipEvent_t take_event() {
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    events.push_back(e);
    return e;
  }
};

uint64_t tiles_of(const svo_serve_request& q) {
  return static_cast<uint64_t>((q.width + 7) / 8) * ((q.height + 7) / 8);
}

int hip_err(hipError_t e, const char* what) {
  return serve_fail(e == hipErrorOutOfMemory ? SVO_OUT_OF_DEVICE_MEMORY : SVO_HIP_ERROR,
                    std::string(what) + ": " + hipGetErrorString(e));
}

void edit_worker_main(Session* s) {
  EditPool& w = s->edits;
  (void)hipSetDevice(s->device);
  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
    std::lock_guard<std::mutex> g(w.mu);
    w.status = SVO_HIP_ERROR;
    w.error = "edit stream creation failed";
  }
  for (;;) {
    EditJob job;
    {
      std::unique_lock<std::mutex> g(w.mu);
      w.cv.wait(g, [&] { return w.stop || !w.todo.empty(); });
      if (w.todo.empty()) break;
      job = std::move(w.todo.front());
      w.todo.pop_front();
    }
    const svo_serve_request& q = s->plan[job.req];
    int st = SVO_HIP_ERROR;
    std::string err = "edit stream missing";
    uint32_t recs = 0, bricks = 0;
    EditTimes et;
    et.start = s->now();
    if (stream) {
      // device order: after the launches that hold earlier renders of the
      // tree, and after the tree's previous patch
      hipError_t e = hipSuccess;
      for (hipEvent_t ev : job.waits)
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, ev, 0);
      if (e == hipSuccess && job.after_edit)
        e = hipStreamWaitEvent(stream, s->tree_ready[q.tree], 0);
      st = apply_edit(s->trees[q.tree]->tree, q);
      if (e != hipSuccess) {
        st = SVO_HIP_ERROR;
        err = std::string("edit ordering: ") + hipGetErrorString(e);
      } else {
        // the layout update and the patch's staging are host work; job.start
        // marks the start of the patch's device work
        // the patch records job.done and the tree's ready event right after
        // its copies (ahead of its staging's release on the host)
        void* const ends[2] = {job.done, s->tree_ready[q.tree]};
        if (st == SVO_OK)
          st = tree_update(s->dtrees[q.tree], s->trees[q.tree], stream, job.start, ends, 2);
        if (st == SVO_OK) (void)svo_tree_update_stats(s->dtrees[q.tree], &recs, &bricks);
        err = st == SVO_OK ? "" : svo_last_error();
      }
      et.applied = s->now();
    }
    std::lock_guard<std::mutex> g(w.mu);
    if (st != SVO_OK && w.status == SVO_OK) {
      w.status = st;
      w.error = "edit failed: " + err;
    }
    w.records += recs;
    w.bricks += bricks;
    w.finished.emplace_back(std::move(job), et);
  }
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
}

void free_slot(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  if (sl.batch) batch_destroy(sl.batch);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  if (sl.hit) (void)hipFree(sl.hit);
  if (sl.albedo) (void)hipFree(sl.albedo);
  if (sl.voxel) (void)hipFree(sl.voxel);
  if (sl.distance) (void)hipFree(sl.distance);
  if (sl.h_digest) (void)hipHostFree(sl.h_digest);
  if (sl.d_digest) (void)hipFree(sl.d_digest);
  sl = Slot();
}

int alloc_slot(Slot& sl, uint64_t cap, uint32_t digest_cap, uint32_t batch_jobs) {
  hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&sl.done);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.hit), cap * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.albedo), cap * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.voxel), cap * 12);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.distance), cap * 4);
  if (e == hipSuccess && digest_cap) {
    e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_digest), digest_cap * sizeof(DigestJob),
                      hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipMalloc(reinterpret_cast<void**>(&sl.d_digest), digest_cap * sizeof(DigestJob));
  }
  if (e != hipSuccess) return hip_err(e, "serving slot allocation");
  sl.cap = cap;
  sl.digest_cap = digest_cap;
  if (batch_jobs) {
    std::string err;
    const int st = batch_create_variant(batch_jobs, batch_variant_serving(), &sl.batch, &err);
    if (st != SVO_OK) return serve_fail(st, err);
  }
  return SVO_OK;
}

// Launches the slot's chunks (already in sl.chunks) and records its event;
// *single is set when a mode-0 launch ran as one plain render.
int launch_slot(Session& s, Slot& sl, bool* single) {
  // Mode 0 runs a launch that holds one whole request as a plain render (the
  // persistent kernel's job table, counters and graph buy nothing for one
  // job); mode 2 uses the persistent kernel for every launch.
  const Chunk& c0 = sl.chunks[0];
  const bool whole = sl.chunks.size() == 1 && c0.first_tile == 0 &&
                     c0.tiles == tiles_of(s.plan[c0.req]);
  const bool batched = s.cfg.mode == 2 || (s.cfg.mode == 0 && !whole);
  *single = s.cfg.mode == 0 && whole;
  std::string err;
  // renders of a tree that has been edited wait, on the device, for its
  // last patch (the host already waited for the edit's host part)
  for (size_t i = 0; i < sl.chunks.size(); ++i) {
    const uint32_t tr = s.plan[sl.chunks[i].req].tree;
    bool seen = false;
    for (size_t k = 0; k < i && !seen; ++k) seen = s.plan[sl.chunks[k].req].tree == tr;
    if (seen || !s.has_ready[tr]) continue;
    const hipError_t e = hipStreamWaitEvent(sl.stream, s.tree_ready[tr], 0);
    if (e != hipSuccess) return hip_err(e, "hipStreamWaitEvent(tree)");
  }
  if (batched) {
    std::vector<BatchJobIn> jobs(sl.chunks.size());
    for (size_t i = 0; i < sl.chunks.size(); ++i) {
      const Chunk& c = sl.chunks[i];
      const svo_serve_request& q = s.plan[c.req];
      BatchJobIn& j = jobs[i];
      j.tree = s.dtrees[q.tree]->dev;
      j.kind = 0;
      j.width = q.width;
      j.height = q.height;
      j.cam = make_camera(V3f{q.viewport.origin[0], q.viewport.origin[1], q.viewport.origin[2]},
                          V3f{q.viewport.direction[0], q.viewport.direction[1],
                              q.viewport.direction[2]},
                          q.viewport.frustum[0], q.viewport.frustum[1], q.viewport.fov, q.width,
                          q.height);
      j.rays = nullptr;
      j.out = svo_hits{};
      j.out.hit = sl.hit + c.offset;
      j.out.albedo = sl.albedo + c.offset;
      j.out.voxel = sl.voxel + 3 * c.offset;
      j.out.distance = sl.distance + c.offset;
      j.first_tile = c.first_tile;
      j.tile_limit = c.tiles;
      j.tile_major = 1;
    }
    const int st = batch_run(sl.batch, jobs.data(), static_cast<uint32_t>(jobs.size()), sl.stream,
                             &err);
    if (st != SVO_OK) return serve_fail(st, err);
  } else {
    const Chunk& c = sl.chunks[0];
    const svo_serve_request& q = s.plan[c.req];
    const Camera cam = make_camera(
        V3f{q.viewport.origin[0], q.viewport.origin[1], q.viewport.origin[2]},
        V3f{q.viewport.direction[0], q.viewport.direction[1], q.viewport.direction[2]},
        q.viewport.frustum[0], q.viewport.frustum[1], q.viewport.fov, q.width, q.height);
    svo_hits out{};
    out.hit = sl.hit;
    out.albedo = sl.albedo;
    out.voxel = sl.voxel;
    out.distance = sl.distance;
    const int st = device_render(*s.dtrees[q.tree]->dev, cam, q.width, q.height, out, sl.stream,
                                 &err);
    if (st != SVO_OK) return serve_fail(st, err);
  }
  if (s.d_digests) {
    uint32_t max_count = 0;
    for (size_t i = 0; i < sl.chunks.size(); ++i) {
      const Chunk& c = sl.chunks[i];
      const svo_serve_request& q = s.plan[c.req];
      DigestJob& d = sl.h_digest[i];
      d.target = s.d_digests + c.req;
      d.hit = sl.hit + c.offset;
      d.albedo = sl.albedo + c.offset;
      d.voxel = sl.voxel + 3 * c.offset;
      d.distance = sl.distance + c.offset;
      d.count = batched ? c.tiles * 64u : q.width * q.height;
      d.tile_major = batched ? 1u : 0u;
      d.first_tile = c.first_tile;
      d.tiles_x = (q.width + 7) / 8;
      d.width = q.width;
      d.height = q.height;
      max_count = std::max(max_count, d.count);
    }
    hipError_t e = hipMemcpyAsync(sl.d_digest, sl.h_digest, sl.chunks.size() * sizeof(DigestJob),
                                  hipMemcpyHostToDevice, sl.stream);
    if (e != hipSuccess) return hip_err(e, "digest table copy");
    const uint32_t gx = std::min<uint32_t>((max_count + 255) / 256, 512);
    hipLaunchKernelGGL(digest_kernel, dim3(std::max(gx, 1u), sl.chunks.size()), dim3(256), 0,
                       sl.stream, sl.d_digest);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_err(e, "digest_kernel");
  }
  const hipError_t e = hipEventRecord(sl.done, sl.stream);
  if (e != hipSuccess) return hip_err(e, "hipEventRecord");
  sl.busy = true;
  return SVO_OK;
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p * static_cast<double>(v.size() - 1);
  const size_t lo = static_cast<size_t>(pos);
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
}

int run_session(Session& s, svo_serve_stats* stats, float* latencies, uint64_t* digests,
                uint64_t cap) {
  const svo_serve_config& cfg = s.cfg;
  const size_t n = s.plan.size();
  const bool batched = cfg.mode != 1;

  // output pools: a batched slot holds max_batch_rays (whole tiles); a plain
  // slot holds the largest request of the plan
  uint64_t max_req = 64;
  for (const svo_serve_request& q : s.plan)
    if (q.kind == 0) max_req = std::max<uint64_t>(max_req, tiles_of(q) * 64);
  const uint64_t pool = batched ? (cfg.max_batch_rays / 64) * 64 : max_req;
  const uint32_t jobs_per_slot = batched ? cfg.max_batch_jobs : 1;
  s.slots.resize(cfg.max_inflight);
  for (Slot& sl : s.slots) {
    const int st = alloc_slot(sl, pool, digests ? jobs_per_slot : 0, batched ? jobs_per_slot : 0);
    if (st != SVO_OK) return st;
  }
  if (digests) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s.d_digests), std::max<size_t>(n, 1) * 8);
    if (e != hipSuccess) return hip_err(e, "digest buffer");
    hipLaunchKernelGGL(zero_u64_kernel, dim3(64), dim3(256), 0, nullptr, s.d_digests,
                       static_cast<uint64_t>(std::max<size_t>(n, 1)));
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_err(e, "digest buffer clear");
  }

  std::vector<double> finish(n, -1.0);
  std::vector<uint8_t> rejected(n, 0);
  std::vector<uint64_t> next_tile(n, 0);
  std::vector<uint32_t> outstanding(n, 0);
  std::vector<uint8_t> editing(s.n_trees, 0);  // host part of an edit in progress
  struct PendingEdit {
    uint32_t req;
    hipEvent_t start, done;
  };
  std::vector<PendingEdit> pending_edits;  // patches queued on the device
  std::vector<uint8_t> blocked(s.n_trees, 0), seen_render(s.n_trees, 0);
  std::vector<uint8_t> started_edit(n, 0), dispatched(n, 0), timed_out(n, 0);
  const double deadline = cfg.deadline_ms * 1e-3;
  std::vector<uint32_t> eligible;
  std::vector<uint64_t> take;
  std::deque<uint32_t> queue;
  size_t next_arrival = 0, done = 0;
  uint64_t launches = 0, chunks_launched = 0, rays_done = 0, single_launches = 0;

  // every tree that an edit of the plan touches gets its patch stagings and
  // its ready event now, so no edit allocates pinned or device memory inside
  // the session
  uint32_t edited_trees = 0;
  s.tree_ready.assign(s.n_trees, nullptr);
  s.has_ready.assign(s.n_trees, 0);
  {
    std::vector<uint8_t> edited(s.n_trees, 0);
    for (const svo_serve_request& q : s.plan)
      if (q.kind == 1) edited[q.tree] = 1;
    for (uint32_t i = 0; i < s.n_trees; ++i) {
      if (!edited[i]) continue;
      ++edited_trees;
      std::string err;
      if (device_tree_reserve_staging(s.dtrees[i]->dev, 1, &err) != SVO_OK)
        return serve_fail(SVO_HIP_ERROR, "edit staging: " + err);
      hipError_t e = hipEventCreateWithFlags(&s.tree_ready[i], hipEventDisableTiming);
      if (e != hipSuccess) return hip_err(e, "hipEventCreate(tree)");
    }
  }
  hipError_t e = hipEventCreate(&s.ref);
  if (e != hipSuccess) return hip_err(e, "hipEventCreate");
  for (uint32_t i = 0; i < 64; ++i) {
    hipEvent_t ev = s.take_event();
    if (!ev) return serve_fail(SVO_HIP_ERROR, "event creation failed");
    s.free_events.push_back(ev);
  }
  std::vector<EditTimes> edit_times(n);
  const uint32_t workers = std::max(1u, std::min(kEditWorkers, edited_trees));
  for (uint32_t i = 0; i < workers; ++i) s.edits.threads.emplace_back(edit_worker_main, &s);
  e = hipEventRecord(s.ref, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(s.ref);
  if (e != hipSuccess) return hip_err(e, "reference event");
  s.t0 = Clock::now();

  int status = SVO_OK;
  bool changed = true;
  // a session that has not drained this long after its last arrival is a bug
  const double drain_limit = (n ? s.plan[n - 1].arrival_s : 0.0) + kDrainSeconds;
  while (done < n && status == SVO_OK) {
    const double t = s.now();
    const size_t done_before = done, arrivals_before = next_arrival;
    const uint64_t launches_before = launches;
    if (t > drain_limit) {
      status = serve_fail(SVO_INVALID_ARGUMENT, "serving session did not drain");
      break;
    }
    while (next_arrival < n && s.plan[next_arrival].arrival_s <= t) {
      const uint32_t r = static_cast<uint32_t>(next_arrival++);
      if (cfg.max_queue && queue.size() >= cfg.max_queue) {
        rejected[r] = 1;
        ++done;
      } else {
        queue.push_back(r);
      }
      changed = true;
    }
    // finished launches
    for (Slot& sl : s.slots) {
      if (!sl.busy) continue;
      e = hipEventQuery(sl.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) {
        status = hip_err(e, "launch failed");
        break;
      }
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, s.ref, sl.done);
      const double tf = e == hipSuccess ? ms * 1e-3 : s.now();
      for (const Chunk& c : sl.chunks) {
        const svo_serve_request& q = s.plan[c.req];
        // chunks of one request may finish out of order on different streams
        finish[c.req] = std::max(finish[c.req], tf);
        if (--outstanding[c.req] == 0 && next_tile[c.req] == tiles_of(q)) {
          rays_done += static_cast<uint64_t>(q.width) * q.height;
          ++done;
        }
      }
      sl.chunks.clear();
      sl.busy = false;
      changed = true;
    }
    // edits whose host part finished: renders of the tree may launch again
    // (they wait for its patch on the device); the edit itself completes
    // when its patch does
    {
      std::lock_guard<std::mutex> g(s.edits.mu);
      if (s.edits.status != SVO_OK) status = serve_fail(s.edits.status, s.edits.error);
      for (auto& f : s.edits.finished) {
        const uint32_t tr = s.plan[f.first.req].tree;
        editing[tr] = 0;
        s.has_ready[tr] = 1;
        edit_times[f.first.req] = f.second;
        for (hipEvent_t ev : f.first.waits) s.free_events.push_back(ev);
        pending_edits.push_back(PendingEdit{f.first.req, f.first.start, f.first.done});
        changed = true;
      }
      s.edits.finished.clear();
    }
    if (status != SVO_OK) break;
    for (size_t i = 0; i < pending_edits.size();) {
      PendingEdit& pe = pending_edits[i];
      e = hipEventQuery(pe.done);
      if (e == hipErrorNotReady) {
        ++i;
        continue;
      }
      if (e != hipSuccess) {
        status = hip_err(e, "edit patch failed");
        break;
      }
      float ms = 0.0f, up = 0.0f;
      e = hipEventElapsedTime(&ms, s.ref, pe.done);
      finish[pe.req] = e == hipSuccess ? ms * 1e-3 : s.now();
      if (hipEventElapsedTime(&up, pe.start, pe.done) == hipSuccess)
        edit_times[pe.req].upload_ms = up;
      s.free_events.push_back(pe.start);
      s.free_events.push_back(pe.done);
      ++done;
      changed = true;
      pending_edits[i] = pending_edits.back();
      pending_edits.pop_back();
    }
    if (status != SVO_OK) break;

    if (changed) {
      changed = false;
      Slot* free = nullptr;
      for (Slot& sl : s.slots)
        if (!sl.busy) {
          free = &sl;
          break;
        }
      std::fill(blocked.begin(), blocked.end(), 0);
      std::fill(seen_render.begin(), seen_render.end(), 0);
      // pass 1, in arrival order: start edits whose tree is idle, and list
      // the renders that may run now (not held behind an edit of their tree)
      eligible.clear();
      const double tnow = s.now();
      for (auto it = queue.begin(); it != queue.end(); ++it) {
        const uint32_t r = *it;
        const svo_serve_request& q = s.plan[r];
        const uint32_t tr = q.tree;
        // a request that has not started by its deadline is dropped
        if (deadline > 0.0 && tnow - q.arrival_s > deadline &&
            (q.kind == 1 || next_tile[r] == 0)) {
          timed_out[r] = 1;
          ++done;
          continue;
        }
        if (blocked[tr] || editing[tr]) {
          blocked[tr] = 1;
          continue;
        }
        if (q.kind == 1) {
          blocked[tr] = 1;
          // an edit starts once every earlier render of its tree has been
          // launched; the launches still running are waited for on the device
          if (!seen_render[tr]) {
            EditJob job;
            job.req = r;
            job.after_edit = s.has_ready[tr] != 0;
            job.start = s.take_event();
            job.done = s.take_event();
            bool ok = job.start && job.done;
            for (Slot& sl : s.slots) {
              if (!ok || !sl.busy) continue;
              bool holds = false;
              for (const Chunk& c : sl.chunks) holds = holds || s.plan[c.req].tree == tr;
              if (!holds) continue;
              hipEvent_t ev = s.take_event();
              ok = ev && hipEventRecord(ev, sl.stream) == hipSuccess;
              if (ev) job.waits.push_back(ev);
            }
            if (!ok) {
              status = serve_fail(SVO_HIP_ERROR, "edit ordering events");
              break;
            }
            editing[tr] = 1;
            started_edit[r] = 1;
            {
              std::lock_guard<std::mutex> g(s.edits.mu);
              s.edits.todo.push_back(std::move(job));
            }
            s.edits.cv.notify_one();
          }
          continue;
        }
        seen_render[tr] = 1;
        eligible.push_back(r);
      }
      // pass 2: tiles of the next launch. Batched: every eligible request
      // first gets an equal share of the launch (at least kMinShareTiles),
      // then what is left goes out in arrival order - so a request that
      // arrives behind a large one starts in the next launch instead of
      // after all of the large one's chunks. Plain: the oldest request.
      if (free && !eligible.empty()) {
        const uint64_t cap_tiles = free->cap / 64;
        const size_t njobs = std::min<size_t>(eligible.size(), jobs_per_slot);
        take.assign(njobs, 0);
        uint64_t budget = cap_tiles;
        if (batched) {
          const uint64_t share = std::max<uint64_t>(kMinShareTiles, cap_tiles / njobs);
          for (size_t i = 0; i < njobs && budget; ++i) {
            const uint32_t r = eligible[i];
            const uint64_t t = std::min({tiles_of(s.plan[r]) - next_tile[r], share, budget});
            take[i] = t;
            budget -= t;
          }
          for (size_t i = 0; i < njobs && budget; ++i) {
            const uint32_t r = eligible[i];
            const uint64_t t = std::min(tiles_of(s.plan[r]) - next_tile[r] - take[i], budget);
            take[i] += t;
            budget -= t;
          }
        } else {
          take[0] = tiles_of(s.plan[eligible[0]]) - next_tile[eligible[0]];
        }
        uint64_t used = 0;
        for (size_t i = 0; i < njobs; ++i) {
          if (take[i] == 0) continue;
          const uint32_t r = eligible[i];
          free->chunks.push_back(Chunk{r, static_cast<uint32_t>(next_tile[r]),
                                       static_cast<uint32_t>(take[i]), used * 64});
          next_tile[r] += take[i];
          used += take[i];
          ++outstanding[r];
          if (next_tile[r] == tiles_of(s.plan[r])) dispatched[r] = 1;
        }
      }
      // drop started edits and fully dispatched renders from the queue
      for (auto it = queue.begin(); it != queue.end();) {
        if (started_edit[*it] || dispatched[*it] || timed_out[*it])
          it = queue.erase(it);
        else
          ++it;
      }
      if (free && !free->chunks.empty()) {
        bool single = false;
        status = launch_slot(s, *free, &single);
        single_launches += single ? 1u : 0u;
        ++launches;
        chunks_launched += free->chunks.size();
        changed = true;  // another slot may be free for the rest of the queue
      }
    }
    if (status != SVO_OK) break;
    // Nothing happened in this pass: wait for the next arrival or completion
    // without spinning on the runtime (a tight hipEventQuery loop contends
    // with the edit worker's HIP calls and delays its uploads).
    bool busy = false;
    for (const Slot& sl : s.slots) busy = busy || sl.busy;
    if (deadline > 0.0 && !queue.empty()) changed = true;  // deadlines pass with time
    const bool progress =
        done != done_before || next_arrival != arrivals_before || launches != launches_before;
    if (!progress) {
      bool edits_running = !pending_edits.empty();
      for (uint8_t v : editing) edits_running = edits_running || v;
      const double wait = next_arrival < n ? s.plan[next_arrival].arrival_s - s.now() : 1.0;
      if (!busy && !edits_running && !changed && wait > 2e-4)
        std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      else if (wait > kPollSeconds)
        std::this_thread::sleep_for(std::chrono::duration<double>(kPollSeconds));
    }
  }

  {
    std::lock_guard<std::mutex> g(s.edits.mu);
    s.edits.stop = true;
  }
  s.edits.cv.notify_all();
  for (std::thread& th : s.edits.threads) th.join();
  s.edits.threads.clear();
  for (Slot& sl : s.slots)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  if (status != SVO_OK) return status;

  if (digests) {
    std::vector<uint64_t> h(n);
    if (n) {
      e = hipMemcpy(h.data(), s.d_digests, n * 8, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_err(e, "digest copy");
    }
    for (size_t i = 0; i < n && i < cap; ++i) digests[i] = s.plan[i].kind == 0 ? h[i] : 0;
  }
  std::vector<double> all, ren, edi;
  double first = n ? s.plan[0].arrival_s : 0.0, last = first;
  uint64_t renders = 0, edits = 0, nrej = 0;
  uint64_t ntimed = 0;
  for (size_t i = 0; i < n; ++i) {
    if (rejected[i]) {
      ++nrej;
      if (latencies && i < cap) latencies[i] = -1.0f;
      continue;
    }
    if (timed_out[i]) {
      ++ntimed;
      if (latencies && i < cap) latencies[i] = -2.0f;
      continue;
    }
    const double lat = std::max(0.0, finish[i] - s.plan[i].arrival_s) * 1e3;
    last = std::max(last, finish[i]);
    all.push_back(lat);
    (s.plan[i].kind == 0 ? ren : edi).push_back(lat);
    (s.plan[i].kind == 0 ? renders : edits) += 1;
    if (latencies && i < cap) latencies[i] = static_cast<float>(lat);
  }
  if (stats) {
    svo_serve_stats& o = *stats;
    o = svo_serve_stats{};
    o.requests = n;
    o.renders = renders;
    o.edits = edits;
    o.rejected = nrej;
    o.timed_out = ntimed;
    o.rays = rays_done;
    o.launches = launches;
    o.single_launches = single_launches;
    o.elapsed_s = last - first;
    o.mrays_per_s = o.elapsed_s > 0 ? static_cast<double>(rays_done) / o.elapsed_s / 1e6 : 0.0;
    o.p50_ms = percentile(all, 0.5);
    o.p99_ms = percentile(all, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : all) {
      sum += v;
      mx = std::max(mx, v);
    }
    o.mean_ms = all.empty() ? 0.0 : sum / static_cast<double>(all.size());
    o.max_ms = mx;
    o.render_p50_ms = percentile(ren, 0.5);
    o.render_p99_ms = percentile(ren, 0.99);
    o.edit_p50_ms = percentile(edi, 0.5);
    o.edit_p99_ms = percentile(edi, 0.99);
    o.mean_jobs_per_launch =
        launches ? static_cast<double>(chunks_launched) / static_cast<double>(launches) : 0.0;
    o.edit_records = s.edits.records;
    o.edit_bricks = s.edits.bricks;
    std::vector<double> wait, apply, upload;
    for (size_t i = 0; i < n; ++i) {
      if (s.plan[i].kind != 1 || rejected[i] || timed_out[i]) continue;
      const EditTimes& et = edit_times[i];
      wait.push_back((et.start - s.plan[i].arrival_s) * 1e3);
      apply.push_back((et.applied - et.start) * 1e3);
      upload.push_back(et.upload_ms);
    }
    o.edit_wait_p99_ms = percentile(wait, 0.99);
    o.edit_apply_p99_ms = percentile(apply, 0.99);
    o.edit_upload_p99_ms = percentile(upload, 0.99);
  }
  return SVO_OK;
}

}  // namespace

}  // namespace svo

extern "C" {

int svo_serve_plan(const svo_serve_config* cfg, const uint32_t* tree_sizes, uint32_t n_trees,
                   svo_serve_request* out, uint64_t cap, uint64_t* count) {
  const int st = svo::check_config(cfg, n_trees);
  if (st != SVO_OK) return st;
  if (!tree_sizes || !count) return svo::serve_fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const std::vector<svo_serve_request> plan = svo::make_plan(*cfg, tree_sizes, n_trees);
  *count = plan.size();
  if (out) std::memcpy(out, plan.data(), std::min<uint64_t>(cap, plan.size()) * sizeof(plan[0]));
  return SVO_OK;
}

int svo_serve_apply_edit(svo_octree* tree, const svo_serve_request* req) {
  if (!tree || !req || req->kind != 1)
    return svo::serve_fail(SVO_INVALID_ARGUMENT, "not an edit request");
  const int st = svo::apply_edit(tree->tree, *req);
  if (st != SVO_OK) return svo::serve_fail(st, "edit position outside the octree");
  return SVO_OK;
}

int svo_serve_run(svo_octree* const* trees, svo_tree* const* dtrees, uint32_t n_trees,
                  const svo_serve_config* cfg, svo_serve_stats* stats, float* latencies_ms,
                  uint64_t* digests, uint64_t cap) {
  int st = svo::check_config(cfg, n_trees);
  if (st != SVO_OK) return st;
  if (!trees || !dtrees) return svo::serve_fail(SVO_INVALID_ARGUMENT, "NULL trees");
  std::vector<uint32_t> sizes(n_trees);
  int device = -1;
  for (uint32_t i = 0; i < n_trees; ++i) {
    if (!trees[i] || !dtrees[i] || !dtrees[i]->dev)
      return svo::serve_fail(SVO_INVALID_ARGUMENT, "missing tree");
    if (dtrees[i]->dev->view.size != static_cast<float>(trees[i]->tree.size()))
      return svo::serve_fail(SVO_INVALID_ARGUMENT, "host and device tree sizes differ");
    if (device >= 0 && dtrees[i]->dev->device != device)
      return svo::serve_fail(SVO_INVALID_ARGUMENT, "trees live on different devices");
    device = dtrees[i]->dev->device;
    sizes[i] = trees[i]->tree.size();
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return svo::hip_err(e, "hipSetDevice");
  svo::Session s;
  s.trees = trees;
  s.dtrees = dtrees;
  s.n_trees = n_trees;
  s.cfg = *cfg;
  s.device = device;
  s.plan = svo::make_plan(*cfg, sizes.data(), n_trees);
  st = svo::run_session(s, stats, latencies_ms, digests, cap);
  if (!s.edits.threads.empty()) {
    {
      std::lock_guard<std::mutex> g(s.edits.mu);
      s.edits.stop = true;
    }
    s.edits.cv.notify_all();
    for (std::thread& th : s.edits.threads) th.join();
  }
  for (svo::Slot& sl : s.slots) svo::free_slot(sl);
  (void)hipDeviceSynchronize();  // patches of a failed session may still be queued
  for (hipEvent_t ev : s.events) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : s.tree_ready)
    if (ev) (void)hipEventDestroy(ev);
  if (s.d_digests) (void)hipFree(s.d_digests);
  if (s.ref) (void)hipEventDestroy(s.ref);
  return st;
}

int svo_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return svo::hip_err(e, "hipMemGetInfo");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return SVO_OK;
}

}  // extern "C"
