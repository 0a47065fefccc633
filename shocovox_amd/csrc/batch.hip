// Batched mode: every tile of every job in one persistent launch.
//
// A job is a viewport render or a ray-array cast on one resident tree. Its
// rays are cut into 64-ray tiles (8x8 pixels for renders), numbered
// contiguously across jobs. The tile space is split into 8 contiguous
// chunks, one per XCD (workgroups are dispatched to the 8 XCDs round-robin),
// so neighbouring tiles - which touch the same nodes and bricks - are served
// by one XCD's L2. Each wave takes the next tile of its XCD's chunk from a
// device counter and, once that chunk is exhausted, steals from the other
// chunks; every wave leaves when all chunks are empty.
//
// The per-run sequence - job table upload from pinned memory, counter reset
// (a fill kernel, not a memset), persistent kernel - is captured once in a
// hipGraph and replayed for each run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "device_common.hpp"

namespace svo {

namespace {

constexpr uint32_t kChunks = 8;  // one per XCD of an MI355X

struct alignas(16) BatchJob {
  uint32_t tree;        // index into the batch's tree table
  uint32_t kind;        // 0 render, 1 cast
  uint32_t width;       // render width or ray count
  uint32_t height;      // render height (1 for casts)
  uint32_t tiles_x;     // render: tiles per row
  uint32_t tile_begin;  // first global tile of the job
  uint32_t tile_count;
  uint32_t pad;
  Camera cam;
  const float* rays;
  HitPtrs out;
};

struct alignas(16) BatchHeader {
  uint32_t njobs;
  uint32_t total_tiles;
  uint32_t chunk_tiles;  // tiles per XCD chunk (the last chunk may be shorter)
  uint32_t pad;
};

struct alignas(128) Counter {
  uint32_t next;
  uint32_t pad[31];
};

// Job of a global tile: the last job whose tile_begin <= tile.
__device__ inline uint32_t find_job(const BatchJob* jobs, uint32_t njobs, uint32_t tile) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].tile_begin <= tile)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(64) void reset_counters_kernel(Counter* counters) {
  if (threadIdx.x < kChunks) counters[threadIdx.x].next = 0u;
}

__global__ __launch_bounds__(256) void batch_kernel(const BatchHeader* __restrict__ hdr,
                                                    const TreeView* __restrict__ trees,
                                                    const BatchJob* __restrict__ jobs,
                                                    Counter* __restrict__ counters) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t njobs = hdr->njobs;
  const uint32_t total = hdr->total_tiles;
  const uint32_t chunk = hdr->chunk_tiles;
  const uint32_t home = blockIdx.x % kChunks;
  for (uint32_t k = 0; k < kChunks; ++k) {
    const uint32_t c = (home + k) % kChunks;
    const uint32_t begin = c * chunk;
    const uint32_t end = min(total, begin + chunk);
    if (begin >= end) continue;
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(&counters[c].next, 1u);
      t = __shfl(t, 0, 64);
      const uint32_t tile = begin + t;
      if (tile >= end) break;
      const uint32_t j = find_job(jobs, njobs, tile);
      const BatchJob& job = jobs[j];
      const TreeView& tv = trees[job.tree];
      const uint32_t local = tile - job.tile_begin;
      if (job.kind == 0) {
        const uint32_t x = (local % job.tiles_x) * 8u + (lane & 7u);
        const uint32_t y = (local / job.tiles_x) * 8u + (lane >> 3);
        if (x < job.width && y < job.height) {
          const V3f d = camera_ray_dir(job.cam, x, y);
          const CastResult r = cast_ray(tv, job.cam.origin, d);
          store_result(tv, job.out, y * job.width + x, r, job.cam.origin);
        }
      } else {
        const uint32_t i = local * 64u + lane;
        if (i < job.width) {
          const float* rr = job.rays + 6ull * i;
          const V3f o{rr[0], rr[1], rr[2]};
          const V3f d{rr[3], rr[4], rr[5]};
          const CastResult r = cast_ray(tv, o, d);
          store_result(tv, job.out, i, r, o);
        }
      }
    }
  }
}

int hip_fail(hipError_t e, const char* what, std::string* err) {
  *err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SVO_OUT_OF_DEVICE_MEMORY : SVO_HIP_ERROR;
}

}  // namespace

struct BatchContext {
  int device = -1;
  uint32_t max_jobs = 0;
  size_t table_bytes = 0;  // header + trees + jobs
  void* host = nullptr;    // pinned staging of the table
  void* dev = nullptr;     // device table
  Counter* counters = nullptr;
  uint32_t grid = 0;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  bool pending = false;
  uint64_t last_tiles = 0;
};

namespace {

size_t trees_offset() { return 256; }
size_t jobs_offset(uint32_t max_jobs) {
  return (trees_offset() + max_jobs * sizeof(TreeView) + 255) & ~size_t(255);
}

int capture(BatchContext* b, std::string* err) {
  hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture", err);
  const char* base = static_cast<const char*>(b->dev);
  (void)hipMemcpyAsync(b->dev, b->host, b->table_bytes, hipMemcpyHostToDevice, b->capture_stream);
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, b->capture_stream, b->counters);
  hipLaunchKernelGGL(batch_kernel, dim3(b->grid), dim3(256), 0, b->capture_stream,
                     reinterpret_cast<const BatchHeader*>(base),
                     reinterpret_cast<const TreeView*>(base + trees_offset()),
                     reinterpret_cast<const BatchJob*>(base + jobs_offset(b->max_jobs)),
                     b->counters);
  e = hipStreamEndCapture(b->capture_stream, &b->graph);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", err);
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", err);
  return SVO_OK;
}

}  // namespace

int batch_create(uint32_t max_jobs, BatchContext** out, std::string* err) {
  *out = nullptr;
  if (max_jobs == 0 || max_jobs > (1u << 20)) {
    *err = "max_jobs must be in [1, 2^20]";
    return SVO_INVALID_ARGUMENT;
  }
  BatchContext* b = new BatchContext();
  b->max_jobs = max_jobs;
  b->table_bytes = jobs_offset(max_jobs) + max_jobs * sizeof(BatchJob);
  hipError_t e = hipGetDevice(&b->device);
  int cus = 0;
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e == hipSuccess) e = hipHostMalloc(&b->host, b->table_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->dev, b->table_bytes);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&b->counters), kChunks * sizeof(Counter));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    const int st = hip_fail(e, "batch_create", err);
    batch_destroy(b);
    return st;
  }
  std::memset(b->host, 0, b->table_bytes);
  // four 256-lane workgroups per CU keep every SIMD at the kernel's occupancy
  b->grid = static_cast<uint32_t>(std::max(cus, 1)) * 4u;
  const int st = capture(b, err);
  if (st != SVO_OK) {
    batch_destroy(b);
    return st;
  }
  *out = b;
  return SVO_OK;
}

void batch_destroy(BatchContext* b) {
  if (!b) return;
  if (b->device >= 0) (void)hipSetDevice(b->device);
  if (b->pending) (void)hipEventSynchronize(b->done);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  if (b->counters) (void)hipFree(b->counters);
  if (b->dev) (void)hipFree(b->dev);
  if (b->host) (void)hipHostFree(b->host);
  delete b;
}

uint64_t batch_tiles(const BatchContext* b) { return b ? b->last_tiles : 0; }

int batch_run(BatchContext* b, const BatchJobIn* jobs, uint32_t n, void* stream, std::string* err) {
  if (n > b->max_jobs) {
    *err = "more jobs than the batch was created for";
    return SVO_INVALID_ARGUMENT;
  }
  // the previous run's graph reads the pinned table: let it finish first
  if (b->pending) {
    const hipError_t e = hipEventSynchronize(b->done);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize", err);
    b->pending = false;
  }
  char* base = static_cast<char*>(b->host);
  BatchHeader* hdr = reinterpret_cast<BatchHeader*>(base);
  TreeView* trees = reinterpret_cast<TreeView*>(base + trees_offset());
  BatchJob* out = reinterpret_cast<BatchJob*>(base + jobs_offset(b->max_jobs));
  std::unordered_map<const void*, uint32_t> tree_index;
  uint64_t tiles = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const BatchJobIn& in = jobs[i];
    if (!in.tree || in.tree->device != b->device) {
      *err = "job tree is missing or lives on another device";
      return SVO_INVALID_ARGUMENT;
    }
    auto it = tree_index.find(in.tree);
    uint32_t ti;
    if (it == tree_index.end()) {
      ti = static_cast<uint32_t>(tree_index.size());
      tree_index.emplace(in.tree, ti);
      trees[ti] = in.tree->view;
    } else {
      ti = it->second;
    }
    BatchJob& j = out[i];
    std::memset(&j, 0, sizeof(j));
    j.tree = ti;
    j.kind = in.kind;
    j.width = in.width;
    j.height = in.kind == 0 ? in.height : 1u;
    if (in.kind == 0) {
      j.tiles_x = (in.width + 7) / 8;
      j.tile_count = j.tiles_x * ((in.height + 7) / 8);
      j.cam = in.cam;
    } else {
      if (in.width && !in.rays) {
        *err = "cast job without rays";
        return SVO_INVALID_ARGUMENT;
      }
      j.tiles_x = 0;
      j.tile_count = (in.width + 63) / 64;
      j.rays = in.rays;
    }
    j.tile_begin = static_cast<uint32_t>(tiles);
    j.out = to_ptrs(in.out);
    tiles += j.tile_count;
    if (tiles >= (1ull << 31)) {
      *err = "too many tiles in one batch";
      return SVO_INVALID_ARGUMENT;
    }
  }
  hdr->njobs = n;
  hdr->total_tiles = static_cast<uint32_t>(tiles);
  hdr->chunk_tiles = static_cast<uint32_t>((tiles + kChunks - 1) / kChunks);
  b->last_tiles = tiles;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch", err);
  e = hipEventRecord(b->done, s);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord", err);
  b->pending = true;
  return SVO_OK;
}

}  // namespace svo
