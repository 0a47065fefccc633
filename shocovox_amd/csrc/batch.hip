// Batched mode: every tile of every job in one persistent launch.
//
// A job is a viewport render or a ray-array cast on one resident tree. Its
// rays are cut into 64-ray tiles (8x8 pixels for renders), numbered
// contiguously across jobs. The tile space is split into 8 contiguous
// chunks, one per XCD (workgroups are dispatched to the 8 XCDs round-robin),
// so neighbouring tiles - which touch the same nodes and bricks - are served
// by one XCD's L2. Each wave takes the next tiles of its XCD's chunk from a
// device counter and, once that chunk is exhausted, steals from the other
// chunks; every wave leaves when all chunks are empty.
//
// Kernel variants: "classic" takes one tile per grab and runs it as one ray
// per lane (with an occupancy hint of 1 or 5 waves per SIMD), "ww" likewise
// with the while-while form of the traversal (cast_ray_ww); the refill
// variants take TPW tiles per grab and run them with lane refill
// (refill.hpp), one job segment at a time.
//
// The per-run sequence - job table upload from pinned memory (only the bytes
// the run uses), counter reset (a fill kernel, not a memset), persistent
// kernel - is a hipGraph built once per context; each run sets the copy's
// source and size and replays it. Two pinned stagings alternate, so the host
// fills the next run's table while the previous run executes; a run waits on
// its stream for the previous run (which may have been issued on another
// stream) before its copy overwrites the device table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include <mutex>

#include "batch.hpp"
#include "device_common.hpp"
#include "handles.hpp"
#include "refill.hpp"

namespace svo {

namespace {

constexpr uint32_t kChunks = 8;  // one per XCD of an MI355X

struct alignas(16) BatchJob {
  uint32_t tree;        // index into the run's tree table
  uint32_t kind;        // 0 render, 1 cast
  uint32_t width;       // render width or ray count
  uint32_t height;      // render height (1 for casts)
  uint32_t tiles_x;     // render: tiles per image row
  uint32_t tile_begin;  // first global tile of the job
  uint32_t tile_count;
  uint32_t first_tile;  // render: first image tile of the job
  uint32_t tile_major;  // 1: outputs stored tile after tile from out[0]
  float viewing_distance;  // > 0: get_by_ray_at_lod with this distance
  uint32_t pad[2];
  Camera cam;
  const float* rays;
  // cast jobs: the traversal order of the rays (lane k of the job's tiles
  // takes ray order[k], order.hip), or NULL for the rays' own order
  const uint32_t* order;
  HitPtrs out;
};

// Table of one run: [header][jobs x njobs][trees x ntrees]
struct alignas(16) BatchHeader {
  uint32_t njobs;
  uint32_t total_tiles;
  uint32_t chunk_tiles;   // tiles per XCD chunk (the last chunk may be shorter)
  uint32_t trees_offset;  // byte offset of the tree table
};
constexpr size_t kJobsOffset = 256;

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

__device__ inline const BatchHeader* header_of(const char* table) {
  return reinterpret_cast<const BatchHeader*>(table);
}
__device__ inline const BatchJob* jobs_of(const char* table) {
  return reinterpret_cast<const BatchJob*>(table + kJobsOffset);
}
__device__ inline const TreeView* trees_of(const char* table) {
  return reinterpret_cast<const TreeView*>(table + header_of(table)->trees_offset);
}

// LOD compiles in get_by_ray_at_lod for jobs with a viewing distance; a
// context launches that form only for runs that hold such jobs, so the plain
// form keeps its smaller register allocation.
// Render jobs are cut into TW x TH pixel tiles (64 pixels, one per lane).
// NODES > 0 stages the first NODES node records of a tree (the top levels of
// the breadth-first numbering) in LDS, where the traversal reads them:
// without SHARED each wave keeps the records of the tree of its current job
// and reloads them when it moves to a job on another tree; with SHARED the
// workgroup keeps those of the tree of the run's first job, loaded once at
// the start (jobs on other trees read global memory).
template <int MIN_WAVES, bool WW = false, bool LOD = false, int TW = 8, int TH = 8, int NODES = 0,
          bool SHARED = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIN_WAVES)))
void batch_kernel(const char* __restrict__ table, Counter* __restrict__ counters) {
  static_assert(TW * TH == 64, "a tile is one pixel per lane");
  static_assert(NODES == 0 || !LOD, "the LOD form reads every record from global memory");
  constexpr int kCaches = NODES > 0 && !SHARED ? 4 : 1;
  __shared__ NodeRec s_nodes[kCaches][NODES > 0 ? NODES : 1];
  const BatchHeader* hdr = header_of(table);
  const BatchJob* jobs = jobs_of(table);
  const TreeView* trees = trees_of(table);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t njobs = hdr->njobs;
  const uint32_t total = hdr->total_tiles;
  const uint32_t chunk = hdr->chunk_tiles;
  const uint32_t home = blockIdx.x % kChunks;
  NodeRec* cache = s_nodes[kCaches > 1 ? (threadIdx.x >> 6) : 0];
  uint32_t cached_tree = 0xFFFFFFFFu;
  uint32_t ncached = 0;
  if constexpr (NODES > 0 && SHARED) {
    if (njobs > 0) {
      const TreeView& tv0 = trees[jobs[0].tree];
      ncached = min(static_cast<uint32_t>(NODES), tv0.node_count);
      const uint4* src = reinterpret_cast<const uint4*>(tv0.nodes);
      uint4* dst = reinterpret_cast<uint4*>(cache);
      for (uint32_t i = threadIdx.x; i < ncached * 4u; i += 256u) dst[i] = src[i];
      cached_tree = jobs[0].tree;
    }
    __syncthreads();
  }
  for (uint32_t k = 0; k < kChunks; ++k) {
    const uint32_t c = (home + k) % kChunks;
    const uint32_t begin = c * chunk;
    const uint32_t end = min(total, begin + chunk);
    if (begin >= end) continue;
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(&counters[c].next, 1u);
      // wave-uniform from here: the job and its tree view load into SGPRs
      t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
      if (t >= end - begin) break;
      const uint32_t tile = begin + t;
      const uint32_t j = __builtin_amdgcn_readfirstlane(find_job(jobs, njobs, tile));
      const BatchJob& job = jobs[j];
      const TreeView& tv = trees[job.tree];
      const uint32_t local = tile - job.tile_begin;
      if constexpr (NODES > 0 && !SHARED) {
        if (job.tree != cached_tree) {
          // the wave's own cache: its lanes write it and then read what the
          // other lanes wrote, so the writes complete before the reads start
          ncached = min(static_cast<uint32_t>(NODES), tv.node_count);
          const uint4* src = reinterpret_cast<const uint4*>(tv.nodes);
          uint4* dst = reinterpret_cast<uint4*>(cache);
          for (uint32_t i = lane; i < ncached * 4u; i += 64u) dst[i] = src[i];
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          cached_tree = job.tree;
        }
      }
      const uint32_t ntop = NODES > 0 && job.tree == cached_tree ? ncached : 0u;
      // the lane's ray; one inlined traversal serves both job kinds
      V3f o, d;
      uint32_t idx;
      bool valid;
      if (job.kind == 0) {
        const uint32_t image_tile = local + job.first_tile;
        const uint32_t x = (image_tile % job.tiles_x) * TW + (lane % TW);
        const uint32_t y = (image_tile / job.tiles_x) * TH + (lane / TW);
        valid = x < job.width && y < job.height;
        o = job.cam.origin;
        d = valid ? camera_ray_dir(job.cam, x, y) : V3f{0.0f, 0.0f, 0.0f};
        idx = job.tile_major ? local * 64u + lane : y * job.width + x;
      } else {
        idx = local * 64u + lane;
        valid = idx < job.width;
        if (valid && job.order) idx = job.order[idx];
        const float* rr = job.rays + 6ull * (valid ? idx : 0u);
        o = V3f{rr[0], rr[1], rr[2]};
        d = V3f{rr[3], rr[4], rr[5]};
      }
      if (valid) {
        CastResult r;
        if constexpr (LOD) {
          if (job.viewing_distance > 0.0f)
            r = cast_ray_lod<true>(tv, tv.ray_lut, o, d, job.viewing_distance);
          else
            r = cast_ray<true>(tv, o, d);
        } else if constexpr (WW) {
          r = cast_ray_ww<true>(tv, tv.ray_lut, o, d, static_cast<const NodeRec*>(cache), ntop);
        } else {
          r = cast_ray<true>(tv, tv.ray_lut, o, d, static_cast<const NodeRec*>(cache), ntop);
        }
        // The origin and the output index are formed again for the store
        // rather than carried through the traversal (four registers fewer
        // there): the empty asm hides that `local` is unchanged, so the
        // compiler recomputes them (a render's origin is the job's camera,
        // a cast's is read again from its ray).
        uint32_t again = local;
        asm volatile("" : "+s"(again));
        if (job.kind == 0) {
          const uint32_t image_tile = again + job.first_tile;
          const uint32_t x = (image_tile % job.tiles_x) * TW + (lane % TW);
          const uint32_t y = (image_tile / job.tiles_x) * TH + (lane / TW);
          o = job.cam.origin;
          idx = job.tile_major ? again * 64u + lane : y * job.width + x;
        } else {
          idx = again * 64u + lane;
          if (job.order) idx = job.order[idx];
          const float* rr = job.rays + 6ull * idx;
          o = V3f{rr[0], rr[1], rr[2]};
        }
        store_result(tv, job.out, idx, r, o);
      }
    }
  }
}

template <int MIN_WAVES, int TPW, int REFILL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIN_WAVES)))
void batch_refill_kernel(const char* __restrict__ table, Counter* __restrict__ counters) {
  const BatchHeader* hdr = header_of(table);
  const BatchJob* jobs = jobs_of(table);
  const TreeView* trees = trees_of(table);
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
      if (lane == 0) t = atomicAdd(&counters[c].next, static_cast<uint32_t>(TPW));
      t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
      if (t >= end - begin) break;
      const uint32_t r0 = begin + t;
      const uint32_t r1 = min(end, r0 + static_cast<uint32_t>(TPW));
      // the grabbed tiles, one job segment at a time (job and tree uniform)
      for (uint32_t tile = r0; tile < r1;) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(find_job(jobs, njobs, tile));
        const BatchJob& job = jobs[j];
        const TreeView& tv = trees[job.tree];
        const uint32_t seg_end = min(r1, job.tile_begin + job.tile_count);
        const uint32_t a = tile - job.tile_begin, b = seg_end - job.tile_begin;
        if (job.kind == 0) {
          const RenderSource src{job.cam, job.width, job.height, job.tiles_x, job.first_tile,
                                 job.tile_major};
          refill_loop<REFILL>(tv, job.out, a * 64u, b * 64u, src);
        } else {
          const ArraySource src{job.rays, job.width};
          refill_loop<REFILL>(tv, job.out, a * 64u, min(b * 64u, job.width), src);
        }
        tile = seg_end;
      }
    }
  }
}

using BatchFn = void (*)(const char*, Counter*);

struct BatchVariant {
  const char* name;
  BatchFn fn;
  BatchFn lod_fn;          // the form runs with get_by_ray_at_lod jobs use
  uint32_t groups_per_cu;  // 256-lane workgroups per CU in the persistent grid
  uint32_t tile_w, tile_h; // pixel tile of render jobs
  bool orders_casts = true;  // its kernel takes a cast job's traversal order
};

// runs that hold get_by_ray_at_lod jobs use the LOD form of the variant's
// tile shape
constexpr BatchFn kLod8x8 = batch_kernel<1, false, true>;
constexpr BatchFn kLod16x4 = batch_kernel<1, false, true, 16, 4>;
constexpr BatchFn kLod4x16 = batch_kernel<1, false, true, 4, 16>;

const BatchVariant kBatchVariants[] = {
    {"classic-w1", batch_kernel<1>, kLod8x8, 4, 8, 8},
    {"classic-w5", batch_kernel<5>, kLod8x8, 4, 8, 8},
    {"r1-t4-f16", batch_refill_kernel<1, 4, 16>, kLod8x8, 4, 8, 8, false},
    {"r1-t8-f16", batch_refill_kernel<1, 8, 16>, kLod8x8, 4, 8, 8, false},
    {"classic-w5-g5", batch_kernel<5>, kLod8x8, 5, 8, 8},
    {"ww-w5", batch_kernel<5, true>, kLod8x8, 4, 8, 8},
    {"ww-w6-g6", batch_kernel<6, true>, kLod8x8, 6, 8, 8},
    {"ww-w5-g5", batch_kernel<5, true>, kLod8x8, 5, 8, 8},
    // tile shape
    {"classic-w5-g5-16x4", batch_kernel<5, false, false, 16, 4>, kLod16x4, 5, 16, 4},
    {"classic-w5-g5-4x16", batch_kernel<5, false, false, 4, 16>, kLod4x16, 5, 4, 16},
    // top node records in LDS: 73 (three levels) per wave, or 585 (four
    // levels) per workgroup (37 KiB: four workgroups per CU)
    {"classic-w5-g5-n73", batch_kernel<5, false, false, 8, 8, 73>, kLod8x8, 5, 8, 8},
    {"classic-w4-g4-s585", batch_kernel<4, false, false, 8, 8, 585, true>, kLod8x8, 4, 8, 8},
    {"ww-w6-g6-n73", batch_kernel<6, true, false, 8, 8, 73>, kLod8x8, 6, 8, 8},
    // six waves per SIMD (at most 80 VGPRs)
    {"classic-w6-g6", batch_kernel<6>, kLod8x8, 6, 8, 8},
    {"classic-w6-g6-n73", batch_kernel<6, false, false, 8, 8, 73>, kLod8x8, 6, 8, 8},
};
constexpr int kBatchVariantCount = sizeof(kBatchVariants) / sizeof(kBatchVariants[0]);
int g_batch_variant = 0;

int hip_fail(hipError_t e, const char* what, std::string* err) {
  *err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SVO_OUT_OF_DEVICE_MEMORY : SVO_HIP_ERROR;
}

constexpr size_t trees_offset_for(uint32_t njobs) {
  return (kJobsOffset + njobs * sizeof(BatchJob) + 63) & ~size_t(63);
}

// Table entry of job `in` on tree slot `ti`, its tiles numbered from `*tiles`
// (advanced past them). Sets *lod when the job is get_by_ray_at_lod.
int fill_job(BatchJob& j, const BatchJobIn& in, uint32_t ti, uint32_t tw, uint32_t th,
             uint64_t* tiles, bool* lod, std::string* err) {
  std::memset(&j, 0, sizeof(j));
  j.tree = ti;
  j.kind = in.kind;
  j.width = in.width;
  j.height = in.kind == 0 ? in.height : 1u;
  if (in.kind == 0) {
    if (in.tile_major && (tw != 8 || th != 8)) {
      *err = "tile-major render jobs need a variant of 8x8 tiles";
      return SVO_INVALID_ARGUMENT;
    }
    j.tiles_x = (in.width + tw - 1) / tw;
    const uint64_t image_tiles = static_cast<uint64_t>(j.tiles_x) * ((in.height + th - 1) / th);
    const uint64_t limit = in.tile_limit ? in.tile_limit : image_tiles - std::min<uint64_t>(
                                                                in.first_tile, image_tiles);
    if (image_tiles >= (1ull << 26) || in.first_tile + limit > image_tiles) {
      *err = "render job tile range outside the image, or image above 2^32 pixels";
      return SVO_INVALID_ARGUMENT;
    }
    j.tile_count = static_cast<uint32_t>(limit);
    j.first_tile = in.first_tile;
    j.tile_major = in.tile_major ? 1u : 0u;
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
  j.tile_begin = static_cast<uint32_t>(*tiles);
  j.out = to_ptrs(in.out);
  if (!(in.viewing_distance >= 0.0f)) {
    *err = "viewing distance must be >= 0 (0: get_by_ray)";
    return SVO_INVALID_ARGUMENT;
  }
  j.viewing_distance = in.viewing_distance;
  *lod = *lod || in.viewing_distance > 0.0f;
  *tiles += j.tile_count;
  if (*tiles >= (1ull << 26)) {
    *err = "too many tiles in one batch (at most 2^26 tiles of 64 rays)";
    return SVO_INVALID_ARGUMENT;
  }
  return SVO_OK;
}

void fill_header(char* base, uint32_t n, uint64_t tiles, size_t trees_offset) {
  BatchHeader* hdr = reinterpret_cast<BatchHeader*>(base);
  hdr->njobs = n;
  hdr->total_tiles = static_cast<uint32_t>(tiles);
  hdr->chunk_tiles = static_cast<uint32_t>((tiles + kChunks - 1) / kChunks);
  hdr->trees_offset = static_cast<uint32_t>(trees_offset);
}

}  // namespace

struct BatchContext {
  int device = -1;
  int variant = 0;
  uint32_t max_jobs = 0;
  size_t table_bytes = 0;  // capacity: header + jobs + trees
  // two pinned stagings of the table, used alternately: a run fills one
  // while the previous run's copy may still be reading the other
  void* host[2] = {nullptr, nullptr};
  uint32_t next = 0;       // staging of the next run
  void* dev = nullptr;     // device table
  Counter* counters = nullptr;
  uint32_t grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphNode_t copy_node = nullptr;
  hipGraphExec_t exec = nullptr;
  // the same sequence with the LOD form of the kernel, built on the first
  // run that holds a job with a viewing distance
  hipGraph_t graph_lod = nullptr;
  hipGraphNode_t copy_node_lod = nullptr;
  hipGraphExec_t exec_lod = nullptr;
  hipEvent_t done[2] = {nullptr, nullptr};  // end of the last run that read host[i]
  bool pending[2] = {false, false};
  uint64_t last_tiles = 0;
  // kernel arguments of the graph's kernel nodes
  const char* arg_table = nullptr;
  Counter* arg_counters = nullptr;
};

namespace {

int build_graph(BatchContext* b, BatchFn fn, hipGraph_t* graph, hipGraphNode_t* copy_node,
                hipGraphExec_t* exec, std::string* err) {
  hipError_t e = hipGraphCreate(graph, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphCreate", err);
  e = hipGraphAddMemcpyNode1D(copy_node, *graph, nullptr, 0, b->dev, b->host[0],
                              b->table_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipGraphAddMemcpyNode1D", err);
  b->arg_table = static_cast<const char*>(b->dev);
  b->arg_counters = b->counters;
  hipGraphNode_t reset_node = nullptr, kernel_node = nullptr;
  void* reset_args[] = {&b->arg_counters};
  hipKernelNodeParams rp{};
  rp.func = reinterpret_cast<void*>(reset_counters_kernel);
  rp.gridDim = dim3(1);
  rp.blockDim = dim3(64);
  rp.sharedMemBytes = 0;
  rp.kernelParams = reset_args;
  rp.extra = nullptr;
  e = hipGraphAddKernelNode(&reset_node, *graph, copy_node, 1, &rp);
  if (e != hipSuccess) return hip_fail(e, "hipGraphAddKernelNode(reset)", err);
  void* args[] = {&b->arg_table, &b->arg_counters};
  hipKernelNodeParams kp{};
  kp.func = reinterpret_cast<void*>(fn);
  kp.gridDim = dim3(b->grid);
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  e = hipGraphAddKernelNode(&kernel_node, *graph, &reset_node, 1, &kp);
  if (e != hipSuccess) return hip_fail(e, "hipGraphAddKernelNode(batch)", err);
  e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", err);
  return SVO_OK;
}

}  // namespace

int batch_variant_count() { return kBatchVariantCount; }
const char* batch_variant_name(int i) {
  return (i >= 0 && i < kBatchVariantCount) ? kBatchVariants[i].name : nullptr;
}
int batch_variant_scratch(int i, uint32_t* bytes, std::string* err) {
  if (i < 0 || i >= kBatchVariantCount) {
    *err = "no such batch variant";
    return SVO_INVALID_ARGUMENT;
  }
  hipFuncAttributes a{};
  const hipError_t e =
      hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kBatchVariants[i].fn));
  if (e != hipSuccess) {
    *err = std::string("hipFuncGetAttributes: ") + hipGetErrorString(e);
    return SVO_HIP_ERROR;
  }
  *bytes = static_cast<uint32_t>(a.localSizeBytes);
  return SVO_OK;
}
bool batch_variant_select(int i) {
  if (i < 0 || i >= kBatchVariantCount) return false;
  g_batch_variant = i;
  return true;
}
int batch_variant_current() { return g_batch_variant; }

int batch_create(uint32_t max_jobs, BatchContext** out, std::string* err) {
  return batch_create_variant(max_jobs, g_batch_variant, out, err);
}

int batch_variant_serving() {
  const BatchVariant& v = kBatchVariants[g_batch_variant];
  if (v.tile_w == 8 && v.tile_h == 8) return g_batch_variant;
  for (int i = 0; i < kBatchVariantCount; ++i)
    if (std::strcmp(kBatchVariants[i].name, "classic-w5-g5") == 0) return i;
  return 0;
}

int batch_create_variant(uint32_t max_jobs, int variant, BatchContext** out, std::string* err) {
  *out = nullptr;
  if (variant < 0 || variant >= kBatchVariantCount) {
    *err = "no such batch variant";
    return SVO_INVALID_ARGUMENT;
  }
  if (max_jobs == 0 || max_jobs > (1u << 20)) {
    *err = "max_jobs must be in [1, 2^20]";
    return SVO_INVALID_ARGUMENT;
  }
  BatchContext* b = new BatchContext();
  b->max_jobs = max_jobs;
  b->variant = variant;
  b->table_bytes = trees_offset_for(max_jobs) + max_jobs * sizeof(TreeView);
  hipError_t e = hipGetDevice(&b->device);
  int cus = 0;
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  for (int i = 0; i < 2 && e == hipSuccess; ++i)
    e = hipHostMalloc(&b->host[i], b->table_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->dev, b->table_bytes);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&b->counters), kChunks * sizeof(Counter));
  for (int i = 0; i < 2 && e == hipSuccess; ++i)
    e = hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    const int st = hip_fail(e, "batch_create", err);
    batch_destroy(b);
    return st;
  }
  std::memset(b->host[0], 0, b->table_bytes);
  std::memset(b->host[1], 0, b->table_bytes);
  // groups_per_cu 256-lane workgroups per CU: that many waves on every SIMD
  b->grid = static_cast<uint32_t>(std::max(cus, 1)) * kBatchVariants[b->variant].groups_per_cu;
  const int st = build_graph(b, kBatchVariants[b->variant].fn, &b->graph, &b->copy_node,
                             &b->exec, err);
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
  for (int i = 0; i < 2; ++i)
    if (b->pending[i]) (void)hipEventSynchronize(b->done[i]);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec_lod) (void)hipGraphExecDestroy(b->exec_lod);
  if (b->graph_lod) (void)hipGraphDestroy(b->graph_lod);
  for (int i = 0; i < 2; ++i)
    if (b->done[i]) (void)hipEventDestroy(b->done[i]);
  if (b->counters) (void)hipFree(b->counters);
  if (b->dev) (void)hipFree(b->dev);
  for (int i = 0; i < 2; ++i)
    if (b->host[i]) (void)hipHostFree(b->host[i]);
  delete b;
}

uint64_t batch_tiles(const BatchContext* b) { return b ? b->last_tiles : 0; }
int batch_context_variant(const BatchContext* b) { return b ? b->variant : -1; }

int batch_run(BatchContext* b, const BatchJobIn* jobs, uint32_t n, void* stream, std::string* err) {
  if (n > b->max_jobs) {
    *err = "more jobs than the batch was created for";
    return SVO_INVALID_ARGUMENT;
  }
  // the run two back copied from this staging: let it finish first (the
  // previous run, on the other staging, may still be in flight)
  const uint32_t slot = b->next;
  if (b->pending[slot]) {
    const hipError_t e = hipEventSynchronize(b->done[slot]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize", err);
    b->pending[slot] = false;
  }
  char* base = static_cast<char*>(b->host[slot]);
  BatchJob* out = reinterpret_cast<BatchJob*>(base + kJobsOffset);
  const size_t trees_offset = trees_offset_for(n);
  TreeView* trees = reinterpret_cast<TreeView*>(base + trees_offset);
  std::unordered_map<const void*, uint32_t> tree_index;
  uint64_t tiles = 0;
  bool any_lod = false;
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
    const int st = fill_job(out[i], in, ti, kBatchVariants[b->variant].tile_w,
                            kBatchVariants[b->variant].tile_h, &tiles, &any_lod, err);
    if (st != SVO_OK) return st;
  }
  // large cast jobs traverse their rays in svo_cast's order (kernels on `s`
  // ahead of the graph; the orders go back to the pool after it)
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<uint32_t*> orders;
  if (kBatchVariants[b->variant].orders_casts)
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t threshold = cast_order_threshold();
      if (jobs[i].kind != 1 || threshold == 0 || jobs[i].width < threshold) continue;
      uint32_t* ord = nullptr;
      const int st = order_rays(jobs[i].rays, jobs[i].width, jobs[i].tree->view.size, &ord, s, err);
      if (st == SVO_OUT_OF_DEVICE_MEMORY) continue;  // this job keeps its rays' own order
      if (st != SVO_OK) {
        for (uint32_t* p : orders) (void)hipFreeAsync(p, s);
        return st;
      }
      if (!ord) continue;
      orders.push_back(ord);
      out[i].order = ord;
    }
  fill_header(base, n, tiles, trees_offset);
  b->last_tiles = tiles;
  const size_t used = trees_offset + tree_index.size() * sizeof(TreeView);
  if (any_lod && !b->exec_lod) {
    const int st = build_graph(b, kBatchVariants[b->variant].lod_fn, &b->graph_lod,
                               &b->copy_node_lod, &b->exec_lod, err);
    if (st != SVO_OK) {
      for (uint32_t* p : orders) (void)hipFreeAsync(p, s);
      return st;
    }
  }
  hipGraphExec_t exec = any_lod ? b->exec_lod : b->exec;
  hipError_t e = hipGraphExecMemcpyNodeSetParams1D(
      exec, any_lod ? b->copy_node_lod : b->copy_node, b->dev, base, used, hipMemcpyHostToDevice);
  if (e == hipSuccess && b->pending[slot ^ 1u]) e = hipStreamWaitEvent(s, b->done[slot ^ 1u], 0);
  if (e == hipSuccess) e = hipGraphLaunch(exec, s);
  for (uint32_t* p : orders) (void)hipFreeAsync(p, s);
  if (e != hipSuccess) return hip_fail(e, "batch launch", err);
  e = hipEventRecord(b->done[slot], s);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord", err);
  b->pending[slot] = true;
  b->next = slot ^ 1u;
  return SVO_OK;
}

// ---------------------------------------------------------------------------
// Batched runs that allocate nothing (svo_batch_run_static). The job table
// and the work counters live in `scratch`, a fixed array in device memory
// that every call on a device shares; the table reaches it through the
// arguments of small upload kernels, so a call makes no allocation, no copy
// from pageable memory and no graph.

namespace {
constexpr uint32_t kStaticMaxJobs = 256;
constexpr size_t kStaticTableBytes =
    trees_offset_for(kStaticMaxJobs) + kStaticMaxJobs * sizeof(TreeView);
constexpr size_t kStaticCountersOffset = (kStaticTableBytes + 255) & ~size_t(255);
constexpr size_t kScratchBytes = kStaticCountersOffset + kChunks * sizeof(Counter);
}  // namespace

// Job table [0, kStaticTableBytes) and XCD work counters of
// svo_batch_run_static; each device holds its own copy for the life of the
// process. Host code never passes this name to a kernel or a copy (on the
// host it does not denote device memory): it reads the device's address of
// the array once per device with hipGetSymbolAddress.
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

// One upload kernel carries up to 3 KiB of the table in its arguments
// (kernel arguments are limited to 4 KiB).
constexpr uint32_t kUploadWords = 768;
struct UploadChunk {
  uint32_t offset_words;
  uint32_t count;
  uint32_t words[kUploadWords];
};

__global__ __launch_bounds__(256) void upload_table_kernel(uint32_t* __restrict__ dst,
                                                           const UploadChunk c) {
  for (uint32_t i = threadIdx.x; i < c.count; i += 256u) dst[c.offset_words + i] = c.words[i];
}

constexpr int kStaticMaxDevices = 64;
struct StaticDevice {
  char* scratch = nullptr;  // the device's address of `scratch`
  uint32_t cus = 0;
};
std::mutex g_static_mu;  // guards g_static and the host image of the table
StaticDevice g_static[kStaticMaxDevices];
alignas(64) char g_static_table[kStaticTableBytes];

// Address of `scratch` on `device` (current), read once per device.
int static_device(int device, StaticDevice** out, std::string* err) {
  StaticDevice& d = g_static[device];
  if (!d.scratch) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", err);
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute", err);
    d.cus = static_cast<uint32_t>(std::max(cus, 1));
    d.scratch = static_cast<char*>(p);
  }
  *out = &d;
  return SVO_OK;
}

int run_static(const svo_job* jobs, uint32_t n, void* stream, std::string* err) {
  if (n > kStaticMaxJobs) {
    *err = "svo_batch_run_static takes at most " + std::to_string(kStaticMaxJobs) +
           " jobs (its job table has a fixed size); use svo_batch_run for more";
    return SVO_INVALID_ARGUMENT;
  }
  if (n && !jobs) {
    *err = "NULL argument";
    return SVO_INVALID_ARGUMENT;
  }
  int device = -1;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice", err);
  if (device < 0 || device >= kStaticMaxDevices) {
    *err = "device index beyond the 64 devices svo_batch_run_static supports";
    return SVO_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> g(g_static_mu);
  StaticDevice* sd = nullptr;
  int st = static_device(device, &sd, err);
  if (st != SVO_OK) return st;
  char* base = g_static_table;
  BatchJob* out = reinterpret_cast<BatchJob*>(base + kJobsOffset);
  const size_t trees_offset = trees_offset_for(n);
  TreeView* trees = reinterpret_cast<TreeView*>(base + trees_offset);
  const DeviceTree* keys[kStaticMaxJobs];
  uint32_t ntrees = 0;
  uint64_t tiles = 0;
  bool any_lod = false;
  for (uint32_t i = 0; i < n; ++i) {
    const svo_job& j = jobs[i];
    if (!j.tree || !j.tree->dev || j.tree->dev->device != device) {
      *err = "job tree is missing or lives on another device";
      return SVO_INVALID_ARGUMENT;
    }
    if (j.kind > 1) {
      *err = "unknown job kind";
      return SVO_INVALID_ARGUMENT;
    }
    BatchJobIn in;
    in.tree = j.tree->dev;
    in.kind = j.kind;
    in.width = j.width;
    in.height = j.height;
    in.rays = j.rays;
    in.out = j.out;
    in.viewing_distance = j.viewing_distance;
    in.cam = Camera{};
    if (j.kind == 0)
      in.cam = make_camera(V3f{j.viewport.origin[0], j.viewport.origin[1], j.viewport.origin[2]},
                           V3f{j.viewport.direction[0], j.viewport.direction[1],
                               j.viewport.direction[2]},
                           j.viewport.frustum[0], j.viewport.frustum[1], j.viewport.fov, j.width,
                           j.height);
    uint32_t ti = 0;
    while (ti < ntrees && keys[ti] != in.tree) ++ti;
    if (ti == ntrees) {
      keys[ntrees++] = in.tree;
      trees[ti] = in.tree->view;
    }
    st = fill_job(out[i], in, ti, kBatchVariants[g_batch_variant].tile_w,
                  kBatchVariants[g_batch_variant].tile_h, &tiles, &any_lod, err);
    if (st != SVO_OK) return st;
  }
  if (tiles == 0) return SVO_OK;  // nothing to cast, nothing to write
  fill_header(base, n, tiles, trees_offset);
  const size_t used = trees_offset + ntrees * sizeof(TreeView);
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* dst = reinterpret_cast<uint32_t*>(sd->scratch);
  const uint32_t words = static_cast<uint32_t>((used + 3) / 4);
  for (uint32_t w0 = 0; w0 < words; w0 += kUploadWords) {
    UploadChunk c;
    c.offset_words = w0;
    c.count = std::min(kUploadWords, words - w0);
    std::memcpy(c.words, base + static_cast<size_t>(w0) * 4, static_cast<size_t>(c.count) * 4);
    hipLaunchKernelGGL(upload_table_kernel, dim3(1), dim3(256), 0, s, dst, c);
  }
  Counter* counters = reinterpret_cast<Counter*>(sd->scratch + kStaticCountersOffset);
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters);
  const BatchVariant& v = kBatchVariants[g_batch_variant];
  const BatchFn fn = any_lod ? v.lod_fn : v.fn;
  const uint32_t grid = sd->cus * v.groups_per_cu;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(256), 0, s, static_cast<const char*>(sd->scratch),
                     counters);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "svo_batch_run_static launch", err);
  return SVO_OK;
}

}  // namespace

}  // namespace svo

extern "C" int svo_batch_run_static(const svo_job* jobs, uint32_t n, void* stream) {
  std::string err;
  const int st = svo::run_static(jobs, n, stream, &err);
  if (st != SVO_OK) svo::set_last_error(err);
  return st;
}
