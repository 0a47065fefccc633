/*
 * libsvo - C ABI of the MI355X sparse-voxel-octree ray-traversal path.
 *
 * The reference (shocovox-rs 0.11.1) exposes this path as Rust methods on
 * `Octree<T>` and as a bevy render plugin. Each entry point below names the
 * reference item it stands in for. Conventions:
 *   - every function returns an int status (svo_status); nothing unwinds
 *     across the ABI; svo_last_error() gives the message of the last failure
 *     on the calling thread;
 *   - positions are voxel coordinates (u32), rays are f32;
 *   - colours are packed like Albedo::from(u32): r<<24 | g<<16 | b<<8 | a;
 *   - device pointers are plain device addresses (hipMalloc'd or from any
 *     allocator on the same device); `stream` is a hipStream_t, NULL for the
 *     default stream;
 *   - casts on one svo_tree are read-only and may run concurrently; edits
 *     (svo_tree_update) need the caller to order them against casts on the
 *     same stream, as `&mut self` does in the reference.
 */
#ifndef SVO_H
#define SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors OctreeError (src/octree/types.rs:8-21) plus device-side failures. */
typedef enum svo_status {
  SVO_OK = 0,
  SVO_INVALID_SIZE = 1,            /* OctreeError::InvalidSize */
  SVO_INVALID_BRICK_DIMENSION = 2, /* OctreeError::InvalidBrickDimension */
  SVO_INVALID_STRUCTURE = 3,       /* OctreeError::InvalidStructure */
  SVO_INVALID_POSITION = 4,        /* OctreeError::InvalidPosition */
  SVO_INVALID_ARGUMENT = 5,
  SVO_OUT_OF_DEVICE_MEMORY = 6,
  SVO_HIP_ERROR = 7,
  SVO_NO_DEVICE = 8
} svo_status;

const char* svo_last_error(void);
/* Library version string. */
const char* svo_version(void);

/* ---------------------------------------------------------------------- */
/* Host octree: Octree<u32> (src/octree/types.rs:169-207)                   */
/* ---------------------------------------------------------------------- */

typedef struct svo_octree svo_octree;

/* OctreeEntry (src/octree/types.rs:23-36). kind: 0 Empty, 1 Visual(albedo),
 * 2 Informative(data), 3 Complex(albedo, data). */
typedef struct svo_entry {
  uint32_t kind;
  uint32_t albedo;
  uint32_t data;
} svo_entry;

/* Octree::new (src/octree/mod.rs:173-205) */
int svo_octree_new(uint32_t size, uint32_t brick_dim, svo_octree** out);
void svo_octree_free(svo_octree* tree);
/* Octree::get_size (src/octree/mod.rs:374-376) */
uint32_t svo_octree_get_size(const svo_octree* tree);
uint32_t svo_octree_get_brick_dim(const svo_octree* tree);
/* Octree::auto_simplify public field (src/octree/types.rs:203) */
int svo_octree_set_auto_simplify(svo_octree* tree, int enabled);
/* Octree::insert (src/octree/update/insert.rs:47-56) */
int svo_octree_insert(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, const svo_entry* e);
/* Octree::insert_at_lod (src/octree/update/insert.rs:63-73) */
int svo_octree_insert_at_lod(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                             uint32_t insert_size, const svo_entry* e);
/* Octree::update (src/octree/update/insert.rs:79-88) */
int svo_octree_update(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, const svo_entry* e);
/* Octree::clear (src/octree/update/clear.rs:48-50) */
int svo_octree_clear(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z);
/* Octree::clear_at_lod (src/octree/update/clear.rs:55-59) */
int svo_octree_clear_at_lod(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                            uint32_t clear_size);
/* Octree::get (src/octree/mod.rs:209-215) */
int svo_octree_get(const svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, svo_entry* out);
/* Batched forms of insert / clear / get over n positions (xyz interleaved):
 * the same calls in index order, status of the first failure returned. */
int svo_octree_insert_many(svo_octree* tree, const uint32_t* xyz, const svo_entry* entries,
                           size_t n);
int svo_octree_clear_many(svo_octree* tree, const uint32_t* xyz, size_t n);
int svo_octree_get_many(const svo_octree* tree, const uint32_t* xyz, size_t n, svo_entry* out);

/* Raw dump of the node model (ObjectPool<NodeContent>, node_children,
 * palettes), for persistence and for independent checking. Call with
 * dump == NULL to learn the sizes, then with buffers of those sizes.
 * content_kind: 0 Nothing 1 Internal 2 Leaf 3 UniformLeaf;
 * children_kind: 0 NoChildren 1 Children 2 OccupancyBitmap;
 * brick_kind: 0 Empty 1 Solid 2 Parted; brick_value is the Solid value or the
 * index (in units of dim^3) of the Parted brick inside `voxels`. */
typedef struct svo_model_sizes {
  uint32_t octree_size;
  uint32_t brick_dim;
  uint32_t pool_len;
  uint32_t parted_count;
  uint32_t color_count;
  uint32_t data_count;
} svo_model_sizes;

typedef struct svo_model_dump {
  uint8_t* reserved;         /* [pool_len] */
  uint8_t* content_kind;     /* [pool_len] */
  uint64_t* internal_bits;   /* [pool_len] */
  uint8_t* children_kind;    /* [pool_len] */
  uint32_t* children;        /* [pool_len * 8] */
  uint64_t* children_bitmap; /* [pool_len] */
  uint8_t* brick_kind;       /* [pool_len * 8] */
  uint32_t* brick_value;     /* [pool_len * 8] */
  uint32_t* voxels;          /* [parted_count * dim^3] */
  uint32_t* colors;          /* [color_count] */
  uint32_t* data;            /* [data_count] */
} svo_model_dump;

int svo_octree_export(const svo_octree* tree, svo_model_sizes* sizes, svo_model_dump* dump);

/* ---------------------------------------------------------------------- */
/* Synthetic scenes (the benchmark workloads), built with the host octree.  */
/* ---------------------------------------------------------------------- */

/* kind 1: heightfield terrain built voxel by voxel with Octree::insert in
 *         x->y->z order (size 128, dim 8 is the CPU-runnable workload);
 * kind 2: 3D value-noise terrain with worm caves and seeded brick forms
 *         (40% Empty, 10% Solid, rest Parted), built bottom-up;
 * kind 3: seeded spheres over a noise heightfield with 50/50 Parted/Solid
 *         leaf bricks, built bottom-up. */
int svo_scene_build(uint32_t kind, uint32_t size, uint32_t brick_dim, uint64_t seed,
                    svo_octree** out);

/* ---------------------------------------------------------------------- */
/* Device-resident tree (replaces OctreeGPUHost::create_new_view,           */
/* src/raytracing/bevy/data.rs:111-166, with full residency)                */
/* ---------------------------------------------------------------------- */

typedef struct svo_tree svo_tree;

/* One 64-byte node record of the device layout. */
typedef struct svo_node_rec {
  uint64_t occ;
  uint32_t meta;
  uint32_t reserved0;
  uint32_t child[8];
  uint32_t reserved1[4];
} svo_node_rec;

/* Host description of a flattened tree: what a binding that owns its own
 * octree (e.g. the Rust `Octree`) fills in to upload it. */
typedef struct svo_tree_desc {
  uint32_t octree_size;
  uint32_t brick_dim;
  uint32_t node_count;
  uint32_t brick_count;
  uint32_t color_count;
  uint32_t top_level_nodes; /* records of the first levels, for LDS staging */
  const svo_node_rec* nodes;
  const uint32_t* brick_pix; /* brick_count * dim^3 palette index values */
  const uint32_t* brick_occ; /* brick_count * ceil(dim^3/32) occupancy words */
  const uint32_t* colors;    /* color_count packed albedo values */
} svo_tree_desc;

/* Flattened layout of a host octree, kept on the host; owned by the handle. */
typedef struct svo_flat svo_flat;
int svo_octree_flatten(const svo_octree* tree, svo_flat** out);
/* Applies the octree's edits since the last sync to the layout in place (the
 * host half of svo_tree_update; drains the octree's log of changed nodes). */
int svo_flat_sync(svo_flat* flat, svo_octree* tree);
int svo_flat_desc(const svo_flat* flat, svo_tree_desc* out);
void svo_flat_free(svo_flat* flat);
/* Point query on the flattened layout (layout round-trip check). */
int svo_flat_get(const svo_flat* flat, const svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                 svo_entry* out);

/* Upload a flattened tree to the current HIP device. */
int svo_tree_create(const svo_tree_desc* desc, svo_tree** out);
/* Flatten + upload in one call; the device tree keeps a host copy of its
 * layout so that later edits can be applied in place. */
int svo_tree_from_octree(svo_octree* tree, svo_tree** out);
/* Bring a device tree up to date with edits made to the host octree since it
 * was created or last updated (OctreeGPUView::reload,
 * src/raytracing/bevy/mod.rs:56-60): only node records and bricks that
 * changed are uploaded (a full upload when the arrays must grow). Consumes the
 * octree's log of changed nodes, so one host octree feeds one device tree.
 * Ordered on `stream`; returns after the upload has completed. */
int svo_tree_update(svo_tree* dtree, svo_octree* tree, void* stream);
/* Records and brick slots written by the last create / update. */
int svo_tree_update_stats(const svo_tree* dtree, uint32_t* records, uint32_t* slots);
void svo_tree_destroy(svo_tree* dtree);
/* Device bytes held by the tree. */
uint64_t svo_tree_device_bytes(const svo_tree* dtree);

/* Per-ray results (device arrays of n entries; any pointer may be NULL).
 * hit      1 where Octree::get_by_ray returns Some
 * pix      palette index value of the hit voxel (0xFFFFFFFF on a miss)
 * albedo   colour of the hit voxel, 0 when it has none
 * voxel    [3n] voxel coordinate of the hit (floor of the impact point
 *          clamped into the hit cell); 0xFFFFFFFF on a miss
 * distance |impact - origin|; +inf on a miss
 * impact   [3n] impact point returned by get_by_ray
 * normal   [3n] impact normal returned by get_by_ray
 * flags    bit 0: iteration guard, bit 1: bitmap index out of range */
typedef struct svo_hits {
  uint32_t* hit;
  uint32_t* pix;
  uint32_t* albedo;
  uint32_t* voxel;
  float* distance;
  float* impact;
  float* normal;
  uint32_t* flags;
} svo_hits;

/* Octree::get_by_ray for n rays (src/raytracing/raytracing_on_cpu.rs:316-318).
 * rays: device array of n x {origin.xyz, direction.xyz} f32. */
int svo_cast(const svo_tree* dtree, const float* rays, uint32_t n, const svo_hits* out,
             void* stream);

/* Viewport (src/raytracing/bevy/types.rs:55-71): frustum = (glass width,
 * glass height, depth), fov = distance of the glass from the eye. */
typedef struct svo_viewport {
  float origin[3];
  float direction[3];
  float frustum[3];
  float fov;
} svo_viewport;

/* Primary rays of a w x h viewport generated on the device (one ray through
 * each pixel centre, up = +y, row 0 at the bottom) and cast; outputs are
 * indexed y*w + x. Replaces the per-frame dispatch of the WGSL `update`
 * entry (assets/shaders/viewport_render.wgsl:1028-1089) with the CPU
 * harness's ray convention (examples/cpu_render.rs:78-114). */
int svo_render(const svo_tree* dtree, const svo_viewport* vp, uint32_t w, uint32_t h,
               const svo_hits* out, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched mode (no counterpart in the reference, which renders one view    */
/* per dispatch): many trees, viewports and ray arrays in one persistent    */
/* launch. Waves pull 64-ray tiles from a device work counter until all     */
/* tiles of all jobs are done. The launch sequence (job-table upload,       */
/* counter reset, persistent kernel) is captured once in a hipGraph and     */
/* replayed for every run.                                                  */
/* ---------------------------------------------------------------------- */

typedef struct svo_batch svo_batch;

/* One job: kind 0 renders `viewport` at width x height (rays generated on
 * the device, as svo_render); kind 1 casts `width` rays from `rays` (as
 * svo_cast). Results go to `out` exactly as the single-job calls write them. */
typedef struct svo_job {
  const svo_tree* tree;
  uint32_t kind;
  uint32_t width;
  uint32_t height;
  svo_viewport viewport;
  const float* rays;
  svo_hits out;
} svo_job;

/* A reusable batch context for up to max_jobs jobs; its device buffers and
 * graph live until svo_batch_destroy. Runs on one context are ordered by
 * the caller; separate contexts may run concurrently on different streams. */
int svo_batch_create(uint32_t max_jobs, svo_batch** out);
void svo_batch_destroy(svo_batch* batch);
/* Launches all jobs as one persistent kernel on `stream` (asynchronous). The
 * caller's job array may be reused as soon as the call returns. */
int svo_batch_run(svo_batch* batch, const svo_job* jobs, uint32_t n, void* stream);
/* Total 64-ray tiles of the last run. */
uint64_t svo_batch_tiles(const svo_batch* batch);

/* Launch variants of the render kernel (no counterpart in the reference,
 * whose WGSL dispatch is fixed at 8x8 workgroups): occupancy hint, placement
 * of the occupancy LUT (global / LDS) and per-wave tile shape. The selection
 * is process-wide and applies to later svo_render calls; every variant
 * computes identical results. */
int svo_render_variant_count(void);
const char* svo_render_variant_name(int variant);
int svo_select_render_variant(int variant);
int svo_current_render_variant(void);

/* The RAY_TO_NODE_OCCUPANCY_BITMASK_LUT the kernels use, [pos*8 + dir]
 * (src/spatial/lut.rs:255-896), regenerated from the reference's generator
 * (lut.rs:32-85). */
int svo_ray_occupancy_lut(uint64_t* out512);

/* Kernel timing (no counterpart in the reference, which has no metrics):
 * while enabled, every svo_cast / svo_render launch is bracketed by HIP
 * events on its stream; svo_collect_kernel_ms waits for them, writes up to
 * `cap` durations in launch order and forgets them. */
int svo_set_timing(int enabled);
uint32_t svo_collect_kernel_ms(float* out_ms, uint32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SVO_H */
