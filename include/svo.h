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
/* Octree::clone (#[derive(Clone)], octree/types.rs:167-168): an independent
 * copy with the same nodes, palettes, MIPs and strategy */
int svo_octree_clone(const svo_octree* tree, svo_octree** out);
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
  uint32_t mip_parted_count; /* Parted MIP bricks */
  uint32_t mips_enabled;     /* MIPMapStrategy::enabled */
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
  uint8_t* mip_kind;         /* [pool_len]: node_mips, as brick_kind */
  uint32_t* mip_value;       /* [pool_len]: Solid value or Parted index in mip_voxels */
  uint32_t* mip_voxels;      /* [mip_parted_count * dim^3] */
} svo_model_dump;

int svo_octree_export(const svo_octree* tree, svo_model_sizes* sizes, svo_model_dump* dump);

/* MIP maps (src/octree/mipmap.rs). Methods are MIPResamplingMethods:
 * 0 BoxFilter, 1 PointFilter, 2 PointFilterBD, 3 Posterize(thr),
 * 4 PosterizeBD(thr); thresholds are clamped to [0, 1]. Where the reference
 * picks among equally frequent colours through a HashMap (PointFilter,
 * Posterize), whose order varies from run to run, the colour met first in
 * sampling order is taken. */
typedef enum svo_mip_method {
  SVO_MIP_BOX_FILTER = 0,
  SVO_MIP_POINT_FILTER = 1,
  SVO_MIP_POINT_FILTER_BD = 2,
  SVO_MIP_POSTERIZE = 3,
  SVO_MIP_POSTERIZE_BD = 4
} svo_mip_method;
/* StrategyUpdater::switch_albedo_mip_maps (mipmap.rs:865-880): turning MIPs
 * on recalculates every MIP of a non-empty tree */
int svo_octree_switch_mip_maps(svo_octree* tree, int enabled);
int svo_octree_mip_maps_enabled(const svo_octree* tree);
/* set_method_at (mipmap.rs:760-765), get_method_at (:754-757) */
int svo_octree_set_mip_method(svo_octree* tree, uint32_t level, uint32_t method, float thr);
int svo_octree_get_mip_method(const svo_octree* tree, uint32_t level, uint32_t* method,
                              float* thr);
/* set_color_similarity_thr_at (:737-742), get_new_color_similarity_at (:729-734) */
int svo_octree_set_mip_color_threshold(svo_octree* tree, uint32_t level, float thr);
int svo_octree_get_mip_color_threshold(const svo_octree* tree, uint32_t level, float* thr);
/* reset (:721-724): the default strategy (disabled) */
int svo_octree_reset_mip_strategy(svo_octree* tree);
/* recalculate_mips (:798-862) */
int svo_octree_recalculate_mips(svo_octree* tree);
/* sample_root_mip (:905-937): octant 8 samples the root's MIP, 0-7 a child's */
int svo_octree_sample_root_mip(const svo_octree* tree, uint32_t octant, uint32_t x, uint32_t y,
                               uint32_t z, svo_entry* out);

/* Persisted trees in the reference's format: the bencode of Octree<u32>
 * that Octree::save / Octree::to_bytes write (octree/mod.rs:136-168,
 * convert/bytecode.rs). MIP maps in a loaded file are dropped (the traversal
 * here runs with MIPs off); saved files carry empty MIPs and a disabled MIP
 * strategy. to_bytes: call with out == NULL to learn the size. */
int svo_octree_to_bytes(const svo_octree* tree, uint8_t* out, size_t cap, size_t* len);
int svo_octree_from_bytes(const uint8_t* data, size_t len, svo_octree** out);
int svo_octree_save(const svo_octree* tree, const char* path);
int svo_octree_load(const char* path, svo_octree** out);

/* MagicaVoxel import: Octree::load_vox_file(filename, brick_dimension)
 * (convert/magicavoxel.rs:259-395). The tree's edge is 2^ceil(log2) of the
 * scene's extent; every voxel that frame 0 of the scene graph places is
 * inserted with its palette colour, converted from MagicaVoxel's
 * right-handed z-up axes to the tree's left-handed y-up ones (x, z, y).
 * Errors: SVO_INVALID_STRUCTURE for a malformed file, one without an RGBA
 * palette chunk or with no model in frame 0; SVO_INVALID_SIZE /
 * SVO_INVALID_BRICK_DIMENSION when no tree of that edge and brick
 * dimension exists; SVO_INVALID_POSITION when a voxel lands outside the
 * tree (each of these is a panic in the reference). */
int svo_octree_from_vox_bytes(const uint8_t* data, size_t len, uint32_t brick_dim,
                              svo_octree** out);
int svo_octree_load_vox(const char* path, uint32_t brick_dim, svo_octree** out);

/* ---------------------------------------------------------------------- */
/* Synthetic scenes (the benchmark workloads), built with the host octree.  */
/* ---------------------------------------------------------------------- */

/* kind 1: heightfield terrain built voxel by voxel with Octree::insert in
 *         x->y->z order (size 128, dim 8 is the CPU-runnable workload);
 * kind 2: 3D value-noise terrain with worm caves and seeded brick forms
 *         (40% Empty, 10% Solid, rest Parted), built bottom-up;
 * kind 3: seeded spheres over a noise heightfield with 50/50 Parted/Solid
 *         leaf bricks, built bottom-up;
 * kind 4: the kind 1 heightfield (same voxels) built bottom-up, for sizes
 *         where inserting voxel by voxel takes too long. */
int svo_scene_build(uint32_t kind, uint32_t size, uint32_t brick_dim, uint64_t seed,
                    svo_octree** out);

/* ---------------------------------------------------------------------- */
/* Device-resident tree (replaces OctreeGPUHost::create_new_view,           */
/* src/raytracing/bevy/data.rs:111-166, with full residency)                */
/* ---------------------------------------------------------------------- */

typedef struct svo_tree svo_tree;

/* One 64-byte node record of the device layout. `mip` holds the form of the
 * node's MIP brick (0 Empty, 1 Solid, 2 Parted) and `mip_word` its Solid value
 * or Parted brick slot. */
typedef struct svo_node_rec {
  uint64_t occ;
  uint32_t meta;
  uint32_t mip;
  uint32_t child[8];
  uint32_t mip_word;
  uint32_t reserved[3];
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
  uint32_t mips_enabled;     /* MIPMapStrategy::enabled: LOD queries consult MIPs */
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
 * changed are uploaded, a changed brick as its changed words, as a fill of
 * its slot plus the words that differ from the fill, or whole, whichever is
 * smallest; arrays that must grow are moved to larger ones on the device.
 * Consumes the octree's log of changed nodes, so one host octree feeds one
 * device tree.
 * Asynchronous on `stream`: the call returns once the patch is queued (it
 * waits on the host only for the update two calls back, whose pinned staging
 * it reuses; a host function queued on `stream` after the patch's copies
 * releases the staging, so no event is kept on `stream` and it may be
 * destroyed after the call), and the host octree may be edited again right
 * away. Casts that
 * must see the update run on `stream` after it or wait for an event recorded
 * there; casts still running on other streams must have finished before the
 * patch runs, which the caller orders (as `&mut self` does in the
 * reference). A full re-upload (a tree created from a description, or a new
 * brick dimension) completes before the call returns. */
int svo_tree_update(svo_tree* dtree, svo_octree* tree, void* stream);
/* Records and brick slots written by the last create / update. */
int svo_tree_update_stats(const svo_tree* dtree, uint32_t* records, uint32_t* slots);
void svo_tree_destroy(svo_tree* dtree);
/* Device bytes held by the tree. */
/* Iteration bound of one ray on this tree (an addition of this library: the
 * reference loop has no bound). A ray that exceeds it ends as a miss with
 * the iteration-guard flag (flags bit 0) set. 0 restores the default of
 * 2^22 stack-loop and DDA iterations, which no well-formed tree reaches;
 * lower values force the guard, for testing failure reporting. */
int svo_tree_set_step_limit(svo_tree* tree, uint32_t steps);
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
 * rays: device array of n x {origin.xyz, direction.xyz} f32. From the
 * threshold of svo_cast_set_order_threshold up, the rays are traversed in the order of
 * (origin cell on a 64^3 grid of the tree, Morton order; direction octant),
 * found by a radix sort on `stream` with scratch from the stream-ordered
 * allocator; each ray's outputs are its own in any order. */
int svo_cast(const svo_tree* dtree, const float* rays, uint32_t n, const svo_hits* out,
             void* stream);
/* No counterpart in the reference (a tuning knob): the ray count from which
 * svo_cast, svo_cast_at_lod and svo_batch_run's cast jobs order the rays (default
 * 2^20; 0 never orders; above 2^31 - 1 rays stay unordered). Process-wide.
 * The sort's scratch comes from the device's default memory pool, which is
 * set to keep freed memory for reuse; without device memory for the sort the
 * rays keep their own order. svo_batch_run_static never orders. */
int svo_cast_set_order_threshold(uint32_t n);

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

/* Octree::get_by_ray_at_lod (src/raytracing/raytracing_on_cpu.rs:325-565)
 * for n rays / for a viewport's primary rays: when the tree's MIP strategy
 * is enabled, a node's MIP brick answers for its subtree once the node's
 * level is below the distance the ray has travelled divided by
 * viewing_distance (> 0). With MIPs disabled the results equal svo_cast /
 * svo_render. */
int svo_cast_at_lod(const svo_tree* dtree, const float* rays, uint32_t n,
                    float viewing_distance, const svo_hits* out, void* stream);
int svo_render_at_lod(const svo_tree* dtree, const svo_viewport* vp, uint32_t w, uint32_t h,
                      float viewing_distance, const svo_hits* out, void* stream);

/* Shaded render to RGBA8, the drop-in for the image the reference draws:
 * out holds w*h u32 pixels (r in the low byte, alpha 255), row 0 at the top.
 * convention 0: examples/cpu_render.rs:78-137 - rays from the eye through
 *   pixel corners of the glass (bottom_left + right*x*pixel_w +
 *   up*y*pixel_h, row h-1-y), hit colour albedo * (1 - (n.L/2 + 0.5)) with
 *   L = normalize(0,-1,1) cast per channel like `as u8`, miss (128,128,128);
 * convention 1: the output texture of OctreeGPUHost::create_new_view, as
 *   viewport_render.wgsl:1028-1089 fills it - rays from the glass point
 *   (x/W8 and 1 - row/H8 across the glass, W8/H8 rounded up to 8) away from
 *   the eye, hit colour albedo * (n.(-0.5,0.5,-0.5)/2 + 0.5), miss 0.25 grey,
 *   unorm8 = floor(clamp(v,0,1)*255 + 0.5).
 * Traversal is get_by_ray as everywhere else; the viewport is read as in
 * svo_render. */
int svo_render_rgba8(const svo_tree* tree, const svo_viewport* viewport, uint32_t width,
                     uint32_t height, uint32_t convention, uint32_t* out, void* stream);

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
 * svo_cast). With viewing_distance > 0 the job is get_by_ray_at_lod (as
 * svo_render_at_lod / svo_cast_at_lod); 0 is get_by_ray; negative or NaN is
 * SVO_INVALID_ARGUMENT. Results go to `out` exactly as the single-job calls
 * write them. */
typedef struct svo_job {
  const svo_tree* tree;
  uint32_t kind;
  uint32_t width;
  uint32_t height;
  svo_viewport viewport;
  const float* rays;
  svo_hits out;
  float viewing_distance;
} svo_job;

/* A reusable batch context for up to max_jobs jobs; its device buffers and
 * graph live until svo_batch_destroy. Runs on one context execute in call
 * order: a run waits, on its stream, for the context's previous run, and the
 * host blocks only when it would refill the job table of the run two calls
 * back before that run's copy has finished. Separate contexts may run
 * concurrently on different streams. The context keeps an event recorded on
 * each of its last two runs' streams: such a stream must not be destroyed
 * before the context's second-next run or svo_batch_destroy (HIP does not
 * support waiting on an event last recorded on a destroyed stream). */
int svo_batch_create(uint32_t max_jobs, svo_batch** out);
void svo_batch_destroy(svo_batch* batch);
/* Launches all jobs as one persistent kernel on `stream` (asynchronous). The
 * caller's job array may be reused as soon as the call returns. */
int svo_batch_run(svo_batch* batch, const svo_job* jobs, uint32_t n, void* stream);
/* Total 64-ray tiles of the last run. */
uint64_t svo_batch_tiles(const svo_batch* batch);
/* svo_batch_run without a context and without allocating: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc). The
 * job table (at most 256 jobs) and the work counters live in a fixed device
 * array of the library (about 64 KiB per device, held until the process
 * ends) that every call on a device shares: calls that overlap on one device
 * race, so the caller orders them (one stream, or events). The table reaches
 * the device through kernel arguments, so the call allocates nothing, copies
 * from no pageable memory and launches on `stream` asynchronously; it uses
 * the kernel variant currently selected (svo_select_batch_variant). More than
 * 256 jobs is SVO_INVALID_ARGUMENT. Results equal svo_batch_run's. */
int svo_batch_run_static(const svo_job* jobs, uint32_t n, void* stream);
/* Kernel variants of batched launches (no counterpart in the reference):
 * "classic" (one 64-ray tile per work grab, one ray per lane) and lane-refill
 * forms (several tiles per grab; finished lanes start the next rays of the
 * grab). A context keeps the variant selected when it was created; results
 * are identical across variants. */
int svo_batch_variant_count(void);
const char* svo_batch_variant_name(int variant);
int svo_select_batch_variant(int variant);
int svo_current_batch_variant(void);
/* Private (scratch) bytes per lane of a variant's kernel build; non-zero
 * means the build spills registers (needs the HIP runtime). */
int svo_batch_variant_scratch(int variant, uint32_t* bytes);

/* ---------------------------------------------------------------------- */
/* Serving loop (no counterpart in the reference): ray-cast requests arrive */
/* as a seeded Poisson process; each picks a tree and a resolution by Zipf; */
/* a share of requests are sparse voxel edits applied in place.             */
/*  mode 0, continuous batching: at every launch boundary the arrived       */
/*    requests are packed, in arrival order, into one batched launch of at  */
/*    most max_batch_rays rays; a large request is cut into 8x8-tile chunks */
/*    that run in consecutive launches, so requests that arrive meanwhile   */
/*    join the next launch; up to max_inflight launches are queued on their */
/*    own streams.                                                          */
/*    A launch that would hold a single whole request runs as a plain       */
/*    render instead (the persistent kernel's set-up buys nothing for one   */
/*    job), so a lightly loaded loop behaves as mode 1.                     */
/*  mode 1, per-request plain launches: each render request is one          */
/*    svo_render launch, in arrival order, on up to max_inflight streams.  */
/*  mode 2: mode 0 with the persistent batched kernel for every launch.    */
/* Both modes keep per-tree order: a render sees exactly the edits that     */
/* arrived before it. Edits run on a pool of host threads (host octree      */
/* edit, then an asynchronous svo_tree_update on the thread's own stream)   */
/* once every earlier render of the tree has been launched; the patch waits */
/* on the device for those launches, and later renders of the tree wait on  */
/* the device for the patch. An edit completes when its patch has run.      */
/* Outputs of a render (hit, albedo, voxel, distance) stay in device memory */
/* owned by the loop until its slot is reused.                              */
/* ---------------------------------------------------------------------- */

typedef struct svo_serve_config {
  double qps;               /* mean arrival rate, requests per second */
  double duration_s;        /* arrivals are generated over [0, duration_s) */
  uint64_t seed;
  uint32_t mode;            /* 0 continuous batching, 1 per-request launches,
                               2 continuous batching, batched kernel only */
  uint32_t max_inflight;    /* launches queued at once (1..16) */
  uint64_t max_batch_rays;  /* ray budget of one batched launch (mode 0) */
  uint32_t max_batch_jobs;  /* chunk budget of one batched launch (mode 0) */
  uint32_t max_queue;       /* arrivals finding this many queued are rejected; 0 = no limit */
  double edit_fraction;     /* share of requests that are edits */
  uint32_t max_edit_voxels; /* an edit inserts or clears 1..max_edit_voxels voxels */
  double zipf_s;            /* Zipf exponent of the tree and resolution choice */
  uint32_t n_resolutions;
  const uint32_t* resolutions; /* n_resolutions (width, height) pairs, most popular first */
  double deadline_ms;       /* requests not started this long after arrival are dropped
                               (timed out; a dropped edit is not applied); 0 = none */
} svo_serve_config;

/* One request of a serving plan. Render viewports look at the tree's centre
 * from outside it (60 degree vertical fov). An edit inserts (random palette
 * colour) or clears edit_voxels uniformly drawn voxels from edit_seed. */
typedef struct svo_serve_request {
  double arrival_s;
  uint32_t kind; /* 0 render, 1 edit */
  uint32_t tree;
  uint32_t width;
  uint32_t height;
  svo_viewport viewport;
  uint32_t edit_voxels;
  uint32_t edit_clear; /* 1 clear, 0 insert */
  uint64_t edit_seed;
} svo_serve_request;

typedef struct svo_serve_stats {
  uint64_t requests, renders, edits, rejected, rays, launches;
  double elapsed_s;         /* first arrival to last completion */
  double mrays_per_s;       /* rays of completed renders / elapsed_s */
  double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests */
  double render_p50_ms, render_p99_ms;    /* render requests only */
  double edit_p50_ms, edit_p99_ms;        /* edit requests only */
  double mean_jobs_per_launch;
  uint64_t edit_records, edit_bricks;     /* node records / brick slots uploaded by edits */
  /* phases of edit requests: arrival to start on a worker (waiting for the
   * earlier renders of the tree to launch), host octree edit and layout
   * update, and the device time of the patch (copy and scatter) */
  double edit_wait_p99_ms, edit_apply_p99_ms, edit_upload_p99_ms;
  uint64_t timed_out;       /* requests dropped by the deadline (latency -2) */
  uint64_t single_launches; /* mode 0: launches that held one whole request and ran as a
                               plain render */
} svo_serve_stats;

/* The request sequence of a serving session (deterministic in cfg and the
 * tree sizes). Call with out == NULL to learn the count. */
int svo_serve_plan(const svo_serve_config* cfg, const uint32_t* tree_sizes, uint32_t n_trees,
                   svo_serve_request* out, uint64_t cap, uint64_t* count);

/* Applies an edit request of a plan to a host octree (what the loop does). */
int svo_serve_apply_edit(svo_octree* tree, const svo_serve_request* req);

/* Runs one serving session in real time over n_trees resident trees (each a
 * host octree and its device tree; edits change both). latencies_ms and
 * digests, when not NULL, receive one value per request of the plan
 * (capacity cap): latency (-1 when rejected, -2 when timed out) and, for
 * renders, a
 * 64-bit digest of the result: the wrapping sum over pixels i = y*w+x of
 * m(...m(m(m(i + 0x9E3779B97F4A7C15) ^ hit) ^ albedo) ... ^ distance_bits)
 * over hit, albedo, voxel x/y/z and the f32 bits of distance, where m is
 * the splitmix64 finaliser. Digests add a small kernel per launch. */
int svo_serve_run(svo_octree* const* trees, svo_tree* const* dtrees, uint32_t n_trees,
                  const svo_serve_config* cfg, svo_serve_stats* stats, float* latencies_ms,
                  uint64_t* digests, uint64_t cap);

/* Free and total bytes of device memory on the current device. */
int svo_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);

/* Admission by free device memory: flattens the octree and uploads it only
 * if its device bytes plus reserve_bytes fit in the free memory (queried
 * when budget_bytes is 0, else budget_bytes stands for it); otherwise
 * returns SVO_OUT_OF_DEVICE_MEMORY without touching the device. */
int svo_tree_admit(svo_octree* tree, uint64_t budget_bytes, uint64_t reserve_bytes,
                   svo_tree** out);

/* Octree::load (octree/mod.rs:161-168) straight into a device tree: parses
 * the file, flattens and uploads it. If host_out is not NULL it receives the
 * host octree (for later edits with svo_tree_update); otherwise the host
 * copy is released once the tree is resident. */
int svo_tree_load(const char* path, svo_octree** host_out, svo_tree** out);

/* Launch variants of the render kernel (no counterpart in the reference,
 * whose WGSL dispatch is fixed at 8x8 workgroups): occupancy hint, placement
 * of the occupancy LUT (global / LDS), per-wave tile shape, top node records
 * in LDS, workgroup order and the arrangement of the traversal loop. The
 * selection is process-wide and applies to later svo_render calls; every
 * variant computes identical results. The initial selection is
 * "w6-slut-8x8-ww", the fastest on the 1080p caves workload. */
int svo_render_variant_count(void);
const char* svo_render_variant_name(int variant);
int svo_select_render_variant(int variant);
int svo_current_render_variant(void);
/* Private (scratch) bytes per lane of a variant's kernel build; non-zero
 * means the build spills registers (needs the HIP runtime). */
int svo_render_variant_scratch(int variant, uint32_t* bytes);

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
