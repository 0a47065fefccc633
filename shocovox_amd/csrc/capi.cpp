// extern "C" boundary of libsvo (include/svo.h).
#include <cstring>
#include <fstream>
#include <iterator>
#include <new>
#include <string>
#include <vector>

#include "../../include/svo.h"
#include "batch.hpp"
#include "bytecode.hpp"
#include "device.hpp"
#include "flatten.hpp"
#include "handles.hpp"
#include "octree.hpp"
#include "vox.hpp"

namespace svo {
Status build_scene(uint32_t kind, uint32_t size, uint32_t brick_dim, uint64_t seed, Octree* out);
}

namespace {

thread_local std::string g_error;

int fail(int status, const std::string& msg) {
  g_error = msg;
  return status;
}

}  // namespace

void svo::set_last_error(const std::string& msg) { g_error = msg; }

namespace {

int from_status(svo::Status s) {
  switch (s) {
    case svo::Status::Ok:
      return SVO_OK;
    case svo::Status::InvalidSize:
      return fail(SVO_INVALID_SIZE, "invalid octree size");
    case svo::Status::InvalidBrickDimension:
      return fail(SVO_INVALID_BRICK_DIMENSION, "invalid brick dimension");
    case svo::Status::InvalidStructure:
      return fail(SVO_INVALID_STRUCTURE, "invalid structure");
    case svo::Status::InvalidPosition:
      return fail(SVO_INVALID_POSITION, "position outside the octree");
  }
  return fail(SVO_INVALID_ARGUMENT, "unknown status");
}

svo::Entry to_entry(const svo_entry* e) {
  svo::Entry r;
  if (!e) return r;
  r.kind = static_cast<svo::EntryKind>(e->kind & 3u);
  r.albedo = e->albedo;
  r.data = e->data;
  return r;
}

void from_entry(const svo::Entry& e, svo_entry* out) {
  out->kind = static_cast<uint32_t>(e.kind);
  out->albedo = e.albedo;
  out->data = e.data;
}

void fill_desc(const svo::FlatTree& f, svo_tree_desc* d) {
  d->octree_size = f.octree_size;
  d->brick_dim = f.brick_dim;
  d->node_count = static_cast<uint32_t>(f.nodes.size());
  d->brick_count = f.brick_count();
  d->color_count = static_cast<uint32_t>(f.colors.size());
  d->top_level_nodes = 0;
  for (uint32_t e : f.level_end) {
    if (e > 2048) break;
    d->top_level_nodes = e;
  }
  d->nodes = reinterpret_cast<const svo_node_rec*>(f.nodes.data());
  d->brick_pix = f.brick_pix.data();
  d->brick_occ = f.brick_occ.data();
  d->colors = f.colors.data();
  d->mips_enabled = f.mips_enabled ? 1u : 0u;
}

}  // namespace

static_assert(sizeof(svo_node_rec) == sizeof(svo::NodeRec), "node record layout");

extern "C" {

const char* svo_last_error(void) { return g_error.c_str(); }
const char* svo_version(void) { return "svo-mi355x 0.1.0"; }

int svo_octree_new(uint32_t size, uint32_t brick_dim, svo_octree** out) {
  if (!out) return fail(SVO_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  svo::Status st;
  svo::Octree t = svo::Octree::create(size, brick_dim, &st);
  if (st != svo::Status::Ok) return from_status(st);
  svo_octree* h = new (std::nothrow) svo_octree{std::move(t)};
  if (!h) return fail(SVO_INVALID_ARGUMENT, "allocation failed");
  *out = h;
  return SVO_OK;
}

void svo_octree_free(svo_octree* tree) { delete tree; }

int svo_octree_clone(const svo_octree* tree, svo_octree** out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = new (std::nothrow) svo_octree{tree->tree};
  if (!*out) return fail(SVO_INVALID_ARGUMENT, "allocation failed");
  return SVO_OK;
}

uint32_t svo_octree_get_size(const svo_octree* tree) { return tree ? tree->tree.size() : 0; }
uint32_t svo_octree_get_brick_dim(const svo_octree* tree) {
  return tree ? tree->tree.brick_dim() : 0;
}

int svo_octree_set_auto_simplify(svo_octree* tree, int enabled) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  tree->tree.auto_simplify = enabled != 0;
  tree->bytes_valid = false;
  return SVO_OK;
}

int svo_octree_insert(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, const svo_entry* e) {
  if (!tree || !e) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  return from_status(tree->tree.insert(x, y, z, to_entry(e)));
}

int svo_octree_insert_at_lod(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                             uint32_t insert_size, const svo_entry* e) {
  if (!tree || !e) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  return from_status(tree->tree.insert_at_lod(x, y, z, insert_size, to_entry(e)));
}

int svo_octree_update(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, const svo_entry* e) {
  if (!tree || !e) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  return from_status(tree->tree.update(x, y, z, to_entry(e)));
}

int svo_octree_clear(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  return from_status(tree->tree.clear(x, y, z));
}

int svo_octree_clear_at_lod(svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                            uint32_t clear_size) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  return from_status(tree->tree.clear_at_lod(x, y, z, clear_size));
}

int svo_octree_get(const svo_octree* tree, uint32_t x, uint32_t y, uint32_t z, svo_entry* out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  from_entry(tree->tree.get(x, y, z), out);
  return SVO_OK;
}

int svo_octree_switch_mip_maps(svo_octree* tree, int enabled) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  tree->tree.switch_albedo_mip_maps(enabled != 0);
  return SVO_OK;
}

int svo_octree_mip_maps_enabled(const svo_octree* tree) {
  return tree && tree->tree.mip_strategy().enabled ? 1 : 0;
}

int svo_octree_set_mip_method(svo_octree* tree, uint32_t level, uint32_t method, float thr) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  if (method > SVO_MIP_POSTERIZE_BD) return fail(SVO_INVALID_ARGUMENT, "unknown MIP method");
  tree->tree.set_mip_method(level, svo::MipMethodSpec{static_cast<svo::MipMethod>(method), thr});
  return SVO_OK;
}

int svo_octree_get_mip_method(const svo_octree* tree, uint32_t level, uint32_t* method,
                              float* thr) {
  if (!tree || !method || !thr) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const svo::MipMethodSpec m = tree->tree.mip_strategy().method_at(level);
  *method = static_cast<uint32_t>(m.kind);
  *thr = m.thr;
  return SVO_OK;
}

int svo_octree_set_mip_color_threshold(svo_octree* tree, uint32_t level, float thr) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  tree->tree.set_mip_color_threshold(level, thr);
  return SVO_OK;
}

int svo_octree_get_mip_color_threshold(const svo_octree* tree, uint32_t level, float* thr) {
  if (!tree || !thr) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const auto& m = tree->tree.mip_strategy().color_thresholds;
  auto it = m.find(level);
  *thr = it == m.end() ? 0.0f : it->second;
  return SVO_OK;
}

int svo_octree_reset_mip_strategy(svo_octree* tree) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  tree->tree.reset_mip_strategy();
  return SVO_OK;
}

int svo_octree_recalculate_mips(svo_octree* tree) {
  if (!tree) return fail(SVO_INVALID_ARGUMENT, "tree is NULL");
  tree->tree.recalculate_mips();
  return SVO_OK;
}

int svo_octree_sample_root_mip(const svo_octree* tree, uint32_t octant, uint32_t x, uint32_t y,
                               uint32_t z, svo_entry* out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const uint32_t d = tree->tree.brick_dim();
  if (octant > 8 || x >= d || y >= d || z >= d)
    return fail(SVO_INVALID_ARGUMENT, "octant above 8 or position outside the brick");
  from_entry(tree->tree.sample_root_mip(octant, x, y, z), out);
  return SVO_OK;
}

int svo_octree_insert_many(svo_octree* tree, const uint32_t* xyz, const svo_entry* entries,
                           size_t n) {
  if (!tree || (n && (!xyz || !entries))) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  int first = SVO_OK;
  for (size_t i = 0; i < n; ++i) {
    const int st = from_status(
        tree->tree.insert(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], to_entry(&entries[i])));
    if (st != SVO_OK && first == SVO_OK) first = st;
  }
  return first;
}

int svo_octree_clear_many(svo_octree* tree, const uint32_t* xyz, size_t n) {
  if (!tree || (n && !xyz)) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  int first = SVO_OK;
  for (size_t i = 0; i < n; ++i) {
    const int st = from_status(tree->tree.clear(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    if (st != SVO_OK && first == SVO_OK) first = st;
  }
  return first;
}

int svo_octree_get_many(const svo_octree* tree, const uint32_t* xyz, size_t n, svo_entry* out) {
  if (!tree || (n && (!xyz || !out))) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  for (size_t i = 0; i < n; ++i)
    from_entry(tree->tree.get(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), &out[i]);
  return SVO_OK;
}

int svo_octree_export(const svo_octree* tree, svo_model_sizes* sizes, svo_model_dump* dump) {
  if (!tree || !sizes) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const svo::Octree& t = tree->tree;
  const svo::NodePool& pool = t.nodes();
  const auto& kids = t.node_children();
  const uint32_t dim = t.brick_dim();
  const size_t n3 = static_cast<size_t>(dim) * dim * dim;
  uint32_t parted = 0;
  for (uint32_t k = 0; k < pool.len(); ++k) {
    const svo::NodeContent& n = pool.get(k);
    const int nb = n.kind == svo::NodeKind::Leaf ? 8 : (n.kind == svo::NodeKind::UniformLeaf ? 1 : 0);
    for (int o = 0; o < nb; ++o)
      if (n.bricks[o].kind == svo::BrickKind::Parted) ++parted;
  }
  sizes->octree_size = t.size();
  sizes->brick_dim = dim;
  sizes->pool_len = pool.len();
  sizes->parted_count = parted;
  sizes->color_count = static_cast<uint32_t>(t.color_palette().size());
  sizes->data_count = static_cast<uint32_t>(t.data_palette().size());
  uint32_t mip_parted = 0;
  for (uint32_t k = 0; k < pool.len(); ++k)
    if (t.node_mip(k).kind == svo::BrickKind::Parted) ++mip_parted;
  sizes->mip_parted_count = mip_parted;
  sizes->mips_enabled = t.mip_strategy().enabled ? 1u : 0u;
  if (!dump) return SVO_OK;
  uint32_t next_mip = 0;
  for (uint32_t k = 0; k < pool.len(); ++k) {
    const svo::Brick& m = t.node_mip(k);
    dump->mip_kind[k] = static_cast<uint8_t>(m.kind);
    dump->mip_value[k] = m.kind == svo::BrickKind::Solid ? m.solid : 0u;
    if (m.kind == svo::BrickKind::Parted) {
      dump->mip_value[k] = next_mip;
      std::memcpy(dump->mip_voxels + static_cast<size_t>(next_mip) * n3, m.parted.data(),
                  n3 * sizeof(uint32_t));
      ++next_mip;
    }
  }
  uint32_t next_parted = 0;
  for (uint32_t k = 0; k < pool.len(); ++k) {
    const svo::NodeContent& n = pool.get(k);
    dump->reserved[k] = pool.reserved(k) ? 1 : 0;
    dump->content_kind[k] = static_cast<uint8_t>(n.kind);
    dump->internal_bits[k] = n.kind == svo::NodeKind::Internal ? n.bits : 0;
    const svo::NodeChildren nc = k < kids.size() ? kids[k] : svo::NodeChildren{};
    dump->children_kind[k] = static_cast<uint8_t>(nc.kind);
    for (int o = 0; o < 8; ++o) dump->children[8 * k + o] = nc.c[o];
    dump->children_bitmap[k] = nc.bitmap;
    const int nb = n.kind == svo::NodeKind::Leaf ? 8 : (n.kind == svo::NodeKind::UniformLeaf ? 1 : 0);
    for (int o = 0; o < 8; ++o) {
      dump->brick_kind[8 * k + o] = 0;
      dump->brick_value[8 * k + o] = 0;
      if (o >= nb) continue;
      const svo::Brick& b = n.bricks[o];
      dump->brick_kind[8 * k + o] = static_cast<uint8_t>(b.kind);
      if (b.kind == svo::BrickKind::Solid) {
        dump->brick_value[8 * k + o] = b.solid;
      } else if (b.kind == svo::BrickKind::Parted) {
        dump->brick_value[8 * k + o] = next_parted;
        std::memcpy(dump->voxels + static_cast<size_t>(next_parted) * n3, b.parted.data(),
                    n3 * sizeof(uint32_t));
        ++next_parted;
      }
    }
  }
  if (sizes->color_count)
    std::memcpy(dump->colors, t.color_palette().data(), sizes->color_count * sizeof(uint32_t));
  if (sizes->data_count)
    std::memcpy(dump->data, t.data_palette().data(), sizes->data_count * sizeof(uint32_t));
  return SVO_OK;
}

int svo_scene_build(uint32_t kind, uint32_t size, uint32_t brick_dim, uint64_t seed,
                    svo_octree** out) {
  if (!out) return fail(SVO_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  svo_octree* h = new (std::nothrow) svo_octree{};
  if (!h) return fail(SVO_INVALID_ARGUMENT, "allocation failed");
  const svo::Status st = svo::build_scene(kind, size, brick_dim, seed, &h->tree);
  if (st != svo::Status::Ok) {
    delete h;
    return from_status(st);
  }
  *out = h;
  return SVO_OK;
}

int svo_octree_flatten(const svo_octree* tree, svo_flat** out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  svo_flat* f = new (std::nothrow) svo_flat{};
  if (!f) return fail(SVO_INVALID_ARGUMENT, "allocation failed");
  f->mirror.build(tree->tree);
  *out = f;
  return SVO_OK;
}

int svo_flat_sync(svo_flat* flat, svo_octree* tree) {
  if (!flat || !tree) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  flat->mirror.apply(tree->tree);
  flat->mirror.mark_synced();
  return SVO_OK;
}

int svo_flat_desc(const svo_flat* flat, svo_tree_desc* out) {
  if (!flat || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  fill_desc(flat->mirror.flat(), out);
  return SVO_OK;
}

void svo_flat_free(svo_flat* flat) { delete flat; }

int svo_flat_get(const svo_flat* flat, const svo_octree* tree, uint32_t x, uint32_t y, uint32_t z,
                 svo_entry* out) {
  if (!flat || !tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const uint32_t pix = svo::flat_get(flat->mirror.flat(), tree->tree, x, y, z);
  from_entry(tree->tree.pix_get_ref(pix), out);
  return SVO_OK;
}

int svo_tree_create(const svo_tree_desc* desc, svo_tree** out) {
  if (!desc || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  svo::DeviceTree* dev = nullptr;
  std::string err;
  const int st = svo::device_tree_create(*desc, &dev, &err);
  if (st != SVO_OK) return fail(st, err);
  *out = new svo_tree{dev};
  return SVO_OK;
}

int svo_tree_from_octree(svo_octree* tree, svo_tree** out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  svo::LayoutMirror mirror;
  (void)tree->tree.take_dirty();
  mirror.build(tree->tree);
  svo_tree_desc d;
  fill_desc(mirror.flat(), &d);
  const int st = svo_tree_create(&d, out);
  if (st != SVO_OK) return st;
  (*out)->mirror = std::move(mirror);
  (*out)->mirror.mark_synced();
  (*out)->has_mirror = true;
  return SVO_OK;
}

int svo_octree_to_bytes(const svo_octree* tree, uint8_t* out, size_t cap, size_t* len) {
  if (!tree || !len) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  // The size query and the copy that follows it encode the tree once: the
  // encoding is kept on the handle until the tree changes.
  if (!tree->bytes_valid || tree->bytes_epoch != tree->tree.edit_epoch()) {
    svo::octree_to_bencode(tree->tree, &tree->bytes);
    tree->bytes_epoch = tree->tree.edit_epoch();
    tree->bytes_valid = true;
  }
  const std::string& b = tree->bytes;
  *len = b.size();
  if (out) {
    if (cap < b.size()) return fail(SVO_INVALID_ARGUMENT, "output buffer too small");
    std::memcpy(out, b.data(), b.size());
    tree->bytes.clear();
    tree->bytes.shrink_to_fit();
    tree->bytes_valid = false;
  }
  return SVO_OK;
}

int svo_octree_from_bytes(const uint8_t* data, size_t len, svo_octree** out) {
  if (!out || (len && !data)) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  svo_octree* t = new (std::nothrow) svo_octree();
  if (!t) return fail(SVO_INVALID_ARGUMENT, "out of host memory");
  std::string err;
  if (!svo::octree_from_bencode(data, len, &t->tree, &err)) {
    delete t;
    return fail(SVO_INVALID_STRUCTURE, "not a serialized octree: " + err);
  }
  *out = t;
  return SVO_OK;
}

int svo_octree_save(const svo_octree* tree, const char* path) {
  if (!tree || !path) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::string b;
  svo::octree_to_bencode(tree->tree, &b);
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f || !f.write(b.data(), static_cast<std::streamsize>(b.size())))
    return fail(SVO_INVALID_ARGUMENT, std::string("cannot write ") + path);
  return SVO_OK;
}

int svo_octree_load(const char* path, svo_octree** out) {
  if (!path || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(SVO_INVALID_ARGUMENT, std::string("cannot read ") + path);
  const std::string b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return svo_octree_from_bytes(reinterpret_cast<const uint8_t*>(b.data()), b.size(), out);
}

int svo_octree_from_vox_bytes(const uint8_t* data, size_t len, uint32_t brick_dim,
                              svo_octree** out) {
  if (!out || (len && !data)) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  svo_octree* t = new (std::nothrow) svo_octree();
  if (!t) return fail(SVO_INVALID_ARGUMENT, "out of host memory");
  std::string err;
  const svo::VoxStatus st = svo::octree_from_vox(data, len, brick_dim, &t->tree, &err);
  if (st != svo::VoxStatus::Ok) {
    delete t;
    const int code = st == svo::VoxStatus::TreeSize   ? SVO_INVALID_SIZE
                     : st == svo::VoxStatus::BrickDim ? SVO_INVALID_BRICK_DIMENSION
                     : st == svo::VoxStatus::Position ? SVO_INVALID_POSITION
                                                      : SVO_INVALID_STRUCTURE;
    return fail(code, "vox import: " + err);
  }
  *out = t;
  return SVO_OK;
}

int svo_octree_load_vox(const char* path, uint32_t brick_dim, svo_octree** out) {
  if (!path || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(SVO_INVALID_ARGUMENT, std::string("cannot read ") + path);
  const std::string b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return svo_octree_from_vox_bytes(reinterpret_cast<const uint8_t*>(b.data()), b.size(),
                                   brick_dim, out);
}

int svo_tree_load(const char* path, svo_octree** host_out, svo_tree** out) {
  if (!out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (host_out) *host_out = nullptr;
  svo_octree* t = nullptr;
  int st = svo_octree_load(path, &t);
  if (st != SVO_OK) return st;
  st = svo_tree_from_octree(t, out);
  if (st != SVO_OK || !host_out) {
    delete t;
    return st;
  }
  *host_out = t;
  return SVO_OK;
}

int svo_tree_admit(svo_octree* tree, uint64_t budget_bytes, uint64_t reserve_bytes,
                   svo_tree** out) {
  if (!tree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  svo::LayoutMirror mirror;
  (void)tree->tree.take_dirty();
  mirror.build(tree->tree);
  svo_tree_desc d;
  fill_desc(mirror.flat(), &d);
  const uint64_t need = svo::device_tree_bytes_for(d);
  uint64_t avail = budget_bytes;
  if (avail == 0) {
    const int st = svo_device_memory(&avail, nullptr);
    if (st != SVO_OK) return st;
  }
  if (need + reserve_bytes > avail)
    return fail(SVO_OUT_OF_DEVICE_MEMORY,
                "tree needs " + std::to_string(need) + " device bytes plus a reserve of " +
                    std::to_string(reserve_bytes) + ", " + std::to_string(avail) + " available");
  const int st = svo_tree_create(&d, out);
  if (st != SVO_OK) return st;
  (*out)->mirror = std::move(mirror);
  (*out)->mirror.mark_synced();
  (*out)->has_mirror = true;
  return SVO_OK;
}

}  // extern "C"

namespace svo {

static_assert(sizeof(LayoutMirror::SlotFill) == 12 && sizeof(LayoutMirror::WordWrite) == 12,
              "patch triples are three packed u32");

int tree_update(svo_tree* dtree, svo_octree* tree, void* stream, void* start_event,
                void* const* end_events, uint32_t n_end) {
  if (!dtree || !tree) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::string err;
  LayoutMirror& m = dtree->mirror;
  if (!dtree->has_mirror) {
    (void)tree->tree.take_dirty();
    m.build(tree->tree);
    dtree->has_mirror = true;
  } else {
    m.apply(tree->tree);
  }
  svo_tree_desc d;
  fill_desc(m.flat(), &d);
  int st = 1;
  if (!m.full_upload()) {
    BrickPatch bp;
    bp.whole = m.whole_slots().data();
    bp.n_whole = m.whole_slots().size();
    bp.fills = reinterpret_cast<const uint32_t*>(m.slot_fills().data());
    bp.n_fills = m.slot_fills().size();
    bp.pix_writes = reinterpret_cast<const uint32_t*>(m.pix_writes().data());
    bp.n_pix = m.pix_writes().size();
    bp.occ_writes = reinterpret_cast<const uint32_t*>(m.occ_writes().data());
    bp.n_occ = m.occ_writes().size();
    st = device_tree_patch(dtree->dev, d, m.dirty_records(), bp, m.colors_synced(), stream,
                           start_event, end_events, n_end, &err);
  }
  if (st == 1) {
    st = start_event ? device_record_event(start_event, stream, &err) : SVO_OK;
    if (st == SVO_OK) st = device_tree_upload(dtree->dev, d, stream, &err);
    for (uint32_t i = 0; i < n_end && st == SVO_OK; ++i)
      st = device_record_event(end_events[i], stream, &err);
  }
  if (st != SVO_OK) return fail(st, err);
  m.mark_synced();
  return SVO_OK;
}

}  // namespace svo

extern "C" {

int svo_tree_update(svo_tree* dtree, svo_octree* tree, void* stream) {
  return svo::tree_update(dtree, tree, stream, nullptr);
}

int svo_tree_update_stats(const svo_tree* dtree, uint32_t* records, uint32_t* slots) {
  if (!dtree) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  if (records) *records = dtree->has_mirror ? dtree->mirror.last_records() : 0;
  if (slots) *slots = dtree->has_mirror ? dtree->mirror.last_slots() : 0;
  return SVO_OK;
}

void svo_tree_destroy(svo_tree* dtree) {
  if (!dtree) return;
  svo::device_tree_destroy(dtree->dev);
  delete dtree;
}

int svo_tree_set_step_limit(svo_tree* dtree, uint32_t steps) {
  if (!dtree || !dtree->dev) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  dtree->dev->step_limit = steps;
  dtree->dev->view.max_steps = steps ? steps : svo::MAX_TRAVERSAL_STEPS;
  return SVO_OK;
}

uint64_t svo_tree_device_bytes(const svo_tree* dtree) {
  return dtree && dtree->dev ? dtree->dev->bytes : 0;
}

// Pixel indices y * w + x are 32-bit in the render kernels: images hold
// fewer than 2^32 pixels.
static bool image_fits(uint32_t w, uint32_t h) {
  return static_cast<uint64_t>(w) * h < (uint64_t(1) << 32);
}
static const char* const kImageTooLarge = "image of 2^32 or more pixels (w * h must be < 2^32)";

int svo_cast(const svo_tree* dtree, const float* rays, uint32_t n, const svo_hits* out,
             void* stream) {
  if (!dtree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::string err;
  const int st = svo::device_cast(*dtree->dev, rays, n, *out, stream, &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

int svo_cast_set_order_threshold(uint32_t n) {
  svo::set_cast_order_threshold(n);
  return SVO_OK;
}

int svo_render(const svo_tree* dtree, const svo_viewport* vp, uint32_t w, uint32_t h,
               const svo_hits* out, void* stream) {
  if (!image_fits(w, h)) return fail(SVO_INVALID_ARGUMENT, kImageTooLarge);
  if (!dtree || !vp || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  const svo::Camera cam = svo::make_camera(
      svo::V3f{vp->origin[0], vp->origin[1], vp->origin[2]},
      svo::V3f{vp->direction[0], vp->direction[1], vp->direction[2]}, vp->frustum[0],
      vp->frustum[1], vp->fov, w, h);
  std::string err;
  const int st = svo::device_render(*dtree->dev, cam, w, h, *out, stream, &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

int svo_cast_at_lod(const svo_tree* dtree, const float* rays, uint32_t n,
                    float viewing_distance, const svo_hits* out, void* stream) {
  if (!dtree || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  if (!(viewing_distance > 0.0f)) return fail(SVO_INVALID_ARGUMENT, "viewing distance must be > 0");
  std::string err;
  const int st = svo::device_cast_lod(*dtree->dev, rays, n, viewing_distance, *out, stream, &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

int svo_render_at_lod(const svo_tree* dtree, const svo_viewport* vp, uint32_t w, uint32_t h,
                      float viewing_distance, const svo_hits* out, void* stream) {
  if (!image_fits(w, h)) return fail(SVO_INVALID_ARGUMENT, kImageTooLarge);
  if (!dtree || !vp || !out) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  if (!(viewing_distance > 0.0f)) return fail(SVO_INVALID_ARGUMENT, "viewing distance must be > 0");
  const svo::Camera cam = svo::make_camera(
      svo::V3f{vp->origin[0], vp->origin[1], vp->origin[2]},
      svo::V3f{vp->direction[0], vp->direction[1], vp->direction[2]}, vp->frustum[0],
      vp->frustum[1], vp->fov, w, h);
  std::string err;
  const int st =
      svo::device_render_lod(*dtree->dev, cam, w, h, viewing_distance, *out, stream, &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

int svo_render_rgba8(const svo_tree* dtree, const svo_viewport* vp, uint32_t w, uint32_t h,
                     uint32_t convention, uint32_t* out, void* stream) {
  if (!image_fits(w, h)) return fail(SVO_INVALID_ARGUMENT, kImageTooLarge);
  if (!dtree || !vp) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::string err;
  const int st = svo::device_render_rgba8(
      *dtree->dev, svo::V3f{vp->origin[0], vp->origin[1], vp->origin[2]},
      svo::V3f{vp->direction[0], vp->direction[1], vp->direction[2]}, vp->frustum[0],
      vp->frustum[1], vp->fov, w, h, convention, out, stream, &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

int svo_batch_create(uint32_t max_jobs, svo_batch** out) {
  if (!out) return fail(SVO_INVALID_ARGUMENT, "out is NULL");
  svo::BatchContext* b = nullptr;
  std::string err;
  const int st = svo::batch_create(max_jobs, &b, &err);
  if (st != SVO_OK) return fail(st, err);
  *out = reinterpret_cast<svo_batch*>(b);
  return SVO_OK;
}

void svo_batch_destroy(svo_batch* batch) {
  svo::batch_destroy(reinterpret_cast<svo::BatchContext*>(batch));
}

int svo_batch_run(svo_batch* batch, const svo_job* jobs, uint32_t n, void* stream) {
  if (!batch || (n && !jobs)) return fail(SVO_INVALID_ARGUMENT, "NULL argument");
  std::vector<svo::BatchJobIn> in(n);
  for (uint32_t i = 0; i < n; ++i) {
    const svo_job& j = jobs[i];
    if (!j.tree || !j.tree->dev) return fail(SVO_INVALID_ARGUMENT, "job without a tree");
    if (j.kind > 1) return fail(SVO_INVALID_ARGUMENT, "unknown job kind");
    svo::BatchJobIn& b = in[i];
    b.tree = j.tree->dev;
    b.kind = j.kind;
    b.width = j.width;
    b.height = j.height;
    b.rays = j.rays;
    b.out = j.out;
    b.viewing_distance = j.viewing_distance;
    if (j.kind == 0)
      b.cam = svo::make_camera(
          svo::V3f{j.viewport.origin[0], j.viewport.origin[1], j.viewport.origin[2]},
          svo::V3f{j.viewport.direction[0], j.viewport.direction[1], j.viewport.direction[2]},
          j.viewport.frustum[0], j.viewport.frustum[1], j.viewport.fov, j.width, j.height);
  }
  std::string err;
  const int st = svo::batch_run(reinterpret_cast<svo::BatchContext*>(batch), in.data(), n, stream,
                                &err);
  if (st != SVO_OK) return fail(st, err);
  return SVO_OK;
}

uint64_t svo_batch_tiles(const svo_batch* batch) {
  return svo::batch_tiles(reinterpret_cast<const svo::BatchContext*>(batch));
}

int svo_batch_variant_count(void) { return svo::batch_variant_count(); }
const char* svo_batch_variant_name(int variant) { return svo::batch_variant_name(variant); }
int svo_select_batch_variant(int variant) {
  if (!svo::batch_variant_select(variant))
    return fail(SVO_INVALID_ARGUMENT, "no such batch variant");
  return SVO_OK;
}
int svo_current_batch_variant(void) { return svo::batch_variant_current(); }
int svo_batch_variant_scratch(int variant, uint32_t* bytes) {
  if (!bytes) return fail(SVO_INVALID_ARGUMENT, "bytes is NULL");
  std::string err;
  const int st = svo::batch_variant_scratch(variant, bytes, &err);
  return st == SVO_OK ? SVO_OK : fail(st, err);
}

int svo_ray_occupancy_lut(uint64_t* out512) {
  if (!out512) return fail(SVO_INVALID_ARGUMENT, "out is NULL");
  svo::generate_ray_occupancy_lut(out512);
  return SVO_OK;
}

int svo_render_variant_count(void) { return svo::render_variant_count(); }

const char* svo_render_variant_name(int variant) { return svo::render_variant_name(variant); }

int svo_render_variant_scratch(int variant, uint32_t* bytes) {
  if (!bytes) return fail(SVO_INVALID_ARGUMENT, "bytes is NULL");
  std::string err;
  const int st = svo::render_variant_scratch(variant, bytes, &err);
  return st == SVO_OK ? SVO_OK : fail(st, err);
}

int svo_select_render_variant(int variant) {
  if (!svo::render_variant_select(variant)) return fail(SVO_INVALID_ARGUMENT, "no such variant");
  return SVO_OK;
}

int svo_current_render_variant(void) { return svo::render_variant_current(); }

int svo_set_timing(int enabled) {
  svo::timing_enable(enabled != 0);
  return SVO_OK;
}

uint32_t svo_collect_kernel_ms(float* out_ms, uint32_t cap) {
  return svo::timing_collect(out_ms, cap);
}

}  // extern "C"
