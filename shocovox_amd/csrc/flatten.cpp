// Host octree -> device layout.  See layout.hpp for the record format.
#include "flatten.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>

namespace svo {

namespace {

inline bool bound_contains(Cube b, V3f p) {
  return p.x >= b.min.x && p.x < b.min.x + b.size && p.y >= b.min.y && p.y < b.min.y + b.size &&
         p.z >= b.min.z && p.z < b.min.z + b.size;
}

}  // namespace

// Runs f(i) for i in [0, n) on up to 16 threads.
template <class F>
void parallel_range(size_t n, F&& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nt = static_cast<unsigned>(
      std::min<size_t>(std::max<size_t>(1, n / 64), std::max(1u, std::min(hw ? hw : 1u, 16u))));
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto& th : pool) th.join();
}

FlatTree flatten(const Octree& tree, bool headroom, std::vector<uint64_t>* versions) {
  FlatTree f;
  f.octree_size = tree.size();
  f.brick_dim = tree.brick_dim();
  f.colors = tree.color_palette();
  f.mips_enabled = tree.mip_strategy().enabled;
  const uint32_t dim = tree.brick_dim();
  const size_t n3 = static_cast<size_t>(dim) * dim * dim;
  const uint32_t words = occ_words_per_brick(dim);
  const NodePool& pool = tree.nodes();
  const auto& kids = tree.node_children();

  // Breadth-first order over keys; `index_of` maps pool key -> record index.
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  std::vector<uint32_t> order;
  std::vector<uint32_t> index_of(pool.len(), kNone);
  order.push_back(Octree::ROOT_NODE_KEY);
  index_of[Octree::ROOT_NODE_KEY] = 0;
  size_t level_begin = 0;
  while (level_begin < order.size()) {
    const size_t level_stop = order.size();
    for (size_t i = level_begin; i < level_stop; ++i) {
      const uint32_t key = order[i];
      const NodeContent& n = pool.get(key);
      if (n.kind != NodeKind::Internal) continue;
      for (uint32_t o = 0; o < 8; ++o) {
        const uint32_t c = key < kids.size() ? kids[key].child(o) : EMPTY_MARKER;
        if (!pool.key_is_valid(c) || index_of[c] != kNone) continue;
        index_of[c] = static_cast<uint32_t>(order.size());
        order.push_back(c);
      }
    }
    f.level_end.push_back(static_cast<uint32_t>(level_stop));
    level_begin = level_stop;
  }

  // Records, with Parted bricks numbered in record order; the brick words
  // are then copied and their occupancy sidecars built in parallel.
  auto with_room = [headroom](size_t n) { return headroom ? 2 * n + 64 : n; };
  f.nodes.reserve(with_room(order.size()));
  f.nodes.resize(order.size());
  std::vector<const Brick*> parted;
  auto encode_brick = [&](const Brick& b, uint32_t octant, NodeRec& r) {
    uint32_t form = FORM_EMPTY;
    uint32_t word = 0;
    if (b.kind == BrickKind::Solid) {
      form = FORM_SOLID;
      word = b.solid;
    } else if (b.kind == BrickKind::Parted) {
      form = FORM_PARTED;
      word = static_cast<uint32_t>(parted.size());
      parted.push_back(&b);
    }
    r.meta |= form << (8u + 2u * octant);
    r.child[octant] = word;
  };
  for (size_t i = 0; i < order.size(); ++i) {
    const uint32_t key = order[i];
    const NodeContent& n = pool.get(key);
    NodeRec& r = f.nodes[i];
    r = NodeRec{};
    for (auto& c : r.child) c = INVALID_INDEX;
    r.occ = tree.stored_occupied_bits(key);
    // the node's MIP brick (its Parted bricks share the brick slots)
    const Brick& mip = tree.node_mip(key);
    if (mip.kind == BrickKind::Solid) {
      r.mip = FORM_SOLID;
      r.mip_word = mip.solid;
    } else if (mip.kind == BrickKind::Parted) {
      r.mip = FORM_PARTED;
      r.mip_word = static_cast<uint32_t>(parted.size());
      parted.push_back(&mip);
    }
    switch (n.kind) {
      case NodeKind::Nothing:
        r.meta = KIND_NOTHING;
        break;
      case NodeKind::Internal:
        r.meta = KIND_INTERNAL;
        for (uint32_t o = 0; o < 8; ++o) {
          const uint32_t c = key < kids.size() ? kids[key].child(o) : EMPTY_MARKER;
          if (pool.key_is_valid(c)) r.child[o] = index_of[c];
        }
        break;
      case NodeKind::Leaf:
        r.meta = KIND_LEAF;
        for (uint32_t o = 0; o < 8; ++o) encode_brick(n.bricks[o], o, r);
        break;
      case NodeKind::UniformLeaf:
        r.meta = KIND_UNIFORM_LEAF;
        encode_brick(n.bricks[0], 0, r);
        break;
    }
  }
  f.brick_pix.reserve(with_room(parted.size()) * n3);
  f.brick_occ.reserve(with_room(parted.size()) * words);
  f.brick_pix.resize(parted.size() * n3);
  f.brick_occ.assign(parted.size() * words, 0u);
  parallel_range(parted.size(), [&](size_t slot) {
    const std::vector<uint32_t>& voxels = parted[slot]->parted;
    uint32_t* pix = f.brick_pix.data() + slot * n3;
    uint32_t* occ = f.brick_occ.data() + slot * words;
    std::memcpy(pix, voxels.data(), n3 * sizeof(uint32_t));
    for (size_t i = 0; i < n3; ++i)
      if (!tree.pix_points_to_empty(voxels[i])) occ[i / 32] |= 1u << (i % 32);
  });
  if (versions) {
    versions->resize(parted.size());
    for (size_t i = 0; i < parted.size(); ++i) (*versions)[i] = parted[i]->version;
  }
  return f;
}

uint32_t flat_get(const FlatTree& f, const Octree& owner, uint32_t x, uint32_t y, uint32_t z) {
  const V3f p{static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)};
  Cube bounds{v3u(0.0f), static_cast<float>(f.octree_size)};
  if (!bound_contains(bounds, p)) return EMPTY_MARKER;
  const uint32_t dim = f.brick_dim;
  const size_t n3 = static_cast<size_t>(dim) * dim * dim;
  auto parted_at = [&](uint32_t slot, Cube bb) {
    const V3f rel = ((p - bb.min) * static_cast<float>(dim)) / bb.size;
    const uint64_t ix = f2usize(roundf(floorf(rel.x)));
    const uint64_t iy = f2usize(roundf(floorf(rel.y)));
    const uint64_t iz = f2usize(roundf(floorf(rel.z)));
    const uint64_t flat = ix + iy * dim + iz * dim * dim;
    const uint32_t v = f.brick_pix[slot * n3 + flat];
    return owner.pix_points_to_empty(v) ? EMPTY_MARKER : v;
  };
  uint32_t idx = 0;
  for (;;) {
    const NodeRec& r = f.nodes[idx];
    switch (node_kind(r.meta)) {
      case KIND_NOTHING:
        return EMPTY_MARKER;
      case KIND_LEAF: {
        const uint32_t o = hash_region(p - bounds.min, bounds.size / 2.0f);
        const uint32_t form = brick_form(r.meta, o);
        if (form == FORM_EMPTY) return EMPTY_MARKER;
        if (form == FORM_SOLID) return r.child[o];
        return parted_at(r.child[o], child_bounds_for(bounds, o));
      }
      case KIND_UNIFORM_LEAF: {
        const uint32_t form = brick_form(r.meta, 0);
        if (form == FORM_EMPTY) return EMPTY_MARKER;
        if (form == FORM_SOLID)
          return owner.pix_points_to_empty(r.child[0]) ? EMPTY_MARKER : r.child[0];
        return parted_at(r.child[0], bounds);
      }
      default: {
        const uint32_t o = hash_region(p - bounds.min, bounds.size / 2.0f);
        if (r.child[o] == INVALID_INDEX) return EMPTY_MARKER;
        idx = r.child[o];
        bounds = child_bounds_for(bounds, o);
      }
    }
  }
}

void generate_ray_occupancy_lut(uint64_t out[64 * 8]) {
  for (int i = 0; i < 64 * 8; ++i) out[i] = 0;
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y)
      for (int z = 0; z < 4; ++z) {
        const int pos = x + 4 * y + 16 * z;
        for (int dx = -1; dx <= 1; dx += 2)
          for (int dy = -1; dy <= 1; dy += 2)
            for (int dz = -1; dz <= 1; dz += 2) {
              const uint32_t dir = hash_direction(V3f{static_cast<float>(dx),
                                                      static_cast<float>(dy),
                                                      static_cast<float>(dz)});
              auto span = [](int c, int d, int* lo, int* hi) {
                int moved = c + d * 4;
                moved = moved < 0 ? 0 : (moved > 3 ? 3 : moved);
                *lo = moved < c ? moved : c;
                *hi = moved > c ? moved : c;
              };
              int x0, x1, y0, y1, z0, z1;
              span(x, dx, &x0, &x1);
              span(y, dy, &y0, &y1);
              span(z, dz, &z0, &z1);
              uint64_t mask = 0;
              for (int bx = x0; bx <= x1; ++bx)
                for (int by = y0; by <= y1; ++by)
                  for (int bz = z0; bz <= z1; ++bz) mask |= 1ull << (bx + 4 * by + 16 * bz);
              out[pos * 8 + dir] = mask;
            }
      }
}

}  // namespace svo

namespace svo {

namespace {
constexpr uint32_t NO_RECORD = 0xFFFFFFFFu;
}

void LayoutMirror::build(const Octree& tree) {
  flat_ = flatten(tree, true, &slot_version_);
  slot_version_.reserve(2 * slot_version_.size() + 64);
  key_to_rec_.reserve(2 * tree.nodes().len() + 64);
  key_to_rec_.assign(tree.nodes().len(), NO_RECORD);
  // recover key -> record from the breadth-first order flatten() used
  std::vector<uint32_t> order{Octree::ROOT_NODE_KEY};
  key_to_rec_[Octree::ROOT_NODE_KEY] = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    const uint32_t key = order[i];
    if (tree.nodes().get(key).kind != NodeKind::Internal) continue;
    for (uint32_t o = 0; o < 8; ++o) {
      const uint32_t c = tree.node_children()[key].child(o);
      if (!tree.nodes().key_is_valid(c) || key_to_rec_[c] != NO_RECORD) continue;
      key_to_rec_[c] = static_cast<uint32_t>(order.size());
      order.push_back(c);
    }
  }
  free_recs_.clear();
  free_slots_.clear();
  dirty_recs_.clear();
  dirty_slots_.clear();
  rec_dirty_flag_.reserve(flat_.nodes.capacity());
  rec_dirty_flag_.assign(flat_.nodes.size(), 0);
  const size_t n3 = static_cast<size_t>(flat_.brick_dim) * flat_.brick_dim * flat_.brick_dim;
  slot_dirty_flag_.reserve(n3 ? flat_.brick_pix.capacity() / n3 : 0);
  slot_dirty_flag_.assign(flat_.brick_count(), 0);
  whole_slots_.clear();
  fills_.clear();
  pix_writes_.clear();
  occ_writes_.clear();
  rewritten_ = false;
  full_ = true;
  colors_synced_ = 0;
  synced_slots_ = 0;
}

uint32_t LayoutMirror::record_for(uint32_t key) {
  if (key >= key_to_rec_.size()) key_to_rec_.resize(static_cast<size_t>(key) + 1, NO_RECORD);
  if (key_to_rec_[key] != NO_RECORD) return key_to_rec_[key];
  uint32_t r;
  if (!free_recs_.empty()) {
    r = free_recs_.back();
    free_recs_.pop_back();
  } else {
    r = static_cast<uint32_t>(flat_.nodes.size());
    NodeRec blank{};
    blank.meta = KIND_NOTHING;
    for (auto& c : blank.child) c = INVALID_INDEX;
    flat_.nodes.push_back(blank);
    rec_dirty_flag_.push_back(0);
  }
  key_to_rec_[key] = r;
  return r;
}

void LayoutMirror::release_slot(uint32_t slot) { free_slots_.push_back(slot); }

void LayoutMirror::release_key(uint32_t key) {
  if (key >= key_to_rec_.size() || key_to_rec_[key] == NO_RECORD) return;
  const uint32_t r = key_to_rec_[key];
  NodeRec& rec = flat_.nodes[r];
  const uint32_t kind = node_kind(rec.meta);
  if (kind == KIND_LEAF || kind == KIND_UNIFORM_LEAF)
    for (uint32_t o = 0; o < 8; ++o)
      if (brick_form(rec.meta, o) == FORM_PARTED) release_slot(rec.child[o]);
  if (rec.mip == FORM_PARTED) release_slot(rec.mip_word);
  rec.mip = FORM_EMPTY;
  rec.mip_word = 0;
  rec.meta = KIND_NOTHING;
  rec.occ = 0;
  key_to_rec_[key] = NO_RECORD;
  if (r != 0) free_recs_.push_back(r);
}

uint32_t LayoutMirror::write_brick(const Octree& tree, const Brick& b, uint32_t old_slot) {
  const uint32_t dim = flat_.brick_dim;
  const size_t n3 = static_cast<size_t>(dim) * dim * dim;
  const uint32_t words = occ_words_per_brick(dim);
  uint32_t slot = old_slot;
  if (slot == NO_RECORD) {
    if (!free_slots_.empty()) {
      slot = free_slots_.back();
      free_slots_.pop_back();
    } else {
      slot = flat_.brick_count();
      flat_.brick_pix.resize(flat_.brick_pix.size() + n3, EMPTY_MARKER);
      flat_.brick_occ.resize(flat_.brick_occ.size() + words, 0u);
      slot_dirty_flag_.push_back(0);
      slot_version_.push_back(0);
    }
  } else if (b.version != 0 && slot_version_[slot] == b.version) {
    return slot;  // the same content as the slot's last write
  }
  slot_version_[slot] = b.version;
  deferred_.emplace_back(slot, b.parted.data());
  return slot;
}

// The slots written by an apply, encoded once each, on several threads when
// they hold many values: each worker takes a contiguous run of them and
// fills its own lists, which are then appended in run order (the same
// patch as one thread would make). Workers touch disjoint slots of the
// mirror and of the dirty flags.
void LayoutMirror::encode_deferred(const Octree& tree) {
  std::vector<std::pair<uint32_t, const uint32_t*>> work;
  {
    std::vector<uint8_t> seen(flat_.brick_count(), 0);
    for (size_t i = deferred_.size(); i-- > 0;)
      if (!seen[deferred_[i].first]) {
        seen[deferred_[i].first] = 1;
        work.push_back(deferred_[i]);
      }
    std::reverse(work.begin(), work.end());
    deferred_.clear();
  }
  if (work.empty()) return;
  const size_t n3 = static_cast<size_t>(flat_.brick_dim) * flat_.brick_dim * flat_.brick_dim;
  size_t workers = 1;
  if (work.size() * n3 >= (size_t(1) << 20))
    workers = std::min<size_t>({work.size(), std::max(1u, std::thread::hardware_concurrency()),
                                size_t(8)});
  std::vector<EncodeOut> outs(workers);
  auto run = [&](size_t w) {
    const size_t begin = work.size() * w / workers, end = work.size() * (w + 1) / workers;
    for (size_t i = begin; i < end; ++i) encode_slot(tree, work[i].first, work[i].second, &outs[w]);
  };
  if (workers == 1) {
    run(0);
  } else {
    std::vector<std::thread> threads;
    for (size_t w = 1; w < workers; ++w) threads.emplace_back(run, w);
    run(0);
    for (std::thread& t : threads) t.join();
  }
  for (EncodeOut& o : outs) {
    dirty_slots_.insert(dirty_slots_.end(), o.dirty.begin(), o.dirty.end());
    whole_slots_.insert(whole_slots_.end(), o.whole.begin(), o.whole.end());
    fills_.insert(fills_.end(), o.fills.begin(), o.fills.end());
    pix_writes_.insert(pix_writes_.end(), o.pix.begin(), o.pix.end());
    occ_writes_.insert(occ_writes_.end(), o.occ.begin(), o.occ.end());
    rewritten_ = rewritten_ || o.rewritten;
  }
}

// Puts a slot's new contents (and their occupancy sidecar) into the mirror
// and lists the cheapest form that brings the device's copy of the slot to
// them. Two passes over the brick: one that builds the sidecar and counts
// the words that differ from the device's copy and from the two candidate
// fill values (the first and the last word), one that writes the mirror
// and emits the chosen form's words.
void LayoutMirror::encode_slot(const Octree& tree, uint32_t slot, const uint32_t* new_pix,
                               EncodeOut* out) {
  const uint32_t dim = flat_.brick_dim;
  const uint32_t n3 = dim * dim * dim;
  const uint32_t words = occ_words_per_brick(dim);
  uint32_t* pix = flat_.brick_pix.data() + static_cast<size_t>(slot) * n3;
  uint32_t* occ = flat_.brick_occ.data() + static_cast<size_t>(slot) * words;
  const bool again = slot_dirty_flag_[slot] != 0;
  // the device holds the mirror's words for a slot it has received and that
  // this patch has not written yet
  const bool known = slot < synced_slots_ && !again;
  out->scratch.assign(words, 0u);
  uint32_t* new_occ = out->scratch.data();
  const uint32_t p0 = new_pix[0], p1 = new_pix[n3 - 1];
  size_t dk = 0, dp0 = 0, dp1 = 0;
  for (uint32_t c = 0; c * 32u < n3; ++c) {
    // one sidecar word per 32 values; the counts of a chunk vectorize, and a
    // chunk of empty markers needs no palette lookups
    const uint32_t* a = new_pix + c * 32u;
    const uint32_t* o = pix + c * 32u;
    const uint32_t m = std::min(32u, n3 - c * 32u);
    uint32_t k = 0, f0 = 0, f1 = 0, empties = 0;
    for (uint32_t j = 0; j < m; ++j) {
      k += a[j] != o[j];
      f0 += a[j] != p0;
      f1 += a[j] != p1;
      empties += a[j] == EMPTY_MARKER;
    }
    dk += k;
    dp0 += f0;
    dp1 += f1;
    if (empties != m) {
      uint32_t bits = 0;
      for (uint32_t j = 0; j < m; ++j)
        if (!tree.pix_points_to_empty(a[j])) bits |= 1u << j;
      new_occ[c] = bits;
    }
  }
  const uint32_t o0 = new_occ[0], o1 = new_occ[words - 1];
  size_t do0 = 0, do1 = 0;
  for (uint32_t i = 0; i < words; ++i) {
    const uint32_t a = new_occ[i];
    dk += a != occ[i];
    do0 += a != o0;
    do1 += a != o1;
  }
  if (known && dk == 0) return;
  if (!again) {
    slot_dirty_flag_[slot] = 1;
    out->dirty.push_back(slot);
  }
  const uint32_t fp = dp0 <= dp1 ? p0 : p1, fo = do0 <= do1 ? o0 : o1;
  const size_t d_fill = std::min(dp0, dp1) + std::min(do0, do1);
  const size_t w = sizeof(WordWrite);
  const size_t whole_cost = (static_cast<size_t>(n3) + words + 1) * sizeof(uint32_t);
  const size_t fill_cost = sizeof(SlotFill) + d_fill * w;
  const size_t known_cost = known ? dk * w : SIZE_MAX;
  if (again) {
    // its earlier form in this patch is superseded: send it whole
    out->whole.push_back(slot);
    out->rewritten = true;
  } else if (whole_cost <= fill_cost && whole_cost <= known_cost) {
    out->whole.push_back(slot);
  } else if (fill_cost < known_cost) {
    out->fills.push_back(SlotFill{slot, fp, fo});
    for (uint32_t i = 0; i < n3; ++i) {
      const uint32_t a = new_pix[i];
      if (a != fp) out->pix.push_back(WordWrite{slot, i, a});
      pix[i] = a;
    }
    for (uint32_t i = 0; i < words; ++i) {
      const uint32_t a = new_occ[i];
      if (a != fo) out->occ.push_back(WordWrite{slot, i, a});
      occ[i] = a;
    }
    return;
  } else {
    for (uint32_t i = 0; i < n3; ++i) {
      const uint32_t a = new_pix[i];
      if (a != pix[i]) {
        out->pix.push_back(WordWrite{slot, i, a});
        pix[i] = a;
      }
    }
    for (uint32_t i = 0; i < words; ++i) {
      const uint32_t a = new_occ[i];
      if (a != occ[i]) {
        out->occ.push_back(WordWrite{slot, i, a});
        occ[i] = a;
      }
    }
    return;
  }
  std::copy(new_pix, new_pix + n3, pix);
  std::copy(new_occ, new_occ + words, occ);
}

// After a slot was written twice in one patch: it is sent whole, so its
// fills and word writes are dropped (and duplicate whole entries merged).
void LayoutMirror::drop_partial_forms() {
  std::vector<uint8_t> whole(flat_.brick_count(), 0);
  std::vector<uint32_t> slots;
  for (uint32_t s : whole_slots_)
    if (!whole[s]) {
      whole[s] = 1;
      slots.push_back(s);
    }
  whole_slots_.swap(slots);
  auto keep_fill = [&](const SlotFill& f) { return whole[f.slot] != 0; };
  auto keep_word = [&](const WordWrite& w) { return whole[w.slot] != 0; };
  fills_.erase(std::remove_if(fills_.begin(), fills_.end(), keep_fill), fills_.end());
  pix_writes_.erase(std::remove_if(pix_writes_.begin(), pix_writes_.end(), keep_word),
                    pix_writes_.end());
  occ_writes_.erase(std::remove_if(occ_writes_.begin(), occ_writes_.end(), keep_word),
                    occ_writes_.end());
  rewritten_ = false;
}

void LayoutMirror::mark_synced() {
  last_recs_ = full_ ? static_cast<uint32_t>(flat_.nodes.size())
                     : static_cast<uint32_t>(dirty_recs_.size());
  last_slots_ = full_ ? flat_.brick_count() : static_cast<uint32_t>(dirty_slots_.size());
  for (uint32_t r : dirty_recs_) rec_dirty_flag_[r] = 0;
  for (uint32_t s : dirty_slots_) slot_dirty_flag_[s] = 0;
  dirty_recs_.clear();
  dirty_slots_.clear();
  whole_slots_.clear();
  fills_.clear();
  pix_writes_.clear();
  occ_writes_.clear();
  rewritten_ = false;
  full_ = false;
  colors_synced_ = static_cast<uint32_t>(flat_.colors.size());
  synced_slots_ = flat_.brick_count();
}

void LayoutMirror::write_record(const Octree& tree, uint32_t key, std::vector<uint32_t>* pending) {
  const uint32_t r = record_for(key);
  const NodeContent& n = tree.nodes().get(key);
  const NodeRec before = flat_.nodes[r];
  NodeRec rec{};
  for (auto& c : rec.child) c = INVALID_INDEX;
  rec.occ = tree.stored_occupied_bits(key);
  const auto& kids = tree.node_children();
  // slots the record held before, per octant
  uint32_t old_slots[8];
  const uint32_t old_kind = node_kind(before.meta);
  uint32_t old_mip_slot = before.mip == FORM_PARTED ? before.mip_word : NO_RECORD;
  const Brick& mip = tree.node_mip(key);
  if (mip.kind == BrickKind::Solid) {
    rec.mip = FORM_SOLID;
    rec.mip_word = mip.solid;
  } else if (mip.kind == BrickKind::Parted) {
    rec.mip = FORM_PARTED;
    rec.mip_word = write_brick(tree, mip, old_mip_slot);
    old_mip_slot = NO_RECORD;
  }
  if (old_mip_slot != NO_RECORD) release_slot(old_mip_slot);
  for (uint32_t o = 0; o < 8; ++o)
    old_slots[o] = ((old_kind == KIND_LEAF || old_kind == KIND_UNIFORM_LEAF) &&
                    brick_form(before.meta, o) == FORM_PARTED)
                       ? before.child[o]
                       : NO_RECORD;
  auto encode = [&](const Brick& b, uint32_t o) {
    if (b.kind == BrickKind::Solid) {
      rec.meta |= FORM_SOLID << (8u + 2u * o);
      rec.child[o] = b.solid;
    } else if (b.kind == BrickKind::Parted) {
      rec.meta |= FORM_PARTED << (8u + 2u * o);
      rec.child[o] = write_brick(tree, b, old_slots[o]);
      old_slots[o] = NO_RECORD;
    }
  };
  switch (n.kind) {
    case NodeKind::Nothing:
      rec.meta = KIND_NOTHING;
      break;
    case NodeKind::Internal:
      rec.meta = KIND_INTERNAL;
      for (uint32_t o = 0; o < 8; ++o) {
        const uint32_t c = key < kids.size() ? kids[key].child(o) : EMPTY_MARKER;
        if (!tree.nodes().key_is_valid(c)) continue;
        const bool had = c < key_to_rec_.size() && key_to_rec_[c] != NO_RECORD;
        rec.child[o] = record_for(c);
        if (!had) pending->push_back(c);
      }
      break;
    case NodeKind::Leaf:
      rec.meta = KIND_LEAF;
      for (uint32_t o = 0; o < 8; ++o) encode(n.bricks[o], o);
      break;
    case NodeKind::UniformLeaf:
      rec.meta = KIND_UNIFORM_LEAF;
      encode(n.bricks[0], 0);
      break;
  }
  for (uint32_t o = 0; o < 8; ++o)
    if (old_slots[o] != NO_RECORD) release_slot(old_slots[o]);
  if (std::memcmp(&rec, &before, sizeof(NodeRec)) != 0 || r >= rec_dirty_flag_.size()) {
    flat_.nodes[r] = rec;
    if (!rec_dirty_flag_[r]) {
      rec_dirty_flag_[r] = 1;
      dirty_recs_.push_back(r);
    }
  }
}

void LayoutMirror::apply(Octree& tree) {
  std::vector<uint32_t> keys = tree.take_dirty();
  flat_.colors = tree.color_palette();
  flat_.mips_enabled = tree.mip_strategy().enabled;
  // released nodes first, so their records and slots can be reused
  for (uint32_t k : keys)
    if (!tree.nodes().key_is_valid(k)) release_key(k);
  std::vector<uint32_t> pending;
  for (uint32_t k : keys)
    if (tree.nodes().key_is_valid(k)) pending.push_back(k);
  std::vector<uint8_t> done(tree.nodes().len(), 0);
  for (size_t i = 0; i < pending.size(); ++i) {
    const uint32_t k = pending[i];
    if (done[k]) continue;
    done[k] = 1;
    write_record(tree, k, &pending);
  }
  encode_deferred(tree);
  if (rewritten_) drop_partial_forms();
}

}  // namespace svo
