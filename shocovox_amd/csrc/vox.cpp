// MagicaVoxel .vox import (src/convert/magicavoxel.rs:259-395).
//
// The file is a RIFF-like chunk tree: "VOX " <version> MAIN{children}. The
// children used here are SIZE/XYZI pairs (one model each, in file order),
// RGBA (the palette), and the scene graph nodes nTRN (transform), nGRP
// (group) and nSHP (shape), which dot_vox keeps in file order and addresses
// by that position. Every other chunk is skipped.
#include "vox.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace svo {
namespace {

using Dict = std::map<std::string, std::string>;

struct Cursor {
  const uint8_t* p;
  const uint8_t* end;
  bool ok = true;
  int32_t i32() {
    if (end - p < 4) {
      ok = false;
      return 0;
    }
    int32_t v;
    std::memcpy(&v, p, 4);
    p += 4;
    return v;
  }
  std::string str() {
    const int32_t n = i32();
    if (!ok || n < 0 || end - p < n) {
      ok = false;
      return {};
    }
    std::string s(reinterpret_cast<const char*>(p), static_cast<size_t>(n));
    p += n;
    return s;
  }
  // DICT: pairs of strings; a repeated key keeps its last value (HashMap insert)
  Dict dict() {
    Dict d;
    const int32_t n = i32();
    if (n < 0) ok = false;
    for (int32_t i = 0; ok && i < n; ++i) {
      std::string k = str();
      std::string v = str();
      if (ok) d[k] = v;
    }
    return d;
  }
};

struct Model {
  int32_t size[3];
  const uint8_t* xyzi;  // 4 bytes per voxel: x, y, z, colour index as stored
  uint32_t count;
};

enum class NodeKind { Transform, Group, Shape };
struct Node {
  NodeKind kind;
  int32_t child = -1;                             // Transform
  std::vector<Dict> frames;                       // Transform
  std::vector<int32_t> children;                  // Group
  std::vector<std::pair<int32_t, Dict>> models;   // Shape
};

struct VoxData {
  std::vector<Model> models;
  std::vector<Node> scenes;
  const uint8_t* palette = nullptr;  // RGBA bytes
  uint32_t palette_len = 0;
};

// Rust's str::parse for integers: optional sign, ASCII digits, in range.
bool parse_int(const std::string& s, bool allow_minus, int64_t lo, int64_t hi, int64_t* out) {
  size_t i = 0;
  bool neg = false;
  if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
    neg = s[i] == '-';
    if (neg && !allow_minus) return false;
    ++i;
  }
  if (i == s.size()) return false;
  const uint64_t limit = neg ? static_cast<uint64_t>(-(lo + 1)) + 1u : static_cast<uint64_t>(hi);
  uint64_t v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const uint64_t d = static_cast<uint64_t>(s[i] - '0');
    if (v > (limit - d) / 10u) return false;
    v = v * 10u + d;
  }
  *out = neg ? (v == 0 ? 0 : -static_cast<int64_t>(v - 1u) - 1) : static_cast<int64_t>(v);
  return true;
}

bool parse_vox(const uint8_t* data, size_t n, VoxData* out, std::string* err) {
  if (n < 20 || std::memcmp(data, "VOX ", 4) != 0 || std::memcmp(data + 8, "MAIN", 4) != 0) {
    *err = "not a .vox file";
    return false;
  }
  int32_t main_content, main_children;
  std::memcpy(&main_content, data + 12, 4);
  std::memcpy(&main_children, data + 16, 4);
  if (main_content < 0 || main_children < 0 ||
      20ull + static_cast<uint64_t>(main_content) + static_cast<uint64_t>(main_children) > n) {
    *err = "truncated MAIN chunk";
    return false;
  }
  const uint8_t* p = data + 20 + main_content;
  const uint8_t* end = p + main_children;
  bool have_size = false;
  int32_t size[3] = {0, 0, 0};
  while (p < end) {
    if (end - p < 12) {
      *err = "truncated chunk header";
      return false;
    }
    int32_t cn, cm;
    std::memcpy(&cn, p + 4, 4);
    std::memcpy(&cm, p + 8, 4);
    if (cn < 0 || cm < 0 || static_cast<uint64_t>(end - p) < 12ull + cn + cm) {
      *err = "truncated chunk";
      return false;
    }
    const uint8_t* c = p + 12;
    Cursor r{c, c + cn};
    const std::string id(reinterpret_cast<const char*>(p), 4);
    if (id == "SIZE") {
      for (int k = 0; k < 3; ++k) size[k] = r.i32();
      have_size = r.ok;
    } else if (id == "XYZI") {
      const int32_t count = r.i32();
      if (!r.ok || count < 0 || (r.end - r.p) / 4 < count) {
        *err = "truncated XYZI chunk";
        return false;
      }
      if (have_size)
        out->models.push_back(Model{{size[0], size[1], size[2]}, r.p,
                                    static_cast<uint32_t>(count)});
    } else if (id == "RGBA") {
      out->palette = c;
      out->palette_len = static_cast<uint32_t>(cn / 4);
    } else if (id == "nTRN") {
      Node nd{NodeKind::Transform};
      (void)r.i32();   // node id
      (void)r.dict();  // node attributes
      nd.child = r.i32();
      (void)r.i32();   // reserved
      (void)r.i32();   // layer
      const int32_t nf = r.i32();
      if (nf < 0) r.ok = false;
      for (int32_t f = 0; r.ok && f < nf; ++f) nd.frames.push_back(r.dict());
      if (!r.ok) {
        *err = "malformed nTRN chunk";
        return false;
      }
      out->scenes.push_back(std::move(nd));
    } else if (id == "nGRP") {
      Node nd{NodeKind::Group};
      (void)r.i32();
      (void)r.dict();
      const int32_t k = r.i32();
      if (k < 0) r.ok = false;
      for (int32_t i = 0; r.ok && i < k; ++i) nd.children.push_back(r.i32());
      if (!r.ok) {
        *err = "malformed nGRP chunk";
        return false;
      }
      out->scenes.push_back(std::move(nd));
    } else if (id == "nSHP") {
      Node nd{NodeKind::Shape};
      (void)r.i32();
      (void)r.dict();
      const int32_t k = r.i32();
      if (k < 0) r.ok = false;
      for (int32_t i = 0; r.ok && i < k; ++i) {
        const int32_t mid = r.i32();
        Dict a = r.dict();
        nd.models.emplace_back(mid, std::move(a));
      }
      if (!r.ok) {
        *err = "malformed nSHP chunk";
        return false;
      }
      out->scenes.push_back(std::move(nd));
    }
    p += 12 + cn + cm;
  }
  return true;
}

struct Mat {
  int32_t m[3][3];
};

const Mat kIdentity{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};

Mat mul(const Mat& a, const Mat& b) {
  Mat o{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      o.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return o;
}

void apply(const Mat& a, const int64_t v[3], int64_t out[3]) {
  for (int r = 0; r < 3; ++r) out[r] = a.m[r][0] * v[0] + a.m[r][1] * v[1] + a.m[r][2] * v[2];
}

struct Visit {
  uint32_t model;
  int64_t t[3];
  Mat rot;
};

// iterate_vox_tree (magicavoxel.rs:104-202) for frame 0. Every stack entry
// keeps an index: a transform pushes its child once and pops when revisited,
// a group pushes its children in order, and a shape reports its frame-0
// models, pops, and advances its parent's index.
bool walk(const VoxData& v, std::vector<Visit>* out, std::string* err) {
  constexpr uint64_t kFrame = 0;
  if (v.scenes.empty() || v.scenes[0].kind != NodeKind::Transform) {
    *err = "the scene root must be a transform node";
    return false;
  }
  struct Item {
    int32_t node;
    int64_t t[3];
    Mat rot;
    uint32_t index;
  };
  std::vector<Item> stack;
  stack.push_back(Item{v.scenes[0].child, {0, 0, 0}, kIdentity, 0});
  while (!stack.empty()) {
    if (stack.size() > v.scenes.size() + 1) {
      *err = "scene graph has a cycle";
      return false;
    }
    Item& e = stack.back();
    if (e.node < 0 || static_cast<size_t>(e.node) >= v.scenes.size()) {
      *err = "scene node index out of range";
      return false;
    }
    const Node& nd = v.scenes[static_cast<size_t>(e.node)];
    if (nd.kind == NodeKind::Transform) {
      if (e.index != 0) {
        stack.pop_back();
        continue;
      }
      if (nd.frames.empty()) {
        *err = "transform node without frames";
        return false;
      }
      const Dict& fr = nd.frames[kFrame < nd.frames.size() ? kFrame : 0];
      Item next{nd.child, {e.t[0], e.t[1], e.t[2]}, kIdentity, 0};
      auto t = fr.find("_t");
      if (t != fr.end()) {
        // "x y z", split on single spaces; values past the third are unused
        std::vector<std::string> parts;
        size_t a = 0;
        for (;;) {
          const size_t b = t->second.find(' ', a);
          parts.push_back(t->second.substr(a, b == std::string::npos ? b : b - a));
          if (b == std::string::npos) break;
          a = b + 1;
        }
        for (const std::string& s : parts) {
          int64_t dummy;
          if (!parse_int(s, true, INT32_MIN, INT32_MAX, &dummy)) {
            *err = "translation is not a list of integers";
            return false;
          }
        }
        if (parts.size() < 3) {
          *err = "translation has fewer than three values";
          return false;
        }
        for (int k = 0; k < 3; ++k) {
          int64_t d;
          parse_int(parts[k], true, INT32_MIN, INT32_MAX, &d);
          next.t[k] += d;
          if (next.t[k] < INT32_MIN || next.t[k] > INT32_MAX) {
            *err = "translation overflows i32";
            return false;
          }
        }
      }
      auto r = fr.find("_r");
      if (r != fr.end()) {
        int64_t b;
        int8_t m[3][3];
        if (!parse_int(r->second, false, 0, 255, &b) || !vox_rotation(static_cast<uint8_t>(b), m)) {
          *err = "invalid rotation";
          return false;
        }
        Mat pm;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) pm.m[i][j] = m[i][j];
        next.rot = mul(e.rot, pm);
      }
      e.index += 1;
      stack.push_back(next);
    } else if (nd.kind == NodeKind::Group) {
      if (e.index < nd.children.size()) {
        Item next{nd.children[e.index], {e.t[0], e.t[1], e.t[2]}, e.rot, 0};
        e.index += 1;
        stack.push_back(next);
      } else {
        stack.pop_back();
      }
    } else {
      for (const auto& m : nd.models) {
        auto f = m.second.find("_f");
        int64_t frame = 0;
        if (f != m.second.end() && !parse_int(f->second, false, 0, INT64_MAX - 1, &frame)) {
          *err = "frame attribute is not an integer";
          return false;
        }
        if (static_cast<uint64_t>(frame) != kFrame) continue;
        if (m.first < 0 || static_cast<size_t>(m.first) >= v.models.size()) {
          *err = "shape names a missing model";
          return false;
        }
        out->push_back(Visit{static_cast<uint32_t>(m.first), {e.t[0], e.t[1], e.t[2]}, e.rot});
      }
      stack.pop_back();
      if (!stack.empty()) stack.back().index += 1;
    }
  }
  return true;
}

// V3c::from(model.size).transformed(orientation) / 2, with i32 division
// (truncation toward zero).
void half_extent(const Model& m, const Mat& rot, int64_t h[3]) {
  const int64_t s[3] = {m.size[0], m.size[1], m.size[2]};
  apply(rot, s, h);
  for (int k = 0; k < 3; ++k) h[k] /= 2;
}

}  // namespace

bool vox_rotation(uint8_t b, int8_t m[3][3]) {
  const int c0 = b & 3, c1 = (b >> 2) & 3;
  const int c2 = ~(c0 ^ c1) & 3;
  if (c0 == 3 || c1 == 3 || c0 == c1) return false;
  std::memset(m, 0, 9);
  const int cols[3] = {c0, c1, c2};
  for (int r = 0; r < 3; ++r) m[r][cols[r]] = (b & (0x10 << r)) ? -1 : 1;
  return true;
}

VoxStatus octree_from_vox(const uint8_t* data, size_t n, uint32_t brick_dim, Octree* out,
                          std::string* err) {
  VoxData v;
  if (!parse_vox(data, n, &v, err)) return VoxStatus::Malformed;
  if (!v.palette) {
    *err = "no RGBA palette chunk (the built-in default palette is not supported)";
    return VoxStatus::Malformed;
  }
  std::vector<Visit> visits;
  if (!walk(v, &visits, err)) return VoxStatus::Malformed;
  if (visits.empty()) {
    *err = "frame 0 of the scene places no model";
    return VoxStatus::Malformed;
  }
  // bounds of every placed model, Rzup (load_vox_file_internal, :281-326)
  int64_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (const Visit& vi : visits) {
    int64_t h[3];
    half_extent(v.models[vi.model], vi.rot, h);
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min({lo[k], vi.t[k] - h[k], vi.t[k] + h[k]});
      hi[k] = std::max({hi[k], vi.t[k] + h[k], vi.t[k] - h[k]});
    }
  }
  // Rzup -> Lyup swaps y and z (convert_coordinate, spatial/math/mod.rs:196-199);
  // the edge is 2^ceil(log2(extent)) in f32 (:266-269)
  int64_t extent = 0;
  for (int k = 0; k < 3; ++k) extent = std::max(extent, hi[k] - lo[k]);
  const float lg = std::ceil(std::log2(static_cast<float>(extent)));
  const uint32_t e = !(lg > 0.0f) ? 0u : static_cast<uint32_t>(lg);
  if (e >= 32) {
    *err = "scene extent too large";
    return VoxStatus::TreeSize;
  }
  Status st;
  Octree t = Octree::create(1u << e, brick_dim, &st);
  if (st != Status::Ok) {
    *err = "no octree of edge " + std::to_string(1u << e) + " and brick dimension " +
           std::to_string(brick_dim);
    return st == Status::InvalidBrickDimension ? VoxStatus::BrickDim : VoxStatus::TreeSize;
  }
  // voxels (load_vox_data_internal, :329-373): a model's lowest corner, less
  // the scene minimum, plus one step back in every axis the orientation
  // reverses, then each rotated voxel offset; inserted in Lyup
  for (const Visit& vi : visits) {
    const Model& m = v.models[vi.model];
    int64_t h[3], base[3];
    half_extent(m, vi.rot, h);
    for (int k = 0; k < 3; ++k) base[k] = vi.t[k] - h[k] - lo[k] + (h[k] < 0 ? -1 : 0);
    for (uint32_t i = 0; i < m.count; ++i) {
      const uint8_t* q = m.xyzi + 4ull * i;
      const int64_t local[3] = {q[0], q[1], q[2]};
      int64_t r[3];
      apply(vi.rot, local, r);
      const int64_t x = base[0] + r[0], y = base[2] + r[2], z = base[1] + r[1];
      const uint32_t ci = q[3] == 0 ? 0u : q[3] - 1u;  // dot_vox stores index - 1
      if (ci >= v.palette_len) {
        *err = "colour index outside the palette";
        return VoxStatus::Malformed;
      }
      const uint8_t* c = v.palette + 4ull * ci;
      Entry en;
      en.kind = EntryKind::Visual;
      en.albedo = (uint32_t(c[0]) << 24) | (uint32_t(c[1]) << 16) | (uint32_t(c[2]) << 8) | c[3];
      // V3c<i32> -> V3c<u32> is an `as` cast: negatives wrap and land outside the tree
      const Status s = t.insert(static_cast<uint32_t>(x), static_cast<uint32_t>(y),
                                static_cast<uint32_t>(z), en);
      if (s != Status::Ok) {
        *err = "voxel outside the tree at (" + std::to_string(x) + ", " + std::to_string(y) +
               ", " + std::to_string(z) + ")";
        return VoxStatus::Position;
      }
    }
  }
  *out = std::move(t);
  return VoxStatus::Ok;
}

}  // namespace svo
