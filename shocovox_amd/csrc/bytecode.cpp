// Bencode reader and writer for Octree<u32> in the reference's layout.
//
// The format (all integers in decimal, `i...e`; strings `len:bytes`; lists
// `l...e`), as src/convert/bytecode.rs and src/object_pool.rs encode it:
//   Octree        l auto_simplify size brick_dim Pool Children* Mips* Colors* Data* Mip e
//   Pool          l first_available l (l reserved NodeContent e)* e e
//   NodeContent   "#" | l "##" bits e | l "###" Brick x8 e | l "##u#" Brick e
//   Brick         "#b" | l "#b#" value e | l "##b#" len value x len "#" e
//   NodeChildren  "##x##" | l "##c##" key x8 e | l "##b##" bitmap e
//   Colors        l (l r g b a e)* e          Data  l value* e
//   Mip           l enabled n (key method)*n m (key threshold)*m e
// The reader walks the bytes once, building the host octree directly.
#include "bytecode.hpp"

#include <cstring>

namespace svo {

namespace {

class Reader {
 public:
  Reader(const uint8_t* d, size_t n) : d_(d), n_(n) {}

  bool fail(const std::string& what) {
    if (err_.empty()) err_ = what + " at byte " + std::to_string(pos_);
    return false;
  }
  const std::string& error() const { return err_; }
  bool ok() const { return err_.empty(); }

  int peek() const { return pos_ < n_ ? d_[pos_] : -1; }
  bool at_list_end() const { return peek() == 'e'; }

  bool list_begin() {
    if (peek() != 'l') return fail("expected a list");
    ++pos_;
    return true;
  }
  bool list_end() {
    if (peek() != 'e') return fail("expected the end of a list");
    ++pos_;
    return true;
  }
  bool integer(uint64_t* v) {
    if (peek() != 'i') return fail("expected an integer");
    ++pos_;
    uint64_t x = 0;
    size_t digits = 0;
    if (peek() == '-') return fail("negative integer");
    while (pos_ < n_ && d_[pos_] >= '0' && d_[pos_] <= '9') {
      const uint64_t nx = x * 10 + (d_[pos_] - '0');
      if (nx / 10 != x) return fail("integer overflow");
      x = nx;
      ++pos_;
      ++digits;
    }
    if (digits == 0 || peek() != 'e') return fail("malformed integer");
    ++pos_;
    *v = x;
    return true;
  }
  bool integer32(uint32_t* v) {
    uint64_t x = 0;
    if (!integer(&x)) return false;
    if (x > 0xFFFFFFFFull) return fail("integer above u32");
    *v = static_cast<uint32_t>(x);
    return true;
  }
  // a byte string; `s` points into the input
  bool bytes(const char** s, size_t* len) {
    size_t l = 0;
    size_t digits = 0;
    while (pos_ < n_ && d_[pos_] >= '0' && d_[pos_] <= '9') {
      l = l * 10 + (d_[pos_] - '0');
      ++pos_;
      if (++digits > 18) return fail("string length too long");
    }
    if (digits == 0 || peek() != ':') return fail("expected a string");
    ++pos_;
    if (l > n_ - pos_) return fail("string runs past the end");
    *s = reinterpret_cast<const char*>(d_ + pos_);
    *len = l;
    pos_ += l;
    return true;
  }
  bool is_string() const { return peek() >= '0' && peek() <= '9'; }
  bool tag(std::string* t) {
    const char* s = nullptr;
    size_t l = 0;
    if (!bytes(&s, &l)) return false;
    t->assign(s, l);
    return true;
  }
  // skips one value of any type; nested lists and dictionaries are tracked
  // with a depth counter, not recursion, so no nesting depth of an input can
  // exhaust the stack
  bool skip() {
    size_t depth = 0;
    do {
      const int c = peek();
      if (c == 'i') {
        ++pos_;
        while (pos_ < n_ && d_[pos_] != 'e') ++pos_;
        if (pos_ >= n_) return fail("unterminated integer");
        ++pos_;
      } else if (c == 'l' || c == 'd') {
        ++pos_;
        ++depth;
      } else if (c == 'e' && depth > 0) {
        ++pos_;
        --depth;
      } else if (c < 0) {
        return fail("unterminated list");
      } else {
        const char* s;
        size_t l;
        if (!bytes(&s, &l)) return false;
      }
    } while (depth > 0);
    return true;
  }
  size_t pos() const { return pos_; }
  size_t size() const { return n_; }

 private:
  const uint8_t* d_;
  size_t n_;
  size_t pos_ = 0;
  std::string err_;
};

bool read_brick(Reader& r, uint32_t n3, Brick* b) {
  if (r.is_string()) {
    std::string t;
    if (!r.tag(&t)) return false;
    if (t != "#b") return r.fail("unknown empty-brick marker '" + t + "'");
    *b = Brick::empty();
    return true;
  }
  if (!r.list_begin()) return false;
  std::string t;
  if (!r.tag(&t)) return false;
  if (t == "#b#") {
    uint32_t v = 0;
    if (!r.integer32(&v)) return false;
    *b = Brick::make_solid(v);
  } else if (t == "##b#") {
    uint64_t len = 0;
    if (!r.integer(&len)) return false;
    if (len != n3) return r.fail("brick of " + std::to_string(len) + " voxels, expected " +
                                 std::to_string(n3));
    std::vector<uint32_t> v(len);
    for (uint64_t i = 0; i < len; ++i)
      if (!r.integer32(&v[i])) return false;
    *b = Brick::make_parted(std::move(v));
  } else {
    return r.fail("unknown brick marker '" + t + "'");
  }
  // the Parted form ends with a "#" marker; anything left is skipped
  while (r.ok() && !r.at_list_end())
    if (!r.skip()) return false;
  return r.list_end();
}

bool read_node(Reader& r, uint32_t n3, NodeContent* n) {
  if (r.is_string()) {
    std::string t;
    if (!r.tag(&t)) return false;
    if (t != "#") return r.fail("unknown node marker '" + t + "'");
    *n = NodeContent::nothing();
    return true;
  }
  if (!r.list_begin()) return false;
  std::string t;
  if (!r.tag(&t)) return false;
  if (t == "##") {
    uint64_t bits = 0;
    if (!r.integer(&bits)) return false;
    *n = NodeContent::internal(bits);
  } else if (t == "###") {
    NodeContent c;
    c.kind = NodeKind::Leaf;
    for (uint32_t o = 0; o < 8; ++o)
      if (!read_brick(r, n3, &c.bricks[o])) return false;
    *n = std::move(c);
  } else if (t == "##u#") {
    Brick b;
    if (!read_brick(r, n3, &b)) return false;
    *n = NodeContent::uniform(std::move(b));
  } else {
    return r.fail("unknown node marker '" + t + "'");
  }
  return r.list_end();
}

bool read_children(Reader& r, NodeChildren* c) {
  *c = NodeChildren{};
  if (r.is_string()) {
    std::string t;
    if (!r.tag(&t)) return false;
    if (t != "##x##") return r.fail("unknown children marker '" + t + "'");
    return true;
  }
  if (!r.list_begin()) return false;
  std::string t;
  if (!r.tag(&t)) return false;
  if (t == "##c##") {
    c->kind = ChildKind::Children;
    for (uint32_t o = 0; o < 8; ++o)
      if (!r.integer32(&c->c[o])) return false;
  } else if (t == "##b##") {
    uint64_t bm = 0;
    if (!r.integer(&bm)) return false;
    *c = NodeChildren::bitmap_of(bm);
  } else {
    return r.fail("unknown children marker '" + t + "'");
  }
  return r.list_end();
}

void put_int(std::string* o, uint64_t v) {
  // "i<decimal>e" formatted in place (no temporary string per integer)
  char buf[24];
  char* e = buf + sizeof(buf);
  char* p = e;
  *--p = 'e';
  do {
    *--p = static_cast<char>('0' + v % 10);
    v /= 10;
  } while (v);
  *--p = 'i';
  o->append(p, static_cast<size_t>(e - p));
}
void put_str(std::string* o, const char* s) {
  const size_t l = std::strlen(s);
  o->append(std::to_string(l));
  o->push_back(':');
  o->append(s, l);
}

void put_brick(std::string* o, const Brick& b) {
  if (b.kind == BrickKind::Empty) {
    put_str(o, "#b");
  } else if (b.kind == BrickKind::Solid) {
    o->push_back('l');
    put_str(o, "#b#");
    put_int(o, b.solid);
    o->push_back('e');
  } else {
    o->push_back('l');
    put_str(o, "##b#");
    put_int(o, b.parted.size());
    for (uint32_t v : b.parted) put_int(o, v);
    put_str(o, "#");
    o->push_back('e');
  }
}

}  // namespace

bool octree_from_bencode(const uint8_t* data, size_t n, Octree* out, std::string* err) {
  Reader r(data, n);
  auto bail = [&]() {
    *err = r.error();
    return false;
  };
  uint64_t auto_simplify = 0, size = 0, dim = 0, first_available = 0;
  if (!r.list_begin() || !r.integer(&auto_simplify) || !r.integer(&size) || !r.integer(&dim))
    return bail();
  if (auto_simplify > 1) {
    r.fail("auto_simplify is not a boolean");
    return bail();
  }
  Status st;
  Octree t = Octree::create(static_cast<uint32_t>(size), static_cast<uint32_t>(dim), &st);
  if (size > 0xFFFFFFFFull || dim > 0xFFFFFFFFull || st != Status::Ok) {
    *err = "invalid octree size " + std::to_string(size) + " or brick dimension " +
           std::to_string(dim);
    return false;
  }
  const uint32_t n3 = static_cast<uint32_t>(dim * dim * dim);
  // ObjectPool<NodeContent>
  std::vector<NodeContent> items;
  std::vector<uint8_t> reserved;
  if (!r.list_begin() || !r.integer(&first_available) || !r.list_begin()) return bail();
  while (r.ok() && !r.at_list_end()) {
    uint64_t res = 0;
    NodeContent c;
    if (!r.list_begin() || !r.integer(&res) || !read_node(r, n3, &c) || !r.list_end())
      return bail();
    if (res > 1) {
      r.fail("reserved flag is not a boolean");
      return bail();
    }
    items.push_back(std::move(c));
    reserved.push_back(static_cast<uint8_t>(res));
  }
  if (!r.list_end() || !r.list_end()) return bail();
  // node_children
  std::vector<NodeChildren> kids;
  if (!r.list_begin()) return bail();
  while (r.ok() && !r.at_list_end()) {
    NodeChildren c;
    if (!read_children(r, &c)) return bail();
    kids.push_back(c);
  }
  if (!r.list_end()) return bail();
  // node_mips
  std::vector<Brick> mips;
  if (!r.list_begin()) return bail();
  while (r.ok() && !r.at_list_end()) {
    Brick b;
    if (!read_brick(r, n3, &b)) return bail();
    mips.push_back(std::move(b));
  }
  if (!r.list_end()) return bail();
  // colour palette
  std::vector<uint32_t> colors;
  if (!r.list_begin()) return bail();
  while (r.ok() && !r.at_list_end()) {
    uint64_t c[4];
    if (!r.list_begin()) return bail();
    for (uint64_t& v : c)
      if (!r.integer(&v)) return bail();
    if (!r.list_end()) return bail();
    if (c[0] > 255 || c[1] > 255 || c[2] > 255 || c[3] > 255) {
      r.fail("colour component above 255");
      return bail();
    }
    colors.push_back(static_cast<uint32_t>(c[0] << 24 | c[1] << 16 | c[2] << 8 | c[3]));
  }
  if (!r.list_end()) return bail();
  // data palette
  std::vector<uint32_t> data_pal;
  if (!r.list_begin()) return bail();
  while (r.ok() && !r.at_list_end()) {
    uint32_t v = 0;
    if (!r.integer32(&v)) return bail();
    data_pal.push_back(v);
  }
  if (!r.list_end()) return bail();
  // MIPMapStrategy (bytecode.rs:437-520): [enabled, n, (level, method)*n,
  // m, (level, threshold * 1000)*m]; methods as bytecode.rs:522-570 encode them
  MipStrategy strategy;
  {
    uint64_t enabled = 0, count = 0;
    if (!r.list_begin() || !r.integer(&enabled) || !r.integer(&count)) return bail();
    if (enabled > 1) {
      r.fail("MIP enabled flag is not a boolean");
      return bail();
    }
    strategy.enabled = enabled != 0;
    strategy.methods.clear();
    for (uint64_t i = 0; i < count; ++i) {
      uint64_t level = 0, code = 0;
      if (!r.integer(&level) || !r.integer(&code)) return bail();
      MipMethodSpec m;
      if (code <= 2) {
        m.kind = static_cast<MipMethod>(code);
      } else if (code >= 3 && code < 1002) {
        m.kind = MipMethod::Posterize;
        m.thr = (static_cast<float>(code) - 3.0f) / 1000.0f;
      } else if (code >= 1003 && code < 2001) {
        m.kind = MipMethod::PosterizeBD;
        m.thr = (static_cast<float>(code) - 1003.0f) / 1000.0f;
      } else {
        r.fail("unknown MIP resampling method " + std::to_string(code));
        return bail();
      }
      strategy.methods[static_cast<uint32_t>(level)] = m;
    }
    if (!r.integer(&count)) return bail();
    strategy.color_thresholds.clear();
    for (uint64_t i = 0; i < count; ++i) {
      uint64_t level = 0;
      uint32_t thr = 0;
      if (!r.integer(&level) || !r.integer32(&thr)) return bail();
      strategy.color_thresholds[static_cast<uint32_t>(level)] = static_cast<float>(thr) / 1000.0f;
    }
    if (!r.list_end()) return bail();
  }
  if (!r.list_end()) return bail();
  if (kids.size() != items.size()) {
    *err = "node and children counts differ (" + std::to_string(items.size()) + " vs " +
           std::to_string(kids.size()) + ")";
    return false;
  }
  if (items.empty() || !reserved[Octree::ROOT_NODE_KEY]) {
    *err = "the root node is missing";
    return false;
  }
  for (size_t k = 0; k < kids.size(); ++k)
    if (kids[k].kind == ChildKind::Children)
      for (uint32_t o = 0; o < 8; ++o)
        if (kids[k].c[o] != EMPTY_MARKER && kids[k].c[o] >= items.size()) {
          *err = "child key out of range";
          return false;
        }
  if (first_available > items.size()) first_available = items.size();
  t.assign_model(std::move(items), std::move(reserved), first_available, std::move(kids),
                 std::move(colors), std::move(data_pal), auto_simplify != 0);
  t.assign_mips(std::move(mips), std::move(strategy));
  *out = std::move(t);
  return true;
}

void octree_to_bencode(const Octree& t, std::string* out) {
  std::string& o = *out;
  o.clear();
  {
    // about 4 characters per stored value: reserve once instead of regrowing
    size_t values = 64;
    const NodePool& pl = t.nodes();
    for (uint32_t k = 0; k < pl.len(); ++k) {
      const NodeContent& n = pl.get(k);
      const int nb = n.kind == NodeKind::Leaf ? 8 : (n.kind == NodeKind::UniformLeaf ? 1 : 0);
      for (int b = 0; b < nb; ++b) values += 4 + n.bricks[b].parted.size();
      values += 4 + t.node_mip(k).parted.size() + 12;
    }
    o.reserve(values * 5);
  }
  o.push_back('l');
  put_int(&o, t.auto_simplify ? 1 : 0);
  put_int(&o, t.size());
  put_int(&o, t.brick_dim());
  const NodePool& pool = t.nodes();
  o.push_back('l');
  put_int(&o, pool.first_available());
  o.push_back('l');
  for (uint32_t k = 0; k < pool.len(); ++k) {
    o.push_back('l');
    put_int(&o, pool.reserved(k) ? 1 : 0);
    const NodeContent& n = pool.get(k);
    switch (n.kind) {
      case NodeKind::Nothing:
        put_str(&o, "#");
        break;
      case NodeKind::Internal:
        o.push_back('l');
        put_str(&o, "##");
        put_int(&o, n.bits);
        o.push_back('e');
        break;
      case NodeKind::Leaf:
        o.push_back('l');
        put_str(&o, "###");
        for (uint32_t b = 0; b < 8; ++b) put_brick(&o, n.bricks[b]);
        o.push_back('e');
        break;
      case NodeKind::UniformLeaf:
        o.push_back('l');
        put_str(&o, "##u#");
        put_brick(&o, n.bricks[0]);
        o.push_back('e');
        break;
    }
    o.push_back('e');
  }
  o.append("ee");
  o.push_back('l');
  for (const NodeChildren& c : t.node_children()) {
    if (c.kind == ChildKind::NoChildren) {
      put_str(&o, "##x##");
    } else if (c.kind == ChildKind::Children) {
      o.push_back('l');
      put_str(&o, "##c##");
      for (uint32_t v : c.c) put_int(&o, v);
      o.push_back('e');
    } else {
      o.push_back('l');
      put_str(&o, "##b##");
      put_int(&o, c.bitmap);
      o.push_back('e');
    }
  }
  o.push_back('e');
  o.push_back('l');  // node_mips: one brick per node
  for (uint32_t k = 0; k < pool.len(); ++k) put_brick(&o, t.node_mip(k));
  o.push_back('e');
  o.push_back('l');
  for (uint32_t c : t.color_palette()) {
    o.push_back('l');
    put_int(&o, c >> 24);
    put_int(&o, (c >> 16) & 0xFFu);
    put_int(&o, (c >> 8) & 0xFFu);
    put_int(&o, c & 0xFFu);
    o.push_back('e');
  }
  o.push_back('e');
  o.push_back('l');
  for (uint32_t v : t.data_palette()) put_int(&o, v);
  o.push_back('e');
  // MIP strategy, levels in ascending order (the reference writes its
  // HashMaps in their iteration order)
  const MipStrategy& ms = t.mip_strategy();
  o.push_back('l');
  put_int(&o, ms.enabled ? 1 : 0);
  put_int(&o, ms.methods.size());
  for (const auto& kv : ms.methods) {
    put_int(&o, kv.first);
    const MipMethodSpec& m = kv.second;
    uint64_t code = static_cast<uint64_t>(m.kind);
    // (threshold * 1000.) as u32: truncation of the f32 product
    if (m.kind == MipMethod::Posterize) code = 3 + f2u32(m.thr * 1000.0f);
    if (m.kind == MipMethod::PosterizeBD) code = 1003 + f2u32(m.thr * 1000.0f);
    put_int(&o, code);
  }
  put_int(&o, ms.color_thresholds.size());
  for (const auto& kv : ms.color_thresholds) {
    put_int(&o, kv.first);
    put_int(&o, f2u32(kv.second * 1000.0f));
  }
  o.push_back('e');
  o.push_back('e');
}

}  // namespace svo
