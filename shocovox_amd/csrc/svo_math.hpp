// Exact single-precision geometry shared by the host octree model, the
// flattener and the gfx950 traversal kernels.
//
// Every function here reproduces the IEEE-754 binary32 operation sequence of
// the reference crate (shocovox-rs 0.11.1) so that host and device produce
// bit-identical trajectories:
//   * no FMA contraction (all sources are built with -ffp-contract=off);
//   * Rust's f32::min / f32::max are NaN-ignoring and, on equal operands,
//     return the receiver (the x86 lowering of llvm.minnum/maxnum:
//     `isnan(a) ? b : (b < a ? b : a)`), spelled out here as selects so the
//     GPU never substitutes v_min_f32 with different signed-zero rules;
//   * f32::signum is +1 for +0.0, -1 for -0.0 and NaN for NaN;
//   * `as` casts from f32 truncate, saturate and map NaN to 0;
//   * conversions V3c<f32> -> V3c<int> use round() (half away from zero);
//   * `x.powf(2.)` is `x * x` (what the optimiser emits for the reference).
// Reference: src/spatial/math/vector.rs:55-81,195-380,
//            src/spatial/raytracing/mod.rs:5-134, src/spatial/mod.rs:31-53,
//            src/spatial/math/mod.rs:11-37, src/spatial/lut.rs:154-253.
#pragma once

#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#define SVO_HD __host__ __device__
#else
#define SVO_HD
#endif

namespace svo {

struct V3f {
  float x, y, z;
};

SVO_HD inline V3f v3(float x, float y, float z) { return V3f{x, y, z}; }
SVO_HD inline V3f v3u(float s) { return V3f{s, s, s}; }
SVO_HD inline V3f operator+(V3f a, V3f b) { return V3f{a.x + b.x, a.y + b.y, a.z + b.z}; }
SVO_HD inline V3f operator-(V3f a, V3f b) { return V3f{a.x - b.x, a.y - b.y, a.z - b.z}; }
SVO_HD inline V3f operator*(V3f a, float s) { return V3f{a.x * s, a.y * s, a.z * s}; }
SVO_HD inline V3f operator/(V3f a, float s) { return V3f{a.x / s, a.y / s, a.z / s}; }

// f32::min / f32::max (reference: vector.rs, raytracing/mod.rs:43-44)
SVO_HD inline float rmin(float a, float b) {
  return (a != a) ? b : ((b < a) ? b : a);
}
SVO_HD inline float rmax(float a, float b) {
  return (a != a) ? b : ((b > a) ? b : a);
}

// f32::signum
SVO_HD inline float rsignum(float x) {
  if (x != x) return x;
  return std::signbit(x) ? -1.0f : 1.0f;
}
// rsignum for a non-NaN x (+-1 by the sign bit); +-1 for NaN as well
SVO_HD inline float sign_one(float x) { return copysignf(1.0f, x); }

// `f as i32` (saturating, NaN -> 0)
// Bit (b mod 32) of w.
SVO_HD inline uint32_t bit_of(uint32_t w, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  // v_bfe_u32 reads the low 5 bits of the offset: one instruction
  return __builtin_amdgcn_ubfe(w, b, 1u);
#else
  return (w >> (b & 31u)) & 1u;
#endif
}

SVO_HD inline int32_t f2i32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // v_cvt_i32_f32 truncates toward zero, clamps out-of-range values to
  // [INT32_MIN, INT32_MAX] and converts NaN to 0: Rust's `as i32` in one
  // instruction (the portable form below compiles to compares and branches)
  int32_t r;
  __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#endif
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return (-2147483647 - 1);
  return static_cast<int32_t>(x);
}

// `s as i32` for a DDA step component. dda_step returns, per axis, either
// +0.0 or signum(direction), i.e. one of {-1, +1, +0, NaN}; on that domain
// this equals f2i32(s) and f2i32(roundf(s)) (NaN -> 0): one conversion
// instruction on the device, two compares on the host.
SVO_HD inline int32_t step_int(float s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return f2i32(s);
#else
  return s == 1.0f ? 1 : (s == -1.0f ? -1 : 0);
#endif
}

// `f as usize` (saturating, NaN -> 0)
SVO_HD inline uint64_t f2usize(float x) {
  if (!(x > 0.0f)) return 0;  // NaN, zero and negatives
  if (x >= 18446744073709551616.0f) return ~0ull;
  return static_cast<uint64_t>(x);
}

// `f as u32` (saturating, NaN -> 0)
SVO_HD inline uint32_t f2u32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // v_cvt_u32_f32: truncation, clamping to [0, UINT32_MAX], NaN -> 0
  uint32_t r;
  __asm__("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#endif
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(x);
}

// V3c<f32>::length: sqrt((x*x + y*y) + z*z)  (vector.rs:75-77)
SVO_HD inline float length(V3f v) { return sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z); }
// V3c<f32>::normalized: three divisions by the length (vector.rs:79-81)
SVO_HD inline V3f normalized(V3f v) { return v / length(v); }

SVO_HD inline V3f cross(V3f a, V3f b) {
  return V3f{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Octant index of a child region: (x>=h) + 2(z>=h) + 4(y>=h)  (math/mod.rs:11-19)
SVO_HD inline uint32_t hash_region(V3f off, float half) {
  return static_cast<uint32_t>(off.x >= half) + static_cast<uint32_t>(off.z >= half) * 2u +
         static_cast<uint32_t>(off.y >= half) * 4u;
}

// Direction octant: hash_region(1 + dir, 1)  (math/mod.rs:22-26)
SVO_HD inline uint32_t hash_direction(V3f d) { return hash_region(v3u(1.0f) + d, 1.0f); }

constexpr uint32_t OOB_OCTANT = 8;

// OCTANT_OFFSET_REGION_LUT (lut.rs:156-197) as bit triples: x=bit0, y=bit1, z=bit2
SVO_HD inline V3f octant_offset(uint32_t octant) {
  // octant = x + 2z + 4y
  return V3f{static_cast<float>(octant & 1u), static_cast<float>((octant >> 2) & 1u),
             static_cast<float>((octant >> 1) & 1u)};
}

// Mask of the 2x2x2 bitmap cells covered by one octant (lut.rs:199-208)
SVO_HD inline uint64_t octant_bitmap_mask(uint32_t octant) {
  const uint64_t base = 0x0000000000330033ull;
  const uint32_t ox = octant & 1u, oz = (octant >> 1) & 1u, oy = (octant >> 2) & 1u;
  return base << (ox * 2u + oy * 8u + oz * 32u);
}

// The reference's OCTANT_STEP_RESULT_LUT (lut.rs:237-253): the octant reached
// from `octant` after a step with the given component signs, 8 = out of bounds.
SVO_HD inline uint32_t step_octant_signs(uint32_t octant, int sx, int sy, int sz) {
  int x = static_cast<int>(octant & 1u) + sx;
  int y = static_cast<int>((octant >> 2) & 1u) + sy;
  int z = static_cast<int>((octant >> 1) & 1u) + sz;
  if (x < 0 || x > 1 || y < 0 || y > 1 || z < 0 || z > 1) return OOB_OCTANT;
  return static_cast<uint32_t>(x + 2 * z + 4 * y);
}

// step_octant (raytracing/mod.rs:68-80): `(step.c as i32).signum()`
SVO_HD inline uint32_t step_octant(uint32_t octant, V3f step) {
  // step_int is already in {-1, 0, 1}, where signum is the identity
  return step_octant_signs(octant, step_int(step.x), step_int(step.y), step_int(step.z));
}

// Cube{min_position, size}  (spatial/mod.rs:31-53)
struct Cube {
  V3f min;
  float size;
};

// (min + offset * half per axis: the offset is 0 or 1, so the product is
// exact and the fused form rounds the same sum once)
SVO_HD inline Cube child_bounds_for(Cube c, uint32_t octant) {
  const float child_size = c.size / 2.0f;
  const V3f o = octant_offset(octant);
  return Cube{V3f{fmaf(o.x, child_size, c.min.x), fmaf(o.y, child_size, c.min.y),
                  fmaf(o.z, child_size, c.min.z)},
              child_size};
}

// Cube::intersect_ray (raytracing/mod.rs:32-61).
// Returns 0: no intersection, 1: origin inside (impact distance None),
// 2: impact at *t (which may be -0.0).
SVO_HD inline int intersect_ray(Cube c, V3f o, V3f d, float* t) {
  const V3f mx = c.min + v3u(c.size);
  const float t1 = (c.min.x - o.x) / d.x;
  const float t2 = (mx.x - o.x) / d.x;
  const float t3 = (c.min.y - o.y) / d.y;
  const float t4 = (mx.y - o.y) / d.y;
  const float t5 = (c.min.z - o.z) / d.z;
  const float t6 = (mx.z - o.z) / d.z;
  const float tmin = rmax(rmax(rmin(t1, t2), rmin(t3, t4)), rmin(t5, t6));
  const float tmax = rmin(rmin(rmax(t1, t2), rmax(t3, t4)), rmax(t5, t6));
  if (tmax < 0.0f || tmin > tmax) return 0;
  if (tmin < 0.0f) return 1;
  *t = tmin;
  return 2;
}

// cube_impact_normal (raytracing/mod.rs:106-134)
SVO_HD inline V3f cube_impact_normal(Cube c, V3f p) {
  const V3f mid = (c.min + v3u(c.size / 2.0f)) - p;
  const float ax = fabsf(mid.x), ay = fabsf(mid.y), az = fabsf(mid.z);
  const float m = rmax(rmax(ax, ay), az);
  const V3f n{ax == m ? -mid.x : 0.0f, ay == m ? -mid.y : 0.0f, az == m ? -mid.z : 0.0f};
  return normalized(n);
}

// get_dda_scale_factors (raytracing_on_cpu.rs:99-112)
SVO_HD inline V3f dda_scale_factors(V3f d) {
  const float zx = d.z / d.x, yx = d.y / d.x;
  const float xy = d.x / d.y, zy = d.z / d.y;
  const float xz = d.x / d.z, yz = d.y / d.z;
  return V3f{sqrtf((1.0f + zx * zx) + yx * yx), sqrtf((xy * xy + 1.0f) + zy * zy),
             sqrtf((xz * xz + 1.0f) + yz * yz)};
}

// c - s * d for a signum s (+1, -1 or NaN): s * d is d negated or not (exact)
// or NaN, so a fused multiply-add, which rounds once, gives exactly the
// reference's product-then-difference (raytracing_on_cpu.rs:136-140).
SVO_HD inline float sub_signed(float c, float s, float d) { return fmaf(-s, d, c); }

// dda_step_to_next_sibling (raytracing_on_cpu.rs:124-152); `sig` is the
// per-ray signum of the direction, hoisted because it never changes.
SVO_HD inline V3f dda_step(V3f dir, V3f sig, V3f* p, Cube b, V3f scale) {
  const V3f diff = *p - b.min;
  const float snx = sub_signed(b.size * rmax(sig.x, 0.0f), sig.x, diff.x);
  const float sny = sub_signed(b.size * rmax(sig.y, 0.0f), sig.y, diff.y);
  const float snz = sub_signed(b.size * rmax(sig.z, 0.0f), sig.z, diff.z);
  const float dx = fabsf(snx * scale.x);
  const float dy = fabsf(sny * scale.y);
  const float dz = fabsf(snz * scale.z);
  // operands are |.| or NaN, where f32::min and fminf (IEEE minNum) agree:
  // NaN yields the other operand and no -0.0 can occur
  const float m = fminf(fminf(dx, dy), dz);
  *p = *p + dir * m;
  return V3f{m == dx ? sig.x : 0.0f, m == dy ? sig.y : 0.0f, m == dz ? sig.z : 0.0f};
}

// dda_step reporting which axes stepped (m == d{x,y,z}, the axes whose step
// component is signum(direction)); the others' components are +0.0.
SVO_HD inline void dda_step_axes(V3f dir, V3f sig, V3f* p, V3f bmin, float size, V3f scale,
                                 bool* ax, bool* ay, bool* az) {
  const V3f diff = *p - bmin;
  const float snx = sub_signed(size * rmax(sig.x, 0.0f), sig.x, diff.x);
  const float sny = sub_signed(size * rmax(sig.y, 0.0f), sig.y, diff.y);
  const float snz = sub_signed(size * rmax(sig.z, 0.0f), sig.z, diff.z);
  const float dx = fabsf(snx * scale.x);
  const float dy = fabsf(sny * scale.y);
  const float dz = fabsf(snz * scale.z);
  const float m = fminf(fminf(dx, dy), dz);
  *p = *p + dir * m;
  *ax = m == dx;
  *ay = m == dy;
  *az = m == dz;
}

// dda_step_axes from the cell's far faces: far = the cell corner the
// direction points to per axis (min + size * max(signum, 0)). The reference's
// size*max(s,0) - s*(p - min) equals far - p on every axis when p - min is
// exact, which holds in the power-of-two layouts (min a multiple of a
// power-of-two size >= 1, coordinates below 2^24, p within about half a
// cell of the cell: min = 0, or Sterbenz's lemma, or p - min in [size, 2 size)
// on the representable grid of size): for s = +1 both are the once-rounded
// real value min + size - p; for s = -1 the reference's is p - min exactly
// and far - p its negation, equal after |.|; for s = NaN every scale factor is
// NaN and both distances are NaN.
// Returns the step length m: when m is not NaN, at least one axis steps
// (fminf returns one of its operands, and m == that operand).
SVO_HD inline float dda_step_far(V3f dir, V3f* p, V3f far, V3f scale, bool* ax, bool* ay,
                                 bool* az) {
  const float dx = fabsf((far.x - p->x) * scale.x);
  const float dy = fabsf((far.y - p->y) * scale.y);
  const float dz = fabsf((far.z - p->z) * scale.z);
  const float m = fminf(fminf(dx, dy), dz);
  *p = *p + dir * m;
  *ax = m == dx;
  *ay = m == dy;
  *az = m == dz;
  return m;
}

// Exact reciprocal of a power of two, 0 for any other value. For s = 2^k the
// products x * 2^-k and the quotient x / 2^k are the same real number, and
// both operations round it once to nearest-even, so `x * pow2_rcp(s)` is
// bit-identical to `x / s` (subnormal results included, as f32 denormals are
// preserved). Reciprocals that would themselves be subnormal are refused.
SVO_HD inline float pow2_rcp(float s) {
  const uint32_t b = __builtin_bit_cast(uint32_t, s);
  const uint32_t e = (b >> 23) & 0xFFu;
  if ((b & 0x807FFFFFu) != 0u || e == 0u || e >= 254u) return 0.0f;
  return __builtin_bit_cast(float, (254u - e) << 23);
}

// x / s, through the exact reciprocal when s is a power of two.
SVO_HD inline float div_by(float x, float s, float rcp) { return rcp != 0.0f ? x * rcp : x / s; }
// div_by for the three components of a vector with one test of the divisor
SVO_HD inline V3f div3_by(V3f v, float s, float rcp) {
  if (rcp != 0.0f) return V3f{v.x * rcp, v.y * rcp, v.z * rcp};
  return V3f{v.x / s, v.y / s, v.z / s};
}

// a - fmodf(a, b) (spatial/math/vector.rs:66-71 as used by the POP at
// raytracing_on_cpu.rs:452-456). For b a power of two >= 1 and 0 < a < 2^24:
// q = floor(a * 2^-k) is exact, q * b is exact, and the reference's
// a - (a - q b) is exactly q b, so floor(a / b) * b is the same value.
// (+-0) - fmodf(+-0, b) is +0.
SVO_HD inline float sub_fmod(float a, float b, float rcp_b) {
  if (a == 0.0f) return 0.0f;
  if (rcp_b != 0.0f && a > 0.0f && a < 16777216.0f && b >= 1.0f) return floorf(a * rcp_b) * b;
  return a - fmodf(a, b);
}

// Palette index value codec (node.rs:354-371)
constexpr uint32_t EMPTY_MARKER = 0xFFFFFFFFu;
constexpr uint32_t NONE16 = 0xFFFFu;
SVO_HD inline uint32_t pix_color_index(uint32_t v) { return v & 0xFFFFu; }
SVO_HD inline uint32_t pix_data_index(uint32_t v) { return (v >> 16) & 0xFFFFu; }

}  // namespace svo
