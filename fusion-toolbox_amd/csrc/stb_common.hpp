// Shared device/host helpers of libslamtb_hip.so (gfx950).
//
// Arithmetic contract shared with the CPU oracle: every function that feeds a
// correspondence decision (transform, normal matrix, projection, rounding,
// the distance and angle gates) is evaluated with FMA contraction disabled,
// IEEE-rounded division and square root, and length-3 sums in the order
// a0 + (a1 + a2). That makes correspondence maps and match counts bit-equal
// to the sequential CPU restatement of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/slamtb_hip.h"

namespace stb {

// ---------------------------------------------------------------------------
// Error reporting (host side). Entry points return STB_OK or an error code
// and leave a message for stb_last_error().
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int check_launch(const char* what);

#define STB_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      ::stb::set_error(__VA_ARGS__);  \
      return STB_ERR_INVALID;         \
    }                                 \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Host: the largest float c in [-1, 1] with fabsf(acosf(c)) >= thr, so that
// "acosf(c) >= thr" equals "-1 <= c <= c*" (acosf is monotone on [-1, 1] and
// NaN elsewhere, where the reference gate lets the pair through). Returns
// -2 when no c qualifies. Computed with the host libm, the same acosf the
// reference's CPU path calls.
float angle_gate_cos_f32(double angle_thresh);
double angle_gate_cos_f64(double angle_thresh);

// ---------------------------------------------------------------------------
// Device math
// ---------------------------------------------------------------------------

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
__device__ __host__ inline V3<T> sub3(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}

// a0*b0 + (a1*b1 + a2*b2): the order of Eigen's fixed-size Vector3 reduction.
template <typename T>
__device__ __host__ inline T dot3(const V3<T>& a, const V3<T>& b) {
#pragma clang fp contract(off)
  const T p0 = a.x * b.x;
  const T p1 = a.y * b.y;
  const T p2 = a.z * b.z;
  return p0 + (p1 + p2);
}

template <typename T>
__device__ __host__ inline V3<T> cross3(const V3<T>& a, const V3<T>& b) {
#pragma clang fp contract(off)
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x};
}

// Rigid transform [R | t] (row-major 3x4) and its inverse-transpose normal
// matrix from cofactors, as include/slamtb/camera.hpp:99-145 defines them.
template <typename T>
struct Rigid {
  T m[12];
  T nm[9];
};

template <typename T>
__device__ __host__ inline void make_normal_matrix(Rigid<T>& r) {
#pragma clang fp contract(off)
  const T* m = r.m;
#define A_(i, j) m[(i)*4 + (j)]
  const T det = +A_(0, 0) * (A_(1, 1) * A_(2, 2) - A_(2, 1) * A_(1, 2)) -
                A_(0, 1) * (A_(1, 0) * A_(2, 2) - A_(1, 2) * A_(2, 0)) +
                A_(0, 2) * (A_(1, 0) * A_(2, 1) - A_(1, 1) * A_(2, 0));
  const T inv = T(1) / det;
  r.nm[0] = (A_(1, 1) * A_(2, 2) - A_(2, 1) * A_(1, 2)) * inv;
  r.nm[3] = -(A_(0, 1) * A_(2, 2) - A_(0, 2) * A_(2, 1)) * inv;
  r.nm[6] = (A_(0, 1) * A_(1, 2) - A_(0, 2) * A_(1, 1)) * inv;
  r.nm[1] = -(A_(1, 0) * A_(2, 2) - A_(1, 2) * A_(2, 0)) * inv;
  r.nm[4] = (A_(0, 0) * A_(2, 2) - A_(0, 2) * A_(2, 0)) * inv;
  r.nm[7] = -(A_(0, 0) * A_(1, 2) - A_(1, 0) * A_(0, 2)) * inv;
  r.nm[2] = (A_(1, 0) * A_(2, 1) - A_(2, 0) * A_(1, 1)) * inv;
  r.nm[5] = -(A_(0, 0) * A_(2, 1) - A_(2, 0) * A_(0, 1)) * inv;
  r.nm[8] = (A_(0, 0) * A_(1, 1) - A_(1, 0) * A_(0, 1)) * inv;
#undef A_
}

template <typename T>
__device__ __host__ inline V3<T> rigid_point(const T* m, const V3<T>& p) {
#pragma clang fp contract(off)
  return V3<T>{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
               m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
               m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}

template <typename T>
__device__ __host__ inline V3<T> rigid_normal(const T* nm, const V3<T>& n) {
#pragma clang fp contract(off)
  return V3<T>{nm[0] * n.x + nm[1] * n.y + nm[2] * n.z,
               nm[3] * n.x + nm[4] * n.y + nm[5] * n.z,
               nm[6] * n.x + nm[7] * n.y + nm[8] * n.z};
}

// fx*x/z + cx (include/slamtb/camera.hpp:53-60).
template <typename T>
__device__ inline void project(T fx, T fy, T cx, T cy, const V3<T>& p, T& u,
                               T& v) {
#pragma clang fp contract(off)
  u = fx * p.x / p.z + cx;
  v = fy * p.y / p.z + cy;
}

// int(round(u)) with the bounds check folded in: round half away from zero,
// then accept only 0 <= r < limit. Non-finite u fails the comparison, which
// is how the reference's INT_MIN conversion fails its bounds check.
template <typename T>
__device__ inline bool round_in_range(T u, int limit, int& out) {
  const T r = round(u);
  if (!(r >= T(0) && r < T(limit))) return false;
  out = int(r);
  return true;
}

// round_in_range for FP32 without a branch or a rounding instruction:
// round(u) (half away from zero) lies in [0, limit) exactly when
// -0.5 < u < limit - 0.5 (NaN fails both), and for such u, with t the
// truncation of u and f = u - t (exact), round(u) = t + (f >= 0.5); for u in
// (-0.5, 0) that is t = 0, f < 0, giving 0 like the -0.0 that round()
// returns. `out` is set whatever the verdict and is meaningful only when it
// is true, so callers use it under a select.
__device__ inline bool round_px(float u, int limit, int& out) {
  const int t = int(u);  // one truncating conversion
  const float f = u - float(t);
  out = t + (f >= 0.5f ? 1 : 0);
  return u > -0.5f && u < float(limit) - 0.5f;
}

// a * b + c for 0 <= a, b < 2^24 whose product fits 32 bits (a full-rate
// 24-bit multiply-add instead of a quarter-rate 32-bit one): pixel offsets of
// levels whose arrays hold fewer than 2^31 bytes.
__device__ inline int mad24(int a, int b, int c) {
  return __umul24(uint32_t(a), uint32_t(b)) + c;
}

template <typename T>
__device__ inline int trunc_int(T x) {
  return int(x);  // callers keep x inside [-1, limit)
}

// ---------------------------------------------------------------------------
// Storage types: maps may be kept as float, bf16 or fp16; all math is FP32.
// ---------------------------------------------------------------------------
template <typename S>
__device__ inline float ldf(const S* p, int64_t i);
template <>
__device__ inline float ldf<float>(const float* p, int64_t i) {
  return p[i];
}
// 16-bit elements are loaded as raw bits and widened in registers: BF16 is
// the high half of an FP32, FP16 converts with one instruction. (Loading them
// through the library's struct types made the compiler wait on every load.)
__device__ inline float bf16_bits_to_float(uint16_t b) {
  return __uint_as_float(uint32_t(b) << 16);
}
template <>
__device__ inline float ldf<__hip_bfloat16>(const __hip_bfloat16* p, int64_t i) {
  return bf16_bits_to_float(reinterpret_cast<const uint16_t*>(p)[i]);
}
template <>
__device__ inline float ldf<__half>(const __half* p, int64_t i) {
  return float(reinterpret_cast<const _Float16*>(p)[i]);
}

template <typename S>
__device__ inline V3<float> ld3f(const S* p, int64_t i) {
  return V3<float>{ldf(p, 3 * i), ldf(p, 3 * i + 1), ldf(p, 3 * i + 2)};
}
template <>
__device__ inline V3<float> ld3f<float>(const float* p, int64_t i) {
  return V3<float>{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

template <typename T>
__device__ inline V3<T> ld3(const T* p, int64_t i) {
  return V3<T>{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

// Load one element of any storage type, widened to its math type.
__device__ inline float ldw(const float* p, int64_t i) { return p[i]; }
__device__ inline double ldw(const double* p, int64_t i) { return p[i]; }
__device__ inline float ldw(const __hip_bfloat16* p, int64_t i) {
  return ldf(p, i);
}
__device__ inline float ldw(const __half* p, int64_t i) { return ldf(p, i); }

// ---------------------------------------------------------------------------
// Correspondence keys.
//
// Reference layout (kept by the drop-in entry point): (index << 32) | bits(d).
// Internal layout of the fused path, from the top bit down:
//   tag (8 bits) | bits(d) (31 bits) | index (25 bits),
// so one 64-bit unsigned atomic min yields, among keys of one tag, the lowest
// distance and among equal distances the lowest source index -- the order of
// the sequential CPU path (distances are non-negative floats, whose bit
// patterns order like their values). A Gauss-Newton loop gives each
// iteration a tag one lower than the last: every key it writes is below
// every key an earlier iteration left, and an owner must match the tag too,
// so the map needs no clearing between iterations (only when the tag runs
// out, every 256 iterations). Source indices must stay below 2^25 - 1.
// ---------------------------------------------------------------------------
constexpr uint64_t kRefSentinel = 0x0fffffff7f800000ull;
constexpr uint64_t kKeySentinel = ~0ull;  // above every key
constexpr int kKeyIdxBits = 25;
constexpr uint64_t kKeyIdxMask = (uint64_t(1) << kKeyIdxBits) - 1;
constexpr int64_t kMaxKeyIndex = int64_t(kKeyIdxMask) - 1;
constexpr uint32_t kFreshTag = 255;  // the tag of a freshly cleared map

__device__ inline uint64_t make_key(float dist, uint32_t idx, uint32_t tag) {
  return (uint64_t(tag) << 56) |
         (uint64_t(__float_as_uint(dist)) << kKeyIdxBits) | uint64_t(idx);
}

__device__ inline bool key_owned(uint64_t key, uint32_t idx, uint32_t tag) {
  return (key & kKeyIdxMask) == idx && uint32_t(key >> 56) == tag;
}

// The tag and map of the iteration after one that used (tag, sel): one tag
// lower on the same map, or, after tag 0, a fresh tag on the other map,
// which that iteration's Jacobian launch swept.
struct MapUse {
  uint32_t tag;
  int sel;
};
__host__ __device__ inline MapUse next_map_use(uint32_t tag, int sel) {
  return tag == 0 ? MapUse{kFreshTag, sel ^ 1} : MapUse{tag - 1, sel};
}

constexpr int kBlock = 256;

// MI355X hands the workgroups of a launch to its 8 XCDs in turn (workgroup i
// to XCD i % 8), and each XCD has an L2 of its own. xcd_block(i, g) numbers
// the g workgroups so that each XCD's share is one contiguous run of work
// items: neighbouring pixels, whose gathers share target cache lines, then
// meet in one L2 instead of being fetched by two. A permutation of [0, g):
// XCD x takes the q or q + 1 consecutive numbers after those of XCDs < x.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int i, int g) {
  const int x = i % kXcds, j = i / kXcds;
  const int q = g / kXcds, r = g % kXcds;
  return x * q + (x < r ? x : r) + j;
}

// The same in chunks of c consecutive work items that rotate over the XCDs
// (chunk k on XCD k % 8): neighbours still share an L2, and work items of
// unequal cost spread over every XCD. Workgroups past the last whole round
// of 8 chunks keep their own numbers, so the map is a permutation of [0, g).
__device__ inline int xcd_chunk(int i, int g, int c) {
  const int full = g / (kXcds * c) * (kXcds * c);
  if (i >= full) return i;
  const int x = i % kXcds, j = i / kXcds;
  return ((j / c) * kXcds + x) * c + j % c;
}

inline int grid_for(int64_t n, int block = kBlock, int max_blocks = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return int(g);
}

}  // namespace stb
