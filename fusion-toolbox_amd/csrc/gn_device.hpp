// Device building blocks of the fused Gauss-Newton path, shared by the
// single-pair kernels (gn.hip) and the batched multi-pair kernels
// (batch.hip):
//   * per-pixel correspondence, Jacobian and information terms over a level
//     descriptor (any map storage type, FP32 math);
//   * the slot reduction: 21 unique JtJ + 6 Jr + r2/count bookkeeping per
//     lane in FP32, recursive halving across the wavefront (permlane swaps
//     and DPP), FP64 across waves through LDS, one FP32 partial per block
//     (FP64 for the adjoint's systems), and a fixed-order FP64 sum of the
//     partials in a following launch (by one block, or by every block of a
//     loop's next correspondence launch);
//   * the per-pair state update: FP64 Cholesky solve and exp_rt_to_matrix,
//     in one thread or redundantly in every lane of one wavefront.
#pragma once

#include "icp_math.hpp"

#include <type_traits>

namespace stb {

constexpr int kSlots = 32;  // 21 JtJ + 6 Jr + r2_geom, r2_feat, n_geom, n_feat, 0
constexpr int kWaves = kBlock / 64;
constexpr int kGroups = kBlock / kSlots;  // slot groups of the final sum
constexpr int kShared = (kGroups > kWaves ? kGroups : kWaves) * kSlots;
// Block size of the launches that sum per-block partials and update poses:
// 32 slot groups share the partial loads.
constexpr int kSolveBlock = 1024;

// Weights of the combined system of slamtb/registration/icp.py:221-234:
// JtJ = gw^2 JtJ_g + fw^2 JtJ_f and Jr = gw Jr_g - fw Jr_f. The per-pixel
// sums leave the weights out: with b = gw (or fw when the geometric term is
// off) and rho = fw / b, a lane accumulates J_g J_g^T + (rho J_f)(rho J_f)^T
// and J_g r_g - (rho J_f) r_f, the feature J picking up rho through the
// divisor of its finite differences (fdiv = rho / (2h)), and the block
// partials scale the JtJ slots by b^2 and the Jr slots by b in FP64. That
// saves the per-pixel J scaling (12 multiplies).
struct TermWeights {
  float fdiv;     // feature finite-difference divisor times rho
  int gon;        // geometric J enters the system (gw != 0)
  float fthr2;    // feature residual threshold squared
  double b;       // weight the partials restore: JtJ slots x b^2, Jr x b
  double gw, fw;  // weights of the residual bookkeeping
};

// The feature finite-difference step h (src/feature_map.hpp:82); the
// derivative divides by 2h as src/feature_map.hpp:57 forms it in FP32.
constexpr float kFeatStep = 0.05f;

inline TermWeights make_weights(const stb_icp_params& P) {
  TermWeights tw;
  const bool g = P.use_geom && P.geom_weight != 0.0;
  tw.b = g ? P.geom_weight : P.feat_weight;
  const double rho = tw.b != 0.0 ? P.feat_weight / tw.b : 0.0;
  tw.fdiv = float(rho * double(1.0f / (2.0f * kFeatStep)));
  tw.gon = g ? 1 : 0;
  tw.fthr2 = P.feat_residual_thresh * P.feat_residual_thresh;
  tw.gw = P.geom_weight;
  tw.fw = P.feat_weight;
  return tw;
}

// ---------------------------------------------------------------------------
// Per-pixel terms
// ---------------------------------------------------------------------------

template <typename S>
struct LevelView {
  const S* __restrict__ sp;
  const S* __restrict__ sn;
  const S* __restrict__ tp;
  const S* __restrict__ tn;
  const S* __restrict__ sf;
  const S* __restrict__ tf;
  const uint8_t* __restrict__ sm;
  const uint8_t* __restrict__ tm;
  int64_t n, HW;
  int H, W, C;
  float fx, fy, cx, cy;
  int il;  // features interleaved ([n, C] / [H, W, C])
};

template <typename S>
__device__ inline LevelView<S> view_of(const stb_icp_level& L) {
  LevelView<S> v;
  v.sp = static_cast<const S*>(L.src_points);
  v.sn = static_cast<const S*>(L.src_normals);
  v.tp = static_cast<const S*>(L.tgt_points);
  v.tn = static_cast<const S*>(L.tgt_normals);
  v.sf = static_cast<const S*>(L.src_feats);
  v.tf = static_cast<const S*>(L.tgt_feats);
  v.sm = L.src_mask;
  v.tm = L.tgt_mask;
  v.n = L.n_src;
  v.H = L.height;
  v.W = L.width;
  v.HW = int64_t(L.height) * L.width;
  v.C = L.channels;
  v.fx = L.fx;
  v.fy = L.fy;
  v.cx = L.cx;
  v.cy = L.cy;
  v.il = L.feat_layout == STB_FEAT_INTERLEAVED;
  return v;
}

// ---------------------------------------------------------------------------
// Gathers through buffer resources
// ---------------------------------------------------------------------------
//
// The fused correspondence and Jacobian passes read their inputs through
// raw buffer resources (points, normals and masks of both frames, each
// channel of the [C, H, W] target features, the [C, n] source features, the
// map): a load is a 32-bit byte offset in a vector register (small constant
// steps as immediates) instead of a 64-bit address of its own, and a load
// whose vector offset lies past the end of the resource returns 0 instead of
// faulting. The scalar offset operand is not range-checked by the hardware,
// so every load here passes 0 there and puts the whole offset in the vector
// operand. The bilinear taps therefore need no clamping: each target channel
// has a resource of its own, a tap past the right edge reads the next row's
// first texel (which bil_mix_res replaces by zero, as bil_mix replaces the
// clamped re-reads of bil_taps) and one past the bottom row returns 0.
// Every array must hold fewer than 2^31 bytes (checked on the host).
constexpr int kRawBufferWord3 = 0x00020000;  // raw 32-bit data format, CDNA

// Whether a level's arrays fit those 32-bit byte offsets.
inline bool level_addressable(const stb_icp_level& L) {
  const int64_t es = L.storage == STB_F32 ? 4 : 2;
  const int64_t lim = (int64_t(1) << 31) - 1;
  const int64_t hw = int64_t(L.height) * L.width;
  return int64_t(L.channels) * hw * es <= lim &&
         int64_t(L.channels) * L.n_src * es <= lim && 3 * hw * es <= lim &&
         3 * L.n_src * es <= lim && 8 * hw <= lim;
}

template <typename S>
__device__ inline __amdgpu_buffer_rsrc_t make_res(const S* p, int64_t elems) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<S*>(p), short(0),
                                           int(elems * int64_t(sizeof(S))),
                                           kRawBufferWord3);
}

template <typename S>
__device__ inline float res_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ inline float res_ld<float>(__amdgpu_buffer_rsrc_t r, int voff,
                                      int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ inline float res_ld<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r,
                                               int voff, int soff) {
  return bf16_bits_to_float(__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
template <>
__device__ inline float res_ld<__half>(__amdgpu_buffer_rsrc_t r, int voff,
                                       int soff) {
  const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
  return float(__builtin_bit_cast(_Float16, b));
}

__device__ inline uint8_t res_u8(__amdgpu_buffer_rsrc_t r, int voff) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, voff, 0, 0);
}

// Element i of an [m, 3] array.
template <typename S>
__device__ inline V3<float> res_ld3(__amdgpu_buffer_rsrc_t r, int i) {
  constexpr int es = int(sizeof(S));
  // 3 * es * i as a shift and a shift-add: an opaque copy of 2i keeps the
  // compiler from folding it back into a quarter-rate 32-bit multiply.
  int i2 = i << 1;
  asm("" : "+v"(i2));
  const int o = (i2 + i) * es;
  return V3<float>{res_ld<S>(r, o, 0), res_ld<S>(r, o + es, 0),
                   res_ld<S>(r, o + 2 * es, 0)};
}

// Buffer resources over a level's frame arrays.
struct LevelRes {
  __amdgpu_buffer_rsrc_t sp, sn, sm, tp, tn, tm;
};

template <typename S>
__device__ inline LevelRes res_of(const LevelView<S>& L) {
  LevelRes r;
  r.sp = make_res(L.sp, 3 * L.n);
  r.sn = make_res(L.sn, 3 * L.n);
  r.sm = make_res(L.sm, L.n);
  r.tp = make_res(L.tp, 3 * L.HW);
  r.tn = make_res(L.tn, 3 * L.HW);
  r.tm = make_res(L.tm, L.HW);
  return r;
}

// Correspondence of source pixel idx (src/correspondence_map.cu:38-61 with
// RobustCorrespondence::Match): transform, project, round, bounds, target
// mask, distance and normal-angle gates, then a 64-bit unsigned atomic min of
// (bits(distance) << 32 | idx) into the map. Each test is written after the
// loads it depends on and reads their results, so the compiler cannot sink a
// load below an earlier test: the source loads issue together, then the
// target loads, and a pixel costs two memory round trips.
// corresp_pixel from the pixel's already loaded source mask, point and
// normal.
template <typename S>
__device__ inline void corresp_pixel_at(const LevelView<S>& L,
                                        const LevelRes& R, const float* m,
                                        const float* nm, int64_t idx,
                                        uint8_t valid_src, V3<float> ps,
                                        V3<float> ns,
                                        unsigned long long* __restrict__ map,
                                        uint32_t tag, float dthr2, float cstar) {
#pragma clang fp contract(off)
  const V3<float> p = rigid_point(m, ps);
  const V3<float> nn = rigid_normal(nm, ns);
  float u, v;
  project(L.fx, L.fy, L.cx, L.cy, p, u, v);
  int ui, vi;
  const bool in_u = round_px(u, L.W, ui);
  const bool in_v = round_px(v, L.H, vi);
  if (!valid_src || !in_u || !in_v) return;
  const int pos = mad24(vi, L.W, ui);
  const uint8_t valid = res_u8(R.tm, pos);
  const V3<float> q = res_ld3<S>(R.tp, pos);
  const V3<float> qn = res_ld3<S>(R.tn, pos);
  const V3<float> d = sub3(q, p);
  const float sq = dot3(d, d);
  const bool keep = (valid != 0) & !(sq > dthr2) &
                    !angle_rejects(gate_cos(nn, qn), cstar);
  if (!keep) return;
  atomicMin(map + pos,
            (unsigned long long)make_key(sqrtf(sq), uint32_t(idx), tag));
}

// corresp_pixel_at for 16-bit storage. Coordinates rounded to 16 bits put
// runs of neighbouring source pixels on one target pixel (BF16 keeps 8
// mantissa bits: ~1.6 cm at 2 m), so lanes of a wavefront would issue
// atomics to one address, which the memory pipeline serialises. Here the
// lanes first take the minimum key over each run of consecutive active
// lanes with the same target pixel (a segmented inclusive scan, six
// steps), and only the last lane of each run issues the atomic: the same
// minimum reaches the map. Every lane that reaches the call takes part;
// lanes outside the exec mask are never read.
template <typename S>
__device__ inline void corresp_pixel_at_runs(
    const LevelView<S>& L, const LevelRes& R, const float* m, const float* nm,
    int64_t idx, uint8_t valid_src, V3<float> ps, V3<float> ns,
    unsigned long long* __restrict__ map, uint32_t tag, float dthr2,
    float cstar) {
#pragma clang fp contract(off)
  const V3<float> p = rigid_point(m, ps);
  const V3<float> nn = rigid_normal(nm, ns);
  float u, v;
  project(L.fx, L.fy, L.cx, L.cy, p, u, v);
  int ui, vi;
  const bool in_u = round_px(u, L.W, ui);
  const bool in_v = round_px(v, L.H, vi);
  const bool inside = valid_src && in_u && in_v;
  const int pos = inside ? mad24(vi, L.W, ui) : 0;
  const uint8_t valid = res_u8(R.tm, pos);
  const V3<float> q = res_ld3<S>(R.tp, pos);
  const V3<float> qn = res_ld3<S>(R.tn, pos);
  const V3<float> d = sub3(q, p);
  const float sq = dot3(d, d);
  const bool keep = inside && (valid != 0) && !(sq > dthr2) &&
                    !angle_rejects(gate_cos(nn, qn), cstar);
  const int lane = threadIdx.x & 63;
  // A lane that keeps nothing gets a target of its own (-1 - lane), so it
  // joins no run.
  const int tgt = keep ? pos : -1 - lane;
  unsigned long long key =
      keep ? (unsigned long long)make_key(sqrtf(sq), uint32_t(idx), tag)
           : ~0ull;
  const unsigned long long exec = __builtin_amdgcn_read_exec();
#pragma unroll
  for (int dl = 1; dl < 64; dl <<= 1) {
    const unsigned long long ok = lane >= dl ? (exec >> (lane - dl)) & 1ull : 0ull;
    const unsigned long long k2 = __shfl_up(key, dl, 64);
    const int t2 = __shfl_up(tgt, dl, 64);
    if (ok && t2 == tgt && k2 < key) key = k2;
  }
  const int tn = __shfl_down(tgt, 1, 64);
  const bool next_ok = lane < 63 && ((exec >> (lane + 1)) & 1ull);
  const bool last = !(next_ok && tn == tgt);
  if (keep && last) atomicMin(map + pos, key);
}

// The per-storage choice: runs merged in the wavefront for 16-bit storage.
template <typename S>
__device__ inline void corresp_pixel_store(
    const LevelView<S>& L, const LevelRes& R, const float* m, const float* nm,
    int64_t idx, uint8_t valid_src, V3<float> ps, V3<float> ns,
    unsigned long long* __restrict__ map, uint32_t tag, float dthr2,
    float cstar) {
  if constexpr (sizeof(S) == 2)
    corresp_pixel_at_runs(L, R, m, nm, idx, valid_src, ps, ns, map, tag, dthr2,
                          cstar);
  else
    corresp_pixel_at(L, R, m, nm, idx, valid_src, ps, ns, map, tag, dthr2,
                     cstar);
}

template <typename S>
__device__ inline void corresp_pixel(const LevelView<S>& L, const LevelRes& R,
                                     const float* m, const float* nm,
                                     int64_t idx,
                                     unsigned long long* __restrict__ map,
                                     uint32_t tag, float dthr2, float cstar) {
  const uint8_t valid_src = res_u8(R.sm, int(idx));
  const V3<float> ps = res_ld3<S>(R.sp, int(idx));
  const V3<float> ns = res_ld3<S>(R.sn, int(idx));
  corresp_pixel_store(L, R, m, nm, idx, valid_src, ps, ns, map, tag, dthr2,
                      cstar);
}

// Adds one term J, r to the JtJ/Jr slots (the weights enter with the block
// partials). BORN: the slots start from this term (the first term of a
// lane's first pixel), so they need no zeroing and no add to 0.
template <bool BORN>
__device__ inline void acc_term(float (&acc)[kSlots], const float j[6], float r) {
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) {
      acc[k] = BORN ? j[a] * j[b] : fmaf(j[a], j[b], acc[k]);
      ++k;
    }
#pragma unroll
  for (int a = 0; a < 6; ++a)
    acc[21 + a] = BORN ? j[a] * r : fmaf(j[a], r, acc[21 + a]);
}

// acc[k] + x, or x for a slot that starts here.
template <bool BORN>
__device__ inline float acc_add(float a, float x) {
  return BORN ? x : a + x;
}

// The four bilinear taps of one sample from the resource of one [H, W]
// channel.
template <typename S>
__device__ inline void res_taps(__amdgpu_buffer_rsrc_t r, int W,
                                const Bil<float>& b, float t[4]) {
  constexpr int es = int(sizeof(S));
  const int o = mad24(b.vi, W, b.ui) * es;
  const int d = o + W * es;
  t[0] = res_ld<S>(r, o, 0);
  t[1] = res_ld<S>(r, o + es, 0);
  t[2] = res_ld<S>(r, d, 0);
  t[3] = res_ld<S>(r, d + es, 0);
}

// 16-bit storage: the low and high halves of a 32-bit word widened to FP32.
template <typename S>
__device__ inline float widen_lo(uint32_t x);
template <typename S>
__device__ inline float widen_hi(uint32_t x);
template <>
__device__ inline float widen_lo<__hip_bfloat16>(uint32_t x) {
  return __uint_as_float(x << 16);
}
template <>
__device__ inline float widen_hi<__hip_bfloat16>(uint32_t x) {
  return __uint_as_float(x & 0xffff0000u);
}
template <>
__device__ inline float widen_lo<__half>(uint32_t x) {
  return float(__builtin_bit_cast(_Float16, uint16_t(x)));
}
template <>
__device__ inline float widen_hi<__half>(uint32_t x) {
  return float(__builtin_bit_cast(_Float16, uint16_t(x >> 16)));
}

// Two horizontally adjacent 16-bit texels at byte offset o from one 8-byte
// load at the 4-byte aligned offset at or below o, clamped to `last8` (the
// channel's bytes - 8, 4-aligned) so that the load ends inside the channel:
// one load instruction instead of two. A texel pair past the channel's end
// (the row below the last) reads shifted data that bil_mix discards.
template <typename S>
__device__ inline void res_pair16(__amdgpu_buffer_rsrc_t r, int o, int last8,
                                  float& a, float& b) {
  const int base = min(o & ~3, last8);
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, base, 0, 0);
  const int sh = min(o - base, 6) * 8;
  const uint32_t lo =
      uint32_t(((uint64_t(w[1]) << 32) | uint64_t(w[0])) >> sh);
  a = widen_lo<S>(lo);
  b = widen_hi<S>(lo);
}

// res_taps for 16-bit storage with the row pairs loaded together; the
// channel must hold an even number of texels (4-byte aligned size), at
// least four.
template <typename S>
__device__ inline void res_taps_pair(__amdgpu_buffer_rsrc_t r, int W, int last8,
                                     const Bil<float>& b, float t[4]) {
  const int o = mad24(b.vi, W, b.ui) * 2;
  res_pair16<S>(r, o, last8, t[0], t[1]);
  res_pair16<S>(r, o + W * 2, last8, t[2], t[3]);
}

// The taps through whichever form the storage and the channel size allow
// (`pairs` is uniform across the launch).
template <typename S>
__device__ inline void res_taps_any(__amdgpu_buffer_rsrc_t r, int W, bool pairs,
                                    int last8, const Bil<float>& b, float t[4]) {
  if constexpr (std::is_same<S, __hip_bfloat16>::value) {
    if (pairs) {
      res_taps_pair<S>(r, W, last8, b, t);
      return;
    }
  }
  res_taps<S>(r, W, b, t);
}

// FP32 taps with wide loads. Buffer loads check the range per dword (a
// dword past the resource's end reads 0, the dwords before it are read),
// so a 2- or 4-texel load near the end of a channel keeps bil_mix_res's
// zero padding.
__device__ inline void res_ld_f2(__amdgpu_buffer_rsrc_t r, int o, float& a,
                                 float& b) {
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, o, 0, 0);
  a = __uint_as_float(w[0]);
  b = __uint_as_float(w[1]);
}

// res_taps<float> with each row's two taps in one 8-byte load.
__device__ inline void res_taps_f32(__amdgpu_buffer_rsrc_t r, int W,
                                    const Bil<float>& b, float t[4]) {
  const int o = mad24(b.vi, W, b.ui) * 4;
  res_ld_f2(r, o, t[0], t[1]);
  res_ld_f2(r, o + W * 4, t[2], t[3]);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// The taps of the last channel's five samples (c0 at (u, v), bu0/bu1 at
// u +- h, bv0/bv1 at v +- h) from four loads instead of ten. The u samples'
// cells lie one column left of, at, or one column right of c0's, so two
// 16-byte loads of rows fv and fv + 1 from column fu - 1 (from column 0
// when fu = 0, where no sample reaches left) cover them; the v samples'
// cells one row above or below, so rows fv - 1 and fv + 2 at columns fu,
// fu + 1 add two 8-byte loads (read only when a sample's cell moved there;
// past the channel's start or end they read 0). Each sample then selects
// its taps from registers.
// The raw words of the last channel's four row windows (res_taps5_f32).
struct Taps5Raw {
  u32x4 a, b;   // rows fv, fv + 1 from column fu - 1 (0 when fu = 0)
  u32x2 up, dn; // rows fv - 1, fv + 2 at columns fu, fu + 1
};

__device__ inline Taps5Raw taps5_load(__amdgpu_buffer_rsrc_t r, int W,
                                      const Bil<float>& c0) {
  const int o = mad24(c0.vi, W, c0.ui) * 4;  // byte offset of (fu, fv)
  const int ow = c0.ui >= 1 ? o - 4 : o;     // the 4-texel window's first column
  Taps5Raw w;
  w.a = __builtin_amdgcn_raw_buffer_load_b128(r, ow, 0, 0);
  w.b = __builtin_amdgcn_raw_buffer_load_b128(r, ow + W * 4, 0, 0);
  w.up = __builtin_amdgcn_raw_buffer_load_b64(r, o - W * 4, 0, 0);
  w.dn = __builtin_amdgcn_raw_buffer_load_b64(r, o + W * 8, 0, 0);
  return w;
}

// Each of the five samples' taps, selected from the windows.
__device__ inline void taps5_select(const Taps5Raw& w, const Bil<float>& c0,
                                    const Bil<float>& bu0, const Bil<float>& bu1,
                                    const Bil<float>& bv0, const Bil<float>& bv1,
                                    float t0[4], float t1[4], float t2[4],
                                    float t3[4], float t4[4]) {
  const bool left = c0.ui >= 1;
  const float up0 = __uint_as_float(w.up[0]), up1 = __uint_as_float(w.up[1]);
  const float dn0 = __uint_as_float(w.dn[0]), dn1 = __uint_as_float(w.dn[1]);
  // Rows fv (A) and fv + 1 (B) at columns fu - 1 .. fu + 2.
  float A[4], B[4];
  A[0] = __uint_as_float(w.a[0]);
  B[0] = __uint_as_float(w.b[0]);
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    A[k] = __uint_as_float(left ? w.a[k] : w.a[k - 1]);
    B[k] = __uint_as_float(left ? w.b[k] : w.b[k - 1]);
  }
  t0[0] = A[1];
  t0[1] = A[2];
  t0[2] = B[1];
  t0[3] = B[2];
  const bool ru = bu0.ui != c0.ui;  // bu0 one column right
  t1[0] = ru ? A[2] : A[1];
  t1[1] = ru ? A[3] : A[2];
  t1[2] = ru ? B[2] : B[1];
  t1[3] = ru ? B[3] : B[2];
  const bool lu = bu1.ui != c0.ui;  // bu1 one column left
  t2[0] = lu ? A[0] : A[1];
  t2[1] = lu ? A[1] : A[2];
  t2[2] = lu ? B[0] : B[1];
  t2[3] = lu ? B[1] : B[2];
  const bool dv = bv0.vi != c0.vi;  // bv0 one row down
  t3[0] = dv ? B[1] : A[1];
  t3[1] = dv ? B[2] : A[2];
  t3[2] = dv ? dn0 : B[1];
  t3[3] = dv ? dn1 : B[2];
  const bool uv = bv1.vi != c0.vi;  // bv1 one row up
  t4[0] = uv ? up0 : A[1];
  t4[1] = uv ? up1 : A[2];
  t4[2] = uv ? A[1] : B[1];
  t4[3] = uv ? A[2] : B[2];
}

__device__ inline void res_taps5_f32(__amdgpu_buffer_rsrc_t r, int W,
                                     const Bil<float>& c0, const Bil<float>& bu0,
                                     const Bil<float>& bu1, const Bil<float>& bv0,
                                     const Bil<float>& bv1, float t0[4],
                                     float t1[4], float t2[4], float t3[4],
                                     float t4[4]) {
  taps5_select(taps5_load(r, W, c0), c0, bu0, bu1, bv0, bv1, t0, t1, t2, t3, t4);
}

// bil_mix for taps read through a resource over one channel: the taps of
// the row below the last lie past the resource's end and read 0, which is
// the value bil_mix selects for them, so only the right-edge select remains
// (whose tap, the next row's first texel, is inside the channel).
__device__ inline float bil_mix_res(int W, const Bil<float>& b, const float t[4]) {
#pragma clang fp contract(off)
  const bool rx = b.ui + 1 < W;
  const float v00 = t[0];
  const float v10 = rx ? t[1] : 0.0f;
  const float v01 = t[2];
  const float v11 = rx ? t[3] : 0.0f;
  const float a = v00 * (1.0f - b.ur) + v10 * b.ur;
  const float c = v01 * (1.0f - b.ur) + v11 * b.ur;
  return a * (1.0f - b.vr) + c * b.vr;
}

// The feature term's J from the derivatives du, dv of the last channel at
// the projection of p (src/icp_jacobian_feature.cu:81-106,
// include/slamtb/camera.hpp:65-78): feature_terms<..., FAST = true>'s
// arithmetic, one hardware reciprocal of z (about 1 ulp) for the four
// divides, no contraction.
__device__ inline void feature_j(float fx, float fy, const V3<float>& p,
                                 float du, float dv, float j[6]) {
#pragma clang fp contract(off)
  const float iz = __builtin_amdgcn_rcpf(p.z);
  const float iz2 = iz * iz;
  const float j00 = fx * iz;
  const float j02 = -p.x * fx * iz2;
  const float j11 = fy * iz;
  const float j12 = -p.y * fy * iz2;
  const V3<float> g{du * j00, dv * j11, du * j02 + dv * j12};
  const V3<float> c = cross3(p, g);
  j[0] = g.x;
  j[1] = g.y;
  j[2] = g.z;
  j[3] = c.x;
  j[4] = c.y;
  j[5] = c.z;
}

// feature_terms<float, S, S, true, NB> (icp_math.hpp) with its taps and
// source values read through buffer resources, the same samples in the same
// arithmetic. (The finite differences subtract samples 0.1 px apart, so a
// last-bit change in a sample moves the derivative by ~1e-5 of its value:
// the samples keep the reference's operation order.) The derivative divisor
// carries the feature weight ratio (TermWeights::fdiv), so J comes out as
// rho J.
template <typename S, int NB>
__device__ inline float feature_terms_res(const LevelView<S>& L, int64_t idx,
                                          const V3<float>& p, float u, float v,
                                          float div, float j[6], float& rlast) {
#pragma clang fp contract(off)
  constexpr int es = int(sizeof(S));
  // The source value of channel c is element c * n + idx of one resource
  // over all channels (idx < n, so the offset stays inside the array).
  const __amdgpu_buffer_rsrc_t rs = make_res(L.sf, int64_t(L.C) * L.n);
  const int sch = int(L.n) * es;  // bytes per source channel
  const int so = int(idx) * es;
  const int H = L.H, W = L.W, C = L.C;
  const float h = kFeatStep;
  const Bil<float> c0 = make_bil(u, v);
  const Bil<float> bu0 = make_bil(u + h, v);
  const Bil<float> bu1 = make_bil(fmax(u - h, 0.0f), v);
  const Bil<float> bv0 = make_bil(u, v + h);
  const Bil<float> bv1 = make_bil(u, fmax(v - h, 0.0f));
  // The last channel's five samples first (see feature_terms).
  const __amdgpu_buffer_rsrc_t rl = make_res(L.tf + int64_t(C - 1) * L.HW, L.HW);
  // BF16 texels load in row pairs when the channel's size is a multiple of
  // 4 bytes (an even number of texels). BF16 coordinates (8 mantissa bits)
  // scatter the projections of neighbouring pixels by a pixel or two, so a
  // wavefront's taps touch more cache lines per load instruction and the
  // loads are bound by the texture addresser; halving the instructions
  // pays for the extraction (1280x720, 32 channels: 96.5 to 76.9 us per
  // Jacobian launch). FP16 keeps one load per texel: its projections are
  // not scattered, the loads are not the bound, and the extraction cost
  // more than it saved (60.8 to 74.2 us).
  const bool pairs = std::is_same<S, __hip_bfloat16>::value && (L.HW & 1) == 0 &&
                     L.HW >= 4;
  const int last8 = int(L.HW) * 2 - 8;
  float t0[4], t1[4], t2[4], t3[4], t4[4];
  if constexpr (std::is_same<S, float>::value) {
    res_taps5_f32(rl, W, c0, bu0, bu1, bv0, bv1, t0, t1, t2, t3, t4);
  } else {
    res_taps_any<S>(rl, W, pairs, last8, c0, t0);
    res_taps_any<S>(rl, W, pairs, last8, bu0, t1);
    res_taps_any<S>(rl, W, pairs, last8, bu1, t2);
    res_taps_any<S>(rl, W, pairs, last8, bv0, t3);
    res_taps_any<S>(rl, W, pairs, last8, bv1, t4);
  }
  const float flast = res_ld<S>(rs, so + (C - 1) * sch, 0);
  // Row pairs past the channel's end read shifted data, not 0: BF16 keeps
  // both selects.
  auto mix = [&](const Bil<float>& b, const float t[4]) {
    if constexpr (std::is_same<S, __hip_bfloat16>::value) return bil_mix(H, W, b, t);
    else return bil_mix_res(W, b, t);
  };
  const float last = mix(c0, t0);
  const float su0 = mix(bu0, t1), su1 = mix(bu1, t2);
  const float sv0 = mix(bv0, t3), sv1 = mix(bv1, t4);
  float sq = 0.0f;
  const int last_c = C - 2 > 0 ? C - 2 : 0;
  for (int c0i = 0; c0i < C - 1; c0i += NB) {
    float taps[NB][4], fs[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int c = c0i + k < last_c ? c0i + k : last_c;
      const __amdgpu_buffer_rsrc_t rc = make_res(L.tf + int64_t(c) * L.HW, L.HW);
      if constexpr (std::is_same<S, float>::value)
        res_taps_f32(rc, W, c0, taps[k]);
      else
        res_taps_any<S>(rc, W, pairs, last8, c0, taps[k]);
      fs[k] = res_ld<S>(rs, so + c * sch, 0);
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const float res = mix(c0, taps[k]) - fs[k];
      sq += (c0i + k < C - 1) ? res * res : 0.0f;
    }
  }
  const float res = last - flast;
  feature_j(L.fx, L.fy, p, (su0 - su1) * div, (sv0 - sv1) * div, j);
  rlast = res;
  sq += res * res;
  return sq;
}

// Interleaved features are read 16 bytes per texel at a time: a chunk of
// CH = 16 / sizeof(S) consecutive channels (4 in FP32, 8 in 16-bit
// storage) from one load at byte offset o (a multiple of 4), widened to
// FP32.
template <typename S>
constexpr int kChunkCh = 16 / int(sizeof(S));

// Channel q of a chunk held as four raw dwords (16-bit channels widened
// when read, so a chunk costs 4 registers per texel whatever the storage).
template <typename S>
__device__ inline float chunk_ch(const uint32_t (&w)[4], int q) {
  if constexpr (std::is_same<S, float>::value)
    return __uint_as_float(w[q]);
  else
    return (q & 1) ? widen_hi<S>(w[q >> 1]) : widen_lo<S>(w[q >> 1]);
}

// chunk_ch for a channel q known only at run time (uniform): a select
// chain over the four dwords, then the half.
template <typename S>
__device__ inline float chunk_ch_at(const uint32_t (&w)[4], int q) {
  const int d = std::is_same<S, float>::value ? q : (q >> 1);
  const uint32_t x = d == 0 ? w[0] : d == 1 ? w[1] : d == 2 ? w[2] : w[3];
  if constexpr (std::is_same<S, float>::value)
    return __uint_as_float(x);
  else
    return (q & 1) ? widen_hi<S>(x) : widen_lo<S>(x);
}

__device__ inline void res_ld16(__amdgpu_buffer_rsrc_t r, int o,
                                uint32_t (&w)[4]) {
  const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0);
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = a[k];
}

// The c0 taps and source values of one chunk of channels, raw.
struct ChunkIL {
  uint32_t t[4][4];  // texels (fu, fv), (fu + 1, fv), (fu, fv + 1), (fu + 1, fv + 1)
  uint32_t s[4];     // source pixel
};

__device__ inline void load_chunk(__amdgpu_buffer_rsrc_t rt,
                                  __amdgpu_buffer_rsrc_t rs, int o00, int o01,
                                  int tb, int so, int co, ChunkIL& k) {
  res_ld16(rt, o00 + co, k.t[0]);
  res_ld16(rt, o00 + tb + co, k.t[1]);
  res_ld16(rt, o01 + co, k.t[2]);
  res_ld16(rt, o01 + tb + co, k.t[3]);
  res_ld16(rs, so + co, k.s);
}

// Adds the chunk's channels cb .. (those below C - 1) of the squared
// residual at sample c0, in channel order.
template <typename S>
__device__ inline void add_chunk(const ChunkIL& k, const Bil<float>& c0, int W,
                                 int cb, int C, float& sq) {
#pragma clang fp contract(off)
#pragma unroll
  for (int q = 0; q < kChunkCh<S>; ++q) {
    const float t[4] = {chunk_ch<S>(k.t[0], q), chunk_ch<S>(k.t[1], q),
                        chunk_ch<S>(k.t[2], q), chunk_ch<S>(k.t[3], q)};
    const float res = bil_mix_res(W, c0, t) - chunk_ch<S>(k.s, q);
    sq += (cb + q < C - 1) ? res * res : 0.0f;
  }
}

// feature_terms_res for interleaved features (LevelView::il): the same
// samples in the same arithmetic. The c0 taps of all channels come in
// chunks of 16 bytes per texel (8 channels of 16-bit storage, 4 of FP32),
// two chunks in flight at a time; the last channel's five samples also read
// the eight other texels of its plus-shaped neighbourhood
// (res_taps5_f32's selection). At most 32 channels; a texel's bytes a
// multiple of 4 (checked on the host). Chunks past the last channel re-read
// the last chunk, unused.
template <typename S>
__device__ inline float feature_terms_il(const LevelView<S>& L, int64_t idx,
                                         const V3<float>& p, float u, float v,
                                         float div, float j[6], float& rlast) {
#pragma clang fp contract(off)
  constexpr int es = int(sizeof(S));
  const int C = L.C, W = L.W;
  const int tb = C * es;  // bytes per texel
  const __amdgpu_buffer_rsrc_t rt = make_res(L.tf, int64_t(C) * L.HW);
  const __amdgpu_buffer_rsrc_t rs = make_res(L.sf, int64_t(C) * L.n);
  const float h = kFeatStep;
  const Bil<float> c0 = make_bil(u, v);
  const Bil<float> bu0 = make_bil(u + h, v);
  const Bil<float> bu1 = make_bil(fmax(u - h, 0.0f), v);
  const Bil<float> bv0 = make_bil(u, v + h);
  const Bil<float> bv1 = make_bil(u, fmax(v - h, 0.0f));
  const int o = mad24(c0.vi, W, c0.ui);  // texel (fu, fv)
  const int so = int(idx) * tb;
  const int lc = (C - 1) * es;
  // The last channel at texel index x (a neighbour left of column 0 or
  // above row 0 reads another texel or 0; it is never selected).
  auto last_at = [&](int x) { return res_ld<S>(rt, x * tb + lc, 0); };
  // Two chunks in flight (the scheduling barriers keep the compiler from
  // hoisting every chunk's loads to the top); chunks past the last channel
  // re-read the last chunk and are not added. The last chunk ends in k[1]
  // (NCH is even), where the last channel's c0 taps and source value are
  // read; its eight other neighbours load with the second-to-last chunk's
  // turn, so they are in flight for the last two chunks only.
  constexpr int CH = kChunkCh<S>;
  constexpr int NCH = 32 / CH;  // chunks of 32 channels
  const int nch = (C + CH - 1) / CH;
  const int o00 = o * tb, o01 = (o + W) * tb;
  ChunkIL k[2];
  load_chunk(rt, rs, o00, o01, tb, so, 0, k[0]);
  load_chunk(rt, rs, o00, o01, tb, so, 16 * (nch > 1 ? 1 : 0), k[1]);
  float sq = 0.0f;
  float A0 = 0.0f, A3 = 0.0f, B0 = 0.0f, B3 = 0.0f;
  float up0 = 0.0f, up1 = 0.0f, dn0 = 0.0f, dn1 = 0.0f;
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    __builtin_amdgcn_sched_barrier(0);
    if (i < nch) add_chunk<S>(k[i & 1], c0, W, i * CH, C, sq);
    __builtin_amdgcn_sched_barrier(0);
    if (i + 2 < NCH)
      load_chunk(rt, rs, o00, o01, tb, so,
                 16 * (i + 2 < nch ? i + 2 : nch - 1), k[i & 1]);
    if (i == NCH - 2) {
      A0 = last_at(o - 1);
      A3 = last_at(o + 2);
      B0 = last_at(o + W - 1);
      B3 = last_at(o + W + 2);
      up0 = last_at(o - W);
      up1 = last_at(o - W + 1);
      dn0 = last_at(o + 2 * W);
      dn1 = last_at(o + 2 * W + 1);
    }
  }
  const int ql = (C - 1) - (nch - 1) * CH;  // the last channel in its chunk
  const float A[4] = {A0, chunk_ch_at<S>(k[1].t[0], ql),
                      chunk_ch_at<S>(k[1].t[1], ql), A3};
  const float B[4] = {B0, chunk_ch_at<S>(k[1].t[2], ql),
                      chunk_ch_at<S>(k[1].t[3], ql), B3};
  const float flast = chunk_ch_at<S>(k[1].s, ql);
  float t0[4], t1[4], t2[4], t3[4], t4[4];
  t0[0] = A[1];
  t0[1] = A[2];
  t0[2] = B[1];
  t0[3] = B[2];
  const bool ru = bu0.ui != c0.ui;
  t1[0] = ru ? A[2] : A[1];
  t1[1] = ru ? A[3] : A[2];
  t1[2] = ru ? B[2] : B[1];
  t1[3] = ru ? B[3] : B[2];
  const bool lu = bu1.ui != c0.ui;
  t2[0] = lu ? A[0] : A[1];
  t2[1] = lu ? A[1] : A[2];
  t2[2] = lu ? B[0] : B[1];
  t2[3] = lu ? B[1] : B[2];
  const bool dv = bv0.vi != c0.vi;
  t3[0] = dv ? B[1] : A[1];
  t3[1] = dv ? B[2] : A[2];
  t3[2] = dv ? dn0 : B[1];
  t3[3] = dv ? dn1 : B[2];
  const bool uv = bv1.vi != c0.vi;
  t4[0] = uv ? up0 : A[1];
  t4[1] = uv ? up1 : A[2];
  t4[2] = uv ? A[1] : B[1];
  t4[3] = uv ? A[2] : B[2];
  const float last = bil_mix_res(W, c0, t0);
  const float su0 = bil_mix_res(W, bu0, t1), su1 = bil_mix_res(W, bu1, t2);
  const float sv0 = bil_mix_res(W, bv0, t3), sv1 = bil_mix_res(W, bv1, t4);
  const float res = last - flast;
  feature_j(L.fx, L.fy, p, (su0 - su1) * div, (sv0 - sv1) * div, j);
  rlast = res;
  sq += res * res;
  return sq;
}

// Point-to-plane and feature terms of source pixel idx when the map names it
// as its target pixel's owner (src/icp_jacobian_geometric.cu:54-85,
// src/icp_jacobian_feature.cu:58-109). The code has no branch: a pixel that
// is not valid, projects outside the target or does not own its target
// pixel reads target pixel 0 (or its clamped feature cell) and adds exact
// zeros (selected, never multiplied in, so NaN or inf in its unused data
// cannot leak into the sums). The owner word, the target point and normal
// and the feature taps are loaded whatever the owner test says, so all of
// them issue in one batch: a pixel costs two memory round trips. The
// projection, the rounding, the owner test and the terms keep the
// reference's arithmetic (the same decisions as the correspondence pass,
// the oracle's per-pixel products). BORN: this is the lane's first pixel,
// the slots start from its terms.
template <typename S, bool GEOM, bool FEAT, bool BORN, int NB = 4,
          bool IL = false>
__device__ inline void jacobian_pixel_at(const LevelView<S>& L,
                                         const LevelRes& R, const float* m,
                                         const unsigned long long* __restrict__ map,
                                         uint32_t tag, int64_t idx, bool valid_src,
                                         V3<float> ps, const TermWeights& tw,
                                         float (&acc)[kSlots]) {
  V3<float> p;
  float u, v;
  {
#pragma clang fp contract(off)
    p = rigid_point(m, ps);
    project(L.fx, L.fy, L.cx, L.cy, p, u, v);
  }
  int ui, vi;
  const bool in_u = round_px(u, L.W, ui);
  const bool in_v = round_px(v, L.H, vi);
  const bool inside = valid_src && in_u && in_v;
  const int pos = inside ? mad24(vi, L.W, ui) : 0;
  const __amdgpu_buffer_rsrc_t rmap = make_res(map, L.HW);
  const uint32_t klo = __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos, 0, 0);
  const uint32_t khi = __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos + 4, 0, 0);
  // The owner and use tests combine with `&`, never `&&`: a short-circuit
  // is a branch, and the compiler sinks into it every load whose value only
  // the owner needs (the target point, the feature taps), which then wait
  // for the key -- a second round trip.
  const bool own =
      inside & key_owned((uint64_t(khi) << 32) | klo, uint32_t(idx), tag);
  if (GEOM) {
    const V3<float> q = res_ld3<S>(R.tp, pos);
    const V3<float> qn = res_ld3<S>(R.tn, pos);
    // The reference's arithmetic (no contraction): the per-pixel products
    // equal the oracle's, so sums differ from it in summation order only.
    const float r = dot3(sub3(q, p), qn);
    const V3<float> c = cross3(p, qn);
    const bool jon = own & (tw.gon != 0);
    const float j[6] = {jon ? qn.x : 0.0f, jon ? qn.y : 0.0f,
                        jon ? qn.z : 0.0f, jon ? c.x : 0.0f,
                        jon ? c.y : 0.0f,  jon ? c.z : 0.0f};
    const float rs = own ? r : 0.0f;
    acc_term<BORN>(acc, j, rs);
    acc[27] = BORN ? rs * rs : fmaf(rs, rs, acc[27]);
    acc[29] = acc_add<BORN>(acc[29], own ? 1.0f : 0.0f);
  }
  if (FEAT) {
    constexpr bool FB = BORN && !GEOM;
    // Lanes outside the target sample cell (0, 0): in range, unused.
    const float uu = inside ? u : 0.0f, vv = inside ? v : 0.0f;
    float jf[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, rl = 0.0f, sq = 0.0f;
    if (IL) {
      // Interleaved (16-bit) features load for owners only: BF16
      // coordinates collapse runs of source pixels onto one target pixel
      // (about half of them own theirs at 1280x720), so skipping the
      // others' 32-channel taps outweighs waiting for the key first.
      if (own) sq = feature_terms_il<S>(L, idx, p, uu, vv, tw.fdiv, jf, rl);
    } else {
      sq = feature_terms_res<S, NB>(L, idx, p, uu, vv, tw.fdiv, jf, rl);
    }
    const bool use = own & !(sq > tw.fthr2);
    float j[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) j[a] = use ? jf[a] : 0.0f;
    acc_term<FB>(acc, j, use ? -rl : 0.0f);
    acc[28] = acc_add<FB>(acc[28], use ? sq : 0.0f);
    acc[30] = acc_add<FB>(acc[30], use ? 1.0f : 0.0f);
  }
  if (BORN) {
    if (!GEOM) acc[27] = acc[29] = 0.0f;
    if (!FEAT) acc[28] = acc[30] = 0.0f;
    acc[31] = 0.0f;
  }
}

// jacobian_pixel_at with the source pixel's mask and point loaded here.
template <typename S, bool GEOM, bool FEAT, bool BORN, int NB = 4,
          bool IL = false>
__device__ inline void jacobian_pixel(const LevelView<S>& L, const LevelRes& R,
                                      const float* m,
                                      const unsigned long long* __restrict__ map,
                                      uint32_t tag, int64_t idx, bool in_run,
                                      const TermWeights& tw,
                                      float (&acc)[kSlots]) {
  const uint8_t valid_src = res_u8(R.sm, int(idx));
  const V3<float> ps = res_ld3<S>(R.sp, int(idx));
  jacobian_pixel_at<S, GEOM, FEAT, BORN, NB, IL>(L, R, m, map, tag, idx,
                                                 in_run & (valid_src != 0), ps,
                                                 tw, acc);
}

// ---------------------------------------------------------------------------
// The Jacobian pixel in two parts, for a software pipeline over a thread's
// pixels (FP32 storage, planar features whose channels other than the last
// fit one batch of NB, or no features)
// ---------------------------------------------------------------------------
//
// A launch whose wavefronts all start together runs every pixel's chain --
// source loads, projection, gathers, terms -- in step across the chip, so
// the terms' VALU work and the gathers' memory time follow each other
// instead of overlapping. jac_issue does everything up to and including the
// issue of a pixel's gathers and consumes none of their results; jac_consume
// does the rest. Issuing the next pixel's gathers before consuming the
// current pixel's keeps them in flight while the VALU works. The arithmetic
// is jacobian_pixel_at's, the same decisions and per-pixel products.
template <int NB>
struct PxGather {
  V3<float> p;   // transformed source point
  float u, v;    // its projection, (0, 0) outside the target
  bool inside;   // valid source pixel projecting into the target
  uint32_t klo, khi;  // the map key at the target pixel
  V3<float> q, qn;    // target point and normal
  Taps5Raw last;      // the last channel's windows
  u32x2 tp[NB][2];    // the other channels' two tap rows
  float fs[NB], flast;  // source feature values
};

template <bool GEOM, bool FEAT, int NB>
__device__ inline void jac_issue(const LevelView<float>& L, const LevelRes& R,
                                 const float* m, __amdgpu_buffer_rsrc_t rmap,
                                 int64_t idx, bool valid_src, V3<float> ps,
                                 PxGather<NB>& X) {
  float u, v;
  {
#pragma clang fp contract(off)
    X.p = rigid_point(m, ps);
    project(L.fx, L.fy, L.cx, L.cy, X.p, u, v);
  }
  int ui, vi;
  const bool in_u = round_px(u, L.W, ui);
  const bool in_v = round_px(v, L.H, vi);
  X.inside = valid_src && in_u && in_v;
  X.u = X.inside ? u : 0.0f;
  X.v = X.inside ? v : 0.0f;
  const int pos = X.inside ? mad24(vi, L.W, ui) : 0;
  X.klo = __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos, 0, 0);
  X.khi = __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos + 4, 0, 0);
  if (GEOM) {
    X.q = res_ld3<float>(R.tp, pos);
    X.qn = res_ld3<float>(R.tn, pos);
  }
  if (FEAT) {
    const int C = L.C;
    const Bil<float> c0 = make_bil(X.u, X.v);
    X.last = taps5_load(make_res(L.tf + int64_t(C - 1) * L.HW, L.HW), L.W, c0);
    const __amdgpu_buffer_rsrc_t rs = make_res(L.sf, int64_t(C) * L.n);
    const int sch = int(L.n) * 4;
    const int so = int(idx) * 4;
    X.flast = res_ld<float>(rs, so + (C - 1) * sch, 0);
    const int last_c = C - 2 > 0 ? C - 2 : 0;
    const int o = mad24(c0.vi, L.W, c0.ui) * 4;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int c = k < last_c ? k : last_c;
      const __amdgpu_buffer_rsrc_t rc = make_res(L.tf + int64_t(c) * L.HW, L.HW);
      X.tp[k][0] = __builtin_amdgcn_raw_buffer_load_b64(rc, o, 0, 0);
      X.tp[k][1] = __builtin_amdgcn_raw_buffer_load_b64(rc, o + L.W * 4, 0, 0);
      X.fs[k] = res_ld<float>(rs, so + c * sch, 0);
    }
  }
}

// The ordering point between two pixels of the pipeline. Scheduling
// barriers order the loads, but the instruction selector places pure
// arithmetic anywhere its operands allow, and it pulled the next pixel's
// first uses of its gathers (and with them the waits for its loads) above
// the current pixel's terms. Here every slot sum and every input of the
// next pixel's terms that is not a gathered word (its point, projection
// and index) pass through empty volatile asm statements, which keep their
// order: the next pixel's terms, and its waits, come after the current
// pixel's terms are in the slots. (Passing the gathered words themselves
// through such a statement made the register allocator copy them, and a
// copy waits for its load.)
template <int NB>
__device__ inline void settle(float (&a)[kSlots], PxGather<NB>& X,
                              uint32_t& idx) {
  asm volatile("" : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]),
               "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]), "+v"(a[9]),
               "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]),
               "+v"(a[15]));
  asm volatile("" : "+v"(a[16]), "+v"(a[17]), "+v"(a[18]), "+v"(a[19]),
               "+v"(a[20]), "+v"(a[21]), "+v"(a[22]), "+v"(a[23]), "+v"(a[24]),
               "+v"(a[25]), "+v"(a[26]), "+v"(a[27]), "+v"(a[28]), "+v"(a[29]),
               "+v"(a[30]), "+v"(a[31]));
  asm volatile("" : "+v"(X.p.x), "+v"(X.p.y), "+v"(X.p.z), "+v"(X.u),
               "+v"(X.v), "+v"(idx));
}

template <bool GEOM, bool FEAT, bool BORN, int NB>
__device__ inline void jac_consume(const LevelView<float>& L,
                                   const PxGather<NB>& X, uint32_t idx,
                                   uint32_t tag, const TermWeights& tw,
                                   float (&acc)[kSlots]) {
  const bool own =
      X.inside & key_owned((uint64_t(X.khi) << 32) | X.klo, idx, tag);
  const V3<float> p = X.p;
  if (GEOM) {
    const float r = dot3(sub3(X.q, p), X.qn);
    const V3<float> c = cross3(p, X.qn);
    const bool jon = own & (tw.gon != 0);
    const float j[6] = {jon ? X.qn.x : 0.0f, jon ? X.qn.y : 0.0f,
                        jon ? X.qn.z : 0.0f, jon ? c.x : 0.0f,
                        jon ? c.y : 0.0f,    jon ? c.z : 0.0f};
    const float rs = own ? r : 0.0f;
    acc_term<BORN>(acc, j, rs);
    acc[27] = BORN ? rs * rs : fmaf(rs, rs, acc[27]);
    acc[29] = acc_add<BORN>(acc[29], own ? 1.0f : 0.0f);
  }
  if (FEAT) {
#pragma clang fp contract(off)
    constexpr bool FB = BORN && !GEOM;
    const int W = L.W, C = L.C;
    const float h = kFeatStep, u = X.u, v = X.v;
    const Bil<float> c0 = make_bil(u, v);
    const Bil<float> bu0 = make_bil(u + h, v);
    const Bil<float> bu1 = make_bil(fmax(u - h, 0.0f), v);
    const Bil<float> bv0 = make_bil(u, v + h);
    const Bil<float> bv1 = make_bil(u, fmax(v - h, 0.0f));
    float t0[4], t1[4], t2[4], t3[4], t4[4];
    taps5_select(X.last, c0, bu0, bu1, bv0, bv1, t0, t1, t2, t3, t4);
    const float last = bil_mix_res(W, c0, t0);
    const float su0 = bil_mix_res(W, bu0, t1), su1 = bil_mix_res(W, bu1, t2);
    const float sv0 = bil_mix_res(W, bv0, t3), sv1 = bil_mix_res(W, bv1, t4);
    float sq = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const float t[4] = {__uint_as_float(X.tp[k][0][0]),
                          __uint_as_float(X.tp[k][0][1]),
                          __uint_as_float(X.tp[k][1][0]),
                          __uint_as_float(X.tp[k][1][1])};
      const float res = bil_mix_res(W, c0, t) - X.fs[k];
      sq += (k < C - 1) ? res * res : 0.0f;
    }
    const float rl = last - X.flast;
    float jf[6];
    feature_j(L.fx, L.fy, p, (su0 - su1) * tw.fdiv, (sv0 - sv1) * tw.fdiv, jf);
    sq += rl * rl;
    const bool use = own & !(sq > tw.fthr2);
    float j[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) j[a] = use ? jf[a] : 0.0f;
    acc_term<FB>(acc, j, use ? -rl : 0.0f);
    acc[28] = acc_add<FB>(acc[28], use ? sq : 0.0f);
    acc[30] = acc_add<FB>(acc[30], use ? 1.0f : 0.0f);
  }
  if (BORN) {
    if (!GEOM) acc[27] = acc[29] = 0.0f;
    if (!FEAT) acc[28] = acc[30] = 0.0f;
    acc[31] = 0.0f;
  }
}

// G^T G of source pixel idx from its matched target point
// (src/icp_jacobian_geometric.cu:117-158), parameter order [rot, trans].
// Two memory round trips, as in the Jacobian pass: the source mask and point
// load together, then the map key and the target point at the projected
// pixel (target pixel 0 for a pixel that is not a valid source or projects
// outside); a pixel that does not own its target pixel adds selected zeros.
// The slots start at +0 and a sum of +0 and a finite term is that term, so
// they come out as if those pixels were skipped.
template <typename S>
__device__ inline void information_pixel(const LevelView<S>& L,
                                         const LevelRes& R, const float* m,
                                         __amdgpu_buffer_rsrc_t rmap,
                                         uint32_t tag, int64_t idx,
                                         float (&acc)[kSlots]) {
#pragma clang fp contract(off)
  const uint8_t vs = res_u8(R.sm, int(idx));
  const V3<float> p = rigid_point(m, res_ld3<S>(R.sp, int(idx)));
  float u, v;
  project(L.fx, L.fy, L.cx, L.cy, p, u, v);
  int ui = 0, vi = 0;
  const bool in_u = round_in_range(u, L.W, ui);
  const bool in_v = round_in_range(v, L.H, vi);
  const bool inside = (vs != 0) & in_u & in_v;
  const int pos = inside ? mad24(vi, L.W, ui) : 0;
  const uint32_t klo = __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos, 0, 0);
  const uint32_t khi =
      __builtin_amdgcn_raw_buffer_load_b32(rmap, 8 * pos + 4, 0, 0);
  const V3<float> q = res_ld3<S>(R.tp, pos);
  const bool own =
      inside & key_owned((uint64_t(khi) << 32) | klo, uint32_t(idx), tag);
  const float r0[6] = {0.f, q.z, -q.y, 1.f, 0.f, 0.f};
  const float r1[6] = {-q.z, 0.f, q.x, 0.f, 1.f, 0.f};
  const float r2[6] = {q.y, -q.x, 0.f, 0.f, 0.f, 1.f};
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) {
      const float g = r0[a] * r0[b] + r1[a] * r1[b] + r2[a] * r2[b];
      acc[k] += own ? g : 0.0f;
      ++k;
    }
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

// The wavefront sum of 32 slots by recursive halving: each step pairs every
// lane with a partner that holds the same slots and differs in one lane bit;
// the lane with the bit set keeps the upper half of the live slots, its
// partner the lower half, and each adds the other's copy. The two upper
// bits go through gfx950's permlane swaps, which exchange the halves in one
// instruction per slot pair with no select; the four lower bits through DPP
// lane patterns that read the partner inside the add (row_ror:8 pairs lanes
// i and i ^ 8 of a row, row_half_mirror i and 7 - i of a half-row, the quad
// permutations i ^ 3 and i ^ 1). The 32 exchanges stay in registers (no
// ds_bpermute round trips through the LDS unit), and the 24 of the upper
// steps need no selects.
constexpr int kDppRowRor8 = 0x128;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppQuad3210 = 0x1B;
constexpr int kDppQuad1032 = 0xB1;

template <int CTRL>
__device__ inline float dpp_read(float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

// Lanes 0..31 keep slots k, lanes 32..63 slots k + 16 (k < 16).
__device__ inline void halve_swap32(float (&v)[kSlots]) {
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const auto r = __builtin_amdgcn_permlane32_swap(
        __float_as_uint(v[k]), __float_as_uint(v[k + 16]), false, false);
    v[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}

// Rows 0 and 2 keep slots k, rows 1 and 3 slots k + 8 (k < 8).
__device__ inline void halve_swap16(float (&v)[kSlots]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const auto r = __builtin_amdgcn_permlane16_swap(
        __float_as_uint(v[k]), __float_as_uint(v[k + 8]), false, false);
    v[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}

template <int CTRL, int H>
__device__ inline void halve_dpp(float (&v)[kSlots], bool up) {
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float send = up ? v[k] : v[k + H];
    const float keep = up ? v[k + H] : v[k];
    v[k] = keep + dpp_read<CTRL>(send);
  }
}

// After this, lane l holds the wavefront total of slot l >> 1 in v[0]. The
// swaps and DPP reads take partner lanes' registers whatever the exec mask
// says, so every lane of the wavefront must be active (block_partial is
// reached by whole blocks of full wavefronts).
__device__ inline void wave_reduce_slots(float (&v)[kSlots], int lane) {
  halve_swap32(v);
  halve_swap16(v);
  halve_dpp<kDppRowRor8, 4>(v, (lane >> 3) & 1);
  halve_dpp<kDppRowHalfMirror, 2>(v, (lane >> 2) & 1);
  halve_dpp<kDppQuad3210, 1>(v, (lane >> 1) & 1);
  v[0] += dpp_read<kDppQuad1032>(v[0]);
}

// Reduces the block's slots in FP64 and stores them at `out` (32 values of
// type P), with no fence: every consumer runs in a later launch (a fence and
// a ticket for a consumer in the same launch cost each block ~20 ns at the
// memory side). The wavefront sums enter FP64 scaled by the weight the
// per-pixel sums left out (TermWeights): b^2 for the 21 JtJ slots, b for the
// 6 Jr slots, 1 for the bookkeeping (b = 1: the slots as accumulated). `sh`
// is LDS of BLOCK / 64 * kSlots doubles (kShared for kBlock). The
// single-pair launches store FP32 partials (P = float): every block of the
// loop's next correspondence launch reads all of them, and half the bytes
// save ~0.6 us of a 14 us iteration at 640x480, while the rounding to FP32
// (6e-8 relative per block) matches that of the FP32 wavefront sums the
// partial is made of. The sums over partials stay FP64.
template <int BLOCK = kBlock, typename P>
__device__ inline void block_partial(float (&acc)[kSlots], P* out,
                                     double* sh, double b = 1.0) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  wave_reduce_slots(acc, lane);
  const int slot = lane >> 1;
  const double scale = slot < 21 ? b * b : (slot < 27 ? b : 1.0);
  if ((lane & 1) == 0) sh[wave * kSlots + slot] = double(acc[0]) * scale;
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += sh[w * kSlots + threadIdx.x];
    out[threadIdx.x] = P(s);
  }
  __syncthreads();
}

// Fixed-order FP64 sum of `count` partials (32 values each) into tot[32] by
// a block of BLOCK threads, the same order whichever block performs it:
// thread group g of BLOCK / 32 takes partials g, g + BLOCK / 32, ... in
// batches of kSumBatch summed pairwise; the two groups of a wavefront are
// added across its halves, and the per-wavefront sums pairwise through LDS.
// Pairwise sums keep the dependent chain short (log2 instead of linear in the
// number of terms): this sum sits on the critical path of every iteration.
// The partials were written by blocks on every XCD, so each load misses
// this XCD's L2: a batch's loads issue before any of its adds. `sh` is LDS of
// BLOCK / 64 * kSlots doubles.
constexpr int kSumBatch = 8;

// One batch of sum_partials' loads: partials b0, b0 + BLOCK / 32, ... of
// this thread's slot. Loads past the end re-read the last partial and are
// dropped, so the batch issues without branches.
template <int BLOCK, typename P>
__device__ inline void load_partial_batch(const P* partials, int count,
                                          int b0, double (&v)[kSumBatch]) {
  constexpr int kG = BLOCK / kSlots;
  const int slot = threadIdx.x & (kSlots - 1);
#pragma unroll
  for (int u = 0; u < kSumBatch; ++u) {
    const int b = b0 + u * kG;
    const double x = partials[int64_t(b < count ? b : count - 1) * kSlots + slot];
    v[u] = b < count ? x : 0.0;  // (FP32 partials widen exactly)
  }
}

// sum_partials with the thread's first batch already loaded (by
// load_partial_batch at b0 = threadIdx.x / 32), so a caller can issue other
// loads behind it: loads return in issue order.
template <int BLOCK, typename P>
__device__ inline void sum_partials_from(double (&v)[kSumBatch],
                                         const P* partials, int count,
                                         double* sh, double* tot) {
  constexpr int kG = BLOCK / kSlots;
  constexpr int kW = BLOCK / 64;
  const int slot = threadIdx.x & (kSlots - 1);
  const int grp = threadIdx.x / kSlots;
  // The first batch outside the loop (all of it dropped when grp >= count),
  // so that its loads are unconditional and stay where the caller put them.
  double s = 0.0;
#pragma unroll
  for (int w = kSumBatch / 2; w > 0; w /= 2)
#pragma unroll
    for (int u = 0; u < w; ++u) v[u] = v[u] + v[u + w];
  s += v[0];
  for (int b0 = grp + kSumBatch * kG; b0 < count; b0 += kSumBatch * kG) {
    load_partial_batch<BLOCK, P>(partials, count, b0, v);
#pragma unroll
    for (int w = kSumBatch / 2; w > 0; w /= 2)
#pragma unroll
      for (int u = 0; u < w; ++u) v[u] = v[u] + v[u + w];
    s += v[0];
  }
  s = s + __shfl_xor(s, 32, 64);  // groups 2w and 2w + 1 of wavefront w
  __syncthreads();
  if ((threadIdx.x & 63) < kSlots) sh[(threadIdx.x >> 6) * kSlots + slot] = s;
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double g[kW];
#pragma unroll
    for (int w = 0; w < kW; ++w) g[w] = sh[w * kSlots + threadIdx.x];
#pragma unroll
    for (int h = kW / 2; h > 0; h /= 2)
#pragma unroll
      for (int w = 0; w < h; ++w) g[w] = g[w] + g[w + h];
    tot[threadIdx.x] = g[0];
  }
  __syncthreads();
}

template <int BLOCK, typename P>
__device__ inline void sum_partials(const P* partials, int count,
                                    double* sh, double* tot) {
  double v[kSumBatch];
  load_partial_batch<BLOCK, P>(partials, count, int(threadIdx.x) / kSlots, v);
  sum_partials_from<BLOCK, P>(v, partials, count, sh, tot);
}

// ---------------------------------------------------------------------------
// Pose state
// ---------------------------------------------------------------------------

__device__ inline void set_pose_floats(stb_icp_state* st) {
  Rigid<float> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) r.m[i * 4 + j] = float(st->transform[i * 4 + j]);
  make_normal_matrix(r);
#pragma unroll
  for (int i = 0; i < 12; ++i) st->rt[i] = r.m[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) st->nm[i] = r.nm[i];
}

// exp_rt_to_matrix in FP64 (src/exp_rt.hpp:27-65, src/exp_rt_to_matrix.cu:22-36):
// Rodrigues rotation of w = x[3:6], translation x[0:3] used as is.
// For |w| <= 1/2 (every Gauss-Newton step of a converging loop) the same
// rotation is formed as R = I + A [w]x + B [w]x^2, with A = sin t / t and
// B = (1 - cos t) / t^2 summed from their Taylor series in t^2 = |w|^2 to the
// term in t^16 (truncation below 1e-22): no square root, division or
// trigonometric range reduction on the solve's serial chain. Larger angles
// take the reference's axis-angle form.
__device__ inline void exp_rt_f64(const double x[6], double U[12]) {
  const double wx = x[3], wy = x[4], wz = x[5];
  const double xx = wx * wx, yy = wy * wy, zz = wz * wz;
  const double sq = xx + (yy + zz);
  double r00, r01, r02, r10, r11, r12, r20, r21, r22;
  if (sq <= 0.25) {
    double p = -1.0 / 355687428096000.0;  // (1 - A) / t^2, highest term first
    p = fma(p, sq, 1.0 / 1307674368000.0);
    p = fma(p, sq, -1.0 / 6227020800.0);
    p = fma(p, sq, 1.0 / 39916800.0);
    p = fma(p, sq, -1.0 / 362880.0);
    p = fma(p, sq, 1.0 / 5040.0);
    p = fma(p, sq, -1.0 / 120.0);
    p = fma(p, sq, 1.0 / 6.0);
    const double A = fma(-p, sq, 1.0);
    double B = 1.0 / 6402373705728000.0;
    B = fma(B, sq, -1.0 / 20922789888000.0);
    B = fma(B, sq, 1.0 / 87178291200.0);
    B = fma(B, sq, -1.0 / 479001600.0);
    B = fma(B, sq, 1.0 / 3628800.0);
    B = fma(B, sq, -1.0 / 40320.0);
    B = fma(B, sq, 1.0 / 720.0);
    B = fma(B, sq, -1.0 / 24.0);
    B = fma(B, sq, 0.5);
    const double bxy = B * wx * wy, bxz = B * wx * wz, byz = B * wy * wz;
    r00 = 1.0 - B * (yy + zz);
    r11 = 1.0 - B * (xx + zz);
    r22 = 1.0 - B * (xx + yy);
    r01 = bxy - A * wz;
    r10 = bxy + A * wz;
    r02 = bxz + A * wy;
    r20 = bxz - A * wy;
    r12 = byz - A * wx;
    r21 = byz + A * wx;
  } else {
    const double th = sqrt(sq);
    const double ax = wx / th, ay = wy / th, az = wz / th;
    double s, c;
    sincos(th, &s, &c);
    const double k = 1.0 - c;
    r00 = k * ax * ax + c;
    r01 = k * ax * ay - s * az;
    r02 = k * ax * az + s * ay;
    r10 = k * ax * ay + s * az;
    r11 = k * ay * ay + c;
    r12 = k * ay * az - s * ax;
    r20 = k * ax * az - s * ay;
    r21 = k * ay * az + s * ax;
    r22 = k * az * az + c;
  }
  U[0] = r00;
  U[1] = r01;
  U[2] = r02;
  U[3] = x[0];
  U[4] = r10;
  U[5] = r11;
  U[6] = r12;
  U[7] = x[1];
  U[8] = r20;
  U[9] = r21;
  U[10] = r22;
  U[11] = x[2];
}

// Index of JtJ[a][c] (a <= c) among the 21 packed slots (upper triangle,
// row-major), and of L[i][j] (j <= i) in a packed lower triangle.
__host__ __device__ constexpr int tri_up(int a, int c) {
  return a * (11 - a) / 2 + c;
}
__host__ __device__ constexpr int tri_lo(int i, int j) {
  return i * (i + 1) / 2 + j;
}
__device__ inline double sym6(const double* tot, int a, int c) {
  return a <= c ? tot[tri_up(a, c)] : tot[tri_up(c, a)];
}

// The tails below are block-cooperative: every thread of the block that
// finished the reduction calls them (tot and w in LDS). Stores to the state
// are spread over threads and every value the serial part needs is read into
// LDS or registers first, so the tail pays one memory round trip, not one
// per element.

// Stores the reduced system of one linearisation in the state.
__device__ inline void store_linearization(stb_icp_state* st,
                                           const double* tot) {
  const int t = threadIdx.x;
  if (t < 36)
    st->last_jtj[t] = sym6(tot, t / 6, t % 6);
  else if (t < 42)
    st->last_jr[t - 36] = tot[21 + t - 36];
  else if (t == 42)
    st->last_geom_r2 = tot[27];
  else if (t == 43)
    st->last_feat_r2 = tot[28];
  else if (t == 44)
    st->last_geom_count = tot[29];
  else if (t == 45)
    st->last_feat_count = tot[30];
}

// One Gauss-Newton update from the reduced totals (slamtb/registration/icp.py:
// 236-256): FP64 Cholesky of JtJ (a non-positive pivot is the reference's
// `except: break` and stops the loop), solve JtJ x = Jr, T <- exp_rt(x) T.
// The factorisation, solve and pose product run in one thread on registers
// (fully unrolled over packed triangles): the serial chain is FP64
// arithmetic only. It runs in a launch of its own, so its register count
// does not bound the occupancy of the per-pixel kernels. `w` is LDS of at
// least 25 doubles whose first 12 hold rows 0..2 of the current pose;
// thread 0 passes the state's iteration count, map_sel and map_epoch, read
// early by the caller so that their latency hides behind the partial sums.
// Advances the map use (next_map_use) whether or not the update succeeds:
// the iteration wrote its map either way.
// The Gauss-Newton update of one summed system (21 JtJ + 6 Jr slots of
// `tot`): Cholesky of JtJ, L y = Jr, L^T x = y, then Tn = exp(x) T (3x4,
// row-major). Returns false on a non-positive pivot (Tn is then not set).
// Straight-line FP64 code with no memory traffic, so it runs at the same
// speed in one thread or, redundantly, in every lane of a wavefront, and
// both give the same bits. The update sits on every loop iteration's
// critical path and its instruction count sets its length (FP64 issues at
// the non-packed FP32 rate), so its products and sums are fused
// multiply-adds (no reference path matches this solve's bits; both entry
// paths share it).
__device__ inline bool gn_update_f64(const double* tot, const double* T,
                                     double Tn[12]) {
  double A[21], b[6];
#pragma unroll
  for (int k = 0; k < 21; ++k) A[k] = tot[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) b[k] = tot[21 + k];
  // L's diagonal enters only through its reciprocals rinv[j] = 1/sqrt(d2):
  // the hardware reciprocal square root refined by two Newton steps
  // (relative error ~1e-16, like sqrt followed by a divide, but a
  // dependent chain of 7 instructions instead of ~25).
  double L[21], rinv[6];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d2 = A[tri_up(j, j)];
#pragma unroll
    for (int q = 0; q < j; ++q) d2 = fma(-L[tri_lo(j, q)], L[tri_lo(j, q)], d2);
    ok = ok && (d2 > 0.0);
    double r = __builtin_amdgcn_rsq(d2);
    r = fma(r * 0.5, fma(-d2 * r, r, 1.0), r);
    r = fma(r * 0.5, fma(-d2 * r, r, 1.0), r);
    rinv[j] = r;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double v = A[tri_up(j, i)];
#pragma unroll
      for (int q = 0; q < j; ++q) v = fma(-L[tri_lo(i, q)], L[tri_lo(j, q)], v);
      L[tri_lo(i, j)] = v * rinv[j];
    }
  }
  if (!ok) return false;
  double y[6], x[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {  // L y = b
    double v = b[i];
#pragma unroll
    for (int q = 0; q < i; ++q) v = fma(-L[tri_lo(i, q)], y[q], v);
    y[i] = v * rinv[i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {  // L^T x = y, the newest x[q] last
    double v = y[i];
#pragma unroll
    for (int q = 5; q > i; --q) v = fma(-L[tri_lo(q, i)], x[q], v);
    x[i] = v * rinv[i];
  }
  double U[12];
  exp_rt_f64(x, U);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double v = fma(U[i * 4 + 2], T[8 + j],
                     fma(U[i * 4 + 1], T[4 + j], U[i * 4 + 0] * T[j]));
      if (j == 3) v += U[i * 4 + 3];
      Tn[i * 4 + j] = v;
    }
  return true;
}

__device__ inline void gn_finalize(stb_icp_state* st, const double* tot,
                                   double gw, double fw, double* w, int iters,
                                   int sel, uint32_t tag) {
  const int t = threadIdx.x;
  store_linearization(st, tot);
  __syncthreads();
  if (t == 0) {
    double T[12], Tn[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) T[k] = w[k];
    const bool ok = gn_update_f64(tot, T, Tn);
    if (ok) {
      Rigid<float> r;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        w[12 + k] = Tn[k];
        r.m[k] = float(Tn[k]);
      }
      make_normal_matrix(r);
#pragma unroll
      for (int i = 0; i < 12; ++i) st->rt[i] = r.m[i];
#pragma unroll
      for (int i = 0; i < 9; ++i) st->nm[i] = r.nm[i];
      const double count = fmax(tot[29], tot[30]);
      st->best_residual = (fw * tot[28] + gw * tot[27]) / count;
      st->best_match_count = count;
      st->iterations = iters + 1;
    } else {
      st->stopped = 1;
    }
    const MapUse nu = next_map_use(tag, sel);
    st->map_epoch = int32_t(nu.tag);
    st->map_sel = nu.sel;
    w[24] = ok ? 1.0 : 0.0;
  }
  __syncthreads();
  if (w[24] != 0.0) {
    if (t < 12) {
      st->transform[t] = w[12 + t];
      st->best_transform[t] = w[12 + t];
    } else if (t >= 64 && t < 100) {
      st->best_jtj[t - 64] = sym6(tot, (t - 64) / 6, (t - 64) % 6);
    }
  }
}

__device__ inline void store_information(stb_icp_state* st, const double* tot) {
  const int t = threadIdx.x;
  if (t < 36) st->information[t] = sym6(tot, t / 6, t % 6);
}

// What finish_partials does with the summed partials.
enum FinishMode { kStoreLinearization = 0, kUpdatePose = 1, kStoreInformation = 2 };

// Sums `count` partials and applies them to the state: a Gauss-Newton update
// (kUpdatePose), the stored linearisation, or the information matrix (the
// partials of information_pixel). Called by every thread of one block of
// BLOCK threads; `sh` is LDS of BLOCK doubles, `tot` of kSlots.
template <int BLOCK, typename P>
__device__ inline void finish_partials(stb_icp_state* st, const P* partials,
                                       int count, int mode, double gw,
                                       double fw, double* sh, double* tot) {
  const int t = threadIdx.x;
  const bool update = mode == kUpdatePose;
  // The state loads of the update issue before the partial loads.
  double row = 0.0;
  int iters = 0, sel = 0;
  uint32_t tag = 0;
  if (update && t < 12) row = st->transform[t];
  if (update && t == 0) {
    iters = st->iterations;
    sel = st->map_sel;
    tag = uint32_t(st->map_epoch);
  }
  sum_partials<BLOCK, P>(partials, count, sh, tot);
  if (update) {
    if (t < 12) sh[t] = row;
    gn_finalize(st, tot, gw, fw, sh, iters, sel, tag);
  } else if (mode == kStoreInformation) {
    store_information(st, tot);
  } else {
    store_linearization(st, tot);
  }
}

// ---------------------------------------------------------------------------
// Pose update inside the next iteration's correspondence launch
// ---------------------------------------------------------------------------
//
// Iteration k's correspondence launch can apply iteration k-1's update
// itself: each of its blocks sums the previous Jacobian launch's partials in
// the same fixed order, solves the 6x6 system and forms the new pose, so no
// launch (and no cross-block ticket) stands between a Jacobian launch and the
// next correspondence pass. Those blocks still read the state, so the update
// goes to a SolveRecord written by block 0; the Jacobian launch of the same
// iteration takes its pose and map from the record, and its block 0 commits
// the record to the state.
struct SolveRecord {
  double tot[kSlots];  // reduced system of the previous iteration
  double T[12];        // rows 0..2 of the updated pose (when ok)
  float rt[12];        // float pose this iteration uses
  float nm[9];         // its normal matrix
  int32_t active;      // 0: the loop had already stopped, nothing to commit
  int32_t ok;          // every Cholesky pivot was positive
  int32_t map_sel;     // correspondence map this iteration uses
  int32_t map_epoch;   // key tag this iteration writes
  int32_t iterations;  // state.iterations after a successful update
  int32_t pad;
};

// The update of a loop iteration from the summed system in LDS, by one
// whole wavefront: every lane reads the system and the previous pose
// (broadcast reads) and evaluates gn_update_f64, so the factorisation waits
// on no cross-lane exchange; lane 0 stores the new pose in FP64 (Tn), in
// FP32 (rtf) and that pose's normal matrix (nmf), once for the whole block.
// Returns whether the factorisation succeeded (Tn, rtf and nmf are left
// unset when not).
__device__ inline bool wave_solve_update(const double* tot, const double* To,
                                         double* Tn, float* rtf, float* nmf) {
  double sys[27], T[12], U[12];
#pragma unroll
  for (int k = 0; k < 27; ++k) sys[k] = tot[k];
#pragma unroll
  for (int k = 0; k < 12; ++k) T[k] = To[k];
  const bool ok = gn_update_f64(sys, T, U);
  if (ok && (threadIdx.x & 63) == 0) {
    Rigid<float> r;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      Tn[k] = U[k];
      r.m[k] = float(U[k]);
      rtf[k] = r.m[k];
    }
    make_normal_matrix(r);
#pragma unroll
    for (int k = 0; k < 9; ++k) nmf[k] = r.nm[k];
  }
  return ok;
}

// What gn_finalize stores, from a SolveRecord (every thread of a block of at
// least 103 threads calls it). Each thread loads the values it stores first,
// so the block pays one memory round trip, not one per field.
__device__ inline void commit_record(stb_icp_state* __restrict__ st,
                                     const SolveRecord* __restrict__ rec,
                                     double gw, double fw) {
  const int t = threadIdx.x;
  const int active = rec->active, ok = rec->ok;
  const double* __restrict__ tot = rec->tot;
  double v0 = 0.0, v1 = 0.0;
  float f0 = 0.0f;
  int i0 = 0, i1 = 0;
  if (t < 36)
    v0 = sym6(tot, t / 6, t % 6);  // last_jtj
  else if (t < 46)
    v0 = tot[21 + t - 36];  // last_jr, r2 and counts (slots 21..30)
  if (t < 12) {
    v1 = rec->T[t];
    f0 = rec->rt[t];
  } else if (t < 21) {
    f0 = rec->nm[t - 12];
  } else if (t >= 64 && t < 100) {
    v1 = sym6(tot, (t - 64) / 6, (t - 64) % 6);
  } else if (t == 100) {  // best_residual and best_match_count
    const double c = fmax(tot[29], tot[30]);
    v0 = (fw * tot[28] + gw * tot[27]) / c;
    v1 = c;
  } else if (t == 101) {
    i0 = rec->iterations;
  } else if (t == 102) {
    i0 = rec->map_sel;
    i1 = rec->map_epoch;
  }
  if (!active) return;
  if (t == 102) {  // the map use advances whether or not the update succeeded
    st->map_sel = i0;
    st->map_epoch = i1;
  }
  if (t < 36)
    st->last_jtj[t] = v0;
  else if (t < 42)
    st->last_jr[t - 36] = v0;
  else if (t == 42)
    st->last_geom_r2 = v0;
  else if (t == 43)
    st->last_feat_r2 = v0;
  else if (t == 44)
    st->last_geom_count = v0;
  else if (t == 45)
    st->last_feat_count = v0;
  if (!ok) {
    if (t == 0) st->stopped = 1;
    return;
  }
  if (t < 12) {
    st->transform[t] = v1;
    st->best_transform[t] = v1;
    st->rt[t] = f0;
  } else if (t < 21) {
    st->nm[t - 12] = f0;
  } else if (t >= 64 && t < 100) {
    st->best_jtj[t - 64] = v1;
  } else if (t == 100) {
    st->best_residual = v0;
    st->best_match_count = v1;
  } else if (t == 101) {
    st->iterations = i0;
  }
}

// Resets one state to pose T0 (row-major 4x4, NULL = identity) or, with
// `chain`, to its own last accepted pose (identity if none).
__device__ inline void reset_state(stb_icp_state* st,
                                   const double* __restrict__ T0, int chain) {
  double T[16];
  const bool from_prev = chain && st->iterations > 0;
  for (int i = 0; i < 16; ++i) {
    const double eye = (i % 5 == 0) ? 1.0 : 0.0;
    if (chain)
      T[i] = from_prev ? st->best_transform[i] : eye;
    else
      T[i] = T0 ? T0[i] : eye;
  }
  for (int i = 0; i < 16; ++i) {
    st->transform[i] = T[i];
    st->best_transform[i] = T[i];
  }
  for (int i = 0; i < 36; ++i) {
    st->best_jtj[i] = 0.0;
    st->last_jtj[i] = 0.0;
    st->information[i] = 0.0;
  }
  for (int i = 0; i < 6; ++i) st->last_jr[i] = 0.0;
  st->best_residual = __builtin_inf();
  st->best_match_count = 0.0;
  st->last_geom_r2 = st->last_feat_r2 = 0.0;
  st->last_geom_count = st->last_feat_count = 0.0;
  st->stopped = 0;
  st->iterations = 0;
  st->map_epoch = kFreshTag;
  st->map_sel = 0;
  set_pose_floats(st);
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace stb
