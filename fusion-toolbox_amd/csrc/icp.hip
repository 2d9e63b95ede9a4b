// ICP Gauss-Newton kernels for gfx950: projective correspondence, point-to-
// plane and feature Jacobians, the 6x6 reduction, FP64 Cholesky solve and the
// exp_rt_to_matrix pose update.
//
// Two families of entry points:
//   * drop-in ones that keep the reference bindings' per-pixel outputs
//     (CorrespondenceMap.compute_correspondence_map,
//     ICPJacobian.estimate_geometric / estimate_feature /
//     get_information_matrix);
//   * the fused Gauss-Newton path: one correspondence kernel and one
//     Jacobian+reduction kernel per iteration. The second kernel accumulates
//     the 21 unique JtJ and 6 Jr terms per lane in FP32, reduces them across
//     the wavefront by recursive halving (32 exchanges for 32 slots), across
//     the workgroup in FP64 through LDS, and writes one FP64 partial per
//     workgroup; the last workgroup to finish (arrival ticket) sums the
//     partials in a fixed order, solves the system and updates the pose on
//     the device. Nothing returns to the host between iterations.
#include "stb_common.hpp"

#include <cmath>

namespace stb {
namespace {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

template <typename T>
struct CamT {
  T fx, fy, cx, cy;
};

template <typename T>
__device__ inline CamT<T> load_cam(const T* __restrict__ k) {
  return CamT<T>{k[0], k[4], k[2], k[5]};
}

// The reference rejects a pair when |acos(c)| >= thr; the host turns thr into
// c* so that the rejection is exactly -1 <= c <= c* (acos of anything outside
// [-1, 1], or of NaN, is NaN and lets the pair through).
template <typename T>
__device__ inline bool angle_rejects(T c, T cstar) {
  return (c >= T(-1)) && (c <= cstar);
}

template <typename T>
__device__ inline T gate_cos(const V3<T>& a, const V3<T>& b) {
#pragma clang fp contract(off)
  return dot3(a, b) / (sqrt(dot3(a, a)) * sqrt(dot3(b, b)));
}

template <typename T>
__device__ inline void rigid_from(const T* __restrict__ rt, Rigid<T>& r) {
#pragma unroll
  for (int i = 0; i < 12; ++i) r.m[i] = rt[i];
  make_normal_matrix(r);
}

// Bilinear sample of one channel (src/feature_map.hpp:8-43): truncating cell
// index, zero padding past the right and bottom edges.
template <typename T>
struct Bil {
  int ui, vi;
  T ur, vr;
};

template <typename T>
__device__ inline Bil<T> make_bil(T u, T v) {
#pragma clang fp contract(off)
  Bil<T> b;
  b.ui = trunc_int(u);
  b.vi = trunc_int(v);
  b.ur = u - T(b.ui);
  b.vr = v - T(b.vi);
  return b;
}

template <typename T, typename S>
__device__ inline T bil_get(const S* __restrict__ ch, int H, int W,
                            const Bil<T>& b) {
#pragma clang fp contract(off)
  const int64_t o = int64_t(b.vi) * W + b.ui;
  const bool rx = b.ui + 1 < W, ry = b.vi + 1 < H;
  const T v00 = T(ldw(ch, o));
  const T v10 = rx ? T(ldw(ch, o + 1)) : T(0);
  const T v01 = ry ? T(ldw(ch, o + W)) : T(0);
  const T v11 = (rx && ry) ? T(ldw(ch, o + W + 1)) : T(0);
  const T a = v00 * (T(1) - b.ur) + v10 * b.ur;
  const T c = v01 * (T(1) - b.ur) + v11 * b.ur;
  return a * (T(1) - b.vr) + c * b.vr;
}

// Feature residuals and the last channel's Jacobian at one matched pixel
// (src/icp_jacobian_feature.cu:81-106, src/feature_map.hpp:45-64,
// include/slamtb/camera.hpp:65-78). Returns the channel-summed squared
// residual; j/rlast describe the last channel (the one the reference keeps).
template <typename T, typename S, typename SF>
__device__ inline T feature_terms(const S* __restrict__ tf, int C, int H, int W,
                                  int64_t HW, const SF* __restrict__ sf,
                                  int64_t n, int64_t idx, T fx, T fy,
                                  const V3<T>& p, T u, T v, T j[6], T& rlast) {
#pragma clang fp contract(off)
  const T h = T(0.05);
  const T div = T(1) / (T(2) * h);
  const Bil<T> c0 = make_bil(u, v);
  T sq = T(0);
  for (int c = 0; c < C - 1; ++c) {
    const T res = bil_get(tf + c * HW, H, W, c0) - T(ldw(sf, c * n + idx));
    sq += res * res;
  }
  const S* ch = tf + int64_t(C - 1) * HW;
  const T res = bil_get(ch, H, W, c0) - T(ldw(sf, int64_t(C - 1) * n + idx));
  const Bil<T> bu0 = make_bil(u + h, v);
  const Bil<T> bu1 = make_bil(fmax(u - h, T(0)), v);
  const Bil<T> bv0 = make_bil(u, v + h);
  const Bil<T> bv1 = make_bil(u, fmax(v - h, T(0)));
  const T du = (bil_get(ch, H, W, bu0) - bil_get(ch, H, W, bu1)) * div;
  const T dv = (bil_get(ch, H, W, bv0) - bil_get(ch, H, W, bv1)) * div;
  const T z2 = p.z * p.z;
  const T j00 = fx / p.z;
  const T j02 = -p.x * fx / z2;
  const T j11 = fy / p.z;
  const T j12 = -p.y * fy / z2;
  const V3<T> g{du * j00, dv * j11, du * j02 + dv * j12};
  const V3<T> c = cross3(p, g);
  j[0] = g.x;
  j[1] = g.y;
  j[2] = g.z;
  j[3] = c.x;
  j[4] = c.y;
  j[5] = c.z;
  rlast = res;
  sq += res * res;
  return sq;
}

// ---------------------------------------------------------------------------
// Drop-in correspondence map (reference key layout, CAS merge).
// ---------------------------------------------------------------------------

__device__ inline void merge_ref(unsigned long long* addr, float dist,
                                 uint32_t idx) {
  const unsigned long long mine =
      (static_cast<unsigned long long>(idx) << 32) |
      static_cast<unsigned long long>(__float_as_uint(dist));
  unsigned long long cur =
      __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const float cd = __uint_as_float(static_cast<uint32_t>(cur & 0xffffffffull));
    const uint32_t ci = static_cast<uint32_t>(cur >> 32);
    const bool better = (dist < cd) || (dist == cd && idx < ci);
    if (!better) return;
    const unsigned long long prev = atomicCAS(addr, cur, mine);
    if (prev == cur) return;
    cur = prev;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    corresp_ref_kernel(const T* __restrict__ sp, const T* __restrict__ sn,
                       const uint8_t* __restrict__ sm, int use_mask, int64_t n,
                       const T* __restrict__ rt, const T* __restrict__ tp,
                       const T* __restrict__ tn, const uint8_t* __restrict__ tm,
                       int H, int W, const T* __restrict__ kcam,
                       unsigned long long* map, T dthr2, T cstar) {
#pragma clang fp contract(off)
  Rigid<T> rig;
  rigid_from(rt, rig);
  const CamT<T> cam = load_cam(kcam);
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    if (use_mask && !sm[idx]) continue;
    const V3<T> p = rigid_point(rig.m, ld3(sp, idx));
    const V3<T> nn = rigid_normal(rig.nm, ld3(sn, idx));
    T u, v;
    project(cam.fx, cam.fy, cam.cx, cam.cy, p, u, v);
    int ui, vi;
    if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) continue;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!tm[pos]) continue;
    const V3<T> q = ld3(tp, pos);
    const V3<T> d = sub3(q, p);
    const T sq = dot3(d, d);
    if (sq > dthr2) continue;
    const V3<T> qn = ld3(tn, pos);
    if (angle_rejects(gate_cos(nn, qn), cstar)) continue;
    merge_ref(map + pos, float(sqrt(sq)), uint32_t(idx));
  }
}

// ---------------------------------------------------------------------------
// Drop-in per-pixel Jacobians.
// ---------------------------------------------------------------------------

template <typename T>
__device__ inline void store_partial(T* __restrict__ jtj, T* __restrict__ jr,
                                     const T j[6], T r) {
#pragma clang fp contract(off)
#pragma unroll
  for (int k = 0; k < 6; ++k) jr[k] = j[k] * r;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) jtj[a * 6 + b] = j[b] * j[a];
}

template <typename T>
__device__ inline void zero_partial(T* __restrict__ jtj, T* __restrict__ jr) {
#pragma unroll
  for (int k = 0; k < 6; ++k) jr[k] = T(0);
#pragma unroll
  for (int k = 0; k < 36; ++k) jtj[k] = T(0);
}

__device__ inline bool owner_ref(const int64_t* __restrict__ map, int64_t pos,
                                 int64_t idx) {
  return int32_t(map[pos] >> 32) == int32_t(idx);
}

// Projects source idx with the 3x4 matrix m and reports its target pixel when
// the reference map names idx as that pixel's owner.
template <typename T>
__device__ inline bool owned_pixel(const T* m, const CamT<T>& cam,
                                   const T* __restrict__ sp, int64_t idx, int H,
                                   int W, const int64_t* __restrict__ map,
                                   V3<T>& p, T& u, T& v, int64_t& pos) {
  p = rigid_point(m, ld3(sp, idx));
  project(cam.fx, cam.fy, cam.cx, cam.cy, p, u, v);
  int ui, vi;
  if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) return false;
  pos = int64_t(vi) * W + ui;
  return owner_ref(map, pos, idx);
}

__global__ void zero_i32_kernel(int32_t* p) { *p = 0; }

template <typename T>
__global__ void __launch_bounds__(kBlock)
    geometric_ref_kernel(const T* __restrict__ sp,
                         const uint8_t* __restrict__ sm, int64_t n,
                         const T* __restrict__ rt,
                         const float* __restrict__ tp,
                         const float* __restrict__ tn, int H, int W,
                         const T* __restrict__ kcam,
                         const int64_t* __restrict__ map, T* __restrict__ JtJ,
                         T* __restrict__ Jr, T* __restrict__ r2,
                         int32_t* __restrict__ count) {
#pragma clang fp contract(off)
  T m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = rt[i];
  const CamT<T> cam = load_cam(kcam);
  int local = 0;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    T* jtj = JtJ + idx * 36;
    T* jr = Jr + idx * 6;
    V3<T> p;
    T u, v;
    int64_t pos = 0;
    const bool ok =
        sm[idx] != 0 && owned_pixel(m, cam, sp, idx, H, W, map, p, u, v, pos);
    if (!ok) {
      zero_partial(jtj, jr);
      r2[idx] = T(0);
      continue;
    }
    const V3<T> q{T(tp[3 * pos]), T(tp[3 * pos + 1]), T(tp[3 * pos + 2])};
    const V3<T> qn{T(tn[3 * pos]), T(tn[3 * pos + 1]), T(tn[3 * pos + 2])};
    const T r = dot3(sub3(q, p), qn);
    const V3<T> c = cross3(p, qn);
    const T j[6] = {qn.x, qn.y, qn.z, c.x, c.y, c.z};
    store_partial(jtj, jr, j, r);
    r2[idx] = r * r;
    ++local;
  }
  if (local) atomicAdd(count, local);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    feature_ref_kernel(const T* __restrict__ sp, const T* __restrict__ sf,
                       const uint8_t* __restrict__ sm, int64_t n,
                       const T* __restrict__ rt, const T* __restrict__ tf,
                       int C, int H, int W, const T* __restrict__ kcam,
                       const int64_t* __restrict__ map, float thr,
                       T* __restrict__ JtJ, T* __restrict__ Jr,
                       T* __restrict__ r2, int32_t* __restrict__ count) {
#pragma clang fp contract(off)
  T m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = rt[i];
  const CamT<T> cam = load_cam(kcam);
  const int64_t HW = int64_t(H) * W;
  const float thr2 = thr * thr;
  int local = 0;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    T* jtj = JtJ + idx * 36;
    T* jr = Jr + idx * 6;
    V3<T> p;
    T u, v;
    int64_t pos = 0;
    bool ok =
        sm[idx] != 0 && owned_pixel(m, cam, sp, idx, H, W, map, p, u, v, pos);
    T j[6], rl = T(0), sq = T(0);
    if (ok) {
      sq = feature_terms<T, T, T>(tf, C, H, W, HW, sf, n, idx, cam.fx, cam.fy,
                                  p, u, v, j, rl);
      ok = !(sq > T(thr2));
    }
    if (!ok) {
      zero_partial(jtj, jr);
      r2[idx] = T(0);
      continue;
    }
    store_partial(jtj, jr, j, rl);
    r2[idx] = sq;
    ++local;
  }
  if (local) atomicAdd(count, local);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    information_ref_kernel(const T* __restrict__ sp,
                           const uint8_t* __restrict__ sm, int64_t n,
                           const T* __restrict__ rt,
                           const float* __restrict__ tp, int H, int W,
                           const T* __restrict__ kcam,
                           const int64_t* __restrict__ map,
                           T* __restrict__ GtG) {
#pragma clang fp contract(off)
  T m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = rt[i];
  const CamT<T> cam = load_cam(kcam);
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    T* g = GtG + idx * 36;
    V3<T> p;
    T u, v;
    int64_t pos = 0;
    const bool ok =
        sm[idx] != 0 && owned_pixel(m, cam, sp, idx, H, W, map, p, u, v, pos);
    if (!ok) {
#pragma unroll
      for (int k = 0; k < 36; ++k) g[k] = T(0);
      continue;
    }
    const T x = T(tp[3 * pos]), y = T(tp[3 * pos + 1]), z = T(tp[3 * pos + 2]);
    const T r0[6] = {T(0), z, -y, T(1), T(0), T(0)};
    const T r1[6] = {-z, T(0), x, T(0), T(1), T(0)};
    const T rr[6] = {y, -x, T(0), T(0), T(0), T(1)};
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b)
        g[a * 6 + b] = r0[a] * r0[b] + r1[a] * r1[b] + rr[a] * rr[b];
  }
}

// ---------------------------------------------------------------------------
// Fused path: reductions.
// ---------------------------------------------------------------------------

constexpr int kSlots = 32;  // 21 JtJ + 6 Jr + r2_geom, r2_feat, n_geom, n_feat, 0
constexpr int kWaves = kBlock / 64;
constexpr int kGroups = kBlock / kSlots;  // slot groups of the final sum
constexpr int kShared = (kGroups > kWaves ? kGroups : kWaves) * kSlots;

// One recursive-halving step: lanes with bit `S` set keep the upper half of
// their H live slots, the others the lower half, and add the partner's copy.
template <int S, int H>
__device__ inline void halve(float (&v)[kSlots], int lane) {
  const bool up = (lane >> S) & 1;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float send = up ? v[k] : v[k + H];
    const float keep = up ? v[k + H] : v[k];
    v[k] = keep + __shfl_xor(send, 1 << S, 64);
  }
}

// After this, lane l < 32 holds the wavefront total of slot bitrev5(l) in v[0]
// (lanes 32..63 hold the same values).
__device__ inline void wave_reduce_slots(float (&v)[kSlots], int lane) {
  halve<0, 16>(v, lane);
  halve<1, 8>(v, lane);
  halve<2, 4>(v, lane);
  halve<3, 2>(v, lane);
  halve<4, 1>(v, lane);
  v[0] += __shfl_xor(v[0], 32, 64);
}

__device__ inline int bitrev5(int l) {
  return ((l & 1) << 4) | ((l & 2) << 2) | (l & 4) | ((l & 8) >> 2) |
         ((l & 16) >> 4);
}

// Adds one weighted term J' = w J, r' = s r to the JtJ / Jr slots.
__device__ inline void acc_term(float (&acc)[kSlots], const float j[6], float w,
                                float r) {
  float jw[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) jw[a] = j[a] * w;
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) {
      acc[k] = fmaf(jw[a], jw[b], acc[k]);
      ++k;
    }
#pragma unroll
  for (int a = 0; a < 6; ++a) acc[21 + a] = fmaf(jw[a], r, acc[21 + a]);
}

// Block-level reduction into FP64 partials and arrival ticket. Returns true in
// every thread of the last block to arrive; `tot` then holds the grid totals
// summed in a fixed order (independent of which block arrived last).
__device__ inline bool block_reduce_and_ticket(float (&acc)[kSlots],
                                               double* __restrict__ partials,
                                               int32_t* ticket,
                                               double* sh /* [kShared] */,
                                               double* tot /* [32] */,
                                               int* s_last) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  wave_reduce_slots(acc, lane);
  if (lane < 32) sh[wave * kSlots + bitrev5(lane)] = double(acc[0]);
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += sh[w * kSlots + threadIdx.x];
    partials[int64_t(blockIdx.x) * kSlots + threadIdx.x] = s;
    __threadfence();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = atomicAdd(ticket, 1);
    *s_last = (t == int(gridDim.x) - 1);
  }
  __syncthreads();
  if (!*s_last) return false;
  __threadfence();
  // Fixed-order FP64 sum: group g of kGroups adds blocks g, g+kGroups, ...
  const int slot = threadIdx.x & (kSlots - 1);
  const int grp = threadIdx.x / kSlots;
  double s = 0.0;
  for (int b = grp; b < int(gridDim.x); b += kGroups)
    s += partials[int64_t(b) * kSlots + slot];
  __syncthreads();
  sh[grp * kSlots + slot] = s;
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double t = 0.0;
#pragma unroll
    for (int g = 0; g < kGroups; ++g) t += sh[g * kSlots + threadIdx.x];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
  return true;
}

// ---------------------------------------------------------------------------
// Fused path: pose state.
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
__device__ inline void exp_rt_f64(const double x[6], double U[12]) {
  const double wx = x[3], wy = x[4], wz = x[5];
  const double sq = wx * wx + (wy * wy + wz * wz);
  const double th = sqrt(sq);
  double ax = wx, ay = wy, az = wz;
  if (sq > 0.0) {
    ax = wx / th;
    ay = wy / th;
    az = wz / th;
  }
  const double s = sin(th), c = cos(th), k = 1.0 - c;
  U[0] = k * ax * ax + c;
  U[1] = k * ax * ay - s * az;
  U[2] = k * ax * az + s * ay;
  U[4] = k * ax * ay + s * az;
  U[5] = k * ay * ay + c;
  U[6] = k * ay * az - s * ax;
  U[8] = k * ax * az - s * ay;
  U[9] = k * ay * az + s * ax;
  U[10] = k * az * az + c;
  U[3] = x[0];
  U[7] = x[1];
  U[11] = x[2];
}

// One Gauss-Newton update from the reduced totals (slamtb/registration/icp.py:
// 236-256): FP64 Cholesky of JtJ (a non-positive pivot is the reference's
// `except: break`), solve JtJ x = Jr, T <- exp_rt(x) T. Runs in one thread
// on matrices kept in LDS (`w` holds 36 + 36 + 6 + 12 doubles) with its
// loops left rolled, so that the kernel's register budget is set by the
// per-pixel loop rather than by this tail.
__device__ inline void gn_finalize(stb_icp_state* st, const double* tot,
                                   double gw, double fw, double* w) {
  double* A = w;
  double* L = w + 36;
  double* x = w + 72;
  double* U = w + 78;
  int k = 0;
#pragma unroll 1
  for (int a = 0; a < 6; ++a)
#pragma unroll 1
    for (int c = a; c < 6; ++c) {
      A[a * 6 + c] = tot[k];
      A[c * 6 + a] = tot[k];
      ++k;
    }
#pragma unroll 1
  for (int i = 0; i < 36; ++i) {
    st->last_jtj[i] = A[i];
    L[i] = 0.0;
  }
#pragma unroll 1
  for (int a = 0; a < 6; ++a) st->last_jr[a] = tot[21 + a];
  st->last_geom_r2 = tot[27];
  st->last_feat_r2 = tot[28];
  st->last_geom_count = tot[29];
  st->last_feat_count = tot[30];

#pragma unroll 1
  for (int j = 0; j < 6; ++j) {
    double s = A[j * 6 + j];
#pragma unroll 1
    for (int q = 0; q < j; ++q) s -= L[j * 6 + q] * L[j * 6 + q];
    if (!(s > 0.0)) {
      st->stopped = 1;
      return;
    }
    const double d = sqrt(s);
    L[j * 6 + j] = d;
#pragma unroll 1
    for (int i = j + 1; i < 6; ++i) {
      double t = A[i * 6 + j];
#pragma unroll 1
      for (int q = 0; q < j; ++q) t -= L[i * 6 + q] * L[j * 6 + q];
      L[i * 6 + j] = t / d;
    }
  }
#pragma unroll 1
  for (int i = 0; i < 6; ++i) {  // L y = b (y kept in x)
    double t = tot[21 + i];
#pragma unroll 1
    for (int q = 0; q < i; ++q) t -= L[i * 6 + q] * x[q];
    x[i] = t / L[i * 6 + i];
  }
#pragma unroll 1
  for (int i = 5; i >= 0; --i) {  // L^T x = y
    double t = x[i];
#pragma unroll 1
    for (int q = i + 1; q < 6; ++q) t -= L[q * 6 + i] * x[q];
    x[i] = t / L[i * 6 + i];
  }
  exp_rt_f64(x, U);
  // T <- U T (rows 0..2; the last row stays [0 0 0 1]); L is free again and
  // stages the new rows.
#pragma unroll 1
  for (int i = 0; i < 3; ++i)
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
      double t = U[i * 4 + 0] * st->transform[0 * 4 + j] +
                 U[i * 4 + 1] * st->transform[1 * 4 + j] +
                 U[i * 4 + 2] * st->transform[2 * 4 + j];
      if (j == 3) t += U[i * 4 + 3];
      L[i * 4 + j] = t;
    }
#pragma unroll 1
  for (int i = 0; i < 12; ++i) {
    st->transform[i] = L[i];
    st->best_transform[i] = L[i];
  }
#pragma unroll 1
  for (int i = 0; i < 36; ++i) st->best_jtj[i] = A[i];
  const double count = fmax(tot[29], tot[30]);
  st->best_residual = (fw * tot[28] + gw * tot[27]) / count;
  st->best_match_count = count;
  st->iterations += 1;
  st->map_sel ^= 1;
  set_pose_floats(st);
}

__global__ void icp_begin_kernel(stb_icp_state* st,
                                 const double* __restrict__ T0, int chain) {
  if (threadIdx.x != 0) return;
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
  st->ticket = 0;
  st->map_sel = 0;
  set_pose_floats(st);
}

// Clears the map that state.map_sel does not name (the spare one).
__global__ void __launch_bounds__(kBlock)
    fill_spare_map_kernel(const stb_icp_state* __restrict__ st,
                          unsigned long long* __restrict__ map0,
                          unsigned long long* __restrict__ map1, int64_t n) {
  unsigned long long* __restrict__ p = st->map_sel ? map0 : map1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = kKeySentinel;
}

__global__ void __launch_bounds__(kBlock)
    fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n,
                    unsigned long long value) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = value;
}

// ---------------------------------------------------------------------------
// Fused path: correspondence (internal key layout, 64-bit atomic min).
// ---------------------------------------------------------------------------

template <typename S>
__global__ void __launch_bounds__(kBlock)
    corresp_fused_kernel(stb_icp_level L, const stb_icp_state* __restrict__ st,
                         unsigned long long* __restrict__ map0,
                         unsigned long long* __restrict__ map1, float dthr2,
                         float cstar, int fixed_map) {
#pragma clang fp contract(off)
  // fixed_map: a one-off pass on the spare map that ignores `stopped`
  // (information, linearisation); otherwise the GN loop's map, the one
  // state.map_sel names.
  if (!fixed_map && st->stopped) return;
  unsigned long long* __restrict__ map =
      (st->map_sel ^ fixed_map) ? map1 : map0;
  float m[12], nm[9];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) nm[i] = st->nm[i];
  const S* __restrict__ sp = static_cast<const S*>(L.src_points);
  const S* __restrict__ sn = static_cast<const S*>(L.src_normals);
  const S* __restrict__ tp = static_cast<const S*>(L.tgt_points);
  const S* __restrict__ tn = static_cast<const S*>(L.tgt_normals);
  const uint8_t* __restrict__ sm = L.src_mask;
  const uint8_t* __restrict__ tm = L.tgt_mask;
  const int H = L.height, W = L.width;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < L.n_src; idx += int64_t(gridDim.x) * blockDim.x) {
    if (!sm[idx]) continue;
    const V3<float> p = rigid_point(m, ld3f(sp, idx));
    const V3<float> nn = rigid_normal(nm, ld3f(sn, idx));
    float u, v;
    project(L.fx, L.fy, L.cx, L.cy, p, u, v);
    int ui, vi;
    if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) continue;
    const int64_t pos = int64_t(vi) * W + ui;
    // Issue the three target loads together; the gates then run on registers.
    const uint8_t valid = tm[pos];
    const V3<float> q = ld3f(tp, pos);
    const V3<float> qn = ld3f(tn, pos);
    if (!valid) continue;
    const V3<float> d = sub3(q, p);
    const float sq = dot3(d, d);
    if (sq > dthr2) continue;
    if (angle_rejects(gate_cos(nn, qn), cstar)) continue;
    atomicMin(map + pos, (unsigned long long)make_key(sqrtf(sq), uint32_t(idx)));
  }
}

// ---------------------------------------------------------------------------
// Fused path: Jacobian + reduction + (last block) solve and pose update.
// ---------------------------------------------------------------------------

struct TermWeights {
  float wg, sg;  // geometric J scale and residual scale
  float wf, sf;  // feature J scale and residual scale
  float fthr2;   // feature residual threshold squared
  double gw, fw; // weights of the residual bookkeeping
};

template <typename S, bool GEOM, bool FEAT>
__global__ void __launch_bounds__(kBlock)
    jacobian_fused_kernel(stb_icp_level L, stb_icp_state* __restrict__ st,
                          unsigned long long* __restrict__ map0,
                          unsigned long long* __restrict__ map1,
                          int64_t map_len, double* __restrict__ partials,
                          TermWeights tw, int do_solve, int fixed_map) {
#pragma clang fp contract(off)
  __shared__ double sh[kShared];
  __shared__ double tot[kSlots];
  __shared__ int s_last;
  // GN loop (fixed_map == 0): read the map state.map_sel names and clear the
  // other one for the next iteration; the solve flips map_sel. Because the
  // choice lives on the device, a captured sequence of any length replays
  // correctly. fixed_map != 0: a one-off pass on the spare map.
  if (!fixed_map && st->stopped) return;
  const int sel = st->map_sel ^ (fixed_map ? 1 : 0);
  const unsigned long long* __restrict__ map = sel ? map1 : map0;
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  const int64_t gtid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (!fixed_map) {
    unsigned long long* __restrict__ next_map = sel ? map0 : map1;
    for (int64_t i = gtid; i < map_len; i += gsize) next_map[i] = kKeySentinel;
  }
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
  const S* __restrict__ sp = static_cast<const S*>(L.src_points);
  const S* __restrict__ tp = static_cast<const S*>(L.tgt_points);
  const S* __restrict__ tn = static_cast<const S*>(L.tgt_normals);
  const S* __restrict__ sf = static_cast<const S*>(L.src_feats);
  const S* __restrict__ tf = static_cast<const S*>(L.tgt_feats);
  const uint8_t* __restrict__ sm = L.src_mask;
  const int H = L.height, W = L.width;
  const int64_t HW = int64_t(H) * W;
  const int64_t n = L.n_src;

  float acc[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;

  for (int64_t idx = gtid; idx < n; idx += gsize) {
    if (!sm[idx]) continue;
    const V3<float> p = rigid_point(m, ld3f(sp, idx));
    float u, v;
    project(L.fx, L.fy, L.cx, L.cy, p, u, v);
    int ui, vi;
    if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) continue;
    const int64_t pos = int64_t(vi) * W + ui;
    if (uint32_t(map[pos]) != uint32_t(idx)) continue;
    if (GEOM) {
      const V3<float> q = ld3f(tp, pos);
      const V3<float> qn = ld3f(tn, pos);
      const float r = dot3(sub3(q, p), qn);
      const V3<float> c = cross3(p, qn);
      const float j[6] = {qn.x, qn.y, qn.z, c.x, c.y, c.z};
      acc_term(acc, j, tw.wg, r * tw.sg);
      acc[27] = fmaf(r, r, acc[27]);
      acc[29] += 1.0f;
    }
    if (FEAT) {
      float j[6], rl;
      const float sq = feature_terms<float, S, S>(tf, L.channels, H, W, HW, sf,
                                                  n, idx, L.fx, L.fy, p, u, v,
                                                  j, rl);
      if (!(sq > tw.fthr2)) {
        acc_term(acc, j, tw.wf, rl * tw.sf);
        acc[28] += sq;
        acc[30] += 1.0f;
      }
    }
  }
  if (!block_reduce_and_ticket(acc, partials, &st->ticket, sh, tot, &s_last))
    return;
  if (threadIdx.x == 0) {
    st->ticket = 0;
    if (do_solve) {
      gn_finalize(st, tot, tw.gw, tw.fw, sh);
    } else {
      for (int a = 0; a < 6; ++a) st->last_jr[a] = tot[21 + a];
      int k = 0;
      for (int a = 0; a < 6; ++a)
        for (int c = a; c < 6; ++c) {
          st->last_jtj[a * 6 + c] = tot[k];
          st->last_jtj[c * 6 + a] = tot[k];
          ++k;
        }
      st->last_geom_r2 = tot[27];
      st->last_feat_r2 = tot[28];
      st->last_geom_count = tot[29];
      st->last_feat_count = tot[30];
    }
  }
}

// Sum of G^T G at the current pose (src/icp_jacobian_geometric.cu:117-158),
// reduced like the Jacobian kernel; rows G0..G2 from the target point with
// parameter order [rotation, translation].
template <typename S>
__global__ void __launch_bounds__(kBlock)
    information_fused_kernel(stb_icp_level L, stb_icp_state* __restrict__ st,
                             const unsigned long long* __restrict__ map0,
                             const unsigned long long* __restrict__ map1,
                             double* __restrict__ partials) {
#pragma clang fp contract(off)
  const unsigned long long* __restrict__ map = st->map_sel ? map0 : map1;
  __shared__ double sh[kShared];
  __shared__ double tot[kSlots];
  __shared__ int s_last;
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
  const S* __restrict__ sp = static_cast<const S*>(L.src_points);
  const S* __restrict__ tp = static_cast<const S*>(L.tgt_points);
  const uint8_t* __restrict__ sm = L.src_mask;
  const int H = L.height, W = L.width;
  float acc[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < L.n_src; idx += int64_t(gridDim.x) * blockDim.x) {
    if (!sm[idx]) continue;
    const V3<float> p = rigid_point(m, ld3f(sp, idx));
    float u, v;
    project(L.fx, L.fy, L.cx, L.cy, p, u, v);
    int ui, vi;
    if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) continue;
    const int64_t pos = int64_t(vi) * W + ui;
    if (uint32_t(map[pos]) != uint32_t(idx)) continue;
    const V3<float> q = ld3f(tp, pos);
    const float r0[6] = {0.f, q.z, -q.y, 1.f, 0.f, 0.f};
    const float r1[6] = {-q.z, 0.f, q.x, 0.f, 1.f, 0.f};
    const float r2[6] = {q.y, -q.x, 0.f, 0.f, 0.f, 1.f};
    int k = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) {
        acc[k] += r0[a] * r0[b] + r1[a] * r1[b] + r2[a] * r2[b];
        ++k;
      }
  }
  if (!block_reduce_and_ticket(acc, partials, &st->ticket, sh, tot, &s_last))
    return;
  if (threadIdx.x == 0) {
    st->ticket = 0;
    int k = 0;
    for (int a = 0; a < 6; ++a)
      for (int c = a; c < 6; ++c) {
        st->information[a * 6 + c] = tot[k];
        st->information[c * 6 + a] = tot[k];
        ++k;
      }
  }
}

// ---------------------------------------------------------------------------
// Workspace layout: [state | map0 | map1 | partials]
// ---------------------------------------------------------------------------

constexpr int kMaxBlocks = 512;

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Workspace {
  stb_icp_state* st;
  unsigned long long* map[2];
  double* partials;
  int64_t map_len;
};

inline size_t ws_bytes(int64_t n_src, int32_t H, int32_t W) {
  const size_t map_len = size_t(H) * size_t(W);
  return align256(sizeof(stb_icp_state)) + 2 * align256(map_len * 8) +
         align256(size_t(kMaxBlocks) * kSlots * sizeof(double));
  (void)n_src;
}

inline Workspace carve(void* base, int32_t H, int32_t W) {
  Workspace w;
  char* p = static_cast<char*>(base);
  w.st = reinterpret_cast<stb_icp_state*>(p);
  p += align256(sizeof(stb_icp_state));
  w.map_len = int64_t(H) * W;
  w.map[0] = reinterpret_cast<unsigned long long*>(p);
  p += align256(size_t(w.map_len) * 8);
  w.map[1] = reinterpret_cast<unsigned long long*>(p);
  p += align256(size_t(w.map_len) * 8);
  w.partials = reinterpret_cast<double*>(p);
  return w;
}

// J scales and residual signs of the combined system of
// slamtb/registration/icp.py:221-234: JtJ = fw^2 JtJ_f + gw^2 JtJ_g and
// Jr = -fw Jr_f + gw Jr_g, i.e. J' = w J with r' = r (geometric) or -r
// (feature).
TermWeights make_weights(const stb_icp_params& P) {
  TermWeights tw;
  tw.wg = float(P.geom_weight);
  tw.sg = 1.0f;
  tw.wf = float(P.feat_weight);
  tw.sf = -1.0f;
  tw.fthr2 = P.feat_residual_thresh * P.feat_residual_thresh;
  tw.gw = P.geom_weight;
  tw.fw = P.feat_weight;
  return tw;
}

int fused_grid(int64_t n) {
  int g = grid_for(n);
  return g > kMaxBlocks ? kMaxBlocks : g;
}

int check_level(const stb_icp_level* L, bool need_feats) {
  STB_REQUIRE(L != nullptr, "level is NULL");
  STB_REQUIRE(L->src_points && L->src_normals && L->src_mask && L->tgt_points &&
                  L->tgt_normals && L->tgt_mask,
              "level has a NULL source/target array");
  STB_REQUIRE(L->n_src > 0 && L->n_src < (int64_t(1) << 31),
              "n_src must be in [1, 2^31), got %lld", (long long)L->n_src);
  STB_REQUIRE(L->height > 0 && L->width > 0, "bad target size %dx%d", L->width,
              L->height);
  STB_REQUIRE(L->storage == STB_F32 || L->storage == STB_BF16 ||
                  L->storage == STB_F16,
              "storage must be STB_F32, STB_BF16 or STB_F16");
  if (need_feats) {
    STB_REQUIRE(L->src_feats && L->tgt_feats && L->channels > 0,
                "feature term needs src_feats, tgt_feats and channels > 0");
  }
  return STB_OK;
}

// Phases of one GN iteration. kCorrespond and kUpdate are the two kernels of
// the loop (on the map state.map_sel names); kLinearize and kInformation are
// one-off passes on the other map, which the loop clears before its next use.
enum Phase { kCorrespond, kUpdate, kLinearize, kInformation };

template <typename S, bool GEOM, bool FEAT>
void launch_jacobian(const stb_icp_level& L, const Workspace& w,
                     const TermWeights& tw, int do_solve, int fixed_map,
                     hipStream_t stream) {
  hipLaunchKernelGGL((jacobian_fused_kernel<S, GEOM, FEAT>),
                     dim3(fused_grid(L.n_src)), dim3(kBlock), 0, stream, L,
                     w.st, w.map[0], w.map[1], w.map_len, w.partials, tw,
                     do_solve, fixed_map);
}

template <typename S>
void launch_phase(const stb_icp_level& L, const stb_icp_params& P,
                  const Workspace& w, Phase phase, hipStream_t stream) {
  const float dthr2 = float(P.distance_thresh * P.distance_thresh);
  const float cstar = angle_gate_cos_f32(P.normal_angle_thresh);
  const int gA = grid_for(L.n_src, kBlock, 4096);
  const int fixed = (phase == kLinearize || phase == kInformation) ? 1 : 0;
  if (fixed) {
    hipLaunchKernelGGL(fill_spare_map_kernel,
                       dim3(grid_for(w.map_len, kBlock, 2048)), dim3(kBlock), 0,
                       stream, w.st, w.map[0], w.map[1], w.map_len);
  }
  if (phase != kUpdate) {
    hipLaunchKernelGGL(corresp_fused_kernel<S>, dim3(gA), dim3(kBlock), 0,
                       stream, L, w.st, w.map[0], w.map[1], dthr2, cstar,
                       fixed);
  }
  if (phase == kCorrespond) return;
  if (phase == kInformation) {
    hipLaunchKernelGGL(information_fused_kernel<S>, dim3(fused_grid(L.n_src)),
                       dim3(kBlock), 0, stream, L, w.st, w.map[0], w.map[1],
                       w.partials);
    return;
  }
  const TermWeights tw = make_weights(P);
  const int solve = phase == kUpdate ? 1 : 0;
  if (P.use_geom && P.use_feat)
    launch_jacobian<S, true, true>(L, w, tw, solve, fixed, stream);
  else if (P.use_geom)
    launch_jacobian<S, true, false>(L, w, tw, solve, fixed, stream);
  else
    launch_jacobian<S, false, true>(L, w, tw, solve, fixed, stream);
}

template <typename S>
int launch_iterations(const stb_icp_level& L, const stb_icp_params& P,
                      const Workspace& w, int count, hipStream_t stream) {
  for (int it = 0; it < count; ++it) {
    launch_phase<S>(L, P, w, kCorrespond, stream);
    launch_phase<S>(L, P, w, kUpdate, stream);
  }
  return check_launch("stb_icp_iterate");
}

}  // namespace
}  // namespace stb

using namespace stb;

// ---------------------------------------------------------------------------
// Drop-in entry points
// ---------------------------------------------------------------------------

extern "C" int stb_compute_correspondence_map(
    int dtype, const void* src_points, const void* src_normals,
    const uint8_t* src_mask, int64_t src_mask_len, int64_t n_src,
    const void* rt_cam, int rt_rows, const void* tgt_points,
    const void* tgt_normals, const uint8_t* tgt_mask, int32_t height,
    int32_t width, const void* kcam, int64_t* corresp_map,
    double distance_thresh, double normal_angle_thresh, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "compute_correspondence_map: dtype must be float32 or float64");
  STB_REQUIRE(rt_rows == 3 || rt_rows == 4, "rt_cam must be [3,4] or [4,4]");
  STB_REQUIRE(n_src >= 0 && n_src < (int64_t(1) << 31), "bad n_src");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  STB_REQUIRE(corresp_map && tgt_points && tgt_normals && tgt_mask && kcam &&
                  rt_cam,
              "NULL argument");
  if (n_src == 0) return STB_OK;
  STB_REQUIRE(src_points && src_normals, "NULL source arrays");
  const int use_mask = (src_mask != nullptr && src_mask_len == n_src) ? 1 : 0;
  const int grid = grid_for(n_src, kBlock, 8192);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32) {
    hipLaunchKernelGGL(corresp_ref_kernel<float>, dim3(grid), dim3(kBlock), 0,
                       s, (const float*)src_points, (const float*)src_normals,
                       src_mask, use_mask, n_src, (const float*)rt_cam,
                       (const float*)tgt_points, (const float*)tgt_normals,
                       tgt_mask, height, width, (const float*)kcam,
                       (unsigned long long*)corresp_map,
                       float(distance_thresh * distance_thresh),
                       angle_gate_cos_f32(normal_angle_thresh));
  } else {
    hipLaunchKernelGGL(corresp_ref_kernel<double>, dim3(grid), dim3(kBlock), 0,
                       s, (const double*)src_points,
                       (const double*)src_normals, src_mask, use_mask, n_src,
                       (const double*)rt_cam, (const double*)tgt_points,
                       (const double*)tgt_normals, tgt_mask, height, width,
                       (const double*)kcam, (unsigned long long*)corresp_map,
                       double(distance_thresh * distance_thresh),
                       angle_gate_cos_f64(normal_angle_thresh));
  }
  return check_launch("compute_correspondence_map");
}

extern "C" int stb_icp_estimate_geometric(
    int dtype, const void* src_points, const uint8_t* src_mask, int64_t n_src,
    const void* rt_cam, int rt_rows, const float* tgt_points,
    const float* tgt_normals, int32_t height, int32_t width, const void* kcam,
    const int64_t* corresp_map, void* JtJ_partial, void* Jr_partial,
    void* squared_residual, int32_t* match_count, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "estimate_geometric: dtype must be float32 or float64");
  STB_REQUIRE(rt_rows == 3 || rt_rows == 4, "rt_cam must be [3,4] or [4,4]");
  STB_REQUIRE(n_src >= 0 && n_src < (int64_t(1) << 31), "bad n_src");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  STB_REQUIRE(match_count && corresp_map && rt_cam && kcam && tgt_points &&
                  tgt_normals,
              "NULL argument");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, s, match_count);
  if (n_src == 0) return check_launch("estimate_geometric");
  STB_REQUIRE(src_points && src_mask && JtJ_partial && Jr_partial &&
                  squared_residual,
              "NULL argument");
  const int grid = grid_for(n_src, kBlock, 8192);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(geometric_ref_kernel<float>, dim3(grid), dim3(kBlock), 0,
                       s, (const float*)src_points, src_mask, n_src,
                       (const float*)rt_cam, tgt_points, tgt_normals, height,
                       width, (const float*)kcam, corresp_map,
                       (float*)JtJ_partial, (float*)Jr_partial,
                       (float*)squared_residual, match_count);
  else
    hipLaunchKernelGGL(geometric_ref_kernel<double>, dim3(grid), dim3(kBlock),
                       0, s, (const double*)src_points, src_mask, n_src,
                       (const double*)rt_cam, tgt_points, tgt_normals, height,
                       width, (const double*)kcam, corresp_map,
                       (double*)JtJ_partial, (double*)Jr_partial,
                       (double*)squared_residual, match_count);
  return check_launch("estimate_geometric");
}

extern "C" int stb_icp_get_information_matrix(
    int dtype, const void* src_points, const uint8_t* src_mask, int64_t n_src,
    const void* rt_cam, int rt_rows, const float* tgt_points, int32_t height,
    int32_t width, const void* kcam, const int64_t* corresp_map,
    void* GtG_partial, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "get_information_matrix: dtype must be float32 or float64");
  STB_REQUIRE(rt_rows == 3 || rt_rows == 4, "rt_cam must be [3,4] or [4,4]");
  STB_REQUIRE(n_src >= 0 && n_src < (int64_t(1) << 31), "bad n_src");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  if (n_src == 0) return STB_OK;
  STB_REQUIRE(src_points && src_mask && rt_cam && tgt_points && kcam &&
                  corresp_map && GtG_partial,
              "NULL argument");
  const int grid = grid_for(n_src, kBlock, 8192);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(information_ref_kernel<float>, dim3(grid), dim3(kBlock),
                       0, s, (const float*)src_points, src_mask, n_src,
                       (const float*)rt_cam, tgt_points, height, width,
                       (const float*)kcam, corresp_map, (float*)GtG_partial);
  else
    hipLaunchKernelGGL(information_ref_kernel<double>, dim3(grid),
                       dim3(kBlock), 0, s, (const double*)src_points, src_mask,
                       n_src, (const double*)rt_cam, tgt_points, height, width,
                       (const double*)kcam, corresp_map, (double*)GtG_partial);
  return check_launch("get_information_matrix");
}

extern "C" int stb_icp_estimate_feature(
    int dtype, const void* src_points, const void* src_feats,
    const uint8_t* src_mask, int64_t n_src, const void* rt_cam, int rt_rows,
    const void* tgt_feats, int32_t channels, int32_t height, int32_t width,
    const void* kcam, const int64_t* corresp_map, float residual_thresh,
    void* JtJ_partial, void* Jr_partial, void* squared_residual,
    int32_t* match_count, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "estimate_feature: dtype must be float32 or float64");
  STB_REQUIRE(rt_rows == 3 || rt_rows == 4, "rt_cam must be [3,4] or [4,4]");
  STB_REQUIRE(n_src >= 0 && n_src < (int64_t(1) << 31), "bad n_src");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  STB_REQUIRE(channels > 0, "estimate_feature: channels must be > 0");
  STB_REQUIRE(match_count && corresp_map && rt_cam && kcam && tgt_feats,
              "NULL argument");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, s, match_count);
  if (n_src == 0) return check_launch("estimate_feature");
  STB_REQUIRE(src_points && src_feats && src_mask && JtJ_partial &&
                  Jr_partial && squared_residual,
              "NULL argument");
  const int grid = grid_for(n_src, kBlock, 8192);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(feature_ref_kernel<float>, dim3(grid), dim3(kBlock), 0,
                       s, (const float*)src_points, (const float*)src_feats,
                       src_mask, n_src, (const float*)rt_cam,
                       (const float*)tgt_feats, channels, height, width,
                       (const float*)kcam, corresp_map, residual_thresh,
                       (float*)JtJ_partial, (float*)Jr_partial,
                       (float*)squared_residual, match_count);
  else
    hipLaunchKernelGGL(feature_ref_kernel<double>, dim3(grid), dim3(kBlock), 0,
                       s, (const double*)src_points, (const double*)src_feats,
                       src_mask, n_src, (const double*)rt_cam,
                       (const double*)tgt_feats, channels, height, width,
                       (const double*)kcam, corresp_map, residual_thresh,
                       (double*)JtJ_partial, (double*)Jr_partial,
                       (double*)squared_residual, match_count);
  return check_launch("estimate_feature");
}

// ---------------------------------------------------------------------------
// Fused Gauss-Newton entry points
// ---------------------------------------------------------------------------

extern "C" size_t stb_icp_workspace_bytes(int64_t n_src, int32_t height,
                                          int32_t width) {
  return ws_bytes(n_src, height, width);
}

extern "C" int stb_icp_begin(void* workspace, size_t workspace_bytes,
                             const double* transform0, int32_t chain,
                             int64_t n_src, int32_t height, int32_t width,
                             void* stream) {
  STB_REQUIRE(workspace != nullptr, "workspace is NULL");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  STB_REQUIRE(workspace_bytes >= ws_bytes(n_src, height, width),
              "workspace too small: %zu < %zu bytes", workspace_bytes,
              ws_bytes(n_src, height, width));
  const Workspace w = carve(workspace, height, width);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(icp_begin_kernel, dim3(1), dim3(64), 0, s, w.st,
                     transform0, chain);
  for (int k = 0; k < 2; ++k)
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(w.map_len, kBlock, 2048)),
                       dim3(kBlock), 0, s, w.map[k], w.map_len,
                       (unsigned long long)kKeySentinel);
  return check_launch("stb_icp_begin");
}

namespace {

// Validates the common arguments of the fused entry points and dispatches on
// the storage type.
int run_fused(const stb_icp_level* level, const stb_icp_params* params,
              void* workspace, size_t workspace_bytes, int phase, int count,
              void* stream, const char* what) {
  STB_REQUIRE(params != nullptr && workspace != nullptr, "%s: NULL argument",
              what);
  const bool jac = phase == kUpdate || phase == kLinearize || phase < 0;
  int rc = check_level(level, jac && params->use_feat != 0);
  if (rc) return rc;
  if (jac) {
    STB_REQUIRE(params->use_geom || params->use_feat,
                "%s: at least one of the geometric and feature terms must be on",
                what);
  }
  STB_REQUIRE(workspace_bytes >=
                  ws_bytes(level->n_src, level->height, level->width),
              "%s: workspace too small", what);
  const Workspace w = carve(workspace, level->height, level->width);
  hipStream_t s = as_stream(stream);
  if (phase < 0) {  // whole iterations
    switch (level->storage) {
      case STB_F32:
        return launch_iterations<float>(*level, *params, w, count, s);
      case STB_BF16:
        return launch_iterations<__hip_bfloat16>(*level, *params, w, count, s);
      default:
        return launch_iterations<__half>(*level, *params, w, count, s);
    }
  }
  switch (level->storage) {
    case STB_F32:
      launch_phase<float>(*level, *params, w, Phase(phase), s);
      break;
    case STB_BF16:
      launch_phase<__hip_bfloat16>(*level, *params, w, Phase(phase), s);
      break;
    default:
      launch_phase<__half>(*level, *params, w, Phase(phase), s);
  }
  return check_launch(what);
}

}  // namespace

extern "C" int stb_icp_iterate(const stb_icp_level* level,
                               const stb_icp_params* params, void* workspace,
                               size_t workspace_bytes, int32_t num_iters,
                               void* stream) {
  STB_REQUIRE(num_iters >= 0, "stb_icp_iterate: num_iters must be >= 0");
  if (num_iters == 0) return STB_OK;
  return run_fused(level, params, workspace, workspace_bytes, -1, num_iters,
                   stream, "stb_icp_iterate");
}

extern "C" int stb_icp_correspond(const stb_icp_level* level,
                                  const stb_icp_params* params,
                                  void* workspace, size_t workspace_bytes,
                                  void* stream) {
  return run_fused(level, params, workspace, workspace_bytes, kCorrespond, 1,
                   stream, "stb_icp_correspond");
}

extern "C" int stb_icp_update(const stb_icp_level* level,
                              const stb_icp_params* params, void* workspace,
                              size_t workspace_bytes, void* stream) {
  return run_fused(level, params, workspace, workspace_bytes, kUpdate, 1,
                   stream, "stb_icp_update");
}

extern "C" int stb_icp_linearize(const stb_icp_level* level,
                                 const stb_icp_params* params, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  return run_fused(level, params, workspace, workspace_bytes, kLinearize, 1,
                   stream, "stb_icp_linearize");
}

extern "C" int stb_icp_information(const stb_icp_level* level,
                                   const stb_icp_params* params,
                                   void* workspace, size_t workspace_bytes,
                                   void* stream) {
  return run_fused(level, params, workspace, workspace_bytes, kInformation, 1,
                   stream, "stb_icp_information");
}
