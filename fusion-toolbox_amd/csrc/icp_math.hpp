// Per-pixel ICP math shared by the drop-in and the fused kernels: camera,
// angle gate, bilinear feature sampling and the feature residual/Jacobian.
// Every helper keeps the reference's IEEE operation order with FMA
// contraction off, so the decisions match the CPU restatement bit for bit.
#pragma once

#include "stb_common.hpp"

namespace stb {

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

// The four taps are loaded unconditionally (a tap past the right or bottom
// edge re-reads an in-range texel and is then replaced by zero), so they
// issue together with no branches between them. The cell (ui, vi) is in
// range for every caller: u and v lie in [-0.5, W - 0.5) x [-0.5, H - 0.5)
// (plus at most h < 0.5), which truncates into the image.
template <typename T, typename S>
__device__ inline void bil_taps(const S* __restrict__ ch, int H, int W,
                                const Bil<T>& b, T t[4]) {
  const int64_t o = int64_t(b.vi) * W + b.ui;
  const bool rx = b.ui + 1 < W, ry = b.vi + 1 < H;
  const int64_t dx = rx ? 1 : 0, dy = ry ? W : 0;
  t[0] = T(ldw(ch, o));
  t[1] = T(ldw(ch, o + dx));
  t[2] = T(ldw(ch, o + dy));
  t[3] = T(ldw(ch, o + dy + dx));
}

template <typename T>
__device__ inline T bil_mix(int H, int W, const Bil<T>& b, const T t[4]) {
#pragma clang fp contract(off)
  const bool rx = b.ui + 1 < W, ry = b.vi + 1 < H;
  const T v00 = t[0];
  const T v10 = rx ? t[1] : T(0);
  const T v01 = ry ? t[2] : T(0);
  const T v11 = (rx && ry) ? t[3] : T(0);
  const T a = v00 * (T(1) - b.ur) + v10 * b.ur;
  const T c = v01 * (T(1) - b.ur) + v11 * b.ur;
  return a * (T(1) - b.vr) + c * b.vr;
}

// Bilinear sample of one channel (src/feature_map.hpp:29-42).
template <typename T, typename S>
__device__ inline T bil_get(const S* __restrict__ ch, int H, int W,
                            const Bil<T>& b) {
  T t[4];
  bil_taps(ch, H, W, b, t);
  return bil_mix(H, W, b, t);
}

// Feature residuals and the last channel's Jacobian at one matched pixel
// (src/icp_jacobian_feature.cu:81-106, src/feature_map.hpp:45-64,
// include/slamtb/camera.hpp:65-78). Returns the channel-summed squared
// residual; j/rlast describe the last channel (the one the reference keeps).
// FAST (the fused path) forms the projection derivatives from one hardware
// reciprocal of z (about 1 ulp) instead of four correctly rounded divides;
// the Jacobian values only enter sums checked to 1e-5, while the decisions
// (projection, gates, owner) keep exact IEEE division everywhere.
template <typename T, typename S, typename SF, bool FAST = false, int NB = 4>
__device__ inline T feature_terms(const S* __restrict__ tf, int C, int H, int W,
                                  int64_t HW, const SF* __restrict__ sf,
                                  int64_t n, int64_t idx, T fx, T fy,
                                  const V3<T>& p, T u, T v, T j[6], T& rlast) {
#pragma clang fp contract(off)
  const T h = T(0.05);
  const T div = T(1) / (T(2) * h);
  const Bil<T> c0 = make_bil(u, v);
  const Bil<T> bu0 = make_bil(u + h, v);
  const Bil<T> bu1 = make_bil(fmax(u - h, T(0)), v);
  const Bil<T> bv0 = make_bil(u, v + h);
  const Bil<T> bv1 = make_bil(u, fmax(v - h, T(0)));
  // The last channel's five samples come first in program order so their
  // loads issue before the other channels' batches wait.
  const S* ch = tf + int64_t(C - 1) * HW;
  const T last = bil_get(ch, H, W, c0);
  const T su0 = bil_get(ch, H, W, bu0), su1 = bil_get(ch, H, W, bu1);
  const T sv0 = bil_get(ch, H, W, bv0), sv1 = bil_get(ch, H, W, bv1);
  const T flast = T(ldw(sf, int64_t(C - 1) * n + idx));
  // Channels 0 .. C-2 in batches of NB (4, or 2 for few channels) whose
  // loads issue together; a batch slot past C-2 re-reads channel C-2 and
  // adds +0, which leaves the running sum (and the reference's summation
  // order) unchanged.
  T sq = T(0);
  const int last_c = C - 2 > 0 ? C - 2 : 0;
  for (int c0i = 0; c0i < C - 1; c0i += NB) {
    // All the batch's taps and source values load before any is used.
    T taps[NB][4], fs[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int c = c0i + k < last_c ? c0i + k : last_c;
      bil_taps(tf + c * HW, H, W, c0, taps[k]);
      fs[k] = T(ldw(sf, c * n + idx));
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const T res = bil_mix(H, W, c0, taps[k]) - fs[k];
      sq += (c0i + k < C - 1) ? res * res : T(0);
    }
  }
  const T res = last - flast;
  const T du = (su0 - su1) * div;
  const T dv = (sv0 - sv1) * div;
  T j00, j02, j11, j12;
  if constexpr (FAST) {
    const T iz = __builtin_amdgcn_rcpf(p.z);
    const T iz2 = iz * iz;
    j00 = fx * iz;
    j02 = -p.x * fx * iz2;
    j11 = fy * iz;
    j12 = -p.y * fy * iz2;
  } else {
    const T z2 = p.z * p.z;
    j00 = fx / p.z;
    j02 = -p.x * fx / z2;
    j11 = fy / p.z;
    j12 = -p.y * fy / z2;
  }
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

}  // namespace stb
