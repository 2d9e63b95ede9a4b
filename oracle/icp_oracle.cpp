// ORACLE — test infrastructure only.
//
// A sequential CPU restatement of slamtb's frame-to-frame ICP Gauss-Newton
// step and its preprocessing, written from the reference's described
// algorithm. Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark load this library, and only as the checker; the product path
// (libslamtb_hip.so) never calls it.
//
// Arithmetic contract (what makes bit-exact comparison possible):
//   * compiled with -ffp-contract=off and without -ffast-math, so every
//     float operation is a single IEEE-rounded add/mul/div/sqrt;
//   * length-3 dot products and squared norms sum as a0 + (a1 + a2), the
//     order of Eigen's fixed-size half-split reduction used by the
//     reference (Eigen redux over Vector3);
//   * the rigid transform sums left to right, m0*p0 + m1*p1 + m2*p2 + m3
//     (include/slamtb/camera.hpp:134-145);
//   * projection is fx*x/z + cx (include/slamtb/camera.hpp:53-60).
//
// Ties in the correspondence map: the sequential loop keeps the first
// writer under a strict "<", i.e. the lowest source index among equal
// distances. The reference's OpenMP CPU path leaves that order to the
// scheduler (src/merge_map.hpp:37-50); the parallel variant here, used only
// for timing, keeps that behaviour.
//
// Parity status of each function (the checks live in tests/):
//   corresp_map, geometric, feature, information: pinned by closed forms,
//     brute force and invariants (test_oracle_*.py);
//   exp_rt / log / exp backward: pinned by Rodrigues closed form, orthogonality,
//     round trip and finite differences;
//   bilateral, normals, erode, downsample: pinned by closed forms on hand
//     inputs (constant depth, planes, spheres, hand bitmaps);
//   gaussian blur 3x3: pinned only against the normalised Gaussian formula;
//     the border rule of the reference's kornia call is parity unpinned;
//   rigid_points / rigid_backward: pinned against matrix products, the
//     inverse-transpose of random matrices and finite differences
//     (test_oracle_matching.py);
//   knn / nn_forward / nn_epsilon / nn_backward / voxel_merge: pinned by
//     numpy brute force (sorted distances), closed forms (equidistant
//     neighbours average, a linear feature field's exact difference
//     quotient) and hand voxel layouts (test_oracle_pcl.py);
//   fpcl_forward / fpcl_backward / project_*: pinned by closed forms
//     (features linear in u and v, whose bilinear sample and central
//     difference are exact), a brute-force matcher over the map, and finite
//     differences of the projection (test_oracle_matching.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
inline V3<T> load3(const T* p, int64_t i) {
  return V3<T>{p[3 * i + 0], p[3 * i + 1], p[3 * i + 2]};
}

template <typename T>
inline V3<T> sub(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}

// Eigen fixed-size Vector3 dot: a0*b0 + (a1*b1 + a2*b2).
template <typename T>
inline T dot3(const V3<T>& a, const V3<T>& b) {
  const T p0 = a.x * b.x;
  const T p1 = a.y * b.y;
  const T p2 = a.z * b.z;
  return p0 + (p1 + p2);
}

template <typename T>
inline T sqnorm3(const V3<T>& a) {
  return dot3(a, a);
}

template <typename T>
inline V3<T> cross3(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x};
}

// RigidTransform: 3x4 matrix plus the inverse-transpose normal matrix built
// from cofactors (include/slamtb/camera.hpp:99-129), in the scalar type.
template <typename T>
struct Rigid {
  T m[12];
  T nm[9];

  Rigid(const T* rt) {
    for (int i = 0; i < 12; ++i) m[i] = rt[i];
    auto A = [&](int r, int c) { return m[r * 4 + c]; };
    const T det = +A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
                  A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                  A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    const T inv = T(1) / det;
    nm[0 * 3 + 0] = (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) * inv;
    nm[1 * 3 + 0] = -(A(0, 1) * A(2, 2) - A(0, 2) * A(2, 1)) * inv;
    nm[2 * 3 + 0] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * inv;
    nm[0 * 3 + 1] = -(A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) * inv;
    nm[1 * 3 + 1] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * inv;
    nm[2 * 3 + 1] = -(A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2)) * inv;
    nm[0 * 3 + 2] = (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1)) * inv;
    nm[1 * 3 + 2] = -(A(0, 0) * A(2, 1) - A(2, 0) * A(0, 1)) * inv;
    nm[2 * 3 + 2] = (A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1)) * inv;
  }

  V3<T> transform(const V3<T>& p) const {
    return V3<T>{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
                 m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                 m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
  }

  V3<T> transform_normal(const V3<T>& n) const {
    return V3<T>{nm[0] * n.x + nm[1] * n.y + nm[2] * n.z,
                 nm[3] * n.x + nm[4] * n.y + nm[5] * n.z,
                 nm[6] * n.x + nm[7] * n.y + nm[8] * n.z};
  }
};

template <typename T>
struct Cam {
  T fx, fy, cx, cy;
  explicit Cam(const T* k) : fx(k[0]), fy(k[4]), cx(k[2]), cy(k[5]) {}
  // KCamera::Project (include/slamtb/camera.hpp:53-60).
  void project(const V3<T>& p, T& u, T& v) const {
    u = fx * p.x / p.z + cx;
    v = fy * p.y / p.z + cy;
  }
};

// int(round(x)) of the reference, with non-finite and out-of-int-range values
// mapped to INT_MIN, which is what x86 cvttss2si produces and which every
// bounds check then rejects.
template <typename T>
inline int round_to_int(T x) {
  const T r = std::round(x);
  if (!(r >= T(-2147483648.0) && r < T(2147483648.0))) return INT32_MIN;
  return int(r);
}

template <typename T>
inline int trunc_to_int(T x) {
  if (!(x > T(-2147483649.0) && x < T(2147483648.0))) return INT32_MIN;
  return int(x);
}

std::mutex g_map_mutex;

inline uint32_t float_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

inline float bits_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// MergeMap<kCPU>::Set (src/merge_map.hpp:37-50): keep the candidate if its
// distance is strictly below the stored one. Packed as (index << 32) | bits.
inline void merge_set(int64_t* map, int64_t pos, float dist, int32_t index) {
  uint64_t* addr = reinterpret_cast<uint64_t*>(map + pos);
  const uint64_t value =
      (uint64_t(uint32_t(index)) << 32) | uint64_t(float_bits(dist));
  const uint64_t cur = *addr;
  if (dist < bits_float(uint32_t(cur & 0xffffffffull))) *addr = value;
}

// CorrespondenceMapKernel::operator() (src/correspondence_map.cu:38-61) with
// RobustCorrespondence::Match (src/correspondence.hpp:69-89) and
// GetVectorsAngle (include/slamtb/math.hpp:14-18).
template <typename T>
void corresp_map(const T* src_pts, const T* src_nrm, const uint8_t* src_mask,
                 int64_t mask_len, int64_t n, const T* rt, const T* tgt_pts,
                 const T* tgt_nrm, const uint8_t* tgt_mask, int H, int W,
                 const T* kcam, int64_t* map, double dist_thresh,
                 double angle_thresh, int parallel) {
  const Rigid<T> rig(rt);
  const Cam<T> cam(kcam);
  const T dthr2 = T(dist_thresh * dist_thresh);
  const T athr = T(angle_thresh);
  const bool use_mask = (mask_len == n);

  auto body = [&](int64_t idx, bool locked) {
    if (use_mask && !src_mask[idx]) return;
    const V3<T> tp = rig.transform(load3(src_pts, idx));
    const V3<T> tn = rig.transform_normal(load3(src_nrm, idx));
    T u, v;
    cam.project(tp, u, v);
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) return;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!tgt_mask[pos]) return;
    const V3<T> q = load3(tgt_pts, pos);
    const T sq = sqnorm3(sub(q, tp));
    if (sq > dthr2) return;
    const V3<T> qn = load3(tgt_nrm, pos);
    const T c = dot3(tn, qn) / (std::sqrt(sqnorm3(tn)) * std::sqrt(sqnorm3(qn)));
    const T angle = std::fabs(std::acos(c));
    if (angle >= athr) return;
    // dist = (target - Tsrc).norm(); the squared norm is the same value.
    const T dist = std::sqrt(sqnorm3(sub(q, tp)));
    if (locked) {
      std::lock_guard<std::mutex> lock(g_map_mutex);
      merge_set(map, pos, float(dist), int32_t(idx));
    } else {
      merge_set(map, pos, float(dist), int32_t(idx));
    }
  };

  if (parallel) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) body(i, true);
  } else {
    for (int64_t i = 0; i < n; ++i) body(i, false);
  }
}

inline bool owns(const int64_t* map, int64_t pos, int64_t idx) {
  // MergeMapAccessor (src/merge_map.hpp:84-98): owner = map >> 32.
  return int32_t(map[pos] >> 32) == int32_t(idx);
}

// SE3ICPJacobian::Compute (src/icp_jacobian_common.hpp:37-61): assigns the
// per-pixel J r and J J^T with J = [n; p x n].
template <typename T>
inline void se3_compute(const V3<T>& tp, const V3<T>& n, T r, T* JtJ, T* Jr) {
  T j[6];
  j[0] = n.x;
  j[1] = n.y;
  j[2] = n.z;
  const V3<T> rt = cross3(tp, n);
  j[3] = rt.x;
  j[4] = rt.y;
  j[5] = rt.z;
  for (int k = 0; k < 6; ++k) Jr[k] = j[k] * r;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) JtJ[a * 6 + b] = j[b] * j[a];
}

// GeometricJacobianKernel::operator() (src/icp_jacobian_geometric.cu:54-85).
// Target maps are float32 whatever the scalar type (:22-23).
template <typename T>
int geometric(const T* src_pts, const uint8_t* src_mask, int64_t n,
              const T* rt, const float* tgt_pts, const float* tgt_nrm, int H,
              int W, const T* kcam, const int64_t* map, T* JtJ, T* Jr, T* r2,
              int parallel) {
  const Rigid<T> rig(rt);
  const Cam<T> cam(kcam);
  int64_t count = 0;
  auto body = [&](int64_t idx) -> int {
    T* jtj = JtJ + idx * 36;
    T* jr = Jr + idx * 6;
    for (int k = 0; k < 36; ++k) jtj[k] = T(0);
    for (int k = 0; k < 6; ++k) jr[k] = T(0);
    r2[idx] = T(0);
    if (!src_mask[idx]) return 0;
    const V3<T> tp = rig.transform(load3(src_pts, idx));
    T u, v;
    cam.project(tp, u, v);
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) return 0;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!owns(map, pos, idx)) return 0;
    const V3<T> q{T(tgt_pts[3 * pos]), T(tgt_pts[3 * pos + 1]),
                  T(tgt_pts[3 * pos + 2])};
    const V3<T> qn{T(tgt_nrm[3 * pos]), T(tgt_nrm[3 * pos + 1]),
                   T(tgt_nrm[3 * pos + 2])};
    const T r = dot3(sub(q, tp), qn);
    se3_compute(tp, qn, r, jtj, jr);
    r2[idx] = r * r;
    return 1;
  };
  if (parallel) {
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t i = 0; i < n; ++i) count += body(i);
  } else {
    for (int64_t i = 0; i < n; ++i) count += body(i);
  }
  return int(count);
}

// BilinearInterp (src/feature_map.hpp:8-43): truncating cell index, zero
// padding past the right and bottom edges.
template <typename T>
struct Bilinear {
  int ui, vi;
  T ur, vr;
  Bilinear(T u, T v) {
    ui = trunc_to_int(u);
    vi = trunc_to_int(v);
    ur = u - T(ui);
    vr = v - T(vi);
  }
  T get(const T* ch, int H, int W) const {
    const T v00 = ch[int64_t(vi) * W + ui];
    const T v10 = (ui + 1 < W) ? ch[int64_t(vi) * W + ui + 1] : T(0);
    const T v01 = (vi + 1 < H) ? ch[int64_t(vi + 1) * W + ui] : T(0);
    const T v11 =
        (ui + 1 < W && vi + 1 < H) ? ch[int64_t(vi + 1) * W + ui + 1] : T(0);
    const T a = v00 * (T(1) - ur) + v10 * ur;
    const T b = v01 * (T(1) - ur) + v11 * ur;
    return a * (T(1) - vr) + b * vr;
  }
};

// FeatureJacobianKernel::operator() (src/icp_jacobian_feature.cu:58-109) with
// BilinearInterpGrad (src/feature_map.hpp:45-64, h = 0.05) and
// KCamera::Dx_Project (include/slamtb/camera.hpp:65-78). Compute() assigns,
// so J^T J and J r keep the last channel while the squared residual sums all.
template <typename T>
int feature(const T* src_pts, const T* src_feats, const uint8_t* src_mask,
            int64_t n, const T* rt, const T* tgt_feats, int C, int H, int W,
            const T* kcam, const int64_t* map, float residual_thresh, T* JtJ,
            T* Jr, T* r2, int parallel) {
  const Rigid<T> rig(rt);
  const Cam<T> cam(kcam);
  const T h = T(0.05);
  const T div = T(1) / (T(2) * h);
  const int64_t HW = int64_t(H) * W;
  int64_t count = 0;
  auto body = [&](int64_t idx) -> int {
    T* jtj = JtJ + idx * 36;
    T* jr = Jr + idx * 6;
    for (int k = 0; k < 36; ++k) jtj[k] = T(0);
    for (int k = 0; k < 6; ++k) jr[k] = T(0);
    r2[idx] = T(0);
    if (!src_mask[idx]) return 0;
    const V3<T> tp = rig.transform(load3(src_pts, idx));
    T u, v;
    cam.project(tp, u, v);
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) return 0;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!owns(map, pos, idx)) return 0;

    const Bilinear<T> c0(u, v);
    const Bilinear<T> bu0(u + h, v);
    const Bilinear<T> bu1(std::max(u - h, T(0)), v);
    const Bilinear<T> bv0(u, v + h);
    const Bilinear<T> bv1(u, std::max(v - h, T(0)));

    T sq = T(0);
    for (int c = 0; c < C; ++c) {
      const T* ch = tgt_feats + c * HW;
      const T res = c0.get(ch, H, W) - src_feats[c * n + idx];
      const T du = (bu0.get(ch, H, W) - bu1.get(ch, H, W)) * div;
      const T dv = (bv0.get(ch, H, W) - bv1.get(ch, H, W)) * div;
      const T z = tp.z;
      const T z2 = z * z;
      const T j00 = cam.fx / z;
      const T j02 = -tp.x * cam.fx / z2;
      const T j11 = cam.fy / z;
      const T j12 = -tp.y * cam.fy / z2;
      const V3<T> g{du * j00, dv * j11, du * j02 + dv * j12};
      se3_compute(tp, g, res, jtj, jr);
      sq += res * res;
    }
    if (sq > residual_thresh * residual_thresh) {
      for (int k = 0; k < 36; ++k) jtj[k] = T(0);
      for (int k = 0; k < 6; ++k) jr[k] = T(0);
      return 0;
    }
    r2[idx] = sq;
    return 1;
  };
  if (parallel) {
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t i = 0; i < n; ++i) count += body(i);
  } else {
    for (int64_t i = 0; i < n; ++i) count += body(i);
  }
  return int(count);
}

// ForwardKernel::operator() of FPCLMatcherOp (src/fpcl_matcher.cu:49-83):
// the rounded projection must land on a target pixel that the map assigns to
// this very source index; then the target point and normal at that pixel and
// the bilinear feature sample at the unrounded projection are copied out.
template <typename T>
void fpcl_forward(const T* tp, const T* tn, const T* tf, int C, int H, int W,
                  const T* sp, int64_t n, const T* kcam, const int64_t* map,
                  T* op, T* on, T* of, uint8_t* mask) {
  const Cam<T> cam(kcam);
  const int64_t HW = int64_t(H) * W;
  for (int64_t i = 0; i < n; ++i) {
    mask[i] = 0;
    T u, v;
    cam.project(load3(sp, i), u, v);
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) continue;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!owns(map, pos, i)) continue;
    mask[i] = 1;
    for (int k = 0; k < 3; ++k) {
      op[3 * i + k] = tp[3 * pos + k];
      on[3 * i + k] = tn[3 * pos + k];
    }
    const Bilinear<T> b(u, v);
    for (int c = 0; c < C; ++c) of[c * n + i] = b.get(tf + c * HW, H, W);
  }
}

// BackwardKernel::operator() of FPCLMatcherOp (src/fpcl_matcher.cu:155-184):
// the feature gradient in (u, v) by central differences of step h
// (BilinearInterpGrad, src/feature_map.hpp:45-64), contracted with the
// incoming gradient over channels, then through Dx_Project
// (include/slamtb/camera.hpp:65-78).
template <typename T>
void fpcl_backward(const T* tf, int C, int H, int W, const T* sp, int64_t n,
                   const uint8_t* mask, const T* dl, const T* kcam, T h,
                   T* dx) {
  const Cam<T> cam(kcam);
  const int64_t HW = int64_t(H) * W;
  const T div = T(1) / (T(2) * h);
  for (int64_t i = 0; i < n; ++i) {
    dx[3 * i] = dx[3 * i + 1] = dx[3 * i + 2] = T(0);
    if (!mask[i]) continue;
    const V3<T> p = load3(sp, i);
    T u, v;
    cam.project(p, u, v);
    // Never false for a mask made by fpcl_forward; it keeps a caller-made
    // mask from sampling outside the image.
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) continue;
    const Bilinear<T> bu0(u + h, v);
    const Bilinear<T> bu1(std::max(u - h, T(0)), v);
    const Bilinear<T> bv0(u, v + h);
    const Bilinear<T> bv1(u, std::max(v - h, T(0)));
    T gu = T(0), gv = T(0);
    for (int c = 0; c < C; ++c) {
      const T* ch = tf + c * HW;
      const T du = (bu0.get(ch, H, W) - bu1.get(ch, H, W)) * div;
      const T dv = (bv0.get(ch, H, W) - bv1.get(ch, H, W)) * div;
      const T g = dl[c * n + i];
      gu += du * g;
      gv += dv * g;
    }
    const T z2 = p.z * p.z;
    const T j00 = cam.fx / p.z;
    const T j02 = -p.x * cam.fx / z2;
    const T j11 = cam.fy / p.z;
    const T j12 = -p.y * cam.fy / z2;
    dx[3 * i] = j00 * gu;
    dx[3 * i + 1] = j11 * gv;
    dx[3 * i + 2] = j02 * gu + j12 * gv;
  }
}

// ProjectKernel and ProjectBackwardKernel (src/camera.cu:22-29, 78-87).
template <typename T>
void project_forward(const T* pts, int64_t n, const T* kcam, T* uv) {
  const Cam<T> cam(kcam);
  for (int64_t i = 0; i < n; ++i) cam.project(load3(pts, i), uv[2 * i], uv[2 * i + 1]);
}

template <typename T>
void project_backward(const T* dy, const T* pts, int64_t n, const T* kcam,
                      T* dx) {
  const Cam<T> cam(kcam);
  for (int64_t i = 0; i < n; ++i) {
    const V3<T> p = load3(pts, i);
    const T z2 = p.z * p.z;
    const T j00 = cam.fx / p.z;
    const T j02 = -p.x * cam.fx / z2;
    const T j11 = cam.fy / p.z;
    const T j12 = -p.y * cam.fy / z2;
    dx[3 * i] = j00 * dy[2 * i];
    dx[3 * i + 1] = j11 * dy[2 * i + 1];
    dx[3 * i + 2] = j02 * dy[2 * i] + j12 * dy[2 * i + 1];
  }
}

// RigidTransformOp (src/camera.cu:134-240): points through [R | t] and
// normals through the cofactor inverse-transpose, and the backward of the
// point transform: dx = R^T dy, dM[r][c] = sum_i dy_i[r] [p_i; 1][c] summed
// sequentially in double.
template <typename T>
void rigid_points(const T* m, const T* in, int64_t n, T* out, int normals) {
  const Rigid<T> rig(m);
  for (int64_t i = 0; i < n; ++i) {
    const V3<T> p = load3(in, i);
    const V3<T> q = normals ? rig.transform_normal(p) : rig.transform(p);
    out[3 * i] = q.x;
    out[3 * i + 1] = q.y;
    out[3 * i + 2] = q.z;
  }
}

template <typename T>
void rigid_backward(const T* m, const T* p, const T* dy, int64_t n, T* dx,
                    T* dM) {
  double acc[12] = {0};
  for (int64_t i = 0; i < n; ++i) {
    const V3<T> g = load3(dy, i);
    const V3<T> q = load3(p, i);
    dx[3 * i] = m[0] * g.x + m[4] * g.y + m[8] * g.z;
    dx[3 * i + 1] = m[1] * g.x + m[5] * g.y + m[9] * g.z;
    dx[3 * i + 2] = m[2] * g.x + m[6] * g.y + m[10] * g.z;
    const double gr[3] = {double(g.x), double(g.y), double(g.z)};
    const double pc[3] = {double(q.x), double(q.y), double(q.z)};
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) acc[r * 4 + c] += gr[r] * pc[c];
      acc[r * 4 + 3] += gr[r];
    }
  }
  for (int k = 0; k < 12; ++k) dM[k] = T(acc[k]);
}

// Exact k nearest neighbours by brute force: squared distance dx*dx + dy*dy +
// dz*dz (left to right), strictly below bound2, ordered by distance then by
// target index; missing slots (inf, m) as the k-d tree of
// slamtb/spatial/kdtree_layer.py:40-60 reports them.
template <typename T>
void knn(const T* tgt, int64_t m, const T* qry, int64_t n, int k, T bound2,
         T* out_d, int64_t* out_i) {
  std::vector<std::pair<T, int64_t>> best;
  for (int64_t q = 0; q < n; ++q) {
    best.clear();
    const V3<T> p = load3(qry, q);
    for (int64_t j = 0; j < m; ++j) {
      const T dx = p.x - tgt[3 * j], dy = p.y - tgt[3 * j + 1],
              dz = p.z - tgt[3 * j + 2];
      const T d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < bound2) best.emplace_back(d2, j);
    }
    std::stable_sort(best.begin(), best.end(),
                     [](const std::pair<T, int64_t>& a,
                        const std::pair<T, int64_t>& b) { return a.first < b.first; });
    for (int s = 0; s < k; ++s) {
      const bool have = s < int(best.size());
      out_d[q * k + s] = have ? best[s].first : std::numeric_limits<T>::infinity();
      out_i[q * k + s] = have ? best[s].second : m;
    }
  }
}

// NearestNeighborsOp (src/nearest_neighbors.cu): ForwardKernel :13-54,
// EpsilonDistancesKernel :86-143, NNGrad + BackwardKernel :166-262.
template <typename T>
void nn_forward(const T* dist, const int64_t* index, int64_t n, int k,
                const T* feats, int C, int64_t m, T* out) {
  for (int64_t q = 0; q < n; ++q) {
    T total = T(0);
    for (int s = 0; s < k; ++s)
      if (dist[q * k + s] < std::numeric_limits<T>::infinity())
        total += T(1) / (T(1) + dist[q * k + s]);
    for (int c = 0; c < C; ++c) {
      T f = T(0);
      for (int s = 0; s < k; ++s) {
        const int64_t j = index[q * k + s];
        if (j >= m) break;
        const T sim = T(1) / (T(1) + dist[q * k + s]);
        const T w = sim / total;
        f += feats[c * m + j] * w;
      }
      out[c * n + q] = f;
    }
  }
}

template <typename T>
void nn_epsilon(const T* tgt, int64_t m, const T* src, const int64_t* index,
                int64_t n, int k, T h, T* eps) {
  for (int64_t row = 0; row < n; ++row) {
    for (int col = 0; col < 6; ++col) {
      T e[3] = {T(0), T(0), T(0)};
      e[col / 2] = (col % 2 == 0) ? h : -h;
      const V3<T> s = load3(src, row);
      const V3<T> x{e[0] + s.x, e[1] + s.y, e[2] + s.z};
      T* out = eps + (row * 6 + col) * k;
      T total = T(0);
      int found = 0;
      for (int j = 0; j < k; ++j) {
        const int64_t t = index[row * k + j];
        if (t > m - 1) break;
        const V3<T> d = sub(load3(tgt, t), x);
        const T sim = T(1) / (T(1) + dot3(d, d));
        out[j] = sim;
        total += sim;
        ++found;
      }
      if (found == 0) {
        out[0] = T(1);
        continue;
      }
      for (int j = 0; j < found; ++j) out[j] = out[j] / total;
    }
  }
}

template <typename T>
void nn_backward(const T* eps, const int64_t* index, int64_t n, int k,
                 const T* feats, int C, int64_t m, const T* dl, T h, T* dx) {
  const T div = T(1) / (T(2) * h);
  for (int64_t q = 0; q < n; ++q) {
    T g3[3] = {T(0), T(0), T(0)};
    for (int c = 0; c < C; ++c) {
      T avg[6];
      for (int w = 0; w < 6; ++w) {
        T a = T(0);
        for (int s = 0; s < k; ++s) {
          const int64_t j = index[q * k + s];
          if (j < m) a += feats[c * m + j] * eps[(q * 6 + w) * k + s];
        }
        avg[w] = a;
      }
      const T g = dl[c * n + q];
      for (int a = 0; a < 3; ++a) g3[a] += (avg[2 * a] - avg[2 * a + 1]) * div * g;
    }
    for (int a = 0; a < 3; ++a) dx[3 * q + a] = g3[a];
  }
}

// SurfelVolume::Merge + ToSurfelCloud (src/surfel_volume.cu:12-123) with
// HashVolume::Hash (include/slamtb/hash_volume.hpp:20-77): sequential sums
// per voxel in point order, scaled by 1 / count; voxels in increasing id.
// Returns the number of voxels; outputs must hold n entries.
int64_t voxel_merge(const float* pts, const float* nrm, const uint8_t* col,
                    const float* conf, const float* rad, const float* feats,
                    int C, int64_t n, float vs, float* o_pts, float* o_nrm,
                    uint8_t* o_col, float* o_conf, float* o_rad,
                    int32_t* o_time, float* o_feats) {
  if (n == 0) return 0;
  float mn[3], mx[3];
  for (int d = 0; d < 3; ++d) mn[d] = mx[d] = pts[d];
  for (int64_t i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], pts[3 * i + d]);
      mx[d] = std::max(mx[d], pts[3 * i + d]);
    }
  const int width = int((mx[0] - mn[0]) / vs);
  const int height = int((mx[1] - mn[1]) / vs);
  std::vector<std::pair<int32_t, int64_t>> ids(n);
  for (int64_t i = 0; i < n; ++i) {
    const int c = int((pts[3 * i] - mn[0]) / vs);
    const int r = int((pts[3 * i + 1] - mn[1]) / vs);
    const int d = int((pts[3 * i + 2] - mn[2]) / vs);
    ids[i] = {d * width * height + r * width + c, i};
  }
  std::stable_sort(ids.begin(), ids.end(),
                   [](const std::pair<int32_t, int64_t>& a,
                      const std::pair<int32_t, int64_t>& b) { return a.first < b.first; });
  std::vector<int64_t> starts;
  for (int64_t a = 0; a < n; ++a)
    if (a == 0 || ids[a].first != ids[a - 1].first) starts.push_back(a);
  const int64_t V = int64_t(starts.size());
  starts.push_back(n);
  for (int64_t v = 0; v < V; ++v) {
    float p[3] = {0, 0, 0}, q[3] = {0, 0, 0}, cc[3] = {0, 0, 0}, cf = 0, r = 0;
    for (int64_t a = starts[v]; a < starts[v + 1]; ++a) {
      const int64_t i = ids[a].second;
      for (int d = 0; d < 3; ++d) {
        p[d] += pts[3 * i + d];
        q[d] += nrm[3 * i + d];
        cc[d] += float(col[3 * i + d]);
      }
      cf += conf[i];
      r += rad[i];
    }
    const float inv = 1.0f / float(starts[v + 1] - starts[v]);
    for (int d = 0; d < 3; ++d) {
      o_pts[3 * v + d] = p[d] * inv;
      o_nrm[3 * v + d] = q[d] * inv;
      o_col[3 * v + d] = uint8_t(cc[d] * inv);
    }
    o_conf[v] = cf * inv;
    o_rad[v] = r * inv;
    o_time[v] = 0;
    for (int ch = 0; ch < C; ++ch) {
      float f = 0;
      for (int64_t a = starts[v]; a < starts[v + 1]; ++a)
        f += feats[ch * n + ids[a].second];
      o_feats[ch * n + v] = f * inv;
    }
  }
  return V;
}

// InformationMatrixKernel::operator() (src/icp_jacobian_geometric.cu:117-158):
// G^T G from the target point, parameter order [rot, trans].
template <typename T>
void information(const T* src_pts, const uint8_t* src_mask, int64_t n,
                 const T* rt, const float* tgt_pts, int H, int W,
                 const T* kcam, const int64_t* map, T* GtG, int parallel) {
  const Rigid<T> rig(rt);
  const Cam<T> cam(kcam);
  auto body = [&](int64_t idx) {
    T* g = GtG + idx * 36;
    for (int k = 0; k < 36; ++k) g[k] = T(0);
    if (!src_mask[idx]) return;
    const V3<T> tp = rig.transform(load3(src_pts, idx));
    T u, v;
    cam.project(tp, u, v);
    const int ui = round_to_int(u), vi = round_to_int(v);
    if (ui < 0 || ui >= W || vi < 0 || vi >= H) return;
    const int64_t pos = int64_t(vi) * W + ui;
    if (!owns(map, pos, idx)) return;
    const T x = T(tgt_pts[3 * pos]), y = T(tgt_pts[3 * pos + 1]),
            z = T(tgt_pts[3 * pos + 2]);
    const T r0[6] = {T(0), z, -y, T(1), T(0), T(0)};
    const T r1[6] = {-z, T(0), x, T(0), T(1), T(0)};
    const T r2[6] = {y, -x, T(0), T(0), T(0), T(1)};
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b)
        g[a * 6 + b] = r0[a] * r0[b] + r1[a] * r1[b] + r2[a] * r2[b];
  };
  if (parallel) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) body(i);
  } else {
    for (int64_t i = 0; i < n; ++i) body(i);
  }
}

// ExpRt::ToMatrix (src/exp_rt.hpp:27-65) with the translation copied raw
// (src/exp_rt_to_matrix.cu:22-36). Rotation by the Rodrigues formula over
// the normalised axis; a zero vector gives the identity.
template <typename T>
void exp_rt_one(const T* x, T* m) {
  const T wx = x[3], wy = x[4], wz = x[5];
  const T sq = wx * wx + (wy * wy + wz * wz);
  const T th = std::sqrt(sq);
  T ax = wx, ay = wy, az = wz;
  if (sq > T(0)) {
    ax = wx / th;
    ay = wy / th;
    az = wz / th;
  }
  const T s = std::sin(th), c = std::cos(th), k = T(1) - c;
  m[0] = k * ax * ax + c;
  m[1] = k * ax * ay - s * az;
  m[2] = k * ax * az + s * ay;
  m[4] = k * ax * ay + s * az;
  m[5] = k * ay * ay + c;
  m[6] = k * ay * az - s * ax;
  m[8] = k * ax * az - s * ay;
  m[9] = k * ay * az + s * ax;
  m[10] = k * az * az + c;
  m[3] = x[0];
  m[7] = x[1];
  m[11] = x[2];
}

// ExpRt::Dx_ToMatrix + BackwardKernel (src/exp_rt.hpp:87-141,
// src/exp_rt_to_matrix.cu:71-99): Gallego-Yezzi derivative of R(w),
// dR/dw_k = skew(v_k) R, v_k = (w w_k + w x (I - R) e_k) / |w|^2, or e_k
// when |w|^2 <= 1e-12.
template <typename T>
void exp_rt_backward_one(const T* dR, const T* x, const T* y, T* out) {
  const T w[3] = {x[3], x[4], x[5]};
  const T wn = w[0] * w[0] + (w[1] * w[1] + w[2] * w[2]);
  T R[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = y[i * 4 + j];
  out[0] = dR[3];
  out[1] = dR[7];
  out[2] = dR[11];
  for (int k = 0; k < 3; ++k) {
    T vk[3];
    if (wn > T(1e-12)) {
      const T e[3] = {(k == 0 ? T(1) : T(0)) - R[0 * 3 + k],
                      (k == 1 ? T(1) : T(0)) - R[1 * 3 + k],
                      (k == 2 ? T(1) : T(0)) - R[2 * 3 + k]};
      const T wxe[3] = {w[1] * e[2] - w[2] * e[1], w[2] * e[0] - w[0] * e[2],
                        w[0] * e[1] - w[1] * e[0]};
      for (int i = 0; i < 3; ++i) vk[i] = (w[i] * w[k] + wxe[i]) / wn;
    } else {
      for (int i = 0; i < 3; ++i) vk[i] = (i == k) ? T(1) : T(0);
    }
    const T S[9] = {T(0), -vk[2], vk[1], vk[2], T(0), -vk[0],
                    -vk[1], vk[0], T(0)};
    T acc = T(0);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T d = T(0);
        for (int l = 0; l < 3; ++l) d += S[i * 3 + l] * R[l * 3 + j];
        acc += d * dR[i * 4 + j];
      }
    out[3 + k] = acc;
  }
}

// MatrixToExpRtOp (src/matrix_to_exp_rt.cu:23-33): log map through the unit
// quaternion of R (Shepperd's branch selection), angle = 2 atan2(|q_v|, |q_w|).
template <typename T>
void matrix_to_exp_rt_one(const T* m, T* out) {
  const T r00 = m[0], r01 = m[1], r02 = m[2];
  const T r10 = m[4], r11 = m[5], r12 = m[6];
  const T r20 = m[8], r21 = m[9], r22 = m[10];
  T qw, qx, qy, qz;
  const T tr = r00 + r11 + r22;
  if (tr > T(0)) {
    T s = std::sqrt(tr + T(1));
    qw = T(0.5) * s;
    s = T(0.5) / s;
    qx = (r21 - r12) * s;
    qy = (r02 - r20) * s;
    qz = (r10 - r01) * s;
  } else {
    int i = 0;
    if (r11 > r00) i = 1;
    if (r22 > (i == 0 ? r00 : r11)) i = 2;
    if (i == 0) {
      T s = std::sqrt(r00 - r11 - r22 + T(1));
      qx = T(0.5) * s;
      s = T(0.5) / s;
      qw = (r21 - r12) * s;
      qy = (r01 + r10) * s;
      qz = (r02 + r20) * s;
    } else if (i == 1) {
      T s = std::sqrt(r11 - r22 - r00 + T(1));
      qy = T(0.5) * s;
      s = T(0.5) / s;
      qw = (r02 - r20) * s;
      qz = (r12 + r21) * s;
      qx = (r10 + r01) * s;
    } else {
      T s = std::sqrt(r22 - r00 - r11 + T(1));
      qz = T(0.5) * s;
      s = T(0.5) / s;
      qw = (r10 - r01) * s;
      qx = (r20 + r02) * s;
      qy = (r21 + r12) * s;
    }
  }
  T nv = std::sqrt(qx * qx + qy * qy + qz * qz);
  T angle = T(0), ax = T(1), ay = T(0), az = T(0);
  if (nv != T(0)) {
    angle = T(2) * std::atan2(nv, std::fabs(qw));
    if (qw < T(0)) nv = -nv;
    ax = qx / nv;
    ay = qy / nv;
    az = qz / nv;
  }
  out[0] = m[3];
  out[1] = m[7];
  out[2] = m[11];
  out[3] = ax * angle;
  out[4] = ay * angle;
  out[5] = az * angle;
}

// BilateralDepthFilterKernel (src/bilateral_depth_filter.cu:35-76) with the
// parameter set-up of :85-87. Float accumulation in window row-major order.
template <typename S>
void bilateral(const S* in, const uint8_t* mask, S* out, int H, int W,
               int filter_width, float sigma_color, float sigma_space,
               float depth_scale) {
  const int hw = filter_width / 2;
  const float inv_c = 1.0 / (sigma_color * sigma_color);
  const float inv_s = 1.0 / (sigma_space * sigma_space);
  const float inv_scale = 1.0 / depth_scale;
  for (int row = 0; row < H; ++row) {
    for (int col = 0; col < W; ++col) {
      out[int64_t(row) * W + col] = S(0);
      if (!mask[int64_t(row) * W + col]) continue;
      const float depth = float(in[int64_t(row) * W + col]) * depth_scale;
      float csum = 0.0f, wsum = 0.0f;
      for (int y = -hw; y <= hw; ++y) {
        const int kr = row + y;
        if (kr < 0 || kr >= H) continue;
        for (int x = -hw; x <= hw; ++x) {
          const int kc = col + x;
          if (kc < 0 || kc >= W) continue;
          if (!mask[int64_t(kr) * W + kc]) continue;
          const float cd = float(in[int64_t(kr) * W + kc]) * depth_scale;
          const float dx = float(col - kc);
          const float dy = float(row - kr);
          const float ss = dx * dx + dy * dy;
          const float dc = depth - cd;
          const float cs = dc * dc;
          const float w = std::exp(-0.5f * (ss * inv_s + cs * inv_c));
          csum += cd * w;
          wsum += w;
        }
      }
      if (wsum > 0.0f) out[int64_t(row) * W + col] = S((csum / wsum) * inv_scale);
    }
  }
}

// CentralDifferencesKernel (src/normals.cu:27-100): per axis, the two-sided
// difference when the squared neighbour distances are within a ratio of 4,
// else the one-sided difference to the closer neighbour; invalid neighbours
// are the zero vector; degenerate normals become (0, 0, -1).
template <typename T>
void normals(const T* xyz, const uint8_t* mask, T* out, int H, int W) {
  for (int row = 0; row < H; ++row) {
    for (int col = 0; col < W; ++col) {
      const int64_t p = int64_t(row) * W + col;
      out[3 * p] = out[3 * p + 1] = out[3 * p + 2] = T(0);
      if (!mask[p]) continue;
      const V3<T> zero{T(0), T(0), T(0)};
      const V3<T> c = load3(xyz, p);
      const V3<T> l = (col > 0 && mask[p - 1]) ? load3(xyz, p - 1) : zero;
      const V3<T> r = (col < W - 1 && mask[p + 1]) ? load3(xyz, p + 1) : zero;
      const V3<T> t = (row > 0 && mask[p - W]) ? load3(xyz, p - W) : zero;
      const V3<T> b = (row < H - 1 && mask[p + W]) ? load3(xyz, p + W) : zero;
      const T thr2 = T(2.f) * T(2.f);
      const T ld = sqnorm3(sub(l, c));
      const T rd = sqnorm3(sub(r, c));
      const T lr = ld / rd;
      V3<T> l2r;
      if (lr < thr2 && lr > T(1.f) / thr2)
        l2r = sub(r, l);
      else if (ld < rd)
        l2r = sub(c, l);
      else
        l2r = sub(r, c);
      const T bd = sqnorm3(sub(b, c));
      const T td = sqnorm3(sub(t, c));
      const T bt = bd / td;
      V3<T> b2t;
      if (bt < thr2 && bt > T(1.f) / thr2)
        b2t = sub(t, b);
      else if (bd < td)
        b2t = sub(c, b);
      else
        b2t = sub(t, c);
      V3<T> nrm = cross3(l2r, b2t);
      const T len = std::sqrt(sqnorm3(nrm));
      if (!(len > T(1e-6f))) {
        nrm = V3<T>{T(0), T(0), T(-1)};
      } else {
        const T s = std::sqrt(sqnorm3(nrm));
        nrm = V3<T>{nrm.x / s, nrm.y / s, nrm.z / s};
      }
      out[3 * p] = nrm.x;
      out[3 * p + 1] = nrm.y;
      out[3 * p + 2] = nrm.z;
    }
  }
}

// DownsampleXYZNearestKernel (src/downsample.cu:25-72): the valid 2x2-block
// point nearest to the block mean (column-bounded, rows unchecked).
template <typename T>
void downsample_xyz(const T* src, const uint8_t* mask, int H, int W,
                    float scale, T* dst, int dh, int dw) {
  const T inv = T(1) / T(scale);
  const int where[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
  for (int r = 0; r < dh; ++r) {
    for (int c = 0; c < dw; ++c) {
      const int cr = trunc_to_int(T(r) * inv), cc = trunc_to_int(T(c) * inv);
      T mx = T(0), my = T(0), mz = T(0);
      int cnt = 0;
      for (int i = 0; i < 4; ++i) {
        const int sr = cr + where[i][0], sc = cc + where[i][1];
        if (sc < 0 || sc >= W) continue;
        // The reference leaves rows unchecked; at scale 0.5 every block row
        // is in range, elsewhere its read past the last row is undefined.
        if (sr < 0 || sr >= H) continue;
        if (!mask[int64_t(sr) * W + sc]) continue;
        const V3<T> p = load3(src, int64_t(sr) * W + sc);
        mx += p.x;
        my += p.y;
        mz += p.z;
        ++cnt;
      }
      mx /= T(cnt);
      my /= T(cnt);
      mz /= T(cnt);
      V3<T> best{T(0), T(0), T(0)};
      T bd = std::numeric_limits<T>::infinity();
      for (int i = 0; i < 4; ++i) {
        const int sr = cr + where[i][0], sc = cc + where[i][1];
        if (sc < 0 || sc >= W) continue;
        if (sr < 0 || sr >= H) continue;
        if (!mask[int64_t(sr) * W + sc]) continue;
        const V3<T> p = load3(src, int64_t(sr) * W + sc);
        const T d = sqnorm3(V3<T>{p.x - mx, p.y - my, p.z - mz});
        if (d < bd) {
          bd = d;
          best = p;
        }
      }
      const int64_t o = int64_t(r) * dw + c;
      dst[3 * o] = best.x;
      dst[3 * o + 1] = best.y;
      dst[3 * o + 2] = best.z;
    }
  }
}

// DownsampleMaskKernel (src/downsample.cu:112-116): top-left sample.
void downsample_mask(const uint8_t* mask, int H, int W, float scale,
                     uint8_t* dst, int dh, int dw) {
  const float inv = 1.0f / scale;
  for (int r = 0; r < dh; ++r)
    for (int c = 0; c < dw; ++c) {
      const int sr = trunc_to_int(float(r) * inv), sc = trunc_to_int(float(c) * inv);
      dst[int64_t(r) * dw + c] =
          (sr >= 0 && sr < H && sc >= 0 && sc < W) ? mask[int64_t(sr) * W + sc] : 0;
    }
}

// DownsampleFeatureKernel (src/downsample.cu:145-168): 2x2 box mean per
// channel, column-bounded.
template <typename T>
void downsample_features(const T* src, int C, int H, int W, float scale,
                         T* dst, int dh, int dw) {
  const T inv = T(1) / T(scale);
  const int where[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
  for (int ch = 0; ch < C; ++ch)
    for (int r = 0; r < dh; ++r)
      for (int c = 0; c < dw; ++c) {
        const int cr = trunc_to_int(T(r) * inv), cc = trunc_to_int(T(c) * inv);
        T sum = T(0);
        int cnt = 0;
        for (int i = 0; i < 4; ++i) {
          const int sr = cr + where[i][0], sc = cc + where[i][1];
          if (sc < 0 || sc >= W) continue;
          if (sr < 0 || sr >= H) continue;
          sum += src[(int64_t(ch) * H + sr) * W + sc];
          ++cnt;
        }
        dst[(int64_t(ch) * dh + r) * dw + c] = sum / T(cnt);
      }
}

// ErodeMaskKernel (src/erode_mask.cu:18-39): the neighbourhood list holds
// (0, 1) twice and never (+1, 0), so the pixel below is not tested.
void erode(const uint8_t* in, uint8_t* out, int H, int W) {
  const int nb[4][2] = {{-1, 0}, {0, 1}, {0, -1}, {0, 1}};
  for (int row = 0; row < H; ++row)
    for (int col = 0; col < W; ++col) {
      const int64_t p = int64_t(row) * W + col;
      if (!in[p]) {
        out[p] = 0;
        continue;
      }
      uint8_t keep = 1;
      for (int k = 0; k < 4 && keep; ++k) {
        const int r = row + nb[k][0], c = col + nb[k][1];
        if (r >= 0 && r < H && c >= 0 && c < W && !in[int64_t(r) * W + c])
          keep = 0;
      }
      out[p] = keep;
    }
}

// Separable normalised 3x3 Gaussian with reflect padding (the border rule of
// a kornia GaussianBlur2d default; that rule is parity unpinned here).
template <typename T>
void gaussian_blur3(const T* src, int C, int H, int W, double sigma, T* dst) {
  const double e = std::exp(-1.0 / (2.0 * sigma * sigma));
  const T w1 = T(e / (1.0 + 2.0 * e));
  const T w0 = T(1.0 / (1.0 + 2.0 * e));
  auto refl = [](int i, int n) {
    if (n == 1) return 0;
    if (i < 0) return -i;
    if (i >= n) return 2 * n - 2 - i;
    return i;
  };
  std::vector<T> tmp(size_t(H) * W);
  for (int ch = 0; ch < C; ++ch) {
    const T* s = src + int64_t(ch) * H * W;
    T* d = dst + int64_t(ch) * H * W;
    for (int r = 0; r < H; ++r)
      for (int c = 0; c < W; ++c)
        tmp[size_t(r) * W + c] = w1 * s[int64_t(r) * W + refl(c - 1, W)] +
                                 w0 * s[int64_t(r) * W + c] +
                                 w1 * s[int64_t(r) * W + refl(c + 1, W)];
    for (int r = 0; r < H; ++r)
      for (int c = 0; c < W; ++c)
        d[int64_t(r) * W + c] = w1 * tmp[size_t(refl(r - 1, H)) * W + c] +
                                w0 * tmp[size_t(r) * W + c] +
                                w1 * tmp[size_t(refl(r + 1, H)) * W + c];
  }
}

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Thread count of the OpenMP loops (the CPU baseline is timed at 1 thread
// and at all of the host's threads).
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

void orc_corresp_map_f32(const float* sp, const float* sn, const uint8_t* sm,
                         int64_t ml, int64_t n, const float* rt,
                         const float* tp, const float* tn, const uint8_t* tm,
                         int H, int W, const float* k, int64_t* map,
                         double dthr, double athr, int parallel) {
  corresp_map<float>(sp, sn, sm, ml, n, rt, tp, tn, tm, H, W, k, map, dthr,
                     athr, parallel);
}
void orc_corresp_map_f64(const double* sp, const double* sn,
                         const uint8_t* sm, int64_t ml, int64_t n,
                         const double* rt, const double* tp, const double* tn,
                         const uint8_t* tm, int H, int W, const double* k,
                         int64_t* map, double dthr, double athr,
                         int parallel) {
  corresp_map<double>(sp, sn, sm, ml, n, rt, tp, tn, tm, H, W, k, map, dthr,
                      athr, parallel);
}

int orc_geometric_f32(const float* sp, const uint8_t* sm, int64_t n,
                      const float* rt, const float* tp, const float* tn,
                      int H, int W, const float* k, const int64_t* map,
                      float* JtJ, float* Jr, float* r2, int parallel) {
  return geometric<float>(sp, sm, n, rt, tp, tn, H, W, k, map, JtJ, Jr, r2,
                          parallel);
}
int orc_geometric_f64(const double* sp, const uint8_t* sm, int64_t n,
                      const double* rt, const float* tp, const float* tn,
                      int H, int W, const double* k, const int64_t* map,
                      double* JtJ, double* Jr, double* r2, int parallel) {
  return geometric<double>(sp, sm, n, rt, tp, tn, H, W, k, map, JtJ, Jr, r2,
                           parallel);
}

int orc_feature_f32(const float* sp, const float* sf, const uint8_t* sm,
                    int64_t n, const float* rt, const float* tf, int C, int H,
                    int W, const float* k, const int64_t* map, float thr,
                    float* JtJ, float* Jr, float* r2, int parallel) {
  return feature<float>(sp, sf, sm, n, rt, tf, C, H, W, k, map, thr, JtJ, Jr,
                        r2, parallel);
}
int orc_feature_f64(const double* sp, const double* sf, const uint8_t* sm,
                    int64_t n, const double* rt, const double* tf, int C,
                    int H, int W, const double* k, const int64_t* map,
                    float thr, double* JtJ, double* Jr, double* r2,
                    int parallel) {
  return feature<double>(sp, sf, sm, n, rt, tf, C, H, W, k, map, thr, JtJ, Jr,
                         r2, parallel);
}

void orc_information_f32(const float* sp, const uint8_t* sm, int64_t n,
                         const float* rt, const float* tp, int H, int W,
                         const float* k, const int64_t* map, float* GtG,
                         int parallel) {
  information<float>(sp, sm, n, rt, tp, H, W, k, map, GtG, parallel);
}
void orc_information_f64(const double* sp, const uint8_t* sm, int64_t n,
                         const double* rt, const float* tp, int H, int W,
                         const double* k, const int64_t* map, double* GtG,
                         int parallel) {
  information<double>(sp, sm, n, rt, tp, H, W, k, map, GtG, parallel);
}

void orc_exp_rt_to_matrix_f64(const double* x, double* m, int64_t B) {
  for (int64_t b = 0; b < B; ++b) exp_rt_one<double>(x + 6 * b, m + 12 * b);
}
void orc_exp_rt_to_matrix_f32(const float* x, float* m, int64_t B) {
  for (int64_t b = 0; b < B; ++b) exp_rt_one<float>(x + 6 * b, m + 12 * b);
}
void orc_exp_rt_backward_f64(const double* dR, const double* x,
                             const double* y, double* out, int64_t B) {
  for (int64_t b = 0; b < B; ++b)
    exp_rt_backward_one<double>(dR + 12 * b, x + 6 * b, y + 12 * b,
                                out + 6 * b);
}
void orc_matrix_to_exp_rt_f64(const double* m, double* out, int64_t B) {
  for (int64_t b = 0; b < B; ++b)
    matrix_to_exp_rt_one<double>(m + 12 * b, out + 6 * b);
}

void orc_bilateral_f32(const float* in, const uint8_t* mask, float* out, int H,
                       int W, int fw, float sc, float ss, float ds) {
  bilateral<float>(in, mask, out, H, W, fw, sc, ss, ds);
}
void orc_bilateral_i32(const int32_t* in, const uint8_t* mask, int32_t* out,
                       int H, int W, int fw, float sc, float ss, float ds) {
  bilateral<int32_t>(in, mask, out, H, W, fw, sc, ss, ds);
}
void orc_normals_f32(const float* xyz, const uint8_t* mask, float* out, int H,
                     int W) {
  normals<float>(xyz, mask, out, H, W);
}
void orc_normals_f64(const double* xyz, const uint8_t* mask, double* out,
                     int H, int W) {
  normals<double>(xyz, mask, out, H, W);
}
void orc_erode(const uint8_t* in, uint8_t* out, int H, int W) {
  erode(in, out, H, W);
}
void orc_downsample_xyz_f32(const float* src, const uint8_t* mask, int H,
                            int W, float scale, float* dst, int dh, int dw) {
  downsample_xyz<float>(src, mask, H, W, scale, dst, dh, dw);
}
void orc_downsample_mask(const uint8_t* mask, int H, int W, float scale,
                         uint8_t* dst, int dh, int dw) {
  downsample_mask(mask, H, W, scale, dst, dh, dw);
}
void orc_downsample_features_f32(const float* src, int C, int H, int W,
                                 float scale, float* dst, int dh, int dw) {
  downsample_features<float>(src, C, H, W, scale, dst, dh, dw);
}
void orc_gaussian_blur3_f32(const float* src, int C, int H, int W,
                            double sigma, float* dst) {
  gaussian_blur3<float>(src, C, H, W, sigma, dst);
}

#define ORC_FPCL(SFX, T)                                                       \
  void orc_fpcl_forward_##SFX(const T* tp, const T* tn, const T* tf, int C,    \
                              int H, int W, const T* sp, int64_t n,           \
                              const T* k, const int64_t* map, T* op, T* on,   \
                              T* of, uint8_t* mask) {                         \
    fpcl_forward<T>(tp, tn, tf, C, H, W, sp, n, k, map, op, on, of, mask);    \
  }                                                                           \
  void orc_fpcl_backward_##SFX(const T* tf, int C, int H, int W, const T* sp, \
                               int64_t n, const uint8_t* mask, const T* dl,   \
                               const T* k, double h, T* dx) {                 \
    fpcl_backward<T>(tf, C, H, W, sp, n, mask, dl, k, T(h), dx);              \
  }                                                                           \
  void orc_project_forward_##SFX(const T* p, int64_t n, const T* k, T* uv) {  \
    project_forward<T>(p, n, k, uv);                                          \
  }                                                                           \
  void orc_project_backward_##SFX(const T* dy, const T* p, int64_t n,         \
                                  const T* k, T* dx) {                        \
    project_backward<T>(dy, p, n, k, dx);                                     \
  }
ORC_FPCL(f32, float)
ORC_FPCL(f64, double)

#define ORC_RIGID(SFX, T)                                                      \
  void orc_rigid_points_##SFX(const T* m, const T* in, int64_t n, T* out,     \
                              int normals) {                                  \
    rigid_points<T>(m, in, n, out, normals);                                  \
  }                                                                           \
  void orc_rigid_backward_##SFX(const T* m, const T* p, const T* dy,          \
                                int64_t n, T* dx, T* dM) {                    \
    rigid_backward<T>(m, p, dy, n, dx, dM);                                   \
  }
ORC_RIGID(f32, float)
ORC_RIGID(f64, double)

#define ORC_PCL(SFX, T)                                                        \
  void orc_knn_##SFX(const T* tgt, int64_t m, const T* qry, int64_t n, int k,  \
                     double bound, T* d, int64_t* i) {                        \
    const T b2 = std::isfinite(bound) ? T(bound) * T(bound)                   \
                                      : std::numeric_limits<T>::infinity();   \
    knn<T>(tgt, m, qry, n, k, b2, d, i);                                      \
  }                                                                           \
  void orc_nn_forward_##SFX(const T* d, const int64_t* i, int64_t n, int k,   \
                            const T* f, int C, int64_t m, T* out) {           \
    nn_forward<T>(d, i, n, k, f, C, m, out);                                  \
  }                                                                           \
  void orc_nn_epsilon_##SFX(const T* tgt, int64_t m, const T* src,            \
                            const int64_t* i, int64_t n, int k, double h,     \
                            T* e) {                                           \
    nn_epsilon<T>(tgt, m, src, i, n, k, T(h), e);                             \
  }                                                                           \
  void orc_nn_backward_##SFX(const T* e, const int64_t* i, int64_t n, int k,  \
                             const T* f, int C, int64_t m, const T* dl,       \
                             double h, T* dx) {                               \
    nn_backward<T>(e, i, n, k, f, C, m, dl, T(h), dx);                        \
  }
ORC_PCL(f32, float)
ORC_PCL(f64, double)

int64_t orc_voxel_merge(const float* pts, const float* nrm, const uint8_t* col,
                        const float* conf, const float* rad, const float* feats,
                        int C, int64_t n, float vs, float* o_pts, float* o_nrm,
                        uint8_t* o_col, float* o_conf, float* o_rad,
                        int32_t* o_time, float* o_feats) {
  return voxel_merge(pts, nrm, col, conf, rad, feats, C, n, vs, o_pts, o_nrm,
                     o_col, o_conf, o_rad, o_time, o_feats);
}

}  // extern "C"
