// Drop-in ICP kernels for gfx950 with the reference bindings' per-pixel
// outputs: CorrespondenceMap.compute_correspondence_map and
// ICPJacobian.estimate_geometric / estimate_feature / get_information_matrix.
// The fused Gauss-Newton path that ICPOdometry uses lives in gn.hip.
#include "icp_math.hpp"

namespace stb {
namespace {

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
