// Differentiable frame-to-point-cloud matching and pinhole projection, the
// reference's FPCLMatcherOp (src/fpcl_matcher.cu:13-229) and ProjectOp
// (src/camera.cu:10-117): per source point, look up the target pixel that the
// correspondence map assigns to it, gather the target point, normal and the
// bilinearly sampled feature vector, and push feature gradients back to the
// source point through the projection.
#include "icp_math.hpp"

namespace stb {
namespace {

// Target index stored in a reference-layout map word (index << 32 | bits).
__device__ inline int32_t map_owner(const int64_t* map, int64_t pos) {
  return int32_t(static_cast<uint64_t>(map[pos]) >> 32);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    fpcl_forward_kernel(const T* __restrict__ tp, const T* __restrict__ tn,
                        const T* __restrict__ tf, int C, int H, int W,
                        const T* __restrict__ sp, int64_t n,
                        const T* __restrict__ kcam,
                        const int64_t* __restrict__ map, T* __restrict__ op,
                        T* __restrict__ on, T* __restrict__ of,
                        uint8_t* __restrict__ mask) {
#pragma clang fp contract(off)
  const CamT<T> cam = load_cam(kcam);
  const int64_t HW = int64_t(H) * W;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const V3<T> p = ld3(sp, idx);
    T u, v;
    project(cam.fx, cam.fy, cam.cx, cam.cy, p, u, v);
    int ui, vi;
    const bool inside = round_in_range(u, W, ui) && round_in_range(v, H, vi);
    const int64_t pos = inside ? int64_t(vi) * W + ui : 0;
    const bool matched = inside && map_owner(map, pos) == int32_t(idx);
    mask[idx] = matched ? 1 : 0;
    if (!matched) {
      // Unmatched rows are written as zeros (the reference leaves them
      // unwritten), so the outputs never carry stale memory.
#pragma unroll
      for (int k = 0; k < 3; ++k) op[3 * idx + k] = on[3 * idx + k] = T(0);
      for (int c = 0; c < C; ++c) of[int64_t(c) * n + idx] = T(0);
      continue;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      op[3 * idx + k] = tp[3 * pos + k];
      on[3 * idx + k] = tn[3 * pos + k];
    }
    const Bil<T> b = make_bil(u, v);
    for (int c = 0; c < C; ++c) of[int64_t(c) * n + idx] = bil_get(tf + c * HW, H, W, b);
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    fpcl_backward_kernel(const T* __restrict__ tf, int C, int H, int W,
                         const T* __restrict__ sp, int64_t n,
                         const uint8_t* __restrict__ mask,
                         const T* __restrict__ dl, const T* __restrict__ kcam,
                         T h, T* __restrict__ dx) {
#pragma clang fp contract(off)
  const CamT<T> cam = load_cam(kcam);
  const int64_t HW = int64_t(H) * W;
  const T div = T(1) / (T(2) * h);
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    if (!mask[idx]) {
      dx[3 * idx] = dx[3 * idx + 1] = dx[3 * idx + 2] = T(0);
      continue;
    }
    const V3<T> p = ld3(sp, idx);
    T u, v;
    project(cam.fx, cam.fy, cam.cx, cam.cy, p, u, v);
    // A mask that came from the forward pass only marks points whose rounded
    // projection is inside the image; anything else (a caller-made mask) is
    // treated as unmatched rather than sampled out of bounds.
    int ui, vi;
    if (!round_in_range(u, W, ui) || !round_in_range(v, H, vi)) {
      dx[3 * idx] = dx[3 * idx + 1] = dx[3 * idx + 2] = T(0);
      continue;
    }
    // src/feature_map.hpp:45-64: central differences of the bilinear sample
    // with step h, the lower samples clamped at 0.
    const Bil<T> bu0 = make_bil(u + h, v);
    const Bil<T> bu1 = make_bil(fmax(u - h, T(0)), v);
    const Bil<T> bv0 = make_bil(u, v + h);
    const Bil<T> bv1 = make_bil(u, fmax(v - h, T(0)));
    T gu = T(0), gv = T(0);
    for (int c = 0; c < C; ++c) {
      const T* ch = tf + c * HW;
      const T du = (bil_get(ch, H, W, bu0) - bil_get(ch, H, W, bu1)) * div;
      const T dv = (bil_get(ch, H, W, bv0) - bil_get(ch, H, W, bv1)) * div;
      const T g = dl[int64_t(c) * n + idx];
      gu += du * g;
      gv += dv * g;
    }
    // include/slamtb/camera.hpp:65-78 (Dx_Project).
    const T z2 = p.z * p.z;
    const T j00 = cam.fx / p.z;
    const T j02 = -p.x * cam.fx / z2;
    const T j11 = cam.fy / p.z;
    const T j12 = -p.y * cam.fy / z2;
    dx[3 * idx] = j00 * gu;
    dx[3 * idx + 1] = j11 * gv;
    dx[3 * idx + 2] = j02 * gu + j12 * gv;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    project_forward_kernel(const T* __restrict__ pts, int64_t n,
                           const T* __restrict__ k, T* __restrict__ uv) {
#pragma clang fp contract(off)
  const CamT<T> cam = load_cam(k);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    T u, v;
    project(cam.fx, cam.fy, cam.cx, cam.cy, ld3(pts, i), u, v);
    uv[2 * i] = u;
    uv[2 * i + 1] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    project_backward_kernel(const T* __restrict__ dy, const T* __restrict__ pts,
                            int64_t n, const T* __restrict__ k,
                            T* __restrict__ dx) {
#pragma clang fp contract(off)
  const CamT<T> cam = load_cam(k);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const V3<T> p = ld3(pts, i);
    const T z2 = p.z * p.z;
    const T j00 = cam.fx / p.z;
    const T j02 = -p.x * cam.fx / z2;
    const T j11 = cam.fy / p.z;
    const T j12 = -p.y * cam.fy / z2;
    const T du = dy[2 * i], dv = dy[2 * i + 1];
    dx[3 * i] = j00 * du;
    dx[3 * i + 1] = j11 * dv;
    dx[3 * i + 2] = j02 * du + j12 * dv;
  }
}

}  // namespace
}  // namespace stb

using namespace stb;

extern "C" int stb_fpcl_match_forward(
    int32_t dtype, const void* tgt_points, const void* tgt_normals,
    const void* tgt_feats, int32_t channels, int32_t height, int32_t width,
    const void* src_points, int64_t n_src, const void* kcam,
    const int64_t* corresp_map, void* out_points, void* out_normals,
    void* out_feats, uint8_t* match_mask, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "FPCLMatcherOp.forward: dtype must be float32 or float64");
  STB_REQUIRE(n_src >= 0 && height > 0 && width > 0 && channels >= 0,
              "FPCLMatcherOp.forward: bad sizes");
  if (n_src == 0) return STB_OK;
  STB_REQUIRE(tgt_points && tgt_normals && src_points && kcam && corresp_map &&
                  out_points && out_normals && match_mask &&
                  (channels == 0 || (tgt_feats && out_feats)),
              "FPCLMatcherOp.forward: NULL argument");
  hipStream_t s = as_stream(stream);
  const int grid = grid_for(n_src);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(fpcl_forward_kernel<float>, dim3(grid), dim3(kBlock), 0,
                       s, static_cast<const float*>(tgt_points),
                       static_cast<const float*>(tgt_normals),
                       static_cast<const float*>(tgt_feats), channels, height,
                       width, static_cast<const float*>(src_points), n_src,
                       static_cast<const float*>(kcam), corresp_map,
                       static_cast<float*>(out_points),
                       static_cast<float*>(out_normals),
                       static_cast<float*>(out_feats), match_mask);
  else
    hipLaunchKernelGGL(fpcl_forward_kernel<double>, dim3(grid), dim3(kBlock), 0,
                       s, static_cast<const double*>(tgt_points),
                       static_cast<const double*>(tgt_normals),
                       static_cast<const double*>(tgt_feats), channels, height,
                       width, static_cast<const double*>(src_points), n_src,
                       static_cast<const double*>(kcam), corresp_map,
                       static_cast<double*>(out_points),
                       static_cast<double*>(out_normals),
                       static_cast<double*>(out_feats), match_mask);
  return check_launch("FPCLMatcherOp.forward");
}

extern "C" int stb_fpcl_match_backward(
    int32_t dtype, const void* tgt_feats, int32_t channels, int32_t height,
    int32_t width, const void* src_points, int64_t n_src,
    const uint8_t* match_mask, const void* dl_feats, const void* kcam,
    double grad_precision, void* dx_points, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "FPCLMatcherOp.backward: dtype must be float32 or float64");
  STB_REQUIRE(n_src >= 0 && height > 0 && width > 0 && channels >= 0,
              "FPCLMatcherOp.backward: bad sizes");
  // With h < 0.5 every shifted sample of an in-image projection truncates
  // to a cell inside the image.
  STB_REQUIRE(grad_precision > 0.0 && grad_precision < 0.5,
              "FPCLMatcherOp.backward: grad_precision must be in (0, 0.5)");
  if (n_src == 0) return STB_OK;
  STB_REQUIRE(src_points && match_mask && kcam && dx_points &&
                  (channels == 0 || (tgt_feats && dl_feats)),
              "FPCLMatcherOp.backward: NULL argument");
  hipStream_t s = as_stream(stream);
  const int grid = grid_for(n_src);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(fpcl_backward_kernel<float>, dim3(grid), dim3(kBlock), 0,
                       s, static_cast<const float*>(tgt_feats), channels,
                       height, width, static_cast<const float*>(src_points),
                       n_src, match_mask, static_cast<const float*>(dl_feats),
                       static_cast<const float*>(kcam), float(grad_precision),
                       static_cast<float*>(dx_points));
  else
    hipLaunchKernelGGL(fpcl_backward_kernel<double>, dim3(grid), dim3(kBlock),
                       0, s, static_cast<const double*>(tgt_feats), channels,
                       height, width, static_cast<const double*>(src_points),
                       n_src, match_mask, static_cast<const double*>(dl_feats),
                       static_cast<const double*>(kcam), grad_precision,
                       static_cast<double*>(dx_points));
  return check_launch("FPCLMatcherOp.backward");
}

extern "C" int stb_project_forward(int32_t dtype, const void* points,
                                   int64_t n, const void* intrinsics,
                                   void* out_uv, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "ProjectOp.forward: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0, "ProjectOp.forward: bad size");
  if (n == 0) return STB_OK;
  STB_REQUIRE(points && intrinsics && out_uv, "ProjectOp.forward: NULL argument");
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(project_forward_kernel<float>, dim3(grid_for(n)),
                       dim3(kBlock), 0, s, static_cast<const float*>(points), n,
                       static_cast<const float*>(intrinsics),
                       static_cast<float*>(out_uv));
  else
    hipLaunchKernelGGL(project_forward_kernel<double>, dim3(grid_for(n)),
                       dim3(kBlock), 0, s, static_cast<const double*>(points),
                       n, static_cast<const double*>(intrinsics),
                       static_cast<double*>(out_uv));
  return check_launch("ProjectOp.forward");
}

extern "C" int stb_project_backward(int32_t dtype, const void* dy_uv,
                                    const void* points, int64_t n,
                                    const void* intrinsics, void* dx_points,
                                    void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "ProjectOp.backward: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0, "ProjectOp.backward: bad size");
  if (n == 0) return STB_OK;
  STB_REQUIRE(dy_uv && points && intrinsics && dx_points,
              "ProjectOp.backward: NULL argument");
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(project_backward_kernel<float>, dim3(grid_for(n)),
                       dim3(kBlock), 0, s, static_cast<const float*>(dy_uv),
                       static_cast<const float*>(points), n,
                       static_cast<const float*>(intrinsics),
                       static_cast<float*>(dx_points));
  else
    hipLaunchKernelGGL(project_backward_kernel<double>, dim3(grid_for(n)),
                       dim3(kBlock), 0, s, static_cast<const double*>(dy_uv),
                       static_cast<const double*>(points), n,
                       static_cast<const double*>(intrinsics),
                       static_cast<double*>(dx_points));
  return check_launch("ProjectOp.backward");
}
