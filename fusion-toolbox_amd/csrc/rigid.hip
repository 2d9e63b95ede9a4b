// Rigid transforms of point and normal arrays on the device: the reference's
// RigidTransformOp.transform_inplace / transform_normals_inplace
// (src/camera.cu:134-240, include/slamtb/camera.hpp:99-166) and the
// forward/backward pair of `RigidTransform @ points` that autograd needs
// (slamtb/camera.py:333-337 computes it with a batched matmul).
//
// The matrix is read from device memory by every thread (3x4 row-major, rows
// 0..2 of a [3,4] or [4,4] tensor), so nothing is copied to the host. Points
// are p' = R p + t summed left to right, normals n' = N n with N the
// cofactor inverse-transpose, exactly as the reference's native RigidTransform.
//
// Backward of p' = M [p; 1]: dp = R^T dp', dM[r][c] = sum_i dp'_i[r] p_i[c]
// and dM[r][3] = sum_i dp'_i[r]. The twelve sums are accumulated in FP64 per
// thread, reduced per block into a caller-provided workspace, and summed in a
// fixed order by a one-block finishing launch (deterministic).
#include "stb_common.hpp"

namespace stb {
namespace {

constexpr int kRigidBlocks = 512;  // reduction blocks of the backward pass
constexpr int kM = 12;

template <typename T>
__device__ inline void load_m(const T* __restrict__ M, T (&m)[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) m[k] = M[k];
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    rigid_points_kernel(const T* __restrict__ M, const T* in, T* out, int64_t n) {
  T m[12];
  load_m(M, m);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const V3<T> p = rigid_point(m, ld3(in, i));
    out[3 * i] = p.x;
    out[3 * i + 1] = p.y;
    out[3 * i + 2] = p.z;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    rigid_normals_kernel(const T* __restrict__ M, const T* in, T* out, int64_t n) {
  Rigid<T> r;
  load_m(M, r.m);
  make_normal_matrix(r);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const V3<T> q = rigid_normal(r.nm, ld3(in, i));
    out[3 * i] = q.x;
    out[3 * i + 1] = q.y;
    out[3 * i + 2] = q.z;
  }
}

// Block sum of twelve FP64 accumulators; thread k < 12 of the block gets the
// block's k-th sum.
__device__ inline double block_sum12(double (&acc)[kM], double* sh, int k) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kM; ++j)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[j] += __shfl_xor(acc[j], o, 64);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < kM; ++j) sh[wave * kM + j] = acc[j];
  __syncthreads();
  double s = 0.0;
  if (k < kM)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += sh[w * kM + k];
  return s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    rigid_points_backward_kernel(const T* __restrict__ M,
                                 const T* __restrict__ p,
                                 const T* __restrict__ dy, T* __restrict__ dx,
                                 double* __restrict__ partials, int64_t n) {
#pragma clang fp contract(off)
  __shared__ double sh[(kBlock / 64) * kM];
  T m[12];
  load_m(M, m);
  double acc[kM];
#pragma unroll
  for (int j = 0; j < kM; ++j) acc[j] = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const V3<T> g = ld3(dy, i);
    const V3<T> q = ld3(p, i);
    if (dx != nullptr) {
      dx[3 * i] = m[0] * g.x + m[4] * g.y + m[8] * g.z;
      dx[3 * i + 1] = m[1] * g.x + m[5] * g.y + m[9] * g.z;
      dx[3 * i + 2] = m[2] * g.x + m[6] * g.y + m[10] * g.z;
    }
    const double gr[3] = {double(g.x), double(g.y), double(g.z)};
    const double pc[3] = {double(q.x), double(q.y), double(q.z)};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[r * 4 + c] += gr[r] * pc[c];
      acc[r * 4 + 3] += gr[r];
    }
  }
  const double s = block_sum12(acc, sh, threadIdx.x);
  if (threadIdx.x < kM) partials[int64_t(blockIdx.x) * kM + threadIdx.x] = s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    rigid_grad_finish_kernel(const double* __restrict__ partials, int blocks,
                             T* __restrict__ dM) {
  __shared__ double sh[(kBlock / 64) * kM];
  double acc[kM];
#pragma unroll
  for (int j = 0; j < kM; ++j) acc[j] = 0.0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x)
#pragma unroll
    for (int j = 0; j < kM; ++j) acc[j] += partials[int64_t(b) * kM + j];
  const double s = block_sum12(acc, sh, threadIdx.x);
  if (threadIdx.x < kM) dM[threadIdx.x] = T(s);
}

template <typename T>
int launch_rigid(bool normals, const void* M, const void* in, void* out,
                 int64_t n, hipStream_t s) {
  const dim3 grid(grid_for(n)), block(kBlock);
  if (normals)
    hipLaunchKernelGGL(rigid_normals_kernel<T>, grid, block, 0, s,
                       static_cast<const T*>(M), static_cast<const T*>(in),
                       static_cast<T*>(out), n);
  else
    hipLaunchKernelGGL(rigid_points_kernel<T>, grid, block, 0, s,
                       static_cast<const T*>(M), static_cast<const T*>(in),
                       static_cast<T*>(out), n);
  return check_launch(normals ? "RigidTransformOp.transform_normals"
                              : "RigidTransformOp.transform");
}

}  // namespace
}  // namespace stb

using namespace stb;

extern "C" int stb_rigid_transform_points(int32_t dtype, const void* matrix,
                                          const void* points, int64_t n,
                                          void* out, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "RigidTransformOp.transform: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0, "RigidTransformOp.transform: bad size");
  if (n == 0) return STB_OK;
  STB_REQUIRE(matrix && points && out, "RigidTransformOp.transform: NULL argument");
  return dtype == STB_F32
             ? launch_rigid<float>(false, matrix, points, out, n, as_stream(stream))
             : launch_rigid<double>(false, matrix, points, out, n, as_stream(stream));
}

extern "C" int stb_rigid_transform_normals(int32_t dtype, const void* matrix,
                                           const void* normals, int64_t n,
                                           void* out, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "RigidTransformOp.transform_normals: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0, "RigidTransformOp.transform_normals: bad size");
  if (n == 0) return STB_OK;
  STB_REQUIRE(matrix && normals && out,
              "RigidTransformOp.transform_normals: NULL argument");
  return dtype == STB_F32
             ? launch_rigid<float>(true, matrix, normals, out, n, as_stream(stream))
             : launch_rigid<double>(true, matrix, normals, out, n, as_stream(stream));
}

extern "C" size_t stb_rigid_backward_workspace_bytes(void) {
  return size_t(kRigidBlocks) * kM * sizeof(double);
}

extern "C" int stb_rigid_transform_backward(int32_t dtype, const void* matrix,
                                            const void* points, int64_t n,
                                            const void* dy, void* dx_points,
                                            void* d_matrix, void* workspace,
                                            size_t workspace_bytes,
                                            void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "RigidTransform backward: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0, "RigidTransform backward: bad size");
  STB_REQUIRE(matrix && d_matrix && workspace && (n == 0 || (points && dy)),
              "RigidTransform backward: NULL argument");
  const int blocks = grid_for(n, kBlock, kRigidBlocks);
  STB_REQUIRE(workspace_bytes >= size_t(blocks) * kM * sizeof(double),
              "RigidTransform backward: workspace too small");
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32) {
    hipLaunchKernelGGL(rigid_points_backward_kernel<float>, dim3(blocks),
                       dim3(kBlock), 0, s, static_cast<const float*>(matrix),
                       static_cast<const float*>(points),
                       static_cast<const float*>(dy),
                       static_cast<float*>(dx_points),
                       static_cast<double*>(workspace), n);
    hipLaunchKernelGGL(rigid_grad_finish_kernel<float>, dim3(1), dim3(kBlock),
                       0, s, static_cast<const double*>(workspace), blocks,
                       static_cast<float*>(d_matrix));
  } else {
    hipLaunchKernelGGL(rigid_points_backward_kernel<double>, dim3(blocks),
                       dim3(kBlock), 0, s, static_cast<const double*>(matrix),
                       static_cast<const double*>(points),
                       static_cast<const double*>(dy),
                       static_cast<double*>(dx_points),
                       static_cast<double*>(workspace), n);
    hipLaunchKernelGGL(rigid_grad_finish_kernel<double>, dim3(1), dim3(kBlock),
                       0, s, static_cast<const double*>(workspace), blocks,
                       static_cast<double*>(d_matrix));
  }
  return check_launch("RigidTransform backward");
}
