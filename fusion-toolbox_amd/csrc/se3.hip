// Batched SE3 helpers: [t; w] -> [R(w) | t] (Rodrigues), its backward pass
// (Gallego-Yezzi derivative of the exponential map) and the log map.
#include "stb_common.hpp"

namespace stb {
namespace {

// ExpRt::ToMatrix with the raw translation (src/exp_rt.hpp:27-65,
// src/exp_rt_to_matrix.cu:22-36).
template <typename T>
__global__ void exp_rt_forward_kernel(const T* __restrict__ x, int64_t B,
                                      T* __restrict__ m) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const T* xi = x + 6 * b;
  T* o = m + 12 * b;
  const T wx = xi[3], wy = xi[4], wz = xi[5];
  const T sq = wx * wx + (wy * wy + wz * wz);
  const T th = sqrt(sq);
  T ax = wx, ay = wy, az = wz;
  if (sq > T(0)) {
    ax = wx / th;
    ay = wy / th;
    az = wz / th;
  }
  const T s = sin(th), c = cos(th), k = T(1) - c;
  o[0] = k * ax * ax + c;
  o[1] = k * ax * ay - s * az;
  o[2] = k * ax * az + s * ay;
  o[4] = k * ax * ay + s * az;
  o[5] = k * ay * ay + c;
  o[6] = k * ay * az - s * ax;
  o[8] = k * ax * az - s * ay;
  o[9] = k * ay * az + s * ax;
  o[10] = k * az * az + c;
  o[3] = xi[0];
  o[7] = xi[1];
  o[11] = xi[2];
}

// ExpRt::Dx_ToMatrix + BackwardKernel (src/exp_rt.hpp:87-141,
// src/exp_rt_to_matrix.cu:71-99): dR/dw_k = skew(v_k) R with
// v_k = (w w_k + w x (I - R) e_k) / |w|^2, or e_k when |w|^2 <= 1e-12.
template <typename T>
__global__ void exp_rt_backward_kernel(const T* __restrict__ dR,
                                       const T* __restrict__ x,
                                       const T* __restrict__ y, int64_t B,
                                       T* __restrict__ out) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const T* g = dR + 12 * b;
  const T* xi = x + 6 * b;
  const T* yi = y + 12 * b;
  T* o = out + 6 * b;
  const T w[3] = {xi[3], xi[4], xi[5]};
  const T wn = w[0] * w[0] + (w[1] * w[1] + w[2] * w[2]);
  T R[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = yi[i * 4 + j];
  o[0] = g[3];
  o[1] = g[7];
  o[2] = g[11];
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
        acc += d * g[i * 4 + j];
      }
    o[3 + k] = acc;
  }
}

// Log map (src/matrix_to_exp_rt.cu:23-33): rotation -> unit quaternion by
// Shepperd's largest-pivot rule, then angle = 2 atan2(|q_v|, |q_w|).
template <typename T>
__global__ void matrix_to_exp_rt_kernel(const T* __restrict__ m, int64_t B,
                                        T* __restrict__ out) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const T* a = m + 12 * b;
  T* o = out + 6 * b;
  const T r00 = a[0], r01 = a[1], r02 = a[2];
  const T r10 = a[4], r11 = a[5], r12 = a[6];
  const T r20 = a[8], r21 = a[9], r22 = a[10];
  T qw, qx, qy, qz;
  const T tr = r00 + r11 + r22;
  if (tr > T(0)) {
    T s = sqrt(tr + T(1));
    qw = T(0.5) * s;
    s = T(0.5) / s;
    qx = (r21 - r12) * s;
    qy = (r02 - r20) * s;
    qz = (r10 - r01) * s;
  } else if (r00 >= r11 && r00 >= r22) {
    T s = sqrt(r00 - r11 - r22 + T(1));
    qx = T(0.5) * s;
    s = T(0.5) / s;
    qw = (r21 - r12) * s;
    qy = (r01 + r10) * s;
    qz = (r02 + r20) * s;
  } else if (r11 >= r22) {
    T s = sqrt(r11 - r22 - r00 + T(1));
    qy = T(0.5) * s;
    s = T(0.5) / s;
    qw = (r02 - r20) * s;
    qz = (r12 + r21) * s;
    qx = (r10 + r01) * s;
  } else {
    T s = sqrt(r22 - r00 - r11 + T(1));
    qz = T(0.5) * s;
    s = T(0.5) / s;
    qw = (r10 - r01) * s;
    qx = (r20 + r02) * s;
    qy = (r21 + r12) * s;
  }
  T nv = sqrt(qx * qx + qy * qy + qz * qz);
  T angle = T(0), ax = T(1), ay = T(0), az = T(0);
  if (nv != T(0)) {
    angle = T(2) * atan2(nv, fabs(qw));
    if (qw < T(0)) nv = -nv;
    ax = qx / nv;
    ay = qy / nv;
    az = qz / nv;
  }
  o[0] = a[3];
  o[1] = a[7];
  o[2] = a[11];
  o[3] = ax * angle;
  o[4] = ay * angle;
  o[5] = az * angle;
}

}  // namespace
}  // namespace stb

using namespace stb;

extern "C" int stb_exp_rt_to_matrix_forward(int dtype, const void* exp_rt,
                                            int64_t batch, void* matrix,
                                            void* stream) {
  STB_REQUIRE(batch >= 0, "bad batch");
  if (batch == 0) return STB_OK;
  STB_REQUIRE(exp_rt && matrix, "exp_rt_to_matrix: NULL argument");
  const int g = grid_for(batch, 64);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(exp_rt_forward_kernel<float>, dim3(g), dim3(64), 0, s,
                       (const float*)exp_rt, batch, (float*)matrix);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(exp_rt_forward_kernel<double>, dim3(g), dim3(64), 0, s,
                       (const double*)exp_rt, batch, (double*)matrix);
  else {
    set_error("exp_rt_to_matrix: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("exp_rt_to_matrix_forward");
}

extern "C" int stb_exp_rt_to_matrix_backward(int dtype, const void* d_matrix,
                                             const void* exp_rt,
                                             const void* y_matrix,
                                             int64_t batch, void* d_exp_rt,
                                             void* stream) {
  STB_REQUIRE(batch >= 0, "bad batch");
  if (batch == 0) return STB_OK;
  STB_REQUIRE(d_matrix && exp_rt && y_matrix && d_exp_rt,
              "exp_rt_to_matrix backward: NULL argument");
  const int g = grid_for(batch, 64);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(exp_rt_backward_kernel<float>, dim3(g), dim3(64), 0, s,
                       (const float*)d_matrix, (const float*)exp_rt,
                       (const float*)y_matrix, batch, (float*)d_exp_rt);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(exp_rt_backward_kernel<double>, dim3(g), dim3(64), 0, s,
                       (const double*)d_matrix, (const double*)exp_rt,
                       (const double*)y_matrix, batch, (double*)d_exp_rt);
  else {
    set_error("exp_rt_to_matrix backward: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("exp_rt_to_matrix_backward");
}

extern "C" int stb_matrix_to_exp_rt_forward(int dtype, const void* matrix,
                                            int64_t batch, void* exp_rt,
                                            void* stream) {
  STB_REQUIRE(batch >= 0, "bad batch");
  if (batch == 0) return STB_OK;
  STB_REQUIRE(matrix && exp_rt, "matrix_to_exp_rt: NULL argument");
  const int g = grid_for(batch, 64);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(matrix_to_exp_rt_kernel<float>, dim3(g), dim3(64), 0, s,
                       (const float*)matrix, batch, (float*)exp_rt);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(matrix_to_exp_rt_kernel<double>, dim3(g), dim3(64), 0,
                       s, (const double*)matrix, batch, (double*)exp_rt);
  else {
    set_error("matrix_to_exp_rt: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("matrix_to_exp_rt_forward");
}
