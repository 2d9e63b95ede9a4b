// Frame preprocessing kernels: bilateral depth filter, central-difference
// normals, mask erosion, pyramid downsampling, 3x3 Gaussian blur of feature
// maps and depth back-projection. One thread per output pixel over a 64x4
// tile (one wavefront per image row segment, so row loads coalesce).
#include "stb_common.hpp"

namespace stb {
namespace {

constexpr int kTX = 64, kTY = 4;

inline dim3 grid2d(int w, int h) {
  return dim3((w + kTX - 1) / kTX, (h + kTY - 1) / kTY);
}

// BilateralDepthFilterKernel (src/bilateral_depth_filter.cu:35-76): float
// accumulation over the (2*hw+1)^2 window in row-major order, skipping masked
// neighbours; exp of -(ds^2/ss^2 + dd^2/sc^2)/2.
template <typename S>
__global__ void __launch_bounds__(kTX* kTY)
    bilateral_kernel(const S* __restrict__ in, const uint8_t* __restrict__ mask,
                     S* __restrict__ out, int H, int W, int hw, float inv_c,
                     float inv_s, float scale, float inv_scale) {
#pragma clang fp contract(off)
  const int col = blockIdx.x * kTX + threadIdx.x;
  const int row = blockIdx.y * kTY + threadIdx.y;
  if (col >= W || row >= H) return;
  const int64_t p = int64_t(row) * W + col;
  if (!mask[p]) {
    out[p] = S(0);
    return;
  }
  const float depth = float(in[p]) * scale;
  float csum = 0.0f, wsum = 0.0f;
  for (int y = -hw; y <= hw; ++y) {
    const int kr = row + y;
    if (kr < 0 || kr >= H) continue;
    for (int x = -hw; x <= hw; ++x) {
      const int kc = col + x;
      if (kc < 0 || kc >= W) continue;
      const int64_t q = int64_t(kr) * W + kc;
      if (!mask[q]) continue;
      const float cd = float(in[q]) * scale;
      const float dx = float(col - kc);
      const float dy = float(row - kr);
      const float ss = dx * dx + dy * dy;
      const float dc = depth - cd;
      const float cs = dc * dc;
      const float w = expf(-0.5f * (ss * inv_s + cs * inv_c));
      csum += cd * w;
      wsum += w;
    }
  }
  out[p] = (wsum > 0.0f) ? S((csum / wsum) * inv_scale) : S(0);
}

// CentralDifferencesKernel (src/normals.cu:27-100).
template <typename T>
__global__ void __launch_bounds__(kTX* kTY)
    normals_kernel(const T* __restrict__ xyz, const uint8_t* __restrict__ mask,
                   T* __restrict__ out, int H, int W) {
#pragma clang fp contract(off)
  const int col = blockIdx.x * kTX + threadIdx.x;
  const int row = blockIdx.y * kTY + threadIdx.y;
  if (col >= W || row >= H) return;
  const int64_t p = int64_t(row) * W + col;
  T* o = out + 3 * p;
  if (!mask[p]) {
    o[0] = o[1] = o[2] = T(0);
    return;
  }
  const V3<T> zero{T(0), T(0), T(0)};
  const V3<T> c = ld3(xyz, p);
  const V3<T> l = (col > 0 && mask[p - 1]) ? ld3(xyz, p - 1) : zero;
  const V3<T> r = (col < W - 1 && mask[p + 1]) ? ld3(xyz, p + 1) : zero;
  const V3<T> t = (row > 0 && mask[p - W]) ? ld3(xyz, p - W) : zero;
  const V3<T> b = (row < H - 1 && mask[p + W]) ? ld3(xyz, p + W) : zero;
  const T thr2 = T(2.f) * T(2.f);
  const V3<T> dl = sub3(l, c), dr = sub3(r, c);
  const T ld = dot3(dl, dl), rd = dot3(dr, dr);
  const T lr = ld / rd;
  V3<T> l2r;
  if (lr < thr2 && lr > T(1.f) / thr2)
    l2r = sub3(r, l);
  else if (ld < rd)
    l2r = sub3(c, l);
  else
    l2r = sub3(r, c);
  const V3<T> db = sub3(b, c), dt = sub3(t, c);
  const T bd = dot3(db, db), td = dot3(dt, dt);
  const T bt = bd / td;
  V3<T> b2t;
  if (bt < thr2 && bt > T(1.f) / thr2)
    b2t = sub3(t, b);
  else if (bd < td)
    b2t = sub3(c, b);
  else
    b2t = sub3(t, c);
  V3<T> n = cross3(l2r, b2t);
  const T len = sqrt(dot3(n, n));
  if (!(len > T(1e-6f))) {
    o[0] = T(0);
    o[1] = T(0);
    o[2] = T(-1);
  } else {
    const T s = sqrt(dot3(n, n));
    o[0] = n.x / s;
    o[1] = n.y / s;
    o[2] = n.z / s;
  }
}

// ErodeMaskKernel (src/erode_mask.cu:18-39); its neighbour list tests
// (-1,0), (0,1), (0,-1) and (0,1) again, never the pixel below.
__global__ void __launch_bounds__(kTX* kTY)
    erode_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                 int H, int W) {
  const int col = blockIdx.x * kTX + threadIdx.x;
  const int row = blockIdx.y * kTY + threadIdx.y;
  if (col >= W || row >= H) return;
  const int64_t p = int64_t(row) * W + col;
  uint8_t keep = in[p] ? 1 : 0;
  if (keep && row > 0 && !in[p - W]) keep = 0;
  if (keep && col + 1 < W && !in[p + 1]) keep = 0;
  if (keep && col > 0 && !in[p - 1]) keep = 0;
  out[p] = keep;
}

// DownsampleXYZNearestKernel (src/downsample.cu:25-72).
template <typename T>
__global__ void __launch_bounds__(kTX* kTY)
    downsample_xyz_kernel(const T* __restrict__ src,
                          const uint8_t* __restrict__ mask, int H, int W,
                          T inv, T* __restrict__ dst, int dh, int dw) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * kTX + threadIdx.x;
  const int r = blockIdx.y * kTY + threadIdx.y;
  if (c >= dw || r >= dh) return;
  const int cr = int(T(r) * inv), cc = int(T(c) * inv);
  V3<T> pts[4];
  bool ok[4];
  T mx = T(0), my = T(0), mz = T(0);
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int sr = cr + (i >> 1), sc = cc + (i & 1);
    ok[i] = sc >= 0 && sc < W && sr >= 0 && sr < H &&
            mask[int64_t(sr) * W + sc];
    if (ok[i]) {
      pts[i] = ld3(src, int64_t(sr) * W + sc);
      mx += pts[i].x;
      my += pts[i].y;
      mz += pts[i].z;
      ++cnt;
    }
  }
  mx /= T(cnt);
  my /= T(cnt);
  mz /= T(cnt);
  V3<T> best{T(0), T(0), T(0)};
  T bd = T(__builtin_inf());
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!ok[i]) continue;
    const V3<T> d{pts[i].x - mx, pts[i].y - my, pts[i].z - mz};
    const T dd = dot3(d, d);
    if (dd < bd) {
      bd = dd;
      best = pts[i];
    }
  }
  T* o = dst + 3 * (int64_t(r) * dw + c);
  o[0] = best.x;
  o[1] = best.y;
  o[2] = best.z;
}

// DownsampleMaskKernel (src/downsample.cu:112-116).
__global__ void __launch_bounds__(kTX* kTY)
    downsample_mask_kernel(const uint8_t* __restrict__ mask, int H, int W,
                           float inv, uint8_t* __restrict__ dst, int dh,
                           int dw) {
  const int c = blockIdx.x * kTX + threadIdx.x;
  const int r = blockIdx.y * kTY + threadIdx.y;
  if (c >= dw || r >= dh) return;
  const int sr = int(float(r) * inv), sc = int(float(c) * inv);
  dst[int64_t(r) * dw + c] =
      (sr < H && sc < W) ? (mask[int64_t(sr) * W + sc] ? 1 : 0) : 0;
}

// DownsampleFeatureKernel (src/downsample.cu:145-168).
template <typename T>
__global__ void __launch_bounds__(kTX* kTY)
    downsample_features_kernel(const T* __restrict__ src, int C, int H, int W,
                               T inv, T* __restrict__ dst, int dh, int dw) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * kTX + threadIdx.x;
  const int r = blockIdx.y * kTY + threadIdx.y;
  if (c >= dw || r >= dh) return;
  const int cr = int(T(r) * inv), cc = int(T(c) * inv);
  for (int ch = 0; ch < C; ++ch) {
    const T* s = src + int64_t(ch) * H * W;
    T sum = T(0);
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int sr = cr + (i >> 1), sc = cc + (i & 1);
      if (sc < 0 || sc >= W || sr < 0 || sr >= H) continue;
      sum += s[int64_t(sr) * W + sc];
      ++cnt;
    }
    dst[(int64_t(ch) * dh + r) * dw + c] = sum / T(cnt);
  }
}

__device__ inline int reflect(int i, int n) {
  if (n == 1) return 0;
  if (i < 0) return -i;
  if (i >= n) return 2 * n - 2 - i;
  return i;
}

// Separable 3x3 Gaussian, horizontal then vertical, reflect borders.
template <typename T>
__global__ void __launch_bounds__(kTX* kTY)
    blur3_kernel(const T* __restrict__ src, int C, int H, int W, T w0, T w1,
                 T* __restrict__ dst) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * kTX + threadIdx.x;
  const int r = blockIdx.y * kTY + threadIdx.y;
  if (c >= W || r >= H) return;
  const int cl = reflect(c - 1, W), cr = reflect(c + 1, W);
  const int ru = reflect(r - 1, H), rd = reflect(r + 1, H);
  for (int ch = 0; ch < C; ++ch) {
    const T* s = src + int64_t(ch) * H * W;
    auto hrow = [&](int rr) {
      const T* row = s + int64_t(rr) * W;
      return w1 * row[cl] + w0 * row[c] + w1 * row[cr];
    };
    const T a = hrow(ru), b = hrow(r), d = hrow(rd);
    dst[(int64_t(ch) * H + r) * W + c] = w1 * a + w0 * b + w1 * d;
  }
}

// depth_image_to_uvz then KCamera.backproject (slamtb/_utils.py:11-37,
// slamtb/camera.py:127-149), in float32 like the torch ops they replace.
template <typename S>
__global__ void __launch_bounds__(kTX* kTY)
    depth_to_points_kernel(const S* __restrict__ depth, int H, int W, float fx,
                           float fy, float cx, float cy, float scale,
                           float bias, float* __restrict__ pts) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * kTX + threadIdx.x;
  const int r = blockIdx.y * kTY + threadIdx.y;
  if (c >= W || r >= H) return;
  const int64_t p = int64_t(r) * W + c;
  const float z = float(depth[p]) * scale + bias;
  float* o = pts + 3 * p;
  o[0] = (float(c) - cx) * z / fx;
  o[1] = (float(r) - cy) * z / fy;
  o[2] = z;
}

template <template <typename> class Launch, typename... Args>
int dispatch_depth(int dtype, Args... args) {
  switch (dtype) {
    case STB_F32:
      return Launch<float>::run(args...);
    case STB_F64:
      return Launch<double>::run(args...);
    case STB_I32:
      return Launch<int32_t>::run(args...);
    case STB_I16:
      return Launch<int16_t>::run(args...);
    case STB_U8:
      return Launch<uint8_t>::run(args...);
    default:
      set_error("unsupported depth dtype %d", dtype);
      return STB_ERR_INVALID;
  }
}

template <typename S>
struct BilateralLaunch {
  static int run(const void* in, const uint8_t* mask, void* out, int H, int W,
                 int hw, float inv_c, float inv_s, float scale,
                 float inv_scale, hipStream_t s) {
    hipLaunchKernelGGL(bilateral_kernel<S>, grid2d(W, H), dim3(kTX, kTY), 0, s,
                       (const S*)in, mask, (S*)out, H, W, hw, inv_c, inv_s,
                       scale, inv_scale);
    return STB_OK;
  }
};

template <typename S>
struct DepthLaunch {
  static int run(const void* depth, int H, int W, float fx, float fy, float cx,
                 float cy, float scale, float bias, float* pts,
                 hipStream_t s) {
    hipLaunchKernelGGL(depth_to_points_kernel<S>, grid2d(W, H), dim3(kTX, kTY),
                       0, s, (const S*)depth, H, W, fx, fy, cx, cy, scale,
                       bias, pts);
    return STB_OK;
  }
};

}  // namespace
}  // namespace stb

using namespace stb;

extern "C" int stb_bilateral_depth_filter(int dtype, const void* input,
                                          const uint8_t* mask, void* result,
                                          int32_t height, int32_t width,
                                          int32_t filter_width,
                                          float sigma_color, float sigma_space,
                                          float depth_scale, void* stream) {
  STB_REQUIRE(input && mask && result, "bilateral_depth_filter: NULL argument");
  STB_REQUIRE(height >= 0 && width >= 0, "bad image size");
  if (height == 0 || width == 0) return STB_OK;
  // Parameter set-up of src/bilateral_depth_filter.cu:85-87 (float operands,
  // double division, float result).
  const int hw = filter_width / 2;
  const float inv_c = 1.0 / (sigma_color * sigma_color);
  const float inv_s = 1.0 / (sigma_space * sigma_space);
  const float inv_scale = 1.0 / depth_scale;
  int rc = dispatch_depth<BilateralLaunch>(dtype, input, mask, result,
                                           int(height), int(width), hw, inv_c,
                                           inv_s, depth_scale, inv_scale,
                                           as_stream(stream));
  if (rc) return rc;
  return check_launch("bilateral_depth_filter");
}

extern "C" int stb_estimate_normals(int dtype, const void* xyz,
                                    const uint8_t* mask, void* out_normals,
                                    int32_t height, int32_t width, int method,
                                    void* stream) {
  STB_REQUIRE(xyz && mask && out_normals, "estimate_normals: NULL argument");
  if (method != 0) {
    set_error("estimate_normals: only CentralDifferences is supported");
    return STB_ERR_UNSUPPORTED;
  }
  if (height == 0 || width == 0) return STB_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(normals_kernel<float>, grid2d(width, height),
                       dim3(kTX, kTY), 0, s, (const float*)xyz, mask,
                       (float*)out_normals, height, width);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(normals_kernel<double>, grid2d(width, height),
                       dim3(kTX, kTY), 0, s, (const double*)xyz, mask,
                       (double*)out_normals, height, width);
  else {
    set_error("estimate_normals: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("estimate_normals");
}

extern "C" int stb_erode_mask(const uint8_t* in_mask, uint8_t* out_mask,
                              int32_t height, int32_t width, void* stream) {
  STB_REQUIRE(in_mask && out_mask, "erode_mask: NULL argument");
  if (height == 0 || width == 0) return STB_OK;
  hipLaunchKernelGGL(erode_kernel, grid2d(width, height), dim3(kTX, kTY), 0,
                     as_stream(stream), in_mask, out_mask, height, width);
  return check_launch("erode_mask");
}

extern "C" int stb_downsample_xyz(int dtype, const void* src,
                                  const uint8_t* mask, int32_t height,
                                  int32_t width, float scale, void* dst,
                                  int32_t dst_height, int32_t dst_width,
                                  int normalize, int method, void* stream) {
  (void)normalize;  // accepted and ignored, as src/downsample.cu:75-98 does
  STB_REQUIRE(src && mask && dst, "downsample_xyz: NULL argument");
  STB_REQUIRE(method == 0, "downsample_xyz: only Nearest is supported");
  STB_REQUIRE(scale > 0.0f, "downsample_xyz: scale must be positive");
  if (dst_height == 0 || dst_width == 0) return STB_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(downsample_xyz_kernel<float>,
                       grid2d(dst_width, dst_height), dim3(kTX, kTY), 0, s,
                       (const float*)src, mask, height, width, 1.0f / scale,
                       (float*)dst, dst_height, dst_width);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(downsample_xyz_kernel<double>,
                       grid2d(dst_width, dst_height), dim3(kTX, kTY), 0, s,
                       (const double*)src, mask, height, width,
                       1.0 / double(scale), (double*)dst, dst_height,
                       dst_width);
  else {
    set_error("downsample_xyz: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("downsample_xyz");
}

extern "C" int stb_downsample_mask(const uint8_t* mask, int32_t height,
                                   int32_t width, float scale, uint8_t* dst,
                                   int32_t dst_height, int32_t dst_width,
                                   void* stream) {
  STB_REQUIRE(mask && dst, "downsample_mask: NULL argument");
  STB_REQUIRE(scale > 0.0f, "downsample_mask: scale must be positive");
  if (dst_height == 0 || dst_width == 0) return STB_OK;
  hipLaunchKernelGGL(downsample_mask_kernel, grid2d(dst_width, dst_height),
                     dim3(kTX, kTY), 0, as_stream(stream), mask, height, width,
                     1.0f / scale, dst, dst_height, dst_width);
  return check_launch("downsample_mask");
}

extern "C" int stb_downsample_features(int dtype, const void* src,
                                       int32_t channels, int32_t height,
                                       int32_t width, float scale, void* dst,
                                       int32_t dst_height, int32_t dst_width,
                                       void* stream) {
  STB_REQUIRE(src && dst, "downsample_features: NULL argument");
  STB_REQUIRE(scale > 0.0f, "downsample_features: scale must be positive");
  if (dst_height == 0 || dst_width == 0 || channels == 0) return STB_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(downsample_features_kernel<float>,
                       grid2d(dst_width, dst_height), dim3(kTX, kTY), 0, s,
                       (const float*)src, channels, height, width,
                       1.0f / scale, (float*)dst, dst_height, dst_width);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(downsample_features_kernel<double>,
                       grid2d(dst_width, dst_height), dim3(kTX, kTY), 0, s,
                       (const double*)src, channels, height, width,
                       1.0 / double(scale), (double*)dst, dst_height,
                       dst_width);
  else {
    set_error("downsample_features: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("downsample_features");
}

extern "C" int stb_gaussian_blur3x3(int dtype, const void* src,
                                    int32_t channels, int32_t height,
                                    int32_t width, double sigma, void* dst,
                                    void* stream) {
  STB_REQUIRE(src && dst, "gaussian_blur3x3: NULL argument");
  STB_REQUIRE(sigma > 0.0, "gaussian_blur3x3: sigma must be positive");
  if (height == 0 || width == 0 || channels == 0) return STB_OK;
  const double e = std::exp(-1.0 / (2.0 * sigma * sigma));
  const double w1 = e / (1.0 + 2.0 * e), w0 = 1.0 / (1.0 + 2.0 * e);
  hipStream_t s = as_stream(stream);
  if (dtype == STB_F32)
    hipLaunchKernelGGL(blur3_kernel<float>, grid2d(width, height),
                       dim3(kTX, kTY), 0, s, (const float*)src, channels,
                       height, width, float(w0), float(w1), (float*)dst);
  else if (dtype == STB_F64)
    hipLaunchKernelGGL(blur3_kernel<double>, grid2d(width, height),
                       dim3(kTX, kTY), 0, s, (const double*)src, channels,
                       height, width, w0, w1, (double*)dst);
  else {
    set_error("gaussian_blur3x3: dtype must be float32 or float64");
    return STB_ERR_INVALID;
  }
  return check_launch("gaussian_blur3x3");
}

extern "C" int stb_depth_to_points(int depth_dtype, const void* depth,
                                   int32_t height, int32_t width, float fx,
                                   float fy, float cx, float cy,
                                   float depth_scale, float depth_bias,
                                   float* points, void* stream) {
  STB_REQUIRE(depth && points, "depth_to_points: NULL argument");
  if (height == 0 || width == 0) return STB_OK;
  int rc = dispatch_depth<DepthLaunch>(depth_dtype, depth, int(height),
                                       int(width), fx, fy, cx, cy, depth_scale,
                                       depth_bias, points, as_stream(stream));
  if (rc) return rc;
  return check_launch("depth_to_points");
}
