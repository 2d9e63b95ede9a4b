// Hand-written adjoint of one Gauss-Newton step, for differentiating the
// estimated pose with respect to the frames and the initial pose.
//
// Forward step k (per pair): A = sum J' J'^T, b = sum J' r' over matched
// pixels at pose T_k (J' = w J, r' = s r as in the forward kernels),
// x = A^-1 b, T_{k+1} = exp_rt(x) T_k. Given x_bar = dL/dx, the solve's
// implicit adjoint is lambda = A^-1 x_bar, b_bar = lambda, A_bar =
// -lambda x^T, so every matched pixel receives
//   J'_bar = lambda r' - lambda (x . J') - x (lambda . J'),  r'_bar = lambda . J'.
// This kernel recomputes J and r per pixel (nothing is stored by the
// forward), pushes J_bar / r_bar through the point-to-plane and feature terms
// into the source points, target points and normals, source and target
// feature maps, and reduces dL/dT_k (the 3x4 pose rows, through
// p' = R p + t) per pair, accumulated in FP64 from the first product. The
// caller supplies lambda and x per pair and resets each pair's pose to T_k
// (stb_icp_batch_begin) before the call; correspondences are rebuilt at T_k
// exactly as the forward built them.
#include "batch_plan.hpp"

namespace stb {
namespace {

constexpr int kPose = 12;  // dL/dT rows 0..2 of the 4x4 pose, row-major

__device__ inline float dot6(const float* a, const float* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] +
         a[5] * b[5];
}

// Bilinear sample with its value, its derivatives in u and v, and the tap
// weights (zero past the right/bottom edge, as the forward samples).
struct BilD {
  int ui, vi;
  float ur, vr;
  float val, du, dv;
  bool rx, ry;
};

__device__ inline BilD bil_full(const float* __restrict__ ch, int H, int W,
                                float u, float v) {
  BilD b;
  b.ui = int(u);
  b.vi = int(v);
  b.ur = u - float(b.ui);
  b.vr = v - float(b.vi);
  const int64_t o = int64_t(b.vi) * W + b.ui;
  b.rx = b.ui + 1 < W;
  b.ry = b.vi + 1 < H;
  const float v00 = ch[o];
  const float v10 = b.rx ? ch[o + 1] : 0.0f;
  const float v01 = b.ry ? ch[o + W] : 0.0f;
  const float v11 = (b.rx && b.ry) ? ch[o + W + 1] : 0.0f;
  const float a = v00 * (1.0f - b.ur) + v10 * b.ur;
  const float c = v01 * (1.0f - b.ur) + v11 * b.ur;
  b.val = a * (1.0f - b.vr) + c * b.vr;
  b.du = (1.0f - b.vr) * (v10 - v00) + b.vr * (v11 - v01);
  b.dv = c - a;
  return b;
}

// Adds g times the sample's tap weights into the channel's gradient.
__device__ inline void bil_scatter(float* __restrict__ ch, int W, const BilD& b,
                                   float g) {
  if (g == 0.0f) return;
  const int64_t o = int64_t(b.vi) * W + b.ui;
  atomicAdd(ch + o, g * (1.0f - b.ur) * (1.0f - b.vr));
  if (b.rx) atomicAdd(ch + o + 1, g * b.ur * (1.0f - b.vr));
  if (b.ry) atomicAdd(ch + o + W, g * (1.0f - b.ur) * b.vr);
  if (b.rx && b.ry) atomicAdd(ch + o + W + 1, g * b.ur * b.vr);
}

// The four tap weights of a sample times g, zero for taps past the
// right/bottom edge (as bil_scatter would skip them).
struct TapW {
  float w00, w10, w01, w11;
};
__device__ inline TapW tap_weights(const BilD& b, float g) {
  TapW t;
  t.w00 = g * (1.0f - b.ur) * (1.0f - b.vr);
  t.w10 = b.rx ? g * b.ur * (1.0f - b.vr) : 0.0f;
  t.w01 = b.ry ? g * (1.0f - b.ur) * b.vr : 0.0f;
  t.w11 = (b.rx && b.ry) ? g * b.ur * b.vr : 0.0f;
  return t;
}

// Scatter of the five samples of the feature Jacobian (the centre (ui, vi)
// cell and the +-h cells, each the centre's or its neighbour along one
// axis) merged in registers over the 4x4 texel patch whose corner is
// (ui - 1, vi - 1): one atomic per touched texel -- four when all five
// samples share the centre cell -- instead of four per sample.
__device__ inline void scatter5(float* __restrict__ ch, int W, const BilD& c0,
                                float g0, const BilD& bu0, float gu0,
                                const BilD& bu1, float gu1, const BilD& bv0,
                                float gv0, const BilD& bv1, float gv1) {
  float P[4][4];
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x) P[y][x] = 0.0f;
  const TapW c = tap_weights(c0, g0);
  P[1][1] += c.w00;
  P[1][2] += c.w10;
  P[2][1] += c.w01;
  P[2][2] += c.w11;
  // Samples displaced along x by +1 (bu0 may) or -1 (bu1 may).
  const TapW a = tap_weights(bu0, gu0);
  const bool ra = bu0.ui != c0.ui;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const float w0 = r ? a.w01 : a.w00, w1 = r ? a.w11 : a.w10;
    P[1 + r][1] += ra ? 0.0f : w0;
    P[1 + r][2] += ra ? w0 : w1;
    P[1 + r][3] += ra ? w1 : 0.0f;
  }
  const TapW b = tap_weights(bu1, gu1);
  const bool lb = bu1.ui != c0.ui;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const float w0 = r ? b.w01 : b.w00, w1 = r ? b.w11 : b.w10;
    P[1 + r][0] += lb ? w0 : 0.0f;
    P[1 + r][1] += lb ? w1 : w0;
    P[1 + r][2] += lb ? 0.0f : w1;
  }
  // Samples displaced along y by +1 (bv0 may) or -1 (bv1 may).
  const TapW d = tap_weights(bv0, gv0);
  const bool dd = bv0.vi != c0.vi;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float w0 = q ? d.w10 : d.w00, w1 = q ? d.w11 : d.w01;
    P[1][1 + q] += dd ? 0.0f : w0;
    P[2][1 + q] += dd ? w0 : w1;
    P[3][1 + q] += dd ? w1 : 0.0f;
  }
  const TapW e = tap_weights(bv1, gv1);
  const bool ue = bv1.vi != c0.vi;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float w0 = q ? e.w10 : e.w00, w1 = q ? e.w11 : e.w01;
    P[0][1 + q] += ue ? w0 : 0.0f;
    P[1][1 + q] += ue ? w1 : w0;
    P[2][1 + q] += ue ? 0.0f : w1;
  }
  const int64_t o = (int64_t(c0.vi) - 1) * W + (c0.ui - 1);
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x)
      // Non-zero entries are inside the image (weights past any edge are
      // zero), so only in-range offsets are dereferenced.
      if (P[y][x] != 0.0f) atomicAdd(ch + (o + int64_t(y) * W + x), P[y][x]);
}

// Block sum of the 12 pose-gradient terms in FP64 (lanes by shuffles, waves
// through LDS), written as one kSlots-wide partial (unused slots zero) for
// the fixed-order sum of adjoint_sum_kernel (a later launch: no fence).
__device__ inline void pose_partial(double (&acc)[kPose], double* out,
                                    double* sh) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kPose; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < kPose; ++k) sh[wave * kPose + k] = acc[k];
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double s = 0.0;
    if (threadIdx.x < kPose)
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += sh[w * kPose + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// J'_bar and r'_bar of one weighted term (see the header comment).
__device__ inline void term_adjoint(const float J[6], float w, float s, float r,
                                    const float* lam, const float* x,
                                    float Jbar[6], float& rbar) {
  float Jw[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) Jw[a] = w * J[a];
  const float rp = s * r;
  const float xJ = dot6(x, Jw), lJ = dot6(lam, Jw);
#pragma unroll
  for (int a = 0; a < 6; ++a)
    Jbar[a] = w * (lam[a] * rp - lam[a] * xJ - x[a] * lJ);
  rbar = s * lJ;
}

template <bool GEOM, bool FEAT>
__device__ inline void adjoint_pixel(const LevelView<float>& L, const float* m,
                                     const unsigned long long* __restrict__ map,
                                     uint32_t tag, int64_t idx,
                                     const TermWeights& tw,
                                     const float* lam, const float* x,
                                     const stb_icp_grads& G,
                                     double (&tacc)[kPose]) {
  if (!L.sm[idx]) return;
  const V3<float> ps = ld3f(L.sp, idx);
  const V3<float> p = rigid_point(m, ps);
  float u, v;
  project(L.fx, L.fy, L.cx, L.cy, p, u, v);
  int ui, vi;
  if (!round_in_range(u, L.W, ui) || !round_in_range(v, L.H, vi)) return;
  const int64_t pos = int64_t(vi) * L.W + ui;
  if (!key_owned(map[pos], uint32_t(idx), tag)) return;
  V3<float> pb{0.0f, 0.0f, 0.0f};  // dL/dp'
  if (GEOM) {
    const V3<float> q = ld3f(L.tp, pos);
    const V3<float> n = ld3f(L.tn, pos);
    const V3<float> d = sub3(q, p);
    const float r = dot3(d, n);
    const V3<float> c = cross3(p, n);
    const float J[6] = {n.x, n.y, n.z, c.x, c.y, c.z};
    float Jb[6], rb;
    term_adjoint(J, tw.wg, tw.sg, r, lam, x, Jb, rb);
    // r = (q - p') . n
    V3<float> nb{d.x * rb + Jb[0], d.y * rb + Jb[1], d.z * rb + Jb[2]};
    const V3<float> wb{Jb[3], Jb[4], Jb[5]};
    // J_w = p' x n: (p' x n) . wb = p' . (n x wb) = n . (wb x p')
    const V3<float> a1 = cross3(n, wb), a2 = cross3(wb, p);
    pb.x += -n.x * rb + a1.x;
    pb.y += -n.y * rb + a1.y;
    pb.z += -n.z * rb + a1.z;
    nb.x += a2.x;
    nb.y += a2.y;
    nb.z += a2.z;
    // The map has one owner per target pixel, so these updates never collide.
    float* tq = G.d_tgt_points + 3 * pos;
    float* tn = G.d_tgt_normals + 3 * pos;
    tq[0] += n.x * rb;
    tq[1] += n.y * rb;
    tq[2] += n.z * rb;
    tn[0] += nb.x;
    tn[1] += nb.y;
    tn[2] += nb.z;
  }
  if (FEAT) {
    const int C = L.C;
    const float* __restrict__ tf = reinterpret_cast<const float*>(L.tf);
    const float* __restrict__ sf = reinterpret_cast<const float*>(L.sf);
    const float h = 0.05f;
    const float div = 1.0f / (2.0f * h);
    float sq = 0.0f;
    for (int c = 0; c < C - 1; ++c) {
      const float res = bil_full(tf + c * L.HW, L.H, L.W, u, v).val -
                        sf[c * L.n + idx];
      sq += res * res;
    }
    const float* ch = tf + int64_t(C - 1) * L.HW;
    const BilD c0 = bil_full(ch, L.H, L.W, u, v);
    const float res = c0.val - sf[int64_t(C - 1) * L.n + idx];
    sq += res * res;
    if (!(sq > tw.fthr2)) {
      const float um = fmaxf(u - h, 0.0f), vm = fmaxf(v - h, 0.0f);
      const BilD bu0 = bil_full(ch, L.H, L.W, u + h, v);
      const BilD bu1 = bil_full(ch, L.H, L.W, um, v);
      const BilD bv0 = bil_full(ch, L.H, L.W, u, v + h);
      const BilD bv1 = bil_full(ch, L.H, L.W, u, vm);
      const float du = (bu0.val - bu1.val) * div;
      const float dv = (bv0.val - bv1.val) * div;
      const float z = p.z, z2 = z * z, z3 = z2 * z;
      const float j00 = L.fx / z, j11 = L.fy / z;
      const float j02 = -p.x * L.fx / z2, j12 = -p.y * L.fy / z2;
      const V3<float> g{du * j00, dv * j11, du * j02 + dv * j12};
      const V3<float> c = cross3(p, g);
      const float J[6] = {g.x, g.y, g.z, c.x, c.y, c.z};
      float Jb[6], rb;
      term_adjoint(J, tw.wf, tw.sf, res, lam, x, Jb, rb);
      const V3<float> wb{Jb[3], Jb[4], Jb[5]};
      // J_w = p' x g: g_bar += wb x p', p'_bar += g x wb
      const V3<float> a2 = cross3(wb, p), a1 = cross3(g, wb);
      const V3<float> gb{Jb[0] + a2.x, Jb[1] + a2.y, Jb[2] + a2.z};
      pb.x += a1.x;
      pb.y += a1.y;
      pb.z += a1.z;
      const float dub = gb.x * j00 + gb.z * j02;
      const float dvb = gb.y * j11 + gb.z * j12;
      const float j00b = gb.x * du, j11b = gb.y * dv;
      const float j02b = gb.z * du, j12b = gb.z * dv;
      pb.x += -j02b * L.fx / z2;
      pb.y += -j12b * L.fy / z2;
      pb.z += -j00b * L.fx / z2 - j11b * L.fy / z2 +
              2.0f * j02b * p.x * L.fx / z3 + 2.0f * j12b * p.y * L.fy / z3;
      // r = I(u, v) - f_s and the finite differences du, dv.
      const float mu = (u - h > 0.0f) ? 1.0f : 0.0f;
      const float mv = (v - h > 0.0f) ? 1.0f : 0.0f;
      const float ub = rb * c0.du + dub * div * (bu0.du - mu * bu1.du) +
                       dvb * div * (bv0.du - bv1.du);
      const float vb = rb * c0.dv + dub * div * (bu0.dv - bu1.dv) +
                       dvb * div * (bv0.dv - mv * bv1.dv);
      float* gch = G.d_tgt_feats + int64_t(C - 1) * L.HW;
      scatter5(gch, L.W, c0, rb, bu0, dub * div, bu1, -dub * div, bv0,
               dvb * div, bv1, -dvb * div);
      G.d_src_feats[int64_t(C - 1) * L.n + idx] -= rb;
      // u = fx x / z + cx, v = fy y / z + cy
      pb.x += ub * L.fx / z;
      pb.y += vb * L.fy / z;
      pb.z += -ub * L.fx * p.x / z2 - vb * L.fy * p.y / z2;
    }
  }
  // p' = R ps + t: ps_bar = R^T pb; dT[i][j] += pb_i ps_j, dT[i][3] += pb_i.
  float* sp = G.d_src_points + 3 * idx;
  sp[0] += m[0] * pb.x + m[4] * pb.y + m[8] * pb.z;
  sp[1] += m[1] * pb.x + m[5] * pb.y + m[9] * pb.z;
  sp[2] += m[2] * pb.x + m[6] * pb.y + m[10] * pb.z;
  const double pbv[3] = {pb.x, pb.y, pb.z};
  const double psv[3] = {ps.x, ps.y, ps.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) tacc[i * 4 + j] += pbv[i] * psv[j];
    tacc[i * 4 + 3] += pbv[i];
  }
}

template <bool GEOM, bool FEAT>
__global__ void __launch_bounds__(kBlock)
    batch_adjoint_kernel(const void* __restrict__ plan,
                         stb_icp_state* __restrict__ states,
                         unsigned long long* __restrict__ maps0,
                         unsigned long long* __restrict__ maps1,
                         const stb_icp_grads* __restrict__ grads,
                         const double* __restrict__ coef,
                         double* __restrict__ partials, TermWeights tw) {
  __shared__ double sh[kShared];
  const PlanView pv = plan_view(plan);
  for (int t = blockIdx.x; t < pv.tiles; t += gridDim.x) {
    const int p = pv.tile_pair[t];
    stb_icp_state* st = states + p;
    const LevelView<float> L = view_of<float>(pv.levels[p]);
    const unsigned long long* map = pair_map(st, maps0, maps1, pv.map_off[p], 0);
    const uint32_t tag = uint32_t(st->map_epoch);
    const stb_icp_grads G = grads[p];
    float lam[6], x[6], m[12];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      lam[a] = float(coef[12 * p + a]);
      x[a] = float(coef[12 * p + 6 + a]);
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
    const int64_t base = int64_t(t - pv.tile_start[p]) * pv.tile_px;
    double tacc[kPose];
#pragma unroll
    for (int k = 0; k < kPose; ++k) tacc[k] = 0.0;
    for (int k = threadIdx.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t idx = base + k;
      if (idx < L.n)
        adjoint_pixel<GEOM, FEAT>(L, m, map, tag, idx, tw, lam, x, G, tacc);
    }
    pose_partial(tacc, partials + int64_t(t) * kSlots, sh);
  }
}

// One block per pair: the fixed-order sum of the pair's tile partials into
// its 12 dL/dT terms.
__global__ void __launch_bounds__(kSolveBlock)
    adjoint_sum_kernel(const void* __restrict__ plan,
                       const double* __restrict__ partials,
                       double* __restrict__ dT) {
  __shared__ double sh[kSolveBlock];
  __shared__ double tot[kSlots];
  const PlanView pv = plan_view(plan);
  const int p = blockIdx.x;
  if (p >= pv.P) return;
  const int first = pv.tile_start[p];
  sum_partials<kSolveBlock>(partials + int64_t(first) * kSlots,
                            pv.tile_start[p + 1] - first, sh, tot);
  if (threadIdx.x < 12) dT[12 * p + threadIdx.x] = tot[threadIdx.x];
}

template <bool GEOM, bool FEAT>
void launch_adjoint(const void* plan, int num_pairs, stb_icp_state* st,
                    const Scratch& sc, int grid, const stb_icp_grads* grads,
                    const double* coef, double* dT, const TermWeights& tw,
                    hipStream_t s) {
  hipLaunchKernelGGL((batch_adjoint_kernel<GEOM, FEAT>), dim3(grid),
                     dim3(kBlock), 0, s, plan, st, sc.map0, sc.map1, grads,
                     coef, sc.partials, tw);
  hipLaunchKernelGGL(adjoint_sum_kernel, dim3(num_pairs), dim3(kSolveBlock), 0,
                     s, plan, sc.partials, dT);
}

}  // namespace

// Correspondence pass of the batched path (defined in batch.hip).
void launch_batch_correspond_f32(const void* plan,
                                 const stb_icp_batch_sizes& z,
                                 const stb_icp_params& P, stb_icp_state* states,
                                 const Scratch& sc, hipStream_t s);

}  // namespace stb

using namespace stb;

extern "C" int stb_icp_batch_adjoint(const void* plan,
                                     const stb_icp_batch_sizes* sizes,
                                     const stb_icp_params* params, void* states,
                                     const stb_icp_grads* grads,
                                     const double* coef, double* dT,
                                     void* scratch, size_t scratch_bytes,
                                     void* stream) {
  STB_REQUIRE(sizes && plan && params && states && grads && coef && dT &&
                  scratch,
              "stb_icp_batch_adjoint: NULL argument");
  STB_REQUIRE(sizes->storage == STB_F32,
              "stb_icp_batch_adjoint: FP32 storage only");
  STB_REQUIRE(params->use_geom || params->use_feat,
              "stb_icp_batch_adjoint: no active term");
  const Scratch sc = carve_scratch(scratch, sizes->num_pairs,
                                   sizes->total_tiles, sizes->total_map);
  STB_REQUIRE(scratch_bytes >= sc.bytes, "stb_icp_batch_adjoint: scratch too small");
  hipStream_t s = as_stream(stream);
  stb_icp_state* st = static_cast<stb_icp_state*>(states);
  launch_batch_correspond_f32(plan, *sizes, *params, st, sc, s);
  const TermWeights tw = make_weights(*params);
  const int cap = params->grid_update > 0 ? params->grid_update : 1024;
  const int grid = sizes->total_tiles < cap ? sizes->total_tiles : cap;
  const int P = sizes->num_pairs;
  if (params->use_geom && params->use_feat)
    launch_adjoint<true, true>(plan, P, st, sc, grid, grads, coef, dT, tw, s);
  else if (params->use_geom)
    launch_adjoint<true, false>(plan, P, st, sc, grid, grads, coef, dT, tw, s);
  else
    launch_adjoint<false, true>(plan, P, st, sc, grid, grads, coef, dT, tw, s);
  return check_launch("stb_icp_batch_adjoint");
}
