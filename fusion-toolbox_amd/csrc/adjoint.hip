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
// p' = R p + t) per pair in FP64. A second kernel recomputes the step's
// normal equations (A, b) the same way, for the caller's solves. The caller
// supplies lambda and x per pair and resets each pair's pose to T_k
// (stb_icp_batch_begin) before the calls; correspondences are rebuilt at T_k
// exactly as the forward built them.
#include "batch_plan.hpp"

namespace stb {
namespace {

constexpr int kPose = 12;  // dL/dT rows 0..2 of the 4x4 pose, row-major

__device__ inline double dot6(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] +
         a[5] * b[5];
}

// The adjoint takes its decisions -- correspondence owner, feature gate,
// bilinear cells, the clamp of the lower samples at 0 -- from the forward's
// FP32 arithmetic at the FP32 pose, and evaluates everything smooth (the
// transformed point, u and v, the bilinear fractions and samples, J, r and
// their adjoints) in FP64 at the FP64 pose. The FP32 rounding of u and v
// (half an ulp is 3e-5 px at u = 600) moves a sample within its cell and
// changes the sampled derivatives by ~1e-4 relative; the backward through
// the GN steps amplifies such errors about 2.5x per step, because dL/dT_k
// is the difference of two nearly cancelling terms (a step undoes most of a
// pose perturbation).

// Bilinear sample in FP64 at (u, v) over the cell (ui, vi) the forward
// chose, with its derivatives in u and v, and the taps' zero padding past the
// right/bottom edge (src/feature_map.hpp:8-64).
struct BilD {
  int ui, vi;
  double ur, vr;
  double val, du, dv;
  bool rx, ry;
};

__device__ inline BilD bil_at(const float* __restrict__ ch, int H, int W,
                              int ui, int vi, double u, double v) {
  BilD b;
  b.ui = ui;
  b.vi = vi;
  b.ur = u - double(ui);
  b.vr = v - double(vi);
  const int64_t o = int64_t(vi) * W + ui;
  b.rx = ui + 1 < W;
  b.ry = vi + 1 < H;
  const double v00 = ch[o];
  const double v10 = b.rx ? double(ch[o + 1]) : 0.0;
  const double v01 = b.ry ? double(ch[o + W]) : 0.0;
  const double v11 = (b.rx && b.ry) ? double(ch[o + W + 1]) : 0.0;
  const double a = v00 * (1.0 - b.ur) + v10 * b.ur;
  const double c = v01 * (1.0 - b.ur) + v11 * b.ur;
  b.val = a * (1.0 - b.vr) + c * b.vr;
  b.du = (1.0 - b.vr) * (v10 - v00) + b.vr * (v11 - v01);
  b.dv = c - a;
  return b;
}

// FP32 bilinear value at (u, v) as the forward's gate computes it.
__device__ inline float bil_val32(const float* __restrict__ ch, int H, int W,
                                  float u, float v) {
  const Bil<float> b = make_bil(u, v);
  float t[4];
  bil_taps(ch, H, W, b, t);
  return bil_mix(H, W, b, t);
}

// The four tap weights of a sample times g, zero for taps past the
// right/bottom edge.
struct TapW {
  double w00, w10, w01, w11;
};
__device__ inline TapW tap_weights(const BilD& b, double g) {
  TapW t;
  t.w00 = g * (1.0 - b.ur) * (1.0 - b.vr);
  t.w10 = b.rx ? g * b.ur * (1.0 - b.vr) : 0.0;
  t.w01 = b.ry ? g * (1.0 - b.ur) * b.vr : 0.0;
  t.w11 = (b.rx && b.ry) ? g * b.ur * b.vr : 0.0;
  return t;
}

// Scatter of the five samples of the feature Jacobian (the centre (ui, vi)
// cell and the +-h cells, each the centre's or its neighbour along one
// axis) merged in registers over the 4x4 texel patch whose corner is
// (ui - 1, vi - 1): one atomic per touched texel -- four when all five
// samples share the centre cell -- instead of four per sample.
__device__ inline void scatter5(float* __restrict__ ch, int W, const BilD& c0,
                                double g0, const BilD& bu0, double gu0,
                                const BilD& bu1, double gu1, const BilD& bv0,
                                double gv0, const BilD& bv1, double gv1) {
  double P[4][4];
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x) P[y][x] = 0.0;
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
    const double w0 = r ? a.w01 : a.w00, w1 = r ? a.w11 : a.w10;
    P[1 + r][1] += ra ? 0.0 : w0;
    P[1 + r][2] += ra ? w0 : w1;
    P[1 + r][3] += ra ? w1 : 0.0;
  }
  const TapW b = tap_weights(bu1, gu1);
  const bool lb = bu1.ui != c0.ui;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const double w0 = r ? b.w01 : b.w00, w1 = r ? b.w11 : b.w10;
    P[1 + r][0] += lb ? w0 : 0.0;
    P[1 + r][1] += lb ? w1 : w0;
    P[1 + r][2] += lb ? 0.0 : w1;
  }
  // Samples displaced along y by +1 (bv0 may) or -1 (bv1 may).
  const TapW d = tap_weights(bv0, gv0);
  const bool dd = bv0.vi != c0.vi;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double w0 = q ? d.w10 : d.w00, w1 = q ? d.w11 : d.w01;
    P[1][1 + q] += dd ? 0.0 : w0;
    P[2][1 + q] += dd ? w0 : w1;
    P[3][1 + q] += dd ? w1 : 0.0;
  }
  const TapW e = tap_weights(bv1, gv1);
  const bool ue = bv1.vi != c0.vi;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double w0 = q ? e.w10 : e.w00, w1 = q ? e.w11 : e.w01;
    P[0][1 + q] += ue ? w0 : 0.0;
    P[1][1 + q] += ue ? w1 : w0;
    P[2][1 + q] += ue ? 0.0 : w1;
  }
  const int64_t o = (int64_t(c0.vi) - 1) * W + (c0.ui - 1);
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x)
      // Non-zero entries are inside the image (weights past any edge are
      // zero), so only in-range offsets are dereferenced.
      if (P[y][x] != 0.0)
        atomicAdd(ch + (o + int64_t(y) * W + x), float(P[y][x]));
}

// J'_bar and r'_bar of one weighted term (see the header comment).
__device__ inline void term_adjoint(const double J[6], double w, double s,
                                    double r, const double* lam,
                                    const double* x, double Jbar[6],
                                    double& rbar) {
  double Jw[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) Jw[a] = w * J[a];
  const double rp = s * r;
  const double xJ = dot6(x, Jw), lJ = dot6(lam, Jw);
#pragma unroll
  for (int a = 0; a < 6; ++a) Jbar[a] = w * (lam[a] * (rp - xJ) - x[a] * lJ);
  rbar = s * lJ;
}

// One matched pixel's terms at the FP64 pose, with the forward's decisions
// (see above): the geometric J and r, and, when the pixel passes the feature
// gate, the feature J and r with the samples and projection derivatives the
// adjoint needs.
struct PixelTerms {
  int64_t pos;
  V3<double> ps, p;  // source point, transformed point
  V3<double> q, n;   // target point and normal
  double Jg[6], rg;
  bool feat;
  BilD c0, bu0, bu1, bv0, bv1;
  bool cu, cv;       // lower samples not clamped at 0
  double du, dv, iz, z2, z3, j00, j11, j02, j12;
  V3<double> g;
  double Jf[6], rf;
};

// Whether source pixel idx is matched at the pose (m: FP32 rows, md: FP64
// rows), and its terms if so.
template <bool GEOM, bool FEAT>
__device__ inline bool pixel_terms(const LevelView<float>& L, const float* m,
                                   const double* md,
                                   const unsigned long long* __restrict__ map,
                                   uint32_t tag, int64_t idx,
                                   const TermWeights& tw, PixelTerms& T) {
  if (!L.sm[idx]) return false;
  const V3<float> ps = ld3f(L.sp, idx);
  // The forward's decisions: FP32 point and projection at the FP32 pose.
  const V3<float> pf = rigid_point(m, ps);
  float uf, vf;
  project(L.fx, L.fy, L.cx, L.cy, pf, uf, vf);
  int ui, vi;
  if (!round_in_range(uf, L.W, ui) || !round_in_range(vf, L.H, vi)) return false;
  T.pos = int64_t(vi) * L.W + ui;
  if (!key_owned(map[T.pos], uint32_t(idx), tag)) return false;
  // The smooth part in FP64 at the FP64 pose.
  T.ps = V3<double>{ps.x, ps.y, ps.z};
  T.p = rigid_point(md, T.ps);
  const V3<double>& p = T.p;
  const double fx = L.fx, fy = L.fy;
  if (GEOM) {
    const V3<float> qf = ld3f(L.tp, T.pos);
    const V3<float> nf = ld3f(L.tn, T.pos);
    T.q = V3<double>{qf.x, qf.y, qf.z};
    T.n = V3<double>{nf.x, nf.y, nf.z};
    T.rg = dot3(sub3(T.q, p), T.n);
    const V3<double> c = cross3(p, T.n);
    T.Jg[0] = T.n.x;
    T.Jg[1] = T.n.y;
    T.Jg[2] = T.n.z;
    T.Jg[3] = c.x;
    T.Jg[4] = c.y;
    T.Jg[5] = c.z;
  }
  T.feat = false;
  if (FEAT) {
    const int C = L.C;
    const float* __restrict__ tf = reinterpret_cast<const float*>(L.tf);
    const float* __restrict__ sf = reinterpret_cast<const float*>(L.sf);
    // The feature gate as the forward evaluates it (FP32).
    float sq = 0.0f;
    for (int c = 0; c < C; ++c) {
      const float res =
          bil_val32(tf + int64_t(c) * L.HW, L.H, L.W, uf, vf) - sf[c * L.n + idx];
      sq += res * res;
    }
    T.feat = !(sq > tw.fthr2);
    if (T.feat) {
      // Cells and clamps of the five samples from the FP32 positions.
      const float hf = 0.05f;
      T.cu = uf - hf > 0.0f;
      T.cv = vf - hf > 0.0f;
      const int u0 = trunc_int(uf), v0 = trunc_int(vf);
      const int up = trunc_int(uf + hf), vp = trunc_int(vf + hf);
      const int um = T.cu ? trunc_int(uf - hf) : 0;
      const int vm = T.cv ? trunc_int(vf - hf) : 0;
      const double h = 0.05, div = 1.0 / (2.0 * h);
      T.iz = 1.0 / p.z;
      const double u = fx * p.x * T.iz + double(L.cx);
      const double v = fy * p.y * T.iz + double(L.cy);
      const double ud = T.cu ? u - h : 0.0, vd = T.cv ? v - h : 0.0;
      const float* ch = tf + int64_t(C - 1) * L.HW;
      T.c0 = bil_at(ch, L.H, L.W, u0, v0, u, v);
      T.bu0 = bil_at(ch, L.H, L.W, up, v0, u + h, v);
      T.bu1 = bil_at(ch, L.H, L.W, um, v0, ud, v);
      T.bv0 = bil_at(ch, L.H, L.W, u0, vp, u, v + h);
      T.bv1 = bil_at(ch, L.H, L.W, u0, vm, u, vd);
      T.rf = T.c0.val - double(sf[int64_t(C - 1) * L.n + idx]);
      T.du = (T.bu0.val - T.bu1.val) * div;
      T.dv = (T.bv0.val - T.bv1.val) * div;
      T.z2 = p.z * p.z;
      T.z3 = T.z2 * p.z;
      T.j00 = fx * T.iz;
      T.j11 = fy * T.iz;
      T.j02 = -p.x * fx / T.z2;
      T.j12 = -p.y * fy / T.z2;
      T.g = V3<double>{T.du * T.j00, T.dv * T.j11, T.du * T.j02 + T.dv * T.j12};
      const V3<double> c = cross3(p, T.g);
      T.Jf[0] = T.g.x;
      T.Jf[1] = T.g.y;
      T.Jf[2] = T.g.z;
      T.Jf[3] = c.x;
      T.Jf[4] = c.y;
      T.Jf[5] = c.z;
    }
  }
  return true;
}

template <bool GEOM, bool FEAT>
__device__ inline void adjoint_pixel(const LevelView<float>& L, const float* m,
                                     const double* md,
                                     const unsigned long long* __restrict__ map,
                                     uint32_t tag, int64_t idx,
                                     const TermWeights& tw, const double* lam,
                                     const double* x, const stb_icp_grads& G,
                                     double (&tacc)[kPose]) {
  PixelTerms T;
  if (!pixel_terms<GEOM, FEAT>(L, m, md, map, tag, idx, tw, T)) return;
  const V3<double>& p = T.p;
  const double fx = L.fx, fy = L.fy;
  const int64_t pos = T.pos;
  V3<double> pb{0.0, 0.0, 0.0};  // dL/dp'
  if (GEOM) {
    const V3<double>& n = T.n;
    const V3<double> d = sub3(T.q, p);
    double Jb[6], rb;
    term_adjoint(T.Jg, tw.gw, 1.0, T.rg, lam, x, Jb, rb);
    // r = (q - p') . n
    V3<double> nb{d.x * rb + Jb[0], d.y * rb + Jb[1], d.z * rb + Jb[2]};
    const V3<double> wb{Jb[3], Jb[4], Jb[5]};
    // J_w = p' x n: (p' x n) . wb = p' . (n x wb) = n . (wb x p')
    const V3<double> a1 = cross3(n, wb), a2 = cross3(wb, p);
    pb.x += -n.x * rb + a1.x;
    pb.y += -n.y * rb + a1.y;
    pb.z += -n.z * rb + a1.z;
    nb.x += a2.x;
    nb.y += a2.y;
    nb.z += a2.z;
    // The map has one owner per target pixel, so these updates never collide.
    float* tq = G.d_tgt_points + 3 * pos;
    float* tn = G.d_tgt_normals + 3 * pos;
    tq[0] += float(n.x * rb);
    tq[1] += float(n.y * rb);
    tq[2] += float(n.z * rb);
    tn[0] += float(nb.x);
    tn[1] += float(nb.y);
    tn[2] += float(nb.z);
  }
  if (FEAT && T.feat) {
    const int C = L.C;
    const double div = 1.0 / (2.0 * 0.05);
    double Jb[6], rb;
    term_adjoint(T.Jf, tw.fw, -1.0, T.rf, lam, x, Jb, rb);
    const V3<double> wb{Jb[3], Jb[4], Jb[5]};
    // J_w = p' x g: g_bar += wb x p', p'_bar += g x wb
    const V3<double> a2 = cross3(wb, p), a1 = cross3(T.g, wb);
    const V3<double> gb{Jb[0] + a2.x, Jb[1] + a2.y, Jb[2] + a2.z};
    pb.x += a1.x;
    pb.y += a1.y;
    pb.z += a1.z;
    const double dub = gb.x * T.j00 + gb.z * T.j02;
    const double dvb = gb.y * T.j11 + gb.z * T.j12;
    const double j00b = gb.x * T.du, j11b = gb.y * T.dv;
    const double j02b = gb.z * T.du, j12b = gb.z * T.dv;
    pb.x += -j02b * fx / T.z2;
    pb.y += -j12b * fy / T.z2;
    pb.z += -j00b * fx / T.z2 - j11b * fy / T.z2 + 2.0 * j02b * p.x * fx / T.z3 +
            2.0 * j12b * p.y * fy / T.z3;
    // r = I(u, v) - f_s and the finite differences du, dv (the lower
    // samples do not move with u or v where they are clamped at 0).
    const double mu = T.cu ? 1.0 : 0.0, mv = T.cv ? 1.0 : 0.0;
    const double ub = rb * T.c0.du + dub * div * (T.bu0.du - mu * T.bu1.du) +
                      dvb * div * (T.bv0.du - T.bv1.du);
    const double vb = rb * T.c0.dv + dub * div * (T.bu0.dv - T.bu1.dv) +
                      dvb * div * (T.bv0.dv - mv * T.bv1.dv);
    float* gch = G.d_tgt_feats + int64_t(C - 1) * L.HW;
    scatter5(gch, L.W, T.c0, rb, T.bu0, dub * div, T.bu1, -dub * div, T.bv0,
             dvb * div, T.bv1, -dvb * div);
    G.d_src_feats[int64_t(C - 1) * L.n + idx] -= float(rb);
    // u = fx x / z + cx, v = fy y / z + cy
    pb.x += ub * fx * T.iz;
    pb.y += vb * fy * T.iz;
    pb.z += -ub * fx * p.x / T.z2 - vb * fy * p.y / T.z2;
  }
  // p' = R ps + t: ps_bar = R^T pb; dT[i][j] += pb_i ps_j, dT[i][3] += pb_i.
  float* sp = G.d_src_points + 3 * idx;
  sp[0] += float(md[0] * pb.x + md[4] * pb.y + md[8] * pb.z);
  sp[1] += float(md[1] * pb.x + md[5] * pb.y + md[9] * pb.z);
  sp[2] += float(md[2] * pb.x + md[6] * pb.y + md[10] * pb.z);
  const double pbv[3] = {pb.x, pb.y, pb.z};
  const double psv[3] = {T.ps.x, T.ps.y, T.ps.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) tacc[i * 4 + j] += pbv[i] * psv[j];
    tacc[i * 4 + 3] += pbv[i];
  }
}

// Adds one weighted term J' = w J, r' = s r to the 27 FP64 system slots
// (21 JtJ upper triangle, row-major, then 6 Jr).
__device__ inline void acc_term64(double* acc, const double J[6], double w,
                                  double s, double r) {
  double Jw[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) Jw[a] = w * J[a];
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) acc[k++] += Jw[a] * Jw[b];
#pragma unroll
  for (int a = 0; a < 6; ++a) acc[21 + a] += Jw[a] * (s * r);
}

// Block sum of N FP64 slots per thread (lanes by shuffles, waves through
// LDS), written as one kSlots-wide partial for a later fixed-order sum.
template <int N>
__device__ inline void block_sum_f64(double (&acc)[N], double* out, double* sh) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) sh[wave * N + k] = acc[k];
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double s = 0.0;
    if (threadIdx.x < N)
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += sh[w * N + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// The normal equations of one step recomputed in FP64 per pixel (same
// decisions as the forward): 21 JtJ + 6 Jr slots per tile.
template <bool GEOM, bool FEAT>
__global__ void __launch_bounds__(kBlock)
    batch_system64_kernel(const void* __restrict__ plan,
                          stb_icp_state* __restrict__ states,
                          unsigned long long* __restrict__ maps0,
                          unsigned long long* __restrict__ maps1,
                          double* __restrict__ partials, TermWeights tw) {
  __shared__ double sh[kWaves * 27];
  const PlanView pv = plan_view(plan);
  for (int t = blockIdx.x; t < pv.tiles; t += gridDim.x) {
    const int p = pv.tile_pair[t];
    stb_icp_state* st = states + p;
    const LevelView<float> L = view_of<float>(pv.levels[p]);
    const unsigned long long* map = pair_map(st, maps0, maps1, pv.map_off[p], 0);
    const uint32_t tag = uint32_t(st->map_epoch);
    double md[12];
    float m[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      m[i] = st->rt[i];
      md[i] = st->transform[i];
    }
    const int64_t base = int64_t(t - pv.tile_start[p]) * pv.tile_px;
    double acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = 0.0;
    for (int k = threadIdx.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t idx = base + k;
      PixelTerms T;
      if (idx < L.n && pixel_terms<GEOM, FEAT>(L, m, md, map, tag, idx, tw, T)) {
        if (GEOM) acc_term64(acc, T.Jg, tw.gw, 1.0, T.rg);
        if (FEAT && T.feat) acc_term64(acc, T.Jf, tw.fw, -1.0, T.rf);
      }
    }
    block_sum_f64<27>(acc, partials + int64_t(t) * kSlots, sh);
  }
}

// One block per pair: the fixed-order sum of the pair's tile partials into
// its first `count` slots.
__global__ void __launch_bounds__(kSolveBlock)
    batch_tile_sum_kernel(const void* __restrict__ plan,
                          const double* __restrict__ partials, int count,
                          double* __restrict__ out) {
  __shared__ double sh[kSolveBlock];
  __shared__ double tot[kSlots];
  const PlanView pv = plan_view(plan);
  const int p = blockIdx.x;
  if (p >= pv.P) return;
  const int first = pv.tile_start[p];
  sum_partials<kSolveBlock>(partials + int64_t(first) * kSlots,
                            pv.tile_start[p + 1] - first, sh, tot);
  if (int(threadIdx.x) < count) out[count * p + threadIdx.x] = tot[threadIdx.x];
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
    double lam[6], x[6], md[12];
    float m[12];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      lam[a] = coef[12 * p + a];
      x[a] = coef[12 * p + 6 + a];
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      m[i] = st->rt[i];
      md[i] = st->transform[i];
    }
    const int64_t base = int64_t(t - pv.tile_start[p]) * pv.tile_px;
    double tacc[kPose];
#pragma unroll
    for (int k = 0; k < kPose; ++k) tacc[k] = 0.0;
    for (int k = threadIdx.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t idx = base + k;
      if (idx < L.n)
        adjoint_pixel<GEOM, FEAT>(L, m, md, map, tag, idx, tw, lam, x, G, tacc);
    }
    block_sum_f64<kPose>(tacc, partials + int64_t(t) * kSlots, sh);
  }
}

template <bool GEOM, bool FEAT>
void launch_adjoint(const void* plan, int num_pairs, stb_icp_state* st,
                    const Scratch& sc, int grid, const stb_icp_grads* grads,
                    const double* coef, double* dT, const TermWeights& tw,
                    hipStream_t s) {
  hipLaunchKernelGGL((batch_adjoint_kernel<GEOM, FEAT>), dim3(grid),
                     dim3(kBlock), 0, s, plan, st, sc.map0, sc.map1, grads,
                     coef, sc.partials, tw);
  hipLaunchKernelGGL(batch_tile_sum_kernel, dim3(num_pairs), dim3(kSolveBlock),
                     0, s, plan, sc.partials, kPose, dT);
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
  STB_REQUIRE(sizes->feat_layout == STB_FEAT_PLANAR,
              "stb_icp_batch_adjoint: channel-planar features only");
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

extern "C" int stb_icp_batch_system64(const void* plan,
                                      const stb_icp_batch_sizes* sizes,
                                      const stb_icp_params* params,
                                      void* states, double* systems,
                                      void* scratch, size_t scratch_bytes,
                                      void* stream) {
  STB_REQUIRE(sizes && plan && params && states && systems && scratch,
              "stb_icp_batch_system64: NULL argument");
  STB_REQUIRE(sizes->storage == STB_F32,
              "stb_icp_batch_system64: FP32 storage only");
  STB_REQUIRE(sizes->feat_layout == STB_FEAT_PLANAR,
              "stb_icp_batch_system64: channel-planar features only");
  STB_REQUIRE(params->use_geom || params->use_feat,
              "stb_icp_batch_system64: no active term");
  const Scratch sc = carve_scratch(scratch, sizes->num_pairs,
                                   sizes->total_tiles, sizes->total_map);
  STB_REQUIRE(scratch_bytes >= sc.bytes, "stb_icp_batch_system64: scratch too small");
  hipStream_t s = as_stream(stream);
  stb_icp_state* st = static_cast<stb_icp_state*>(states);
  launch_batch_correspond_f32(plan, *sizes, *params, st, sc, s);
  const TermWeights tw = make_weights(*params);
  const int cap = params->grid_update > 0 ? params->grid_update : 1024;
  const int grid = sizes->total_tiles < cap ? sizes->total_tiles : cap;
  if (params->use_geom && params->use_feat)
    hipLaunchKernelGGL((batch_system64_kernel<true, true>), dim3(grid),
                       dim3(kBlock), 0, s, plan, st, sc.map0, sc.map1,
                       sc.partials, tw);
  else if (params->use_geom)
    hipLaunchKernelGGL((batch_system64_kernel<true, false>), dim3(grid),
                       dim3(kBlock), 0, s, plan, st, sc.map0, sc.map1,
                       sc.partials, tw);
  else
    hipLaunchKernelGGL((batch_system64_kernel<false, true>), dim3(grid),
                       dim3(kBlock), 0, s, plan, st, sc.map0, sc.map1,
                       sc.partials, tw);
  hipLaunchKernelGGL(batch_tile_sum_kernel, dim3(sizes->num_pairs),
                     dim3(kSolveBlock), 0, s, plan, sc.partials, 27, systems);
  return check_launch("stb_icp_batch_system64");
}
