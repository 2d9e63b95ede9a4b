// Fused single-pair Gauss-Newton path (ICPOdometry): per iteration a
// correspondence kernel, a Jacobian kernel that leaves one FP64 partial per
// block, and a one-block kernel that sums the partials in a fixed order,
// solves the 6x6 system and updates the pose on the device. (A last-block
// ticket inside the Jacobian launch would cost every block a device-scope
// fence and an atomic on one shared counter, about 20 ns per block in all --
// 10 us for 512 blocks, several times the cost of the third launch.)
#include "gn_device.hpp"

namespace stb {
namespace {

// Device map choice: the GN loop (fixed_map == 0) uses the map state.map_sel
// names; one-off passes (fixed_map == 1: information, linearisation) use the
// other, "spare" map, which the loop clears before its next use. Because the
// choice lives on the device, a captured sequence of any length replays
// correctly.
__device__ inline unsigned long long* pick_map(const stb_icp_state* st,
                                               unsigned long long* map0,
                                               unsigned long long* map1,
                                               int fixed_map) {
  return (st->map_sel ^ fixed_map) ? map1 : map0;
}

template <typename S>
__global__ void __launch_bounds__(kBlock)
    corresp_fused_kernel(stb_icp_level Ld, const stb_icp_state* __restrict__ st,
                         unsigned long long* __restrict__ map0,
                         unsigned long long* __restrict__ map1, float dthr2,
                         float cstar, int fixed_map) {
  // No early exit on state.stopped: a stopped loop's remaining launches
  // only rewrite maps that nothing reads again (the solve launch checks the
  // flag and leaves the state alone), and not testing it first lets the
  // pose, the map choice and the pixel data load together.
  unsigned long long* __restrict__ map = pick_map(st, map0, map1, fixed_map);
  float m[12], nm[9];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) nm[i] = st->nm[i];
  const LevelView<S> L = view_of<S>(Ld);
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < L.n;
       idx += int64_t(gridDim.x) * blockDim.x)
    corresp_pixel(L, m, nm, idx, map, dthr2, cstar);
}

// Per-pixel Jacobians of one iteration; every block leaves one FP64 partial
// for solve_fused_kernel (the next launch), so no block waits on another.
// With ONE_PX each thread owns exactly one pixel (levels that fit the grid
// cap): the slot accumulators are then born from that pixel's terms instead
// of living across a loop (~84 VGPRs instead of ~116). Larger levels use the
// grid-stride loop.
template <typename S, bool GEOM, bool FEAT, bool ONE_PX, int NB>
__global__ void __launch_bounds__(kBlock)
    jacobian_fused_kernel(stb_icp_level Ld, stb_icp_state* __restrict__ st,
                          unsigned long long* __restrict__ map0,
                          unsigned long long* __restrict__ map1,
                          int64_t map_len, double* __restrict__ partials,
                          TermWeights tw, int fixed_map) {
  __shared__ double sh[kShared];
  // No early exit on state.stopped (see corresp_fused_kernel): the solve
  // launch ignores the partials of a stopped loop.
  const unsigned long long* __restrict__ map =
      pick_map(st, map0, map1, fixed_map);
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  const int64_t gtid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
  const LevelView<S> L = view_of<S>(Ld);
  float acc[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;
  if (ONE_PX) {
    if (gtid < L.n) jacobian_pixel<S, GEOM, FEAT, NB>(L, m, map, gtid, tw, acc);
  } else {
    for (int64_t idx = gtid; idx < L.n; idx += gsize)
      jacobian_pixel<S, GEOM, FEAT, NB>(L, m, map, idx, tw, acc);
  }
  if (!fixed_map) {  // clear the other map for the next iteration
    unsigned long long* __restrict__ next = pick_map(st, map0, map1, 1);
    for (int64_t i = gtid; i < map_len; i += gsize) next[i] = kKeySentinel;
  }
  block_partial(acc, partials + int64_t(blockIdx.x) * kSlots, sh, false);
}

// The fixed-order sum of the Jacobian launch's `count` partials, then the
// pose update (do_solve) or the stored linearisation. kSolveGroups blocks
// each sum a contiguous range of partials into a group partial; the last to
// arrive adds the groups in order and finishes -- a handful of fences and
// tickets instead of one per Jacobian block, and many CUs pulling partials.
__global__ void __launch_bounds__(kBlock)
    solve_fused_kernel(stb_icp_state* __restrict__ st,
                       const double* __restrict__ partials, int count,
                       double* __restrict__ groups, double gw, double fw,
                       int do_solve, int fixed_map) {
  __shared__ double sh[kShared];
  __shared__ double tot[kSlots];
  __shared__ int s_last;
  if (!fixed_map && st->stopped) return;
  const int G = int(gridDim.x), g = int(blockIdx.x);
  const int lo = int(int64_t(count) * g / G);
  const int hi = int(int64_t(count) * (g + 1) / G);
  sum_partials<kBlock>(partials + int64_t(lo) * kSlots, hi - lo, sh, tot);
  if (threadIdx.x < kSlots) {
    groups[g * kSlots + threadIdx.x] = tot[threadIdx.x];
    __threadfence();
  }
  __syncthreads();
  if (!arrive_last(&st->ticket, G - 1, &s_last)) return;
  if (threadIdx.x == 0) st->ticket = 0;
  finish_partials<kBlock>(st, groups, G, do_solve, gw, fw, sh, tot);
}

template <typename S>
__global__ void __launch_bounds__(kBlock)
    information_fused_kernel(stb_icp_level Ld, stb_icp_state* __restrict__ st,
                             unsigned long long* __restrict__ map0,
                             unsigned long long* __restrict__ map1,
                             double* __restrict__ partials) {
  __shared__ double sh[kShared];
  __shared__ double tot[kSlots];
  __shared__ int s_last;
  const unsigned long long* __restrict__ map = pick_map(st, map0, map1, 1);
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
  const LevelView<S> L = view_of<S>(Ld);
  float acc[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < L.n;
       idx += int64_t(gridDim.x) * blockDim.x)
    information_pixel(L, m, map, idx, acc);
  block_partial(acc, partials + int64_t(blockIdx.x) * kSlots, sh, true);
  if (!arrive_last(&st->ticket, int(gridDim.x) - 1, &s_last)) return;
  sum_partials<kBlock>(partials, int(gridDim.x), sh, tot);
  if (threadIdx.x == 0) st->ticket = 0;
  store_information(st, tot);
}

__global__ void icp_begin_kernel(stb_icp_state* st,
                                 const double* __restrict__ T0, int chain) {
  if (threadIdx.x == 0) reset_state(st, T0, chain);
}

__global__ void __launch_bounds__(kBlock)
    fill_maps_kernel(unsigned long long* __restrict__ map0,
                     unsigned long long* __restrict__ map1, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    map0[i] = kKeySentinel;
    map1[i] = kKeySentinel;
  }
}

// Clears the spare map (the one state.map_sel does not name).
__global__ void __launch_bounds__(kBlock)
    fill_spare_map_kernel(const stb_icp_state* __restrict__ st,
                          unsigned long long* __restrict__ map0,
                          unsigned long long* __restrict__ map1, int64_t n) {
  unsigned long long* __restrict__ p = pick_map(st, map0, map1, 1);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = kKeySentinel;
}

// ---------------------------------------------------------------------------
// Workspace layout: [state | map0 | map1 | partials]
// ---------------------------------------------------------------------------

constexpr int kDefaultUpdateBlocks = 1024;
constexpr int kMaxUpdateBlocks = 4096;
constexpr int kInformationBlocks = 256;
constexpr int kSolveGroups = 16;
constexpr int kDefaultCorrespBlocks = 4096;

struct Workspace {
  stb_icp_state* st;
  unsigned long long* map[2];
  double* partials;
  double* groups;
  int64_t map_len;
};

size_t ws_bytes(int32_t H, int32_t W) {
  const size_t map_len = size_t(H) * size_t(W);
  return align256(sizeof(stb_icp_state)) + 2 * align256(map_len * 8) +
         align256(size_t(kMaxUpdateBlocks) * kSlots * sizeof(double)) +
         align256(size_t(kSolveGroups) * kSlots * sizeof(double));
}

Workspace carve(void* base, int32_t H, int32_t W) {
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
  p += align256(size_t(kMaxUpdateBlocks) * kSlots * sizeof(double));
  w.groups = reinterpret_cast<double*>(p);
  return w;
}

// Jacobian grid: one pixel per thread when that fits the cap (grid_update,
// or kDefaultUpdateBlocks), else `cap` blocks looping over pixels. The
// kernel is bound by vector-instruction issue once all its loads are in
// flight, and at 640x480 1024 looping blocks beat 1200 one-pixel blocks
// (7.3 against 7.9 us): each block's reduction is shared by more pixels.
int update_grid(const stb_icp_params& P, int64_t n, bool* one_px) {
  int cap = P.grid_update > 0 ? P.grid_update : kDefaultUpdateBlocks;
  if (cap > kMaxUpdateBlocks) cap = kMaxUpdateBlocks;
  const int g = grid_for(n);
  *one_px = g <= cap;
  return g > cap ? cap : g;
}

int corresp_grid(const stb_icp_params& P, int64_t n) {
  const int cap = P.grid_corresp > 0 ? P.grid_corresp : kDefaultCorrespBlocks;
  return grid_for(n, kBlock, cap);
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
// the loop; kLinearize and kInformation are one-off passes on the spare map.
enum Phase { kCorrespond, kUpdate, kLinearize, kInformation, kJacobians };

template <typename S, bool GEOM, bool FEAT, int NB>
void launch_jacobian_nb(const stb_icp_level& L, int grid, bool one_px,
                        const Workspace& w, const TermWeights& tw,
                        int fixed_map, hipStream_t stream) {
  if (one_px)
    hipLaunchKernelGGL((jacobian_fused_kernel<S, GEOM, FEAT, true, NB>),
                       dim3(grid), dim3(kBlock), 0, stream, L, w.st, w.map[0],
                       w.map[1], w.map_len, w.partials, tw, fixed_map);
  else
    hipLaunchKernelGGL((jacobian_fused_kernel<S, GEOM, FEAT, false, NB>),
                       dim3(grid), dim3(kBlock), 0, stream, L, w.st, w.map[0],
                       w.map[1], w.map_len, w.partials, tw, fixed_map);
}

template <typename S, bool GEOM, bool FEAT>
void launch_jacobian(const stb_icp_level& L, const stb_icp_params& P,
                     const Workspace& w, const TermWeights& tw, int do_solve,
                     int fixed_map, hipStream_t stream) {
  bool one_px = false;
  const int grid = update_grid(P, L.n_src, &one_px);
  // Feature channels other than the last load in batches of four, or of two
  // when there are fewer than four of them (three-channel colour: no
  // padding slots).
  if (FEAT && L.channels - 1 < 4)
    launch_jacobian_nb<S, GEOM, FEAT, 2>(L, grid, one_px, w, tw, fixed_map, stream);
  else
    launch_jacobian_nb<S, GEOM, FEAT, 4>(L, grid, one_px, w, tw, fixed_map, stream);
  if (do_solve < 0) return;  // the Jacobian launch alone
  const int groups = grid < kSolveGroups ? grid : kSolveGroups;
  hipLaunchKernelGGL(solve_fused_kernel, dim3(groups), dim3(kBlock), 0, stream,
                     w.st, w.partials, grid, w.groups, tw.gw, tw.fw, do_solve,
                     fixed_map);
}

template <typename S>
void launch_phase(const stb_icp_level& L, const stb_icp_params& P,
                  const Workspace& w, Phase phase, hipStream_t stream) {
  const float dthr2 = float(P.distance_thresh * P.distance_thresh);
  const float cstar = angle_gate_cos_f32(P.normal_angle_thresh);
  const int fixed =
      (phase == kLinearize || phase == kInformation || phase == kJacobians) ? 1
                                                                            : 0;
  if (fixed && phase != kJacobians) {
    hipLaunchKernelGGL(fill_spare_map_kernel,
                       dim3(grid_for(w.map_len, kBlock, 2048)), dim3(kBlock), 0,
                       stream, w.st, w.map[0], w.map[1], w.map_len);
  }
  if (phase != kUpdate && phase != kJacobians) {
    hipLaunchKernelGGL(corresp_fused_kernel<S>, dim3(corresp_grid(P, L.n_src)),
                       dim3(kBlock), 0, stream, L, w.st, w.map[0], w.map[1],
                       dthr2, cstar, fixed);
  }
  if (phase == kCorrespond) return;
  if (phase == kInformation) {
    // A one-off pass reduced by its last block: every block pays a fence and
    // a ticket, so a few hundred blocks loop over the pixels.
    hipLaunchKernelGGL(information_fused_kernel<S>,
                       dim3(grid_for(L.n_src, kBlock, kInformationBlocks)),
                       dim3(kBlock), 0, stream, L, w.st, w.map[0], w.map[1],
                       w.partials);
    return;
  }
  const TermWeights tw = make_weights(P);
  const int solve = phase == kUpdate ? 1 : (phase == kJacobians ? -1 : 0);
  if (P.use_geom && P.use_feat)
    launch_jacobian<S, true, true>(L, P, w, tw, solve, fixed, stream);
  else if (P.use_geom)
    launch_jacobian<S, true, false>(L, P, w, tw, solve, fixed, stream);
  else
    launch_jacobian<S, false, true>(L, P, w, tw, solve, fixed, stream);
}

template <typename S>
void launch_iterations(const stb_icp_level& L, const stb_icp_params& P,
                       const Workspace& w, int count, hipStream_t stream) {
  for (int it = 0; it < count; ++it) {
    launch_phase<S>(L, P, w, kCorrespond, stream);
    launch_phase<S>(L, P, w, kUpdate, stream);
  }
}

// Validates the common arguments of the fused entry points and dispatches on
// the storage type. phase < 0 runs `count` whole iterations.
int run_fused(const stb_icp_level* level, const stb_icp_params* params,
              void* workspace, size_t workspace_bytes, int phase, int count,
              void* stream, const char* what) {
  STB_REQUIRE(params != nullptr && workspace != nullptr, "%s: NULL argument",
              what);
  const bool jac = phase == kUpdate || phase == kLinearize ||
                   phase == kJacobians || phase < 0;
  int rc = check_level(level, jac && params->use_feat != 0);
  if (rc) return rc;
  if (jac) {
    STB_REQUIRE(params->use_geom || params->use_feat,
                "%s: at least one of the geometric and feature terms must be on",
                what);
  }
  STB_REQUIRE(workspace_bytes >= ws_bytes(level->height, level->width),
              "%s: workspace too small", what);
  const Workspace w = carve(workspace, level->height, level->width);
  hipStream_t s = as_stream(stream);
  switch (level->storage) {
    case STB_F32:
      if (phase < 0)
        launch_iterations<float>(*level, *params, w, count, s);
      else
        launch_phase<float>(*level, *params, w, Phase(phase), s);
      break;
    case STB_BF16:
      if (phase < 0)
        launch_iterations<__hip_bfloat16>(*level, *params, w, count, s);
      else
        launch_phase<__hip_bfloat16>(*level, *params, w, Phase(phase), s);
      break;
    default:
      if (phase < 0)
        launch_iterations<__half>(*level, *params, w, count, s);
      else
        launch_phase<__half>(*level, *params, w, Phase(phase), s);
  }
  return check_launch(what);
}

}  // namespace
}  // namespace stb

using namespace stb;

extern "C" size_t stb_icp_workspace_bytes(int64_t n_src, int32_t height,
                                          int32_t width) {
  (void)n_src;
  return ws_bytes(height, width);
}

extern "C" int stb_icp_begin(void* workspace, size_t workspace_bytes,
                             const double* transform0, int32_t chain,
                             int64_t n_src, int32_t height, int32_t width,
                             void* stream) {
  (void)n_src;
  STB_REQUIRE(workspace != nullptr, "workspace is NULL");
  STB_REQUIRE(height > 0 && width > 0, "bad target size");
  STB_REQUIRE(workspace_bytes >= ws_bytes(height, width),
              "workspace too small: %zu < %zu bytes", workspace_bytes,
              ws_bytes(height, width));
  const Workspace w = carve(workspace, height, width);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(icp_begin_kernel, dim3(1), dim3(64), 0, s, w.st,
                     transform0, chain);
  hipLaunchKernelGGL(fill_maps_kernel, dim3(grid_for(w.map_len, kBlock, 2048)),
                     dim3(kBlock), 0, s, w.map[0], w.map[1], w.map_len);
  return check_launch("stb_icp_begin");
}

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

extern "C" int stb_icp_jacobians(const stb_icp_level* level,
                                 const stb_icp_params* params, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  return run_fused(level, params, workspace, workspace_bytes, kJacobians, 1,
                   stream, "stb_icp_jacobians");
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
