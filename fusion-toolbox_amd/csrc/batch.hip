// Batched Gauss-Newton over many frame pairs of uneven resolution.
//
// A batch plan (built on the host by stb_icp_batch_plan, uploaded by the
// caller) lists one level descriptor per pair and cuts every pair's source
// pixels into equal tiles. Each iteration is three launches whatever the pair
// count or resolutions: correspondence and Jacobians are launches over the
// flattened tile list of all pairs (by default one workgroup per tile; with a
// grid cap a workgroup walks tiles t = b, b + gridDim.x, ..., where b is its
// XCD-chunked number xcd_chunk(blockIdx.x, gridDim.x, kTileChunk); every
// Jacobian tile leaves one partial, summed over the tile in FP64 and stored
// in FP32), then one workgroup per pair sums that pair's partials in FP64 in
// tile order, solves its 6x6 system in FP64 and updates its pose. All pairs
// advance one iteration per three launches with no host involvement, and a
// whole multi-level loop can be captured in one hipGraph.
#include "batch_plan.hpp"

#include <cstring>
#include <vector>

namespace stb {
namespace {

// Consecutive tiles per XCD chunk (xcd_chunk): 64 mixed pairs ran 8,255
// pairs/s with tiles dealt out one by one, 8,320 with chunks of 4 to 16, and
// 7,920 with each XCD taking one contiguous eighth of the tile list.
constexpr int kTileChunk = 8;

template <typename S>
__global__ void __launch_bounds__(kBlock)
    batch_corresp_kernel(const void* __restrict__ plan,
                         const stb_icp_state* __restrict__ states,
                         unsigned long long* __restrict__ maps0,
                         unsigned long long* __restrict__ maps1, float dthr2,
                         float cstar, int fixed_map) {
  const PlanView pv = plan_view(plan);
  for (int t = xcd_chunk(blockIdx.x, gridDim.x, kTileChunk); t < pv.tiles;
       t += gridDim.x) {
    const int p = pv.tile_pair[t];
    const stb_icp_state* st = states + p;
    if (!fixed_map && st->stopped) continue;
    const LevelView<S> L = view_of<S>(pv.levels[p]);
    const LevelRes R = res_of(L);
    float m[12], nm[9];
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) nm[i] = st->nm[i];
    unsigned long long* map =
        pair_map(st, maps0, maps1, pv.map_off[p], fixed_map);
    const uint32_t tag = fixed_map ? kFreshTag : uint32_t(st->map_epoch);
    const int64_t base = int64_t(t - pv.tile_start[p]) * pv.tile_px;
    for (int k = threadIdx.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t idx = base + k;
      if (idx < L.n) corresp_pixel(L, R, m, nm, idx, map, tag, dthr2, cstar);
    }
  }
}

template <typename S, bool GEOM, bool FEAT, int NB, bool IL = false>
__global__ void __launch_bounds__(kBlock, 4)
    batch_jacobian_kernel(const void* __restrict__ plan,
                          stb_icp_state* __restrict__ states,
                          unsigned long long* __restrict__ maps0,
                          unsigned long long* __restrict__ maps1,
                          float* __restrict__ partials, TermWeights tw,
                          int fixed_map) {
  __shared__ double sh[kShared];
  const PlanView pv = plan_view(plan);
  for (int t = xcd_chunk(blockIdx.x, gridDim.x, kTileChunk); t < pv.tiles;
       t += gridDim.x) {
    const int p = pv.tile_pair[t];
    const LevelView<S> L = view_of<S>(pv.levels[p]);
    const LevelRes R = res_of(L);
    const int first = pv.tile_start[p];
    const int ntiles = pv.tile_start[p + 1] - first;
    const int64_t base = int64_t(t - first) * pv.tile_px;
    // The tile's first source pixel loads before the pair's state is read
    // (the state was written by the previous launch and misses this XCD's
    // L2); for a stopped pair the loads are simply unused.
    int64_t idx = base + threadIdx.x;
    uint8_t vs = res_u8(R.sm, int(idx));
    V3<float> ps = res_ld3<S>(R.sp, int(idx));
    __builtin_amdgcn_sched_barrier(0);
    stb_icp_state* st = states + p;
    if (!fixed_map && st->stopped) continue;
    const int64_t off = pv.map_off[p];
    const unsigned long long* map = pair_map(st, maps0, maps1, off, fixed_map);
    const uint32_t tag = fixed_map ? kFreshTag : uint32_t(st->map_epoch);
    if (!fixed_map && tag == 0) {
      // The pair's last tag on this map: clear this tile's share of the
      // other map, which the next iteration uses with a fresh tag; the
      // pair's last tile also clears target pixels beyond the source count.
      unsigned long long* next = pair_map(st, maps0, maps1, off, 1);
      int64_t end = base + pv.tile_px;
      if (t - first == ntiles - 1 && end < L.HW) end = L.HW;
      if (end > L.HW) end = L.HW;
      for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x)
        next[i] = kKeySentinel;
    }
    float m[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
    float acc[kSlots];
    // The next pixel's source mask and point load while this pixel's gathers
    // and terms run, so a pixel waits for one memory round trip instead of
    // two. Past the tile (another tile's pixel) or the arrays' ends (the
    // buffer loads return 0) the prefetch is simply unused. The first pixel
    // starts the slots (every lane takes part; lanes past the tile or the
    // pair's pixels add zeros).
    {
      const int64_t nxt = idx + blockDim.x;
      const uint8_t vs_n = res_u8(R.sm, int(nxt));
      const V3<float> ps_n = res_ld3<S>(R.sp, int(nxt));
      const bool ok = (int(threadIdx.x) < pv.tile_px) & (idx < L.n) & (vs != 0);
      jacobian_pixel_at<S, GEOM, FEAT, true, NB, IL>(L, R, m, map, tag, idx,
                                                     ok, ps, tw, acc);
      idx = nxt;
      vs = vs_n;
      ps = ps_n;
    }
    for (int k = threadIdx.x + blockDim.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t nxt = idx + blockDim.x;
      const uint8_t vs_n = res_u8(R.sm, int(nxt));
      const V3<float> ps_n = res_ld3<S>(R.sp, int(nxt));
      if (idx < L.n)
        jacobian_pixel_at<S, GEOM, FEAT, false, NB, IL>(L, R, m, map, tag,
                                                        idx, vs != 0, ps, tw,
                                                        acc);
      idx = nxt;
      vs = vs_n;
      ps = ps_n;
    }
    block_partial(acc, partials + int64_t(t) * kSlots, sh, tw.b);
  }
}

// batch_jacobian_kernel as a software pipeline over a thread's four pixels
// of a tile of kPipeTile pixels (FP32 storage; planar features whose
// channels other than the last fit one batch of NB, or none): jac_issue /
// jac_consume as in jacobian_pipe_kernel, each next pixel's gathers issued
// before the current pixel's terms, so that they are in flight while the
// VALU works instead of after it. The same per-pixel products and
// decisions; only the order of the FP32 lane sums differs from
// batch_jacobian_kernel's. Pixels past the pair's source count evaluate as
// invalid (their buffer loads return 0). Two pixels' gathered data in
// registers (132 VGPRs with three-channel colour) leave three wavefronts per
// SIMD; 64 mixed pairs ran 9,031-9,067 pairs/s this way against 8,837-8,853
// with batch_jacobian_kernel, 8,670-8,700 when bounded to four wavefronts
// (four registers spilled), and 8,880-8,900 with the third and fourth
// pixels' source loads issued later, behind the gathers.
constexpr int kPipeTile = 4 * kBlock;

template <bool GEOM, bool FEAT, int NB>
__global__ void __launch_bounds__(kBlock, 3)
    batch_jacobian_pipe_kernel(const void* __restrict__ plan,
                               stb_icp_state* __restrict__ states,
                               unsigned long long* __restrict__ maps0,
                               unsigned long long* __restrict__ maps1,
                               float* __restrict__ partials, TermWeights tw,
                               int fixed_map) {
  __shared__ double sh[kShared];
  const PlanView pv = plan_view(plan);
  for (int t = xcd_chunk(blockIdx.x, gridDim.x, kTileChunk); t < pv.tiles;
       t += gridDim.x) {
    const int p = pv.tile_pair[t];
    const LevelView<float> L = view_of<float>(pv.levels[p]);
    const LevelRes R = res_of(L);
    const int first = pv.tile_start[p];
    const int ntiles = pv.tile_start[p + 1] - first;
    const int64_t base = int64_t(t - first) * kPipeTile;
    // All four pixels' source data first, before the pair's state (written
    // by the previous launch, an L2 miss) is read.
    const int64_t i0 = base + threadIdx.x, i1 = i0 + kBlock, i2 = i1 + kBlock,
                  i3 = i2 + kBlock;
    const uint8_t vs0 = res_u8(R.sm, int(i0)), vs1 = res_u8(R.sm, int(i1));
    const uint8_t vs2 = res_u8(R.sm, int(i2)), vs3 = res_u8(R.sm, int(i3));
    const V3<float> ps0 = res_ld3<float>(R.sp, int(i0));
    const V3<float> ps1 = res_ld3<float>(R.sp, int(i1));
    const V3<float> ps2 = res_ld3<float>(R.sp, int(i2));
    const V3<float> ps3 = res_ld3<float>(R.sp, int(i3));
    __builtin_amdgcn_sched_barrier(0);
    stb_icp_state* st = states + p;
    if (!fixed_map && st->stopped) continue;
    const int64_t off = pv.map_off[p];
    const unsigned long long* map = pair_map(st, maps0, maps1, off, fixed_map);
    const uint32_t tag = fixed_map ? kFreshTag : uint32_t(st->map_epoch);
    if (!fixed_map && tag == 0) {
      // As in batch_jacobian_kernel: this tile's share of the other map.
      unsigned long long* next = pair_map(st, maps0, maps1, off, 1);
      int64_t end = base + kPipeTile;
      if (t - first == ntiles - 1 && end < L.HW) end = L.HW;
      if (end > L.HW) end = L.HW;
      for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x)
        next[i] = kKeySentinel;
    }
    float m[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
    const __amdgpu_buffer_rsrc_t rmap = make_res(map, L.HW);
    PxGather<NB> A, B;
    float acc[kSlots];
    jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, i0, (i0 < L.n) & (vs0 != 0), ps0, A);
    __builtin_amdgcn_sched_barrier(0);
    jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, i1, (i1 < L.n) & (vs1 != 0), ps1, B);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, true, NB>(L, A, uint32_t(i0), tag, tw, acc);
    uint32_t k1 = uint32_t(i1);
    settle(acc, B, k1);
    __builtin_amdgcn_sched_barrier(0);
    jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, i2, (i2 < L.n) & (vs2 != 0), ps2, A);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, false, NB>(L, B, k1, tag, tw, acc);
    uint32_t k2 = uint32_t(i2);
    settle(acc, A, k2);
    __builtin_amdgcn_sched_barrier(0);
    jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, i3, (i3 < L.n) & (vs3 != 0), ps3, B);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, false, NB>(L, A, k2, tag, tw, acc);
    uint32_t k3 = uint32_t(i3);
    settle(acc, B, k3);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, false, NB>(L, B, k3, tag, tw, acc);
    block_partial(acc, partials + int64_t(t) * kSlots, sh, tw.b);
  }
}

// One block per pair: the fixed-order sum of the pair's tile partials, then
// its pose update, stored linearisation or information matrix (`mode`, a
// FinishMode).
__global__ void __launch_bounds__(kSolveBlock)
    batch_solve_kernel(const void* __restrict__ plan,
                       stb_icp_state* __restrict__ states,
                       const float* __restrict__ partials, double gw,
                       double fw, int mode, int fixed_map) {
  __shared__ double sh[kSolveBlock];
  __shared__ double tot[kSlots];
  const PlanView pv = plan_view(plan);
  const int p = blockIdx.x;
  if (p >= pv.P) return;
  stb_icp_state* st = states + p;
  if (!fixed_map && st->stopped) return;
  const int first = pv.tile_start[p];
  finish_partials<kSolveBlock>(st, partials + int64_t(first) * kSlots,
                  pv.tile_start[p + 1] - first, mode, gw, fw, sh, tot);
}

template <typename S>
__global__ void __launch_bounds__(kBlock)
    batch_information_kernel(const void* __restrict__ plan,
                             stb_icp_state* __restrict__ states,
                             unsigned long long* __restrict__ maps0,
                             unsigned long long* __restrict__ maps1,
                             float* __restrict__ partials) {
  __shared__ double sh[kShared];
  const PlanView pv = plan_view(plan);
  for (int t = xcd_chunk(blockIdx.x, gridDim.x, kTileChunk); t < pv.tiles;
       t += gridDim.x) {
    const int p = pv.tile_pair[t];
    stb_icp_state* st = states + p;
    const LevelView<S> L = view_of<S>(pv.levels[p]);
    const unsigned long long* map =
        pair_map(st, maps0, maps1, pv.map_off[p], 1);
    const int64_t base = int64_t(t - pv.tile_start[p]) * pv.tile_px;
    float m[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
    float acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;
    const LevelRes R = res_of(L);
    const __amdgpu_buffer_rsrc_t rmap = make_res(map, L.HW);
    for (int k = threadIdx.x; k < pv.tile_px; k += blockDim.x) {
      const int64_t idx = base + k;
      if (idx < L.n) information_pixel(L, R, m, rmap, kFreshTag, idx, acc);
    }
    block_partial(acc, partials + int64_t(t) * kSlots, sh);
  }
}

__global__ void batch_begin_kernel(const void* __restrict__ plan,
                                   stb_icp_state* __restrict__ states,
                                   const double* __restrict__ T0, int chain) {
  const PlanView pv = plan_view(plan);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pv.P) return;
  reset_state(states + p, T0 ? T0 + 16 * p : nullptr, chain);
}

// Clears map entries [0, total) of both maps (spare == 0) or, per pair
// (blockIdx.y), the pair's spare map (spare == 1).
__global__ void __launch_bounds__(kBlock)
    batch_fill_kernel(const void* __restrict__ plan,
                      const stb_icp_state* __restrict__ states,
                      unsigned long long* __restrict__ maps0,
                      unsigned long long* __restrict__ maps1, int64_t total,
                      int spare) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t start = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (!spare) {
    for (int64_t i = start; i < total; i += stride) {
      maps0[i] = kKeySentinel;
      maps1[i] = kKeySentinel;
    }
    return;
  }
  const PlanView pv = plan_view(plan);
  const int p = blockIdx.y;
  if (p >= pv.P) return;
  const stb_icp_level& L = pv.levels[p];
  const int64_t hw = int64_t(L.height) * L.width;
  unsigned long long* m = pair_map(states + p, maps0, maps1, pv.map_off[p], 1);
  for (int64_t i = start; i < hw; i += stride) m[i] = kKeySentinel;
}

int check_sizes(const stb_icp_batch_sizes* z) {
  STB_REQUIRE(z != nullptr, "batch sizes are NULL");
  STB_REQUIRE(z->num_pairs > 0 && z->total_tiles > 0 && z->tile_px > 0,
              "empty batch plan");
  STB_REQUIRE(z->storage == STB_F32 || z->storage == STB_BF16 ||
                  z->storage == STB_F16,
              "bad batch storage");
  return STB_OK;
}

int grid_tiles(const stb_icp_params* P, int tiles, int dflt) {
  const int cap = (P && P->grid_update > 0) ? P->grid_update : dflt;
  return tiles < cap ? tiles : cap;
}

template <typename S, bool GEOM, bool FEAT>
void launch_batch_jacobian(const void* plan, int num_pairs, int max_channels,
                           int feat_layout, int tile_px, stb_icp_state* states,
                           const Scratch& sc, int grid, const TermWeights& tw,
                           int do_solve, int fixed, hipStream_t s) {
  // The pipelined kernel for FP32 tiles of kPipeTile pixels whose features
  // (planar) have at most four channels besides the last; otherwise batches
  // of two feature channels when no pair has four besides its last, and
  // interleaved features in chunks of eight.
  const bool pipe = std::is_same<S, float>::value && tile_px == kPipeTile &&
                    (!FEAT || (feat_layout == STB_FEAT_PLANAR &&
                               max_channels - 1 <= 4));
  if (pipe && (!FEAT || max_channels - 1 <= 2))
    hipLaunchKernelGGL((batch_jacobian_pipe_kernel<GEOM, FEAT, 2>), dim3(grid),
                       dim3(kBlock), 0, s, plan, states, sc.map0, sc.map1,
                       sc.partials32, tw, fixed);
  else if (pipe)
    hipLaunchKernelGGL((batch_jacobian_pipe_kernel<GEOM, FEAT, 4>), dim3(grid),
                       dim3(kBlock), 0, s, plan, states, sc.map0, sc.map1,
                       sc.partials32, tw, fixed);
  else if (FEAT && feat_layout == STB_FEAT_INTERLEAVED)
    hipLaunchKernelGGL((batch_jacobian_kernel<S, GEOM, FEAT, 4, true>),
                       dim3(grid), dim3(kBlock), 0, s, plan, states, sc.map0,
                       sc.map1, sc.partials32, tw, fixed);
  else if (FEAT && max_channels - 1 < 4)
    hipLaunchKernelGGL((batch_jacobian_kernel<S, GEOM, FEAT, 2>), dim3(grid),
                       dim3(kBlock), 0, s, plan, states, sc.map0, sc.map1,
                       sc.partials32, tw, fixed);
  else
    hipLaunchKernelGGL((batch_jacobian_kernel<S, GEOM, FEAT, 4>), dim3(grid),
                       dim3(kBlock), 0, s, plan, states, sc.map0, sc.map1,
                       sc.partials32, tw, fixed);
  hipLaunchKernelGGL(batch_solve_kernel, dim3(num_pairs), dim3(kSolveBlock), 0,
                     s, plan, states, sc.partials32, tw.gw, tw.fw, do_solve,
                     fixed);
}

enum BatchPhase { kIterate, kLinearize, kInformation };

template <typename S>
void launch_batch(const void* plan, const stb_icp_batch_sizes& z,
                  const stb_icp_params& P, stb_icp_state* states,
                  const Scratch& sc, BatchPhase phase, int count,
                  hipStream_t s) {
  const float dthr2 = float(P.distance_thresh * P.distance_thresh);
  const float cstar = angle_gate_cos_f32(P.normal_angle_thresh);
  // By default one block per tile in both launches (64 mixed pairs, tiles of
  // 1024 pixels: 8,220 pairs/s, against 7,880 with 4,096 blocks looping over
  // the tiles and 7,680 with 1,024).
  const int gc = (P.grid_corresp > 0) ? (z.total_tiles < P.grid_corresp
                                             ? z.total_tiles
                                             : P.grid_corresp)
                                      : z.total_tiles;
  const int gj = grid_tiles(&P, z.total_tiles, z.total_tiles);
  const TermWeights tw = make_weights(P);
  const int fixed = phase == kIterate ? 0 : 1;
  const int reps = phase == kIterate ? count : 1;
  for (int it = 0; it < reps; ++it) {
    if (fixed) {
      dim3 g(int(grid_for(1 << 20, kBlock, 256)), z.num_pairs);
      hipLaunchKernelGGL(batch_fill_kernel, g, dim3(kBlock), 0, s, plan,
                         states, sc.map0, sc.map1, z.total_map, 1);
    }
    hipLaunchKernelGGL(batch_corresp_kernel<S>, dim3(gc), dim3(kBlock), 0, s,
                       plan, states, sc.map0, sc.map1, dthr2, cstar, fixed);
    if (phase == kInformation) {
      hipLaunchKernelGGL(batch_information_kernel<S>, dim3(gj), dim3(kBlock), 0,
                         s, plan, states, sc.map0, sc.map1, sc.partials32);
      hipLaunchKernelGGL(batch_solve_kernel, dim3(z.num_pairs),
                         dim3(kSolveBlock), 0, s, plan, states, sc.partials32,
                         0.0, 0.0, int(kStoreInformation), 1);
      continue;
    }
    const int solve = phase == kIterate ? 1 : 0;
    if (P.use_geom && P.use_feat)
      launch_batch_jacobian<S, true, true>(plan, z.num_pairs, z.max_channels,
                                           z.feat_layout, z.tile_px, states, sc,
                                           gj, tw, solve, fixed, s);
    else if (P.use_geom)
      launch_batch_jacobian<S, true, false>(plan, z.num_pairs, z.max_channels,
                                            z.feat_layout, z.tile_px, states,
                                            sc, gj, tw, solve, fixed, s);
    else
      launch_batch_jacobian<S, false, true>(plan, z.num_pairs, z.max_channels,
                                            z.feat_layout, z.tile_px, states,
                                            sc, gj, tw, solve, fixed, s);
  }
}

int run_batch(const void* plan, const stb_icp_batch_sizes* z,
              const stb_icp_params* params, void* states, void* scratch,
              size_t scratch_bytes, BatchPhase phase, int count, void* stream,
              const char* what) {
  int rc = check_sizes(z);
  if (rc) return rc;
  STB_REQUIRE(plan && params && states && scratch, "%s: NULL argument", what);
  if (phase != kInformation) {
    STB_REQUIRE(params->use_geom || params->use_feat,
                "%s: at least one of the geometric and feature terms must be on",
                what);
  }
  const Scratch sc =
      carve_scratch(scratch, z->num_pairs, z->total_tiles, z->total_map);
  STB_REQUIRE(scratch_bytes >= sc.bytes, "%s: scratch too small (%zu < %zu)",
              what, scratch_bytes, sc.bytes);
  stb_icp_state* st = static_cast<stb_icp_state*>(states);
  hipStream_t s = as_stream(stream);
  switch (z->storage) {
    case STB_F32:
      launch_batch<float>(plan, *z, *params, st, sc, phase, count, s);
      break;
    case STB_BF16:
      launch_batch<__hip_bfloat16>(plan, *z, *params, st, sc, phase, count, s);
      break;
    default:
      launch_batch<__half>(plan, *z, *params, st, sc, phase, count, s);
  }
  return check_launch(what);
}

}  // namespace

void launch_batch_correspond_f32(const void* plan,
                                 const stb_icp_batch_sizes& z,
                                 const stb_icp_params& P, stb_icp_state* states,
                                 const Scratch& sc, hipStream_t s) {
  const float dthr2 = float(P.distance_thresh * P.distance_thresh);
  const float cstar = angle_gate_cos_f32(P.normal_angle_thresh);
  const int gc = z.total_tiles;  // one block per tile
  hipLaunchKernelGGL(batch_corresp_kernel<float>, dim3(gc), dim3(kBlock), 0, s,
                     plan, states, sc.map0, sc.map1, dthr2, cstar, 0);
}

}  // namespace stb

using namespace stb;

extern "C" int stb_icp_batch_plan(int32_t num_pairs,
                                  const stb_icp_level* levels, int32_t tile_px,
                                  void* plan_host, size_t plan_capacity,
                                  stb_icp_batch_sizes* sizes) {
  STB_REQUIRE(num_pairs > 0 && levels != nullptr && sizes != nullptr,
              "stb_icp_batch_plan: bad arguments");
  STB_REQUIRE(tile_px > 0 && tile_px % kBlock == 0,
              "stb_icp_batch_plan: tile_px must be a positive multiple of %d",
              kBlock);
  const int32_t storage = levels[0].storage;
  const int32_t layout = levels[0].feat_layout;
  std::vector<int32_t> start(size_t(num_pairs) + 1, 0);
  std::vector<int64_t> off(size_t(num_pairs), 0);
  int64_t tiles = 0, maps = 0;
  int32_t max_channels = 0;
  for (int p = 0; p < num_pairs; ++p) {
    const stb_icp_level& L = levels[p];
    if (L.channels > max_channels) max_channels = L.channels;
    STB_REQUIRE(L.src_points && L.src_normals && L.src_mask && L.tgt_points &&
                    L.tgt_normals && L.tgt_mask,
                "stb_icp_batch_plan: pair %d has a NULL array", p);
    STB_REQUIRE(L.n_src > 0 && L.height > 0 && L.width > 0,
                "stb_icp_batch_plan: pair %d is empty", p);
    STB_REQUIRE(L.n_src <= kMaxKeyIndex,
                "stb_icp_batch_plan: pair %d has more than %lld source pixels",
                p, (long long)kMaxKeyIndex);
    STB_REQUIRE(L.storage == storage,
                "stb_icp_batch_plan: all pairs must share one storage type");
    STB_REQUIRE(L.feat_layout == layout,
                "stb_icp_batch_plan: all pairs must share one feature layout");
    STB_REQUIRE(layout == STB_FEAT_PLANAR ||
                    (layout == STB_FEAT_INTERLEAVED && storage != STB_F32 &&
                     L.channels <= 32 && L.channels % 2 == 0),
                "stb_icp_batch_plan: pair %d: feat_layout must be planar, or "
                "interleaved with 16-bit storage and an even channel count "
                "of at most 32", p);
    STB_REQUIRE(level_addressable(L),
                "stb_icp_batch_plan: every array of pair %d must hold fewer "
                "than 2^31 bytes", p);
    start[p] = int32_t(tiles);
    off[p] = maps;
    tiles += (L.n_src + tile_px - 1) / tile_px;
    maps += int64_t(L.height) * L.width;
    STB_REQUIRE(tiles < (int64_t(1) << 30), "stb_icp_batch_plan: too many tiles");
  }
  start[num_pairs] = int32_t(tiles);
  const PlanLayout l = plan_layout(num_pairs, int(tiles));
  const Scratch sc = carve_scratch(nullptr, num_pairs, int(tiles), maps);
  sizes->num_pairs = num_pairs;
  sizes->total_tiles = int32_t(tiles);
  sizes->tile_px = tile_px;
  sizes->storage = storage;
  sizes->total_map = maps;
  sizes->plan_bytes = l.bytes;
  sizes->scratch_bytes = sc.bytes;
  sizes->max_channels = max_channels;
  sizes->feat_layout = layout;
  if (plan_host == nullptr) return STB_OK;  // size query
  STB_REQUIRE(plan_capacity >= l.bytes, "stb_icp_batch_plan: buffer too small");
  char* b = static_cast<char*>(plan_host);
  std::memset(b, 0, l.bytes);
  PlanHeader h{num_pairs, int32_t(tiles), tile_px, storage, maps, 0};
  std::memcpy(b, &h, sizeof(h));
  std::memcpy(b + l.levels, levels, sizeof(stb_icp_level) * size_t(num_pairs));
  std::memcpy(b + l.map_off, off.data(), sizeof(int64_t) * size_t(num_pairs));
  std::memcpy(b + l.tile_start, start.data(),
              sizeof(int32_t) * size_t(num_pairs + 1));
  int32_t* tp = reinterpret_cast<int32_t*>(b + l.tile_pair);
  for (int p = 0; p < num_pairs; ++p)
    for (int32_t t = start[p]; t < start[p + 1]; ++t) tp[t] = p;
  return STB_OK;
}

extern "C" int stb_icp_batch_begin(const void* plan,
                                   const stb_icp_batch_sizes* sizes,
                                   void* states, const double* transforms0,
                                   int32_t chain, void* scratch,
                                   size_t scratch_bytes, void* stream) {
  int rc = check_sizes(sizes);
  if (rc) return rc;
  STB_REQUIRE(plan && states && scratch, "stb_icp_batch_begin: NULL argument");
  const Scratch sc = carve_scratch(scratch, sizes->num_pairs,
                                   sizes->total_tiles, sizes->total_map);
  STB_REQUIRE(scratch_bytes >= sc.bytes, "stb_icp_batch_begin: scratch too small");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(batch_begin_kernel, dim3((sizes->num_pairs + 63) / 64),
                     dim3(64), 0, s, plan, static_cast<stb_icp_state*>(states),
                     transforms0, chain);
  hipLaunchKernelGGL(batch_fill_kernel,
                     dim3(grid_for(sizes->total_map, kBlock, 4096)),
                     dim3(kBlock), 0, s, plan,
                     static_cast<const stb_icp_state*>(states), sc.map0,
                     sc.map1, sizes->total_map, 0);
  return check_launch("stb_icp_batch_begin");
}

extern "C" int stb_icp_batch_iterate(const void* plan,
                                     const stb_icp_batch_sizes* sizes,
                                     const stb_icp_params* params, void* states,
                                     void* scratch, size_t scratch_bytes,
                                     int32_t num_iters, void* stream) {
  STB_REQUIRE(num_iters >= 0, "stb_icp_batch_iterate: num_iters must be >= 0");
  if (num_iters == 0) return STB_OK;
  return run_batch(plan, sizes, params, states, scratch, scratch_bytes, kIterate,
                   num_iters, stream, "stb_icp_batch_iterate");
}

extern "C" int stb_icp_batch_linearize(const void* plan,
                                       const stb_icp_batch_sizes* sizes,
                                       const stb_icp_params* params,
                                       void* states, void* scratch,
                                       size_t scratch_bytes, void* stream) {
  return run_batch(plan, sizes, params, states, scratch, scratch_bytes,
                   kLinearize, 1, stream, "stb_icp_batch_linearize");
}

extern "C" int stb_icp_batch_information(const void* plan,
                                         const stb_icp_batch_sizes* sizes,
                                         const stb_icp_params* params,
                                         void* states, void* scratch,
                                         size_t scratch_bytes, void* stream) {
  return run_batch(plan, sizes, params, states, scratch, scratch_bytes,
                   kInformation, 1, stream, "stb_icp_batch_information");
}
