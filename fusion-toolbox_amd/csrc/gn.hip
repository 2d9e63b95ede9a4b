// Fused single-pair Gauss-Newton path (ICPOdometry): per iteration a
// correspondence kernel and a Jacobian kernel that leaves one partial per
// block (summed over the block in FP64, stored in FP32). The partials are
// summed in a fixed order, the 6x6 system solved and the pose updated on the
// device, by the next iteration's correspondence launch (every block
// redundantly, see SolveRecord) or, after the last iteration and in the
// split entry points, by a small launch of its own. (A
// last-block ticket inside the Jacobian launch would cost every block a
// device-scope fence and an atomic on one shared counter, about 20 ns per
// block in all -- 10 us for 512 blocks.)
#include "gn_device.hpp"

#include <mutex>

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
  const uint32_t tag = fixed_map ? kFreshTag : uint32_t(st->map_epoch);
  float m[12], nm[9];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) nm[i] = st->nm[i];
  const LevelView<S> L = view_of<S>(Ld);
  const LevelRes R = res_of(L);
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  for (int64_t idx = int64_t(blk) * blockDim.x + threadIdx.x; idx < L.n;
       idx += int64_t(gridDim.x) * blockDim.x)
    corresp_pixel(L, R, m, nm, idx, map, tag, dthr2, cstar);
}

// The correspondence launch of iteration k > 0 of a loop: every block first
// sums iteration k-1's `count` partials (sum_partials, the same order in
// every block), solves and updates the pose in one wavefront, then builds
// the map at the new pose. Block 0 leaves the outcome in `rec` for the
// Jacobian launch of the same iteration, which commits it to the state.
// Every load that does not depend on the new pose -- the thread's first
// source pixel, the previous pose, the partials, the state flags -- issues
// before the first wait, so the update costs one memory round trip plus the
// factorisation; barriers do not wait for outstanding global loads. With
// TWO_PX (runs longer than a block) the thread's second source pixel loads
// up front too.
template <typename S, bool TWO_PX>
__global__ void __launch_bounds__(kSolveBlock)
    corresp_solve_kernel(stb_icp_level Ld, const stb_icp_state* __restrict__ st,
                         SolveRecord* __restrict__ rec,
                         const float* __restrict__ partials, int count,
                         unsigned long long* __restrict__ map0,
                         unsigned long long* __restrict__ map1, float dthr2,
                         float cstar) {
#pragma clang fp contract(off)
  __shared__ double sh[kSolveBlock];
  __shared__ double tot[kSlots];
  __shared__ double To[12], Tn[12];
  __shared__ float rtf[12], nmf[9];
  __shared__ int s_ok;
  const int t = threadIdx.x;
  const LevelView<S> L = view_of<S>(Ld);
  // A contiguous run of pixels per block (see jacobian_fused_kernel).
  const int64_t chunk = (L.n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = int64_t(xcd_block(blockIdx.x, gridDim.x)) * chunk;
  const int64_t hi = lo + chunk < L.n ? lo + chunk : L.n;
  const int64_t idx = lo + t;
  // The partials first: the sum waits only for them, the loads below stay
  // in flight through it (loads return in issue order).
  double pv[kSumBatch];
  load_partial_batch<kSolveBlock>(partials, count, t / kSlots, pv);
  // Then what the solve uses (the previous pose row and the words stopped,
  // iterations, map_epoch, map_sel, consecutive, in one vector load, lane l
  // of each wavefront taking word l % 4: a uniform scalar load would be
  // issued where first used, after the factorisation), then the source.
  const double row = st->transform[t < 12 ? t : 0];
  const int32_t flag = (&st->stopped)[t & 3];
  // Keeps the scheduler from hoisting the source loads above these.
  __builtin_amdgcn_sched_barrier(0);
  // Unconditional, so that no branch waits for them (past the end of the
  // arrays the buffer loads return 0).
  const LevelRes R = res_of(L);
  const uint8_t vs = res_u8(R.sm, int(idx));
  const V3<float> ps = res_ld3<S>(R.sp, int(idx));
  const V3<float> ns = res_ld3<S>(R.sn, int(idx));
  // The thread's second pixel, or its first again (see
  // jacobian_fused_kernel). Only with TWO_PX: loads return in issue order,
  // so needless ones would delay the partials.
  const int64_t idx1 = idx + kSolveBlock;
  uint8_t vs1 = 0;
  V3<float> ps1 = {0.0f, 0.0f, 0.0f}, ns1 = {0.0f, 0.0f, 0.0f};
  if (TWO_PX) {
    const int64_t j1 = idx1 < hi ? idx1 : idx;
    vs1 = res_u8(R.sm, int(j1));
    ps1 = res_ld3<S>(R.sp, int(j1));
    ns1 = res_ld3<S>(R.sn, int(j1));
  }
  sum_partials_from<kSolveBlock>(pv, partials, count, sh, tot);
  if (t < 12) To[t] = row;
  __syncthreads();
  if (t < 64) {
    const bool good = wave_solve_update(tot, To, Tn, rtf, nmf);
    if (t == 0) s_ok = good;
  }
  const int stopped = __shfl(flag, 0, 64);
  const int iters = __shfl(flag, 1, 64);
  const uint32_t tag = uint32_t(__shfl(flag, 2, 64));
  const int sel = __shfl(flag, 3, 64);
  __syncthreads();
  const bool ok = !stopped && s_ok != 0;
  Rigid<float> r;
  if (ok) {
#pragma unroll
    for (int i = 0; i < 12; ++i) r.m[i] = rtf[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) r.nm[i] = nmf[i];
  } else {  // stopped before, or stopping now: the pose stays
#pragma unroll
    for (int i = 0; i < 12; ++i) r.m[i] = st->rt[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) r.nm[i] = st->nm[i];
  }
  // This iteration's map use: the next after the last iteration's, unless the
  // loop had already stopped (then nothing advances; the same keys are
  // written again).
  const MapUse use = stopped ? MapUse{tag, sel} : next_map_use(tag, sel);
  if (blockIdx.x == 0) {
    if (!stopped && t < kSlots) rec->tot[t] = tot[t];
    if (ok && t < 12) rec->T[t] = Tn[t];
    if (t == 0) {
#pragma unroll
      for (int i = 0; i < 12; ++i) rec->rt[i] = r.m[i];
#pragma unroll
      for (int i = 0; i < 9; ++i) rec->nm[i] = r.nm[i];
      rec->active = !stopped;
      rec->ok = ok;
      rec->map_sel = use.sel;
      rec->map_epoch = int32_t(use.tag);
      rec->iterations = iters + 1;
    }
  }
  unsigned long long* __restrict__ map = use.sel ? map1 : map0;
  if (idx < hi)
    corresp_pixel_store(L, R, r.m, r.nm, idx, vs, ps, ns, map, use.tag, dthr2,
                        cstar);
  if (TWO_PX && idx1 < hi)
    corresp_pixel_store(L, R, r.m, r.nm, idx1, vs1, ps1, ns1, map, use.tag,
                        dthr2, cstar);
  for (int64_t i = TWO_PX ? idx1 + kSolveBlock : idx1; i < hi; i += kSolveBlock)
    corresp_pixel(L, R, r.m, r.nm, i, map, use.tag, dthr2, cstar);
}

// Per-pixel Jacobians of one iteration; every block leaves one partial
// (block_partial) for the launch that sums them, so no block waits on
// another. With ONE_PX each thread owns at most one pixel (levels that fit
// the grid cap); larger levels give block `blk` the blk-th run of `chunk`
// pixels (chunk computed on the host). With `rec` (iterations after the
// first of a loop) the pose and map come from the record of this
// iteration's correspondence launch, and block 0 commits that record to the
// state. The source loads issue
// before any read of the state or the record: those were written by the
// previous launch, miss this XCD's L2, and would otherwise hold the first
// memory round trip back by one of their own.
template <typename S, bool GEOM, bool FEAT, bool ONE_PX, int NB, int BLOCK,
          bool IL = false>
__global__ void __launch_bounds__(BLOCK)
    jacobian_fused_kernel(stb_icp_level Ld, stb_icp_state* __restrict__ st,
                          const SolveRecord* __restrict__ rec,
                          unsigned long long* __restrict__ map0,
                          unsigned long long* __restrict__ map1,
                          int64_t map_len, float* __restrict__ partials,
                          TermWeights tw, int fixed_map, int64_t chunk) {
  __shared__ double sh[(BLOCK / 64) * kSlots];
  // Block `blk` takes the blk-th run of pixels and leaves the blk-th partial
  // (XCD-contiguous runs; the partials keep the image order).
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const LevelView<S> L = view_of<S>(Ld);
  const LevelRes R = res_of(L);
  // Each block takes a contiguous run of pixels, so the pixels beyond one
  // per thread spread over every block instead of doubling the work of the
  // first ones (with one 1024-thread block per CU, a grid-stride loop left
  // 44 CUs with twice the work at 640x480). A thread's first two pixels load
  // their source data together, so its second pixel does not wait a round
  // trip behind its first (past the end of the arrays the buffer loads
  // return 0); a thread without a second pixel re-reads its first instead of
  // touching the next block's run.
  const int64_t lo = ONE_PX ? int64_t(blk) * BLOCK : int64_t(blk) * chunk;
  const int64_t hi = ONE_PX ? L.n : (lo + chunk < L.n ? lo + chunk : L.n);
  const int64_t i0 = lo + threadIdx.x, i1 = i0 + BLOCK;
  const int64_t j1 = i1 < hi ? i1 : i0;
  const uint8_t vs0 = res_u8(R.sm, int(i0));
  const V3<float> ps0 = res_ld3<S>(R.sp, int(i0));
  uint8_t vs1 = 0;
  V3<float> ps1 = {0.0f, 0.0f, 0.0f};
  if (!ONE_PX) {
    vs1 = res_u8(R.sm, int(j1));
    ps1 = res_ld3<S>(R.sp, int(j1));
  }
  __builtin_amdgcn_sched_barrier(0);
  // No early exit on state.stopped (see corresp_fused_kernel): the solve
  // ignores the partials of a stopped loop. The record's or the state's
  // words through selected pointers: independent loads in one round trip.
  const int32_t* __restrict__ selp = rec ? &rec->map_sel : &st->map_sel;
  const int32_t* __restrict__ tagp = rec ? &rec->map_epoch : &st->map_epoch;
  const float* __restrict__ rt = rec ? rec->rt : st->rt;
  const int sel = *selp;
  const uint32_t tag = fixed_map ? kFreshTag : uint32_t(*tagp);
  const unsigned long long* __restrict__ map = (sel ^ fixed_map) ? map1 : map0;
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = rt[i];
  float acc[kSlots];
  jacobian_pixel_at<S, GEOM, FEAT, true, NB, IL>(L, R, m, map, tag, i0,
                                             (i0 < hi) & (vs0 != 0), ps0, tw,
                                             acc);
  if (!ONE_PX) {
    if (i1 < hi)
      jacobian_pixel_at<S, GEOM, FEAT, false, NB, IL>(L, R, m, map, tag, i1,
                                                  vs1 != 0, ps1, tw, acc);
    for (int64_t idx = i1 + BLOCK; idx < hi; idx += BLOCK)
      jacobian_pixel<S, GEOM, FEAT, false, NB, IL>(L, R, m, map, tag, idx, true,
                                               tw, acc);
  }
  if (!fixed_map && tag == 0) {
    // The last tag of this map: sweep the other one, which the next
    // iteration uses with a fresh tag (one-off passes may have left it set).
    unsigned long long* __restrict__ next = sel ? map0 : map1;
    const int64_t mchunk = (map_len + gridDim.x - 1) / gridDim.x;
    const int64_t mlo = int64_t(blk) * mchunk;
    const int64_t mhi = mlo + mchunk < map_len ? mlo + mchunk : map_len;
    for (int64_t i = mlo + threadIdx.x; i < mhi; i += BLOCK) next[i] = kKeySentinel;
  }
  block_partial<BLOCK>(acc, partials + int64_t(blk) * kSlots, sh, tw.b);
  // Last, so that its round trip does not delay block 0's pixels.
  if (rec && blockIdx.x == 0) commit_record(st, rec, tw.gw, tw.fw);
}

// jacobian_fused_kernel as a software pipeline over up to three pixels per
// thread (FP32 storage, planar features whose channels other than the last
// fit one batch of NB, or none). All three pixels' source data load first;
// the second pixel's gathers issue before the first pixel's terms, and in
// wavefronts with a third pixel its gathers issue before the second's
// terms, so that a pixel's gathers are in flight while the VALU works on
// the one before. Blocks of 512 threads leave two wavefronts per SIMD, each
// with room in its registers for two pixels' gathered data. The block
// partials keep their meaning (each block sums its run of pixels); only the
// order of the FP32 lane sums differs from jacobian_fused_kernel's.
// Requires chunk <= 3 * BLOCK (pipe_fits, on the host, takes runs of at
// most 2.5 * BLOCK).
template <bool GEOM, bool FEAT, int NB, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    jacobian_pipe_kernel(stb_icp_level Ld, stb_icp_state* __restrict__ st,
                         const SolveRecord* __restrict__ rec,
                         unsigned long long* __restrict__ map0,
                         unsigned long long* __restrict__ map1,
                         int64_t map_len, float* __restrict__ partials,
                         TermWeights tw, int fixed_map, int64_t chunk) {
  __shared__ double sh[(BLOCK / 64) * kSlots];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const LevelView<float> L = view_of<float>(Ld);
  const LevelRes R = res_of(L);
  const int64_t lo = int64_t(blk) * chunk;
  const int64_t hi = lo + chunk < L.n ? lo + chunk : L.n;
  const int64_t i0 = lo + threadIdx.x, i1 = i0 + BLOCK, i2 = i1 + BLOCK;
  // A thread without a second or third pixel re-reads its first (a cache
  // hit) and evaluates it as invalid.
  const int64_t j1 = i1 < hi ? i1 : i0, j2 = i2 < hi ? i2 : i0;
  const uint8_t vs0 = res_u8(R.sm, int(i0));
  const V3<float> ps0 = res_ld3<float>(R.sp, int(i0));
  const uint8_t vs1 = res_u8(R.sm, int(j1));
  const V3<float> ps1 = res_ld3<float>(R.sp, int(j1));
  const uint8_t vs2 = res_u8(R.sm, int(j2));
  const V3<float> ps2 = res_ld3<float>(R.sp, int(j2));
  __builtin_amdgcn_sched_barrier(0);
  const int32_t* __restrict__ selp = rec ? &rec->map_sel : &st->map_sel;
  const int32_t* __restrict__ tagp = rec ? &rec->map_epoch : &st->map_epoch;
  const float* __restrict__ rt = rec ? rec->rt : st->rt;
  const int sel = *selp;
  const uint32_t tag = fixed_map ? kFreshTag : uint32_t(*tagp);
  const unsigned long long* __restrict__ map = (sel ^ fixed_map) ? map1 : map0;
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = rt[i];
  const __amdgpu_buffer_rsrc_t rmap = make_res(map, L.HW);
  PxGather<NB> X0, X1;
  // Scheduling barriers keep each pixel's loads together and in order: the
  // waits of a pixel's terms then leave the later pixels' loads in flight.
  jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, i0, (i0 < hi) & (vs0 != 0), ps0, X0);
  __builtin_amdgcn_sched_barrier(0);
  jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, j1, (i1 < hi) & (vs1 != 0), ps1, X1);
  __builtin_amdgcn_sched_barrier(0);
  float acc[kSlots];
  jac_consume<GEOM, FEAT, true, NB>(L, X0, uint32_t(i0), tag, tw, acc);
  uint32_t k1 = uint32_t(j1);
  settle(acc, X1, k1);
  __builtin_amdgcn_sched_barrier(0);
  // Whether any lane of this wavefront has a third pixel (uniform). Each
  // branch is straight-line, so its waits count only its own loads.
  const int64_t wave0 =
      int64_t(__builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u);
  if (lo + wave0 + 2 * BLOCK < hi) {
    PxGather<NB> X2;
    jac_issue<GEOM, FEAT, NB>(L, R, m, rmap, j2, (i2 < hi) & (vs2 != 0), ps2,
                              X2);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, false, NB>(L, X1, k1, tag, tw, acc);
    uint32_t k2 = uint32_t(j2);
    settle(acc, X2, k2);
    __builtin_amdgcn_sched_barrier(0);
    jac_consume<GEOM, FEAT, false, NB>(L, X2, k2, tag, tw, acc);
  } else {
    jac_consume<GEOM, FEAT, false, NB>(L, X1, k1, tag, tw, acc);
  }
  if (!fixed_map && tag == 0) {
    unsigned long long* __restrict__ next = sel ? map0 : map1;
    const int64_t mchunk = (map_len + gridDim.x - 1) / gridDim.x;
    const int64_t mlo = int64_t(blk) * mchunk;
    const int64_t mhi = mlo + mchunk < map_len ? mlo + mchunk : map_len;
    for (int64_t i = mlo + threadIdx.x; i < mhi; i += BLOCK) next[i] = kKeySentinel;
  }
  block_partial<BLOCK>(acc, partials + int64_t(blk) * kSlots, sh, tw.b);
  if (rec && blockIdx.x == 0) commit_record(st, rec, tw.gw, tw.fw);
}

// The fixed-order sum of `count` partials by one block -- the order every
// correspondence launch that applies an update uses too, so a loop and the
// split entry points give the same bits -- then the pose update, the stored
// linearisation or the information matrix (`mode`, a FinishMode).
__global__ void __launch_bounds__(kSolveBlock)
    solve_fused_kernel(stb_icp_state* __restrict__ st,
                       const float* __restrict__ partials, int count,
                       double gw, double fw, int mode, int fixed_map) {
  __shared__ double sh[kSolveBlock];
  __shared__ double tot[kSlots];
  if (!fixed_map && st->stopped) return;
  finish_partials<kSolveBlock>(st, partials, count, mode, gw, fw, sh, tot);
}

// G^T G of the matched pixels at the current pose over the spare map, one
// partial per block; solve_fused_kernel (kStoreInformation) sums them.
template <typename S>
__global__ void __launch_bounds__(kBlock)
    information_fused_kernel(stb_icp_level Ld, stb_icp_state* __restrict__ st,
                             unsigned long long* __restrict__ map0,
                             unsigned long long* __restrict__ map1,
                             float* __restrict__ partials) {
  __shared__ double sh[kShared];
  const unsigned long long* __restrict__ map = pick_map(st, map0, map1, 1);
  float m[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) m[i] = st->rt[i];
  const LevelView<S> L = view_of<S>(Ld);
  const LevelRes R = res_of(L);
  const __amdgpu_buffer_rsrc_t rmap = make_res(map, L.HW);
  float acc[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) acc[k] = 0.0f;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  for (int64_t idx = int64_t(blk) * blockDim.x + threadIdx.x; idx < L.n;
       idx += int64_t(gridDim.x) * blockDim.x)
    information_pixel(L, R, m, rmap, kFreshTag, idx, acc);
  block_partial(acc, partials + int64_t(blk) * kSlots, sh);
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
// Workspace layout: [state | map0 | map1 | partials | record]
// ---------------------------------------------------------------------------

// Jacobian launch shape. By default at most kDefaultUpdateBlocks partials,
// so that every block of the next correspondence launch sums them (32 KB)
// cheaply: levels above kWidePixels use runs of pixels per block of
// kSolveBlock threads (the same occupancy as four times as many blocks of
// 256) or, for FP32 levels whose runs fit 2.5 pixels per thread, the
// pipelined launch with blocks of kPipeBlock (at 640x480: 256 blocks of
// 512 threads); smaller levels use blocks of kBlock. An explicit
// grid_update keeps blocks of kBlock with that cap.
constexpr int kDefaultUpdateBlocks = 256;
constexpr int64_t kWidePixels = 2 * 65536;
constexpr int kMaxUpdateBlocks = 4096;
constexpr int kInformationBlocks = 1024;
constexpr int kDefaultCorrespBlocks = 4096;
// Blocks of the correspondence launch that also applies the previous update
// (kSolveBlock threads each; every block sums the partials): one per CU.
constexpr int kCorrespSolveBlocks = 256;

struct Workspace {
  stb_icp_state* st;
  unsigned long long* map[2];
  float* partials;  // FP32 block partials (block_partial)
  SolveRecord* rec;
  int64_t map_len;
};

size_t ws_bytes(int32_t H, int32_t W) {
  const size_t map_len = size_t(H) * size_t(W);
  return align256(sizeof(stb_icp_state)) + 2 * align256(map_len * 8) +
         align256(size_t(kMaxUpdateBlocks) * kSlots * sizeof(float)) +
         align256(sizeof(SolveRecord));
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
  w.partials = reinterpret_cast<float*>(p);
  p += align256(size_t(kMaxUpdateBlocks) * kSlots * sizeof(float));
  w.rec = reinterpret_cast<SolveRecord*>(p);
  return w;
}

// Jacobian grid: one pixel per thread when that fits the cap, else `cap`
// blocks looping over pixels. The kernel is bound by vector-instruction
// issue once all its loads are in flight, and at 640x480 looping blocks beat
// one-pixel blocks (7.3 against 7.9 us for 1024 against 1200 blocks of
// 256, with the one-pixel-per-thread kernel): each block's reduction is
// shared by more pixels.
struct JacShape {
  int block;
  int grid;
  bool one_px;
};

JacShape jac_shape(const stb_icp_params& P, int64_t n) {
  JacShape s;
  int cap = kDefaultUpdateBlocks;
  s.block = kBlock;
  if (P.grid_update > 0)
    cap = P.grid_update > kMaxUpdateBlocks ? kMaxUpdateBlocks : P.grid_update;
  else if (n > kWidePixels)
    s.block = kSolveBlock;
  const int g = grid_for(n, s.block);
  s.one_px = g <= cap && s.block == kBlock;
  s.grid = g > cap ? cap : g;
  return s;
}

int corresp_grid(const stb_icp_params& P, int64_t n) {
  const int cap = P.grid_corresp > 0 ? P.grid_corresp : kDefaultCorrespBlocks;
  return grid_for(n, kBlock, cap);
}

int corresp_solve_grid(const stb_icp_params& P, int64_t n) {
  const int cap = P.grid_corresp > 0 ? P.grid_corresp : kCorrespSolveBlocks;
  return grid_for(n, kSolveBlock, cap);
}

int check_level(const stb_icp_level* L, bool need_feats) {
  STB_REQUIRE(L != nullptr, "level is NULL");
  STB_REQUIRE(L->src_points && L->src_normals && L->src_mask && L->tgt_points &&
                  L->tgt_normals && L->tgt_mask,
              "level has a NULL source/target array");
  STB_REQUIRE(L->n_src > 0 && L->n_src <= kMaxKeyIndex,
              "n_src must be in [1, %lld], got %lld", (long long)kMaxKeyIndex,
              (long long)L->n_src);
  STB_REQUIRE(L->height > 0 && L->width > 0, "bad target size %dx%d", L->width,
              L->height);
  STB_REQUIRE(L->storage == STB_F32 || L->storage == STB_BF16 ||
                  L->storage == STB_F16,
              "storage must be STB_F32, STB_BF16 or STB_F16");
  STB_REQUIRE(level_addressable(*L),
              "every level array (points, normals, masks, features, map) must "
              "hold fewer than 2^31 bytes");
  if (need_feats) {
    STB_REQUIRE(L->src_feats && L->tgt_feats && L->channels > 0,
                "feature term needs src_feats, tgt_feats and channels > 0");
  }
  STB_REQUIRE(L->feat_layout == STB_FEAT_PLANAR ||
                  L->feat_layout == STB_FEAT_INTERLEAVED,
              "feat_layout must be STB_FEAT_PLANAR or STB_FEAT_INTERLEAVED");
  if (L->feat_layout == STB_FEAT_INTERLEAVED && need_feats) {
    STB_REQUIRE(L->storage != STB_F32 && L->channels <= 32 &&
                    L->channels % 2 == 0,
                "interleaved features need 16-bit storage and an even channel "
                "count of at most 32, got storage %d with %d channels",
                L->storage, L->channels);
  }
  return STB_OK;
}

// Phases of one GN iteration. kCorrespond and kUpdate are the two kernels of
// the loop; kLinearize and kInformation are one-off passes on the spare map.
enum Phase { kCorrespond, kUpdate, kLinearize, kInformation, kJacobians };

template <typename S, bool GEOM, bool FEAT, int NB, bool IL = false>
void launch_jacobian_nb(const stb_icp_level& L, const JacShape& js,
                        const Workspace& w, const SolveRecord* rec,
                        const TermWeights& tw, int fixed_map,
                        hipStream_t stream) {
  const int64_t chunk = (L.n_src + js.grid - 1) / js.grid;
  if (js.block == kSolveBlock)
    hipLaunchKernelGGL(
        (jacobian_fused_kernel<S, GEOM, FEAT, false, NB, kSolveBlock, IL>),
        dim3(js.grid), dim3(kSolveBlock), 0, stream, L, w.st, rec, w.map[0],
        w.map[1], w.map_len, w.partials, tw, fixed_map, chunk);
  else if (js.one_px)
    hipLaunchKernelGGL((jacobian_fused_kernel<S, GEOM, FEAT, true, NB, kBlock, IL>),
                       dim3(js.grid), dim3(kBlock), 0, stream, L, w.st, rec,
                       w.map[0], w.map[1], w.map_len, w.partials, tw,
                       fixed_map, chunk);
  else
    hipLaunchKernelGGL((jacobian_fused_kernel<S, GEOM, FEAT, false, NB, kBlock, IL>),
                       dim3(js.grid), dim3(kBlock), 0, stream, L, w.st, rec,
                       w.map[0], w.map[1], w.map_len, w.partials, tw,
                       fixed_map, chunk);
}

// The pipelined Jacobian kernel's block, and whether it takes a launch:
// FP32 storage, the wide-level shape (kSolveBlock blocks looping over runs of
// pixels) with runs of at most 2.5 pixels per thread, and planar features
// whose channels other than the last fit one batch of four. Longer runs give
// nearly every wavefront a third pixel, whose terms have no loads left to
// overlap:
// at 720x540 (runs of 1519) the pipelined launch took 7.02 us against 6.61
// for jacobian_fused_kernel, at 640x480 (runs of 1200) 5.93 against 6.06.
constexpr int kPipeBlock = 512;

bool pipe_fits(const stb_icp_level& L, const JacShape& js, bool feat) {
  if (L.storage != STB_F32 || js.block != kSolveBlock || js.one_px) return false;
  if (feat && (L.feat_layout != STB_FEAT_PLANAR || L.channels - 1 > 4))
    return false;
  const int64_t chunk = (L.n_src + js.grid - 1) / js.grid;
  return chunk <= 5 * kPipeBlock / 2;
}

template <bool GEOM, bool FEAT, int NB>
void launch_jacobian_pipe(const stb_icp_level& L, const JacShape& js,
                          const Workspace& w, const SolveRecord* rec,
                          const TermWeights& tw, int fixed_map,
                          hipStream_t stream) {
  const int64_t chunk = (L.n_src + js.grid - 1) / js.grid;
  hipLaunchKernelGGL((jacobian_pipe_kernel<GEOM, FEAT, NB, kPipeBlock>),
                     dim3(js.grid), dim3(kPipeBlock), 0, stream, L, w.st, rec,
                     w.map[0], w.map[1], w.map_len, w.partials, tw, fixed_map,
                     chunk);
}

// The Jacobian launch, then (do_solve >= 0) the launch that sums its
// partials and applies them: a pose update (1) or the stored linearisation
// (0). do_solve < 0: the Jacobian launch alone.
template <typename S, bool GEOM, bool FEAT>
void launch_jacobian(const stb_icp_level& L, const stb_icp_params& P,
                     const Workspace& w, const SolveRecord* rec,
                     const TermWeights& tw, int do_solve, int fixed_map,
                     hipStream_t stream) {
  const JacShape js = jac_shape(P, L.n_src);
  // Feature channels other than the last load in batches of four, or of two
  // when there are fewer than four of them (three-channel colour: no
  // padding slots); interleaved features in chunks of eight. The pipelined
  // launch loads them in one batch: of two when there are at most two, else
  // of four.
  if (std::is_same<S, float>::value && pipe_fits(L, js, FEAT)) {
    if (!FEAT || L.channels - 1 <= 2)
      launch_jacobian_pipe<GEOM, FEAT, 2>(L, js, w, rec, tw, fixed_map, stream);
    else
      launch_jacobian_pipe<GEOM, FEAT, 4>(L, js, w, rec, tw, fixed_map, stream);
  } else if (FEAT && L.feat_layout == STB_FEAT_INTERLEAVED)
    launch_jacobian_nb<S, GEOM, FEAT, 4, true>(L, js, w, rec, tw, fixed_map,
                                               stream);
  else if (FEAT && L.channels - 1 < 4)
    launch_jacobian_nb<S, GEOM, FEAT, 2>(L, js, w, rec, tw, fixed_map, stream);
  else
    launch_jacobian_nb<S, GEOM, FEAT, 4>(L, js, w, rec, tw, fixed_map, stream);
  if (do_solve < 0) return;
  hipLaunchKernelGGL(solve_fused_kernel, dim3(1), dim3(kSolveBlock), 0, stream,
                     w.st, w.partials, js.grid, tw.gw, tw.fw, do_solve,
                     fixed_map);
}

template <typename S>
void launch_jacobian_terms(const stb_icp_level& L, const stb_icp_params& P,
                           const Workspace& w, const SolveRecord* rec,
                           int do_solve, int fixed_map, hipStream_t stream) {
  const TermWeights tw = make_weights(P);
  if (P.use_geom && P.use_feat)
    launch_jacobian<S, true, true>(L, P, w, rec, tw, do_solve, fixed_map, stream);
  else if (P.use_geom)
    launch_jacobian<S, true, false>(L, P, w, rec, tw, do_solve, fixed_map, stream);
  else
    launch_jacobian<S, false, true>(L, P, w, rec, tw, do_solve, fixed_map, stream);
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
    const int grid = grid_for(L.n_src, kBlock, kInformationBlocks);
    hipLaunchKernelGGL(information_fused_kernel<S>, dim3(grid), dim3(kBlock),
                       0, stream, L, w.st, w.map[0], w.map[1], w.partials);
    hipLaunchKernelGGL(solve_fused_kernel, dim3(1), dim3(kSolveBlock), 0,
                       stream, w.st, w.partials, grid, 0.0, 0.0,
                       int(kStoreInformation), 1);
    return;
  }
  const int solve = phase == kUpdate ? 1 : (phase == kJacobians ? -1 : 0);
  launch_jacobian_terms<S>(L, P, w, nullptr, solve, fixed, stream);
}

// `count` iterations in 2 * count + 1 launches: the first iteration's
// correspondence and Jacobian launches read the state; each later
// correspondence launch applies the previous iteration's update (SolveRecord)
// and its Jacobian launch commits it; one solve launch applies the last.
template <typename S>
void launch_iterations(const stb_icp_level& L, const stb_icp_params& P,
                       const Workspace& w, int count, hipStream_t stream) {
  const float dthr2 = float(P.distance_thresh * P.distance_thresh);
  const float cstar = angle_gate_cos_f32(P.normal_angle_thresh);
  const int partials = jac_shape(P, L.n_src).grid;
  const int cgrid = corresp_solve_grid(P, L.n_src);
  const bool two_px = (L.n_src + cgrid - 1) / cgrid > kSolveBlock;
  for (int it = 0; it < count; ++it) {
    if (it == 0)
      hipLaunchKernelGGL(corresp_fused_kernel<S>,
                         dim3(corresp_grid(P, L.n_src)), dim3(kBlock), 0,
                         stream, L, w.st, w.map[0], w.map[1], dthr2, cstar, 0);
    else if (two_px)
      hipLaunchKernelGGL((corresp_solve_kernel<S, true>), dim3(cgrid),
                         dim3(kSolveBlock), 0, stream, L, w.st, w.rec,
                         w.partials, partials, w.map[0], w.map[1], dthr2,
                         cstar);
    else
      hipLaunchKernelGGL((corresp_solve_kernel<S, false>), dim3(cgrid),
                         dim3(kSolveBlock), 0, stream, L, w.st, w.rec,
                         w.partials, partials, w.map[0], w.map[1], dthr2,
                         cstar);
    launch_jacobian_terms<S>(L, P, w, it ? w.rec : nullptr,
                             it + 1 == count ? 1 : -1, 0, stream);
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

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point
// ---------------------------------------------------------------------------
//
// Every GPU the library runs on holds one copy of this array until the
// process ends. In host code `scratch` is only a stand-in for the symbol:
// its device address is obtained once per device with hipGetSymbolAddress
// (with that device current) and only that address is ever passed to
// kernels.
constexpr size_t kStaticScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of the current device's copy of `scratch`.
int static_scratch(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return STB_ERR_LAUNCH;
  }
  STB_REQUIRE(dev >= 0 && dev < kMaxScratchDevices,
              "device %d is beyond the %d devices the static scratch serves",
              dev, kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return STB_ERR_LAUNCH;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return STB_OK;
}

__global__ void copy_state_kernel(const uint64_t* __restrict__ src,
                                  uint64_t* __restrict__ dst, int words) {
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

}  // namespace

extern "C" size_t stb_icp_noalloc_scratch_bytes(void) {
  return kStaticScratchBytes;
}

extern "C" int stb_icp_estimate_noalloc(const stb_icp_level* level,
                                        const stb_icp_params* params,
                                        const double* transform0,
                                        int32_t num_iters,
                                        int32_t with_information,
                                        stb_icp_state* out_state,
                                        void* stream) {
  const char* what = "stb_icp_estimate_noalloc";
  STB_REQUIRE(level != nullptr && params != nullptr && out_state != nullptr,
              "%s: NULL argument", what);
  STB_REQUIRE(num_iters >= 0, "%s: num_iters must be >= 0", what);
  STB_REQUIRE(level->height > 0 && level->width > 0, "%s: bad target size %dx%d",
              what, level->width, level->height);
  const size_t need = ws_bytes(level->height, level->width);
  STB_REQUIRE(need <= kStaticScratchBytes,
              "%s: a %dx%d target needs %zu bytes of scratch, more than the "
              "%zu bytes this entry point holds (targets up to about 1280x720 "
              "pixels); use stb_icp_begin / stb_icp_iterate with a workspace",
              what, level->width, level->height, need, kStaticScratchBytes);
  void* ws = nullptr;
  int rc = static_scratch(&ws);
  if (rc) return rc;
  rc = stb_icp_begin(ws, kStaticScratchBytes, transform0, 0, level->n_src,
                     level->height, level->width, stream);
  if (rc) return rc;
  if (num_iters > 0) {
    rc = stb_icp_iterate(level, params, ws, kStaticScratchBytes, num_iters,
                         stream);
    if (rc) return rc;
  }
  if (with_information && params->use_geom) {
    rc = stb_icp_information(level, params, ws, kStaticScratchBytes, stream);
    if (rc) return rc;
  }
  static_assert(sizeof(stb_icp_state) % 8 == 0, "state is copied in 8-byte words");
  hipLaunchKernelGGL(copy_state_kernel, dim3(1), dim3(256), 0, as_stream(stream),
                     static_cast<const uint64_t*>(ws),
                     reinterpret_cast<uint64_t*>(out_state),
                     int(sizeof(stb_icp_state) / 8));
  return check_launch(what);
}

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
