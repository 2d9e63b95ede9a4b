// Batch plan shared by the batched Gauss-Newton kernels (batch.hip) and the
// batched adjoint (adjoint.hip): per-pair level descriptors, map offsets and
// the flattened tile list, plus the device scratch layout.
#pragma once

#include "gn_device.hpp"

namespace stb {

struct PlanHeader {
  int32_t num_pairs;
  int32_t total_tiles;
  int32_t tile_px;
  int32_t storage;
  int64_t total_map;
  int64_t reserved;
};

// Plan image: header | levels[P] | map_off[P] (int64) | tile_start[P + 1]
// (int32) | tile_pair[total_tiles] (int32), each part 64-byte aligned.
struct PlanLayout {
  size_t levels, map_off, tile_start, tile_pair, bytes;
};

__host__ __device__ inline size_t align64(size_t x) {
  return (x + 63) & ~size_t(63);
}

__host__ __device__ inline PlanLayout plan_layout(int P, int tiles) {
  PlanLayout l;
  l.levels = align64(sizeof(PlanHeader));
  l.map_off = l.levels + align64(sizeof(stb_icp_level) * size_t(P));
  l.tile_start = l.map_off + align64(sizeof(int64_t) * size_t(P));
  l.tile_pair = l.tile_start + align64(sizeof(int32_t) * size_t(P + 1));
  l.bytes = l.tile_pair + align64(sizeof(int32_t) * size_t(tiles));
  return l;
}

struct PlanView {
  int P, tiles, tile_px;
  const stb_icp_level* levels;
  const int64_t* map_off;
  const int32_t* tile_start;
  const int32_t* tile_pair;
};

__device__ inline PlanView plan_view(const void* plan) {
  const char* b = static_cast<const char*>(plan);
  const PlanHeader* h = reinterpret_cast<const PlanHeader*>(b);
  const PlanLayout l = plan_layout(h->num_pairs, h->total_tiles);
  PlanView v;
  v.P = h->num_pairs;
  v.tiles = h->total_tiles;
  v.tile_px = h->tile_px;
  v.levels = reinterpret_cast<const stb_icp_level*>(b + l.levels);
  v.map_off = reinterpret_cast<const int64_t*>(b + l.map_off);
  v.tile_start = reinterpret_cast<const int32_t*>(b + l.tile_start);
  v.tile_pair = reinterpret_cast<const int32_t*>(b + l.tile_pair);
  return v;
}

// Scratch: map0[total_map] | map1[total_map] | partials[tiles][32]. The
// forward launches keep FP32 tile partials (the per-pair solve reads all of
// a pair's tiles every iteration), the adjoint's re-summed systems FP64: the
// same region seen as either.
struct Scratch {
  unsigned long long* map0;
  unsigned long long* map1;
  double* partials;
  float* partials32;
  size_t bytes;
};

inline Scratch carve_scratch(void* base, int P, int tiles, int64_t total_map) {
  (void)P;
  Scratch s;
  char* p = static_cast<char*>(base);
  s.map0 = reinterpret_cast<unsigned long long*>(p);
  p += align256(size_t(total_map) * 8);
  s.map1 = reinterpret_cast<unsigned long long*>(p);
  p += align256(size_t(total_map) * 8);
  s.partials = reinterpret_cast<double*>(p);
  s.partials32 = reinterpret_cast<float*>(p);
  p += align256(size_t(tiles) * kSlots * sizeof(double));
  s.bytes = size_t(p - static_cast<char*>(base));
  return s;
}

__device__ inline unsigned long long* pair_map(const stb_icp_state* st,
                                               unsigned long long* maps0,
                                               unsigned long long* maps1,
                                               int64_t off, int spare) {
  return ((st->map_sel ^ spare) ? maps1 : maps0) + off;
}

}  // namespace stb
