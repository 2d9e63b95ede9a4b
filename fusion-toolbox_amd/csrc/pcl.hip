// Point-cloud matching: exact k-nearest-neighbour search, the reference's
// nearest-neighbour feature interpolation with its finite-difference
// backward (src/nearest_neighbors.cu:13-295, the ops behind KDTreeLayer,
// slamtb/spatial/kdtree_layer.py:12-107), and the voxel merge of surfel
// clouds (SurfelVolume, src/surfel_volume.cu:12-123,
// include/slamtb/hash_volume.hpp:12-83).
//
// The neighbour search replaces the reference's host k-d tree (pykdtree,
// approximate with eps = 0.5) with an exact brute-force search on the device:
// each block stages tiles of target points in LDS and every thread keeps the
// k nearest of its query point in registers, sorted by squared distance and,
// among equal distances, by target index. Neighbours at or beyond the
// distance bound are not reported; missing slots hold (inf, m) like the
// k-d tree's. With a finite bound the same search runs over a uniform grid
// of cells at least as wide as the bound (27 cells around each query), with
// the same result.
#include "stb_common.hpp"

#include <atomic>

namespace stb {
namespace {

constexpr int kTile = 256;

template <typename T>
__device__ inline T sq_dist(const V3<T>& a, T bx, T by, T bz) {
#pragma clang fp contract(off)
  const T dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
  return dx * dx + dy * dy + dz * dz;
}

// Top-K of one query, K a power of two >= k; only the first k are written.
template <typename T, int K>
__global__ void __launch_bounds__(kBlock)
    knn_kernel(const T* __restrict__ tgt, int64_t m, const T* __restrict__ qry,
               int64_t n, int k, T bound2, T* __restrict__ out_d,
               int64_t* __restrict__ out_i) {
#pragma clang fp contract(off)
  __shared__ T tile[3 * kTile];
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = q < n;
  const V3<T> p = active ? ld3(qry, q) : V3<T>{T(0), T(0), T(0)};
  const T inf = T(INFINITY);
  T d[K];
  int64_t id[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    d[s] = inf;
    id[s] = m;
  }
  for (int64_t base = 0; base < m; base += kTile) {
    const int count = int(m - base < kTile ? m - base : kTile);
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * count; t += blockDim.x)
      tile[t] = tgt[3 * base + t];
    __syncthreads();
    if (!active) continue;
    for (int t = 0; t < count; ++t) {
      const T d2 = sq_dist(p, tile[3 * t], tile[3 * t + 1], tile[3 * t + 2]);
      // Targets arrive in increasing index order, so a strict comparison
      // keeps the lower index first among equal distances.
      if (!(d2 < bound2) || !(d2 < d[K - 1])) continue;
      const int64_t j = base + t;
#pragma unroll
      for (int s = K - 1; s > 0; --s) {
        const bool shift = d[s - 1] > d2;
        const bool here = !shift && d[s] > d2;
        d[s] = shift ? d[s - 1] : (here ? d2 : d[s]);
        id[s] = shift ? id[s - 1] : (here ? j : id[s]);
      }
      if (d[0] > d2) {
        d[0] = d2;
        id[0] = j;
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s < k) {
      out_d[q * k + s] = d[s];
      out_i[q * k + s] = id[s];
    }
  }
}

// Cell of a point in a uniform grid (origin lo, edge `cell`, nx*ny*nz
// cells): its integer coordinates, which may lie outside the grid for a
// query point; non-finite points get INT_MIN.
template <typename T>
__device__ inline void grid_coord(const V3<T>& p, T lox, T loy, T loz, T inv,
                                  int& cx, int& cy, int& cz) {
  const T fx = floor((p.x - lox) * inv), fy = floor((p.y - loy) * inv),
          fz = floor((p.z - loz) * inv);
  const bool ok = fx > T(-2e9) && fx < T(2e9) && fy > T(-2e9) && fy < T(2e9) &&
                  fz > T(-2e9) && fz < T(2e9);
  cx = ok ? int(fx) : INT_MIN;
  cy = ok ? int(fy) : INT_MIN;
  cz = ok ? int(fz) : INT_MIN;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    grid_cells_kernel(const T* __restrict__ pts, int64_t m, T lox, T loy, T loz,
                      T inv, int nx, int ny, int nz, int32_t* __restrict__ cells) {
  const int32_t invalid = nx * ny * nz;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += int64_t(gridDim.x) * blockDim.x) {
    int cx, cy, cz;
    grid_coord(ld3(pts, i), lox, loy, loz, inv, cx, cy, cz);
    const bool in = cx >= 0 && cx < nx && cy >= 0 && cy < ny && cz >= 0 && cz < nz;
    cells[i] = in ? (cz * ny + cy) * nx + cx : invalid;
  }
}

// One grid entry: the target point and its original index, 4 values of T
// (the index's bits in the fourth float for T = float, its value for double),
// so a candidate is one 16- or 32-byte load.
template <typename T>
struct GridEntry;

template <>
struct GridEntry<float> {
  __device__ static inline void load(const float* e, int64_t a, V3<float>& p,
                                     int32_t& j) {
    const float4 v = reinterpret_cast<const float4*>(e)[a];
    p = V3<float>{v.x, v.y, v.z};
    j = __float_as_int(v.w);
  }
};

template <>
struct GridEntry<double> {
  __device__ static inline void load(const double* e, int64_t a, V3<double>& p,
                                     int32_t& j) {
    const double2 u = reinterpret_cast<const double2*>(e)[2 * a];
    const double2 v = reinterpret_cast<const double2*>(e)[2 * a + 1];
    p = V3<double>{u.x, u.y, v.x};
    j = int32_t(v.y);
  }
};

// Inserts (d2, j) into the (distance, index)-ordered top-K list (d, id) if
// it is within the bound and precedes the list's last entry.
template <typename T, int K>
__device__ inline void topk_insert(T (&d)[K], int32_t (&id)[K], T d2, int32_t j,
                                   T bound2) {
  if (!(d2 < bound2)) return;
  if (!(d2 < d[K - 1] || (d2 == d[K - 1] && j < id[K - 1]))) return;
#pragma unroll
  for (int s = K - 1; s > 0; --s) {
    const bool shift = d[s - 1] > d2 || (d[s - 1] == d2 && id[s - 1] > j);
    const bool here = !shift && (d[s] > d2 || (d[s] == d2 && id[s] > j));
    d[s] = shift ? d[s - 1] : (here ? d2 : d[s]);
    id[s] = shift ? id[s - 1] : (here ? j : id[s]);
  }
  if (d[0] > d2 || (d[0] == d2 && id[0] > j)) {
    d[0] = d2;
    id[0] = j;
  }
}

// The same top-K as knn_kernel, over the 27 grid cells around the query
// (cells are at least the distance bound wide, so they hold every target
// within it). Candidates arrive out of index order, so ties compare
// (distance, index) and the result equals the brute-force search's. Lane
// group t takes query order[t] (order NULL: query t): with the queries
// ordered by grid cell, the lanes of a wavefront scan the same cells, so
// their candidate loads coalesce into the same lines and their loops run the
// same trip counts.
template <typename T, int K, int L>
__global__ void __launch_bounds__(kBlock)
    knn_grid_kernel(const T* __restrict__ entries, const int64_t* __restrict__ start,
                    int nx, int ny, int nz, T lox, T loy, T loz, T inv, int64_t m,
                    const T* __restrict__ qry, int64_t n,
                    const int64_t* __restrict__ order, int k, T bound2,
                    T* __restrict__ out_d, int64_t* __restrict__ out_i) {
#pragma clang fp contract(off)
  // L consecutive lanes share a query: lane `sub` scans candidates sub,
  // sub + L, ... of each run and keeps its own top K; the lanes then merge
  // their lists. The top K of the union of disjoint candidate sets under the
  // total order (distance, index) is the same whichever lane saw what.
  const int64_t gt = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t t = gt / L;
  const int sub = int(gt % L);
  if (t >= n) return;  // whole groups of L lanes: n * L threads are needed
  const int64_t q = order ? order[t] : t;
  const V3<T> p = ld3(qry, q);
  const T inf = T(INFINITY);
  // Indices in 32 bits (m < 2^31): half the registers and compares of the
  // insertion network.
  T d[K];
  int32_t id[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    d[s] = inf;
    id[s] = int32_t(m);
  }
  int cx, cy, cz;
  grid_coord(p, lox, loy, loz, inv, cx, cy, cz);
  if (cx != INT_MIN) {
    // A neighbour cell whose nearest face lies at least the bound away holds
    // no candidate (the test d2 < bound2 is strict): in units of cells, the
    // query sits at r in [0, 1) of its cell, a face of the cell at offset -1
    // is r away, one at +1 is 1 - r. Each face moves towards the query by
    // `slack` cells first, more than the rounding of any point's cell
    // coordinate, so a skipped cell never holds a point the scan would take.
    const T rx = (p.x - lox) * inv - T(cx), ry = (p.y - loy) * inv - T(cy),
            rz = (p.z - loz) * inv - T(cz);
    const T eps = sizeof(T) == 4 ? T(1.1920929e-7) : T(2.220446049250313e-16);
    const T slack = T(1e-3) + T(8) * eps * T(max(nx, max(ny, nz)));
    const T b2 = bound2 * inv * inv;
    auto gap2 = [slack](T r, int off) {
      const T g = (off < 0 ? r : (off > 0 ? T(1) - r : T(0))) - slack;
      return g > T(0) ? g * g : T(0);
    };
    const T gl = gap2(rx, -1), gr = gap2(rx, 1);
    for (int z = max(cz - 1, 0); z <= min(cz + 1, nz - 1); ++z)
      for (int y = max(cy - 1, 0); y <= min(cy + 1, ny - 1); ++y) {
        const T gyz = gap2(rz, z - cz) + gap2(ry, y - cy);
        if (!(gyz < b2)) continue;
        // The x-neighbours of a row are adjacent cells: one run of entries.
        const int x0 = max(cx - (gyz + gl < b2 ? 1 : 0), 0);
        const int x1 = min(cx + (gyz + gr < b2 ? 1 : 0), nx - 1);
        if (x0 > x1) continue;
        const int64_t row = (int64_t(z) * ny + y) * nx;
        const int64_t lo = start[row + x0], hi = start[row + x1 + 1];
        // Candidates in batches of four whose loads issue together; a slot
        // past the run re-reads the run's last entry with an infinite
        // distance, which never inserts.
        for (int64_t a = lo + sub; a < hi; a += 4 * L) {
          T d2s[4];
          int32_t js[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool in = a + u * L < hi;
            V3<T> t;
            GridEntry<T>::load(entries, in ? a + u * L : hi - 1, t, js[u]);
            const T d2 = sq_dist(p, t.x, t.y, t.z);
            d2s[u] = in ? d2 : inf;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) topk_insert<T, K>(d, id, d2s[u], js[u], bound2);
        }
      }
  }
  // Butterfly merge: each step snapshots the partner's list, then inserts
  // it (the lists cover disjoint candidates).
#pragma unroll
  for (int off = 1; off < L; off *= 2) {
    T pd[K];
    int32_t pj[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
      pd[s] = __shfl_xor(d[s], off, 64);
      pj[s] = __shfl_xor(id[s], off, 64);
    }
#pragma unroll
    for (int s = 0; s < K; ++s) topk_insert<T, K>(d, id, pd[s], pj[s], bound2);
  }
  if (sub != 0) return;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (s < k) {
      out_d[q * k + s] = d[s];
      out_i[q * k + s] = int64_t(id[s]);
    }
  }
}

// ForwardKernel (src/nearest_neighbors.cu:13-54): features averaged over the
// neighbours with weights 1 / (1 + d), normalised by their sum over the
// finite distances; the per-channel loop stops at the first missing index.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    nn_forward_kernel(const T* __restrict__ dist, const int64_t* __restrict__ index,
                      int64_t n, int k, const T* __restrict__ feats, int C,
                      int64_t m, T* __restrict__ out) {
#pragma clang fp contract(off)
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n;
       q += int64_t(gridDim.x) * blockDim.x) {
    const T* dq = dist + q * k;
    const int64_t* iq = index + q * k;
    T total = T(0);
    for (int s = 0; s < k; ++s)
      if (dq[s] < T(INFINITY)) total += T(1) / (T(1) + dq[s]);
    for (int c = 0; c < C; ++c) {
      T f = T(0);
      for (int s = 0; s < k; ++s) {
        const int64_t j = iq[s];
        if (j >= m) break;
        const T sim = T(1) / (T(1) + dq[s]);
        const T w = sim / total;
        f += feats[int64_t(c) * m + j] * w;
      }
      out[int64_t(c) * n + q] = f;
    }
  }
}

// EpsilonDistancesKernel (src/nearest_neighbors.cu:86-143): for each of the
// six +-h offsets of the source point along x, y, z, the normalised weights
// 1 / (1 + |target - (source + offset)|^2) of its neighbours. Layout
// [n, 6, k]; slots past the first missing index are left as they were.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    nn_epsilon_kernel(const T* __restrict__ tgt, int64_t m,
                      const T* __restrict__ src, const int64_t* __restrict__ index,
                      int64_t n, int k, T h, T* __restrict__ eps) {
#pragma clang fp contract(off)
  const int64_t total_rc = n * 6;
  for (int64_t rc = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       rc < total_rc; rc += int64_t(gridDim.x) * blockDim.x) {
    const int64_t row = rc / 6;
    const int col = int(rc - row * 6);
    const T ex = col == 0 ? h : (col == 1 ? -h : T(0));
    const T ey = col == 2 ? h : (col == 3 ? -h : T(0));
    const T ez = col == 4 ? h : (col == 5 ? -h : T(0));
    const V3<T> s = ld3(src, row);
    const V3<T> x{ex + s.x, ey + s.y, ez + s.z};
    const int64_t* iq = index + row * k;
    T* e = eps + (row * 6 + col) * k;
    T total = T(0);
    int found = 0;
    for (int j = 0; j < k; ++j) {
      const int64_t t = iq[j];
      if (t > m - 1) break;
      const V3<T> d = sub3(ld3(tgt, t), x);
      const T sim = T(1) / (T(1) + dot3(d, d));
      e[j] = sim;
      total += sim;
      ++found;
    }
    if (found == 0) {
      e[0] = T(1);
      continue;
    }
    for (int j = 0; j < found; ++j) e[j] = e[j] / total;
  }
}

// BackwardKernel + NNGrad (src/nearest_neighbors.cu:166-262): per channel the
// central difference of the weighted feature average along x, y and z,
// contracted with the incoming feature gradient.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    nn_backward_kernel(const T* __restrict__ eps, const int64_t* __restrict__ index,
                       int64_t n, int k, const T* __restrict__ feats, int C,
                       int64_t m, const T* __restrict__ dl, T h,
                       T* __restrict__ dx) {
#pragma clang fp contract(off)
  const T div = T(1) / (T(2) * h);
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n;
       q += int64_t(gridDim.x) * blockDim.x) {
    const int64_t* iq = index + q * k;
    const T* e = eps + q * 6 * k;
    T gx = T(0), gy = T(0), gz = T(0);
    for (int c = 0; c < C; ++c) {
      T avg[6];
      for (int w = 0; w < 6; ++w) {
        T a = T(0);
        for (int s = 0; s < k; ++s) {
          const int64_t j = iq[s];
          if (j < m) a += feats[int64_t(c) * m + j] * e[w * k + s];
        }
        avg[w] = a;
      }
      const T g = dl[int64_t(c) * n + q];
      gx += (avg[0] - avg[1]) * div * g;
      gy += (avg[2] - avg[3]) * div * g;
      gz += (avg[4] - avg[5]) * div * g;
    }
    dx[3 * q] = gx;
    dx[3 * q + 1] = gy;
    dx[3 * q + 2] = gz;
  }
}

// HashVolume::Hash (include/slamtb/hash_volume.hpp:67-77): float offsets
// divided by the voxel size and truncated, combined in int arithmetic with
// the truncated volume dimensions.
__global__ void __launch_bounds__(kBlock)
    voxel_ids_kernel(const float* __restrict__ pts, int64_t n, float min_x,
                     float min_y, float min_z, float vs, int width, int height,
                     int32_t* __restrict__ ids) {
#pragma clang fp contract(off)
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int col = int((pts[3 * i] - min_x) / vs);
    const int row = int((pts[3 * i + 1] - min_y) / vs);
    const int dep = int((pts[3 * i + 2] - min_z) / vs);
    ids[i] = dep * width * height + row * width + col;
  }
}

// AccumKernel + AverageKernel + ToCloudKernel (src/surfel_volume.cu:12-123):
// one thread per voxel sums its members in point order (they are contiguous
// in `order`, a stable sort by voxel id) and scales by 1 / count. Times are
// never accumulated by the reference, so they come out 0; colours are
// truncated to uint8.
__global__ void __launch_bounds__(kBlock)
    voxel_merge_kernel(const float* __restrict__ pts, const float* __restrict__ nrm,
                       const uint8_t* __restrict__ col, const float* __restrict__ conf,
                       const float* __restrict__ rad, const float* __restrict__ feats,
                       int C, int64_t n, const int64_t* __restrict__ order,
                       const int64_t* __restrict__ starts, int64_t V,
                       float* __restrict__ o_pts, float* __restrict__ o_nrm,
                       uint8_t* __restrict__ o_col, float* __restrict__ o_conf,
                       float* __restrict__ o_rad, int32_t* __restrict__ o_time,
                       float* __restrict__ o_feats) {
#pragma clang fp contract(off)
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += int64_t(gridDim.x) * blockDim.x) {
    const int64_t lo = starts[v], hi = starts[v + 1];
    float p[3] = {0, 0, 0}, q[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    float cf = 0, r = 0;
    for (int64_t a = lo; a < hi; ++a) {
      const int64_t i = order[a];
      for (int d = 0; d < 3; ++d) {
        p[d] += pts[3 * i + d];
        q[d] += nrm[3 * i + d];
        c[d] += float(col[3 * i + d]);
      }
      cf += conf[i];
      r += rad[i];
    }
    const float inv = 1.0f / float(hi - lo);
    for (int d = 0; d < 3; ++d) {
      o_pts[3 * v + d] = p[d] * inv;
      o_nrm[3 * v + d] = q[d] * inv;
      o_col[3 * v + d] = uint8_t(c[d] * inv);
    }
    o_conf[v] = cf * inv;
    o_rad[v] = r * inv;
    o_time[v] = 0;
    for (int ch = 0; ch < C; ++ch) {
      float f = 0;
      for (int64_t a = lo; a < hi; ++a) f += feats[int64_t(ch) * n + order[a]];
      o_feats[int64_t(ch) * V + v] = f * inv;
    }
  }
}

template <typename T, int K>
void launch_knn(const void* tgt, int64_t m, const void* qry, int64_t n, int k,
                double bound, void* d, int64_t* i, hipStream_t s) {
  const T b2 = std::isfinite(bound) ? T(bound) * T(bound) : T(INFINITY);
  hipLaunchKernelGGL((knn_kernel<T, K>), dim3(grid_for(n)), dim3(kBlock), 0, s,
                     static_cast<const T*>(tgt), m, static_cast<const T*>(qry), n,
                     k, b2, static_cast<T*>(d), i);
}

struct GridArgs {
  const void* entries;
  const int64_t* start;
  int nx, ny, nz;
  double lo[3], cell;
};

// Lanes per query of the grid search: with fewer queries than it takes to
// keep the CUs' memory pipelines busy one lane per query, each query's
// candidates are split over kKnnLanes lanes.
constexpr int kKnnLanes = 4;
// The kNN launches give each query its own thread (no grid-stride loop):
// grid_for's 2^20 blocks of kBlock threads bound the queries per call.
constexpr int64_t kMaxKnnQueries = int64_t(1) << 28;
// Above this many queries per call each query takes one lane (enough
// threads to fill the CUs without the split). Settable for tests through
// stb_knn_grid_set_lane_threshold, so that small calls can take the
// one-lane instantiation too.
std::atomic<int64_t> g_knn_one_lane_queries{int64_t(1) << 20};

template <typename T, int K, int L>
void launch_knn_grid_lanes(const GridArgs& g, int64_t m, const void* qry,
                           int64_t n, const int64_t* order, int k, double bound,
                           void* d, int64_t* i, hipStream_t s) {
  hipLaunchKernelGGL((knn_grid_kernel<T, K, L>), dim3(grid_for(n * L)),
                     dim3(kBlock), 0, s, static_cast<const T*>(g.entries),
                     g.start, g.nx, g.ny, g.nz, T(g.lo[0]), T(g.lo[1]),
                     T(g.lo[2]), T(1) / T(g.cell), m, static_cast<const T*>(qry),
                     n, order, k, T(bound) * T(bound), static_cast<T*>(d), i);
}

template <typename T, int K>
void launch_knn_grid(const GridArgs& g, int64_t m, const void* qry, int64_t n,
                     const int64_t* order, int k, double bound, void* d,
                     int64_t* i, hipStream_t s) {
  if (n <= g_knn_one_lane_queries.load(std::memory_order_relaxed))
    launch_knn_grid_lanes<T, K, kKnnLanes>(g, m, qry, n, order, k, bound, d, i, s);
  else
    launch_knn_grid_lanes<T, K, 1>(g, m, qry, n, order, k, bound, d, i, s);
}

template <typename T>
void knn_grid_dispatch(const GridArgs& g, int64_t m, const void* qry, int64_t n,
                       const int64_t* order, int k, double bound, void* d,
                       int64_t* i, hipStream_t s) {
  if (k <= 1) launch_knn_grid<T, 1>(g, m, qry, n, order, k, bound, d, i, s);
  else if (k <= 2) launch_knn_grid<T, 2>(g, m, qry, n, order, k, bound, d, i, s);
  else if (k <= 4) launch_knn_grid<T, 4>(g, m, qry, n, order, k, bound, d, i, s);
  else if (k <= 8) launch_knn_grid<T, 8>(g, m, qry, n, order, k, bound, d, i, s);
  else launch_knn_grid<T, 16>(g, m, qry, n, order, k, bound, d, i, s);
}

template <typename T>
void grid_cells_launch(const void* pts, int64_t m, const double* lo, double cell,
                       int nx, int ny, int nz, int32_t* cells, hipStream_t s) {
  hipLaunchKernelGGL(grid_cells_kernel<T>, dim3(grid_for(m)), dim3(kBlock), 0, s,
                     static_cast<const T*>(pts), m, T(lo[0]), T(lo[1]), T(lo[2]),
                     T(1) / T(cell), nx, ny, nz, cells);
}

template <typename T>
void knn_dispatch(const void* tgt, int64_t m, const void* qry, int64_t n, int k,
                  double bound, void* d, int64_t* i, hipStream_t s) {
  if (k <= 1) launch_knn<T, 1>(tgt, m, qry, n, k, bound, d, i, s);
  else if (k <= 2) launch_knn<T, 2>(tgt, m, qry, n, k, bound, d, i, s);
  else if (k <= 4) launch_knn<T, 4>(tgt, m, qry, n, k, bound, d, i, s);
  else if (k <= 8) launch_knn<T, 8>(tgt, m, qry, n, k, bound, d, i, s);
  else launch_knn<T, 16>(tgt, m, qry, n, k, bound, d, i, s);
}

}  // namespace
}  // namespace stb

using namespace stb;

#define STB_FLOAT_DISPATCH(dtype, FN, ...)           \
  do {                                               \
    if ((dtype) == STB_F32) FN<float>(__VA_ARGS__);   \
    else FN<double>(__VA_ARGS__);                    \
  } while (0)

extern "C" int stb_knn(int32_t dtype, const void* target, int64_t m,
                       const void* query, int64_t n, int32_t k,
                       double distance_upper_bound, void* out_sq_dist,
                       int64_t* out_index, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "knn: dtype must be float32 or float64");
  STB_REQUIRE(m >= 0 && n >= 0, "knn: bad sizes");
  STB_REQUIRE(k >= 1 && k <= 16, "knn: k must be in [1, 16], got %d", k);
  STB_REQUIRE(n <= kMaxKnnQueries,
              "knn: at most 2^28 queries per call (one thread each), got %lld",
              (long long)n);
  STB_REQUIRE(distance_upper_bound > 0.0, "knn: distance_upper_bound must be > 0");
  if (n == 0) return STB_OK;
  STB_REQUIRE(query && out_sq_dist && out_index && (m == 0 || target),
              "knn: NULL argument");
  STB_FLOAT_DISPATCH(dtype, knn_dispatch, target, m, query, n, k,
                     distance_upper_bound, out_sq_dist, out_index,
                     as_stream(stream));
  return check_launch("knn");
}

extern "C" int stb_grid_cells(int32_t dtype, const void* points, int64_t m,
                              double lo_x, double lo_y, double lo_z, double cell,
                              int32_t nx, int32_t ny, int32_t nz,
                              int32_t* out_cells, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "grid_cells: dtype must be float32 or float64");
  STB_REQUIRE(m >= 0 && cell > 0.0 && nx > 0 && ny > 0 && nz > 0 &&
                  int64_t(nx) * ny * nz < (int64_t(1) << 31),
              "grid_cells: bad grid");
  if (m == 0) return STB_OK;
  STB_REQUIRE(points && out_cells, "grid_cells: NULL argument");
  const double lo[3] = {lo_x, lo_y, lo_z};
  STB_FLOAT_DISPATCH(dtype, grid_cells_launch, points, m, lo, cell, nx, ny, nz,
                     out_cells, as_stream(stream));
  return check_launch("grid_cells");
}

extern "C" int stb_knn_grid(int32_t dtype, const void* entries,
                            const int64_t* cell_start,
                            int32_t nx, int32_t ny, int32_t nz, double lo_x,
                            double lo_y, double lo_z, double cell, int64_t m,
                            const void* query, int64_t n,
                            const int64_t* query_order, int32_t k,
                            double distance_upper_bound, void* out_sq_dist,
                            int64_t* out_index, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "knn_grid: dtype must be float32 or float64");
  STB_REQUIRE(k >= 1 && k <= 16, "knn_grid: k must be in [1, 16], got %d", k);
  STB_REQUIRE(n <= kMaxKnnQueries,
              "knn_grid: at most 2^28 queries per call (one thread each), got %lld",
              (long long)n);
  STB_REQUIRE(m >= 0 && n >= 0 && nx > 0 && ny > 0 && nz > 0,
              "knn_grid: bad sizes");
  STB_REQUIRE(std::isfinite(distance_upper_bound) && distance_upper_bound > 0.0 &&
                  cell >= distance_upper_bound,
              "knn_grid: the cell (%g) must be at least the finite distance "
              "bound (%g)", cell, distance_upper_bound);
  if (n == 0) return STB_OK;
  STB_REQUIRE(query && out_sq_dist && out_index && cell_start &&
                  (m == 0 || entries),
              "knn_grid: NULL argument");
  STB_REQUIRE(m < (int64_t(1) << 31),
              "knn_grid: at most 2^31 - 1 targets (31-bit indices), got %lld",
              (long long)m);
  GridArgs g{entries, cell_start, nx, ny, nz, {lo_x, lo_y, lo_z}, cell};
  STB_FLOAT_DISPATCH(dtype, knn_grid_dispatch, g, m, query, n, query_order, k,
                     distance_upper_bound, out_sq_dist, out_index,
                     as_stream(stream));
  return check_launch("knn_grid");
}

template <typename T>
static void nn_forward_launch(const void* d, const int64_t* i, int64_t n, int k,
                              const void* f, int C, int64_t m, void* out,
                              hipStream_t s) {
  hipLaunchKernelGGL(nn_forward_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     static_cast<const T*>(d), i, n, k, static_cast<const T*>(f),
                     C, m, static_cast<T*>(out));
}

extern "C" int stb_nn_features_forward(int32_t dtype, const void* nn_distances,
                                       const int64_t* nn_index, int64_t n,
                                       int32_t k, const void* features,
                                       int32_t channels, int64_t m,
                                       void* out_features, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "NearestNeighborsOp.forward: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0 && k >= 1 && channels >= 0 && m >= 0,
              "NearestNeighborsOp.forward: bad sizes");
  if (n == 0 || channels == 0) return STB_OK;
  STB_REQUIRE(nn_distances && nn_index && features && out_features,
              "NearestNeighborsOp.forward: NULL argument");
  STB_FLOAT_DISPATCH(dtype, nn_forward_launch, nn_distances, nn_index, n, k,
                     features, channels, m, out_features, as_stream(stream));
  return check_launch("NearestNeighborsOp.forward");
}

template <typename T>
static void nn_eps_launch(const void* tgt, int64_t m, const void* src,
                          const int64_t* i, int64_t n, int k, double h, void* e,
                          hipStream_t s) {
  hipLaunchKernelGGL(nn_epsilon_kernel<T>, dim3(grid_for(n * 6)), dim3(kBlock),
                     0, s, static_cast<const T*>(tgt), m,
                     static_cast<const T*>(src), i, n, k, T(h),
                     static_cast<T*>(e));
}

extern "C" int stb_nn_epsilon_distances(int32_t dtype, const void* target_xyz,
                                        int64_t m, const void* source_xyz,
                                        const int64_t* nn_index, int64_t n,
                                        int32_t k, double h,
                                        void* epsilon_distances, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "NearestNeighborsOp.compute_epsilon_distances: dtype must be "
              "float32 or float64");
  STB_REQUIRE(n >= 0 && k >= 1 && m >= 0,
              "NearestNeighborsOp.compute_epsilon_distances: bad sizes");
  if (n == 0) return STB_OK;
  STB_REQUIRE(source_xyz && nn_index && epsilon_distances && (m == 0 || target_xyz),
              "NearestNeighborsOp.compute_epsilon_distances: NULL argument");
  STB_FLOAT_DISPATCH(dtype, nn_eps_launch, target_xyz, m, source_xyz, nn_index,
                     n, k, h, epsilon_distances, as_stream(stream));
  return check_launch("NearestNeighborsOp.compute_epsilon_distances");
}

template <typename T>
static void nn_bwd_launch(const void* e, const int64_t* i, int64_t n, int k,
                          const void* f, int C, int64_t m, const void* dl,
                          double h, void* dx, hipStream_t s) {
  hipLaunchKernelGGL(nn_backward_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0,
                     s, static_cast<const T*>(e), i, n, k,
                     static_cast<const T*>(f), C, m, static_cast<const T*>(dl),
                     T(h), static_cast<T*>(dx));
}

extern "C" int stb_nn_features_backward(int32_t dtype, const void* epsilon_distances,
                                        const int64_t* nn_index, int64_t n,
                                        int32_t k, const void* features,
                                        int32_t channels, int64_t m,
                                        const void* dl_features, double h,
                                        void* dl_xyz, void* stream) {
  STB_REQUIRE(dtype == STB_F32 || dtype == STB_F64,
              "NearestNeighborsOp.backward: dtype must be float32 or float64");
  STB_REQUIRE(n >= 0 && k >= 1 && channels >= 0 && m >= 0 && h > 0.0,
              "NearestNeighborsOp.backward: bad sizes");
  if (n == 0) return STB_OK;
  STB_REQUIRE(epsilon_distances && nn_index && dl_xyz &&
                  (channels == 0 || (features && dl_features)),
              "NearestNeighborsOp.backward: NULL argument");
  STB_FLOAT_DISPATCH(dtype, nn_bwd_launch, epsilon_distances, nn_index, n, k,
                     features, channels, m, dl_features, h, dl_xyz,
                     as_stream(stream));
  return check_launch("NearestNeighborsOp.backward");
}

extern "C" int stb_voxel_ids(const float* points, int64_t n, float min_x,
                             float min_y, float min_z, float voxel_size,
                             int32_t width, int32_t height, int32_t* out_ids,
                             void* stream) {
  STB_REQUIRE(n >= 0 && voxel_size > 0.0f, "voxel_ids: bad arguments");
  if (n == 0) return STB_OK;
  STB_REQUIRE(points && out_ids, "voxel_ids: NULL argument");
  hipLaunchKernelGGL(voxel_ids_kernel, dim3(grid_for(n)), dim3(kBlock), 0,
                     as_stream(stream), points, n, min_x, min_y, min_z,
                     voxel_size, width, height, out_ids);
  return check_launch("voxel_ids");
}

extern "C" int stb_voxel_merge(const float* points, const float* normals,
                               const uint8_t* colors, const float* confidences,
                               const float* radii, const float* features,
                               int32_t channels, int64_t n, const int64_t* order,
                               const int64_t* starts, int64_t num_voxels,
                               float* out_points, float* out_normals,
                               uint8_t* out_colors, float* out_confidences,
                               float* out_radii, int32_t* out_times,
                               float* out_features, void* stream) {
  STB_REQUIRE(n >= 0 && num_voxels >= 0 && channels >= 0, "voxel_merge: bad sizes");
  if (num_voxels == 0) return STB_OK;
  STB_REQUIRE(points && normals && colors && confidences && radii && order &&
                  starts && out_points && out_normals && out_colors &&
                  out_confidences && out_radii && out_times &&
                  (channels == 0 || (features && out_features)),
              "voxel_merge: NULL argument");
  hipLaunchKernelGGL(voxel_merge_kernel, dim3(grid_for(num_voxels)), dim3(kBlock),
                     0, as_stream(stream), points, normals, colors, confidences,
                     radii, features, channels, n, order, starts, num_voxels,
                     out_points, out_normals, out_colors, out_confidences,
                     out_radii, out_times, out_features);
  return check_launch("voxel_merge");
}

extern "C" int64_t stb_knn_grid_set_lane_threshold(int64_t max_split_queries) {
  return g_knn_one_lane_queries.exchange(max_split_queries < 0 ? 0 : max_split_queries);
}
