// Median-split binary tree over segment midpoints, built on the GPU element for element like the
// reference (vortex_tree.py:51-115 serial, :170-306 level build + renumbering).
//
// Level-synchronous build.  For every node of a level with more than `leaf` segments:
//   1. box of its midpoints -> split axis (longest extent, the lower axis on ties) and split flag
//      (extent > 0), exactly the reference's comparisons;
//   2. stable sort of the node's range of `order` by the midpoint coordinate on that axis.  A stable sort
//      by key equals an ordinary sort by the unique pair (key, position), so the sort runs on
//      (order-preserving 64-bit key of the coordinate with -0.0 folded into +0.0, position) pairs:
//      bitonic networks in LDS for tiles of up to 4,096 elements, merge-path merges between tiles;
//   3. the children [a, (a+b)//2) and [(a+b)//2, b).
// The host keeps the node table (integer bookkeeping of a few arrays per level) and numbers the nodes
// like the reference's _renumber at the end.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree.h"

namespace ofw {

static constexpr int TILE = 4096;       // elements per LDS sort tile
static constexpr int SORT_THREADS = 256;

__device__ __forceinline__ uint64_t ordered_key(double x) {
  // order-preserving map of a finite double to uint64, with -0.0 == +0.0 (np.argsort compares with <)
  uint64_t u = (x == 0.0) ? 0ull : (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ void k_iota(int m, int32_t* __restrict__ order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) order[i] = i;
}

// ------------------------------------------------------------------ node boxes
struct Box {
  double lo[3], hi[3];
};

__device__ __forceinline__ void box_wave_reduce(double* v) {
  // v[0..2] mins, v[3..5] maxes; butterfly over the 64 lanes (min/max are exact in any order)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double o = __shfl_xor(v[a], off, 64);
      v[a] = o < v[a] ? o : v[a];
      double q = __shfl_xor(v[3 + a], off, 64);
      v[3 + a] = q > v[3 + a] ? q : v[3 + a];
    }
  }
}

// the reference's choice (vortex_tree.py:87-94): axis 0 first, a later axis only if strictly longer
__device__ __forceinline__ void choose_axis(const double* v, int8_t* axis_out, int8_t* split_out) {
  int axis = 0;
  double extent = v[3] - v[0];
  for (int c = 1; c < 3; ++c) {
    double e = v[3 + c] - v[c];
    if (e > extent) {
      extent = e;
      axis = c;
    }
  }
  *axis_out = (int8_t)axis;
  *split_out = (int8_t)(extent > 0.0 ? 1 : 0);
}

// one wavefront per node (nodes of at most TILE segments)
__global__ __launch_bounds__(256) void k_box_small(int count, const int32_t* __restrict__ list,
                                                    const int32_t* __restrict__ clo, const int32_t* __restrict__ chi,
                                                    const int32_t* __restrict__ order, const double* __restrict__ mx,
                                                    const double* __restrict__ my, const double* __restrict__ mz,
                                                    int8_t* __restrict__ axis, int8_t* __restrict__ split) {
  const int w = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (w >= count) return;
  const int node = list[w];
  const int a = clo[node], b = chi[node];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int p = a + lane; p < b; p += 64) {
    const int s = order[p];
    const double c[3] = {mx[s], my[s], mz[s]};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[q] = c[q] < v[q] ? c[q] : v[q];
      v[3 + q] = c[q] > v[3 + q] ? c[q] : v[3 + q];
    }
  }
  box_wave_reduce(v);
  if (lane == 0) choose_axis(v, &axis[node], &split[node]);
}

// big nodes: one workgroup per tile, a partial box per tile
__global__ __launch_bounds__(256) void k_box_tiles(const int32_t* __restrict__ tile_node,
                                                    const int32_t* __restrict__ tile_p0,
                                                    const int32_t* __restrict__ tile_p1,
                                                    const int32_t* __restrict__ order, const double* __restrict__ mx,
                                                    const double* __restrict__ my, const double* __restrict__ mz,
                                                    double* __restrict__ partial) {
  __shared__ double red[4][6];
  const int t = blockIdx.x;
  const int p0 = tile_p0[t], p1 = tile_p1[t];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int s = order[p];
    const double c[3] = {mx[s], my[s], mz[s]};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[q] = c[q] < v[q] ? c[q] : v[q];
      v[3 + q] = c[q] > v[3 + q] ? c[q] : v[3 + q];
    }
  }
  box_wave_reduce(v);
  const int wv = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) red[wv][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / 64); ++k)
      for (int q = 0; q < 3; ++q) {
        v[q] = red[k][q] < v[q] ? red[k][q] : v[q];
        v[3 + q] = red[k][3 + q] > v[3 + q] ? red[k][3 + q] : v[3 + q];
      }
    for (int q = 0; q < 6; ++q) partial[6 * (size_t)t + q] = v[q];
  }
}

__global__ void k_box_combine(int count, const int32_t* __restrict__ list, const int32_t* __restrict__ first_tile,
                              const int32_t* __restrict__ tiles, const double* __restrict__ partial,
                              int8_t* __restrict__ axis, int8_t* __restrict__ split) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int t = first_tile[i]; t < first_tile[i] + tiles[i]; ++t)
    for (int q = 0; q < 3; ++q) {
      double lo = partial[6 * (size_t)t + q], hi = partial[6 * (size_t)t + 3 + q];
      v[q] = lo < v[q] ? lo : v[q];
      v[3 + q] = hi > v[3 + q] ? hi : v[3 + q];
    }
  choose_axis(v, &axis[list[i]], &split[list[i]]);
}

// ------------------------------------------------------------------ sorting
__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t pa, uint64_t kb, uint32_t pb) {
  return ka < kb || (ka == kb && pa < pb);
}

// Bitonic sort of (key, pos) in LDS over `len` elements padded to the power of two `width`.
__device__ void lds_bitonic(uint64_t* key, uint32_t* pos, int width) {
  for (int size = 2; size <= width; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < (width >> 1); i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t ka = key[lo], kb = key[hi];
        const uint32_t pa = pos[lo], pb = pos[hi];
        const bool swap = up ? pair_less(kb, pb, ka, pa) : pair_less(ka, pa, kb, pb);
        if (swap) {
          key[lo] = kb;
          key[hi] = ka;
          pos[lo] = pb;
          pos[hi] = pa;
        }
      }
      __syncthreads();
    }
  }
}

// Sort one node's whole range (at most TILE elements) in LDS and write the new order in place.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_small(const int32_t* __restrict__ list,
                                                             const int32_t* __restrict__ clo,
                                                             const int32_t* __restrict__ chi,
                                                             const int8_t* __restrict__ axis_of,
                                                             int32_t* __restrict__ order,
                                                             const double* __restrict__ mx,
                                                             const double* __restrict__ my,
                                                             const double* __restrict__ mz) {
  __shared__ uint64_t key[TILE];
  __shared__ uint32_t pos[TILE];
  __shared__ int32_t seg[TILE];
  const int node = list[blockIdx.x];
  const int a = clo[node], b = chi[node];
  const int len = b - a;
  int width = 1;
  while (width < len) width <<= 1;
  const int axis = axis_of[node];
  const double* coord = axis == 0 ? mx : (axis == 1 ? my : mz);
  for (int i = threadIdx.x; i < width; i += blockDim.x) {
    if (i < len) {
      const int s = order[a + i];
      seg[i] = s;
      key[i] = ordered_key(coord[s]);
      pos[i] = (uint32_t)i;
    } else {
      key[i] = ~0ull;
      pos[i] = 0xffffffffu;
    }
  }
  __syncthreads();
  lds_bitonic(key, pos, width);
  for (int i = threadIdx.x; i < len; i += blockDim.x) order[a + i] = seg[pos[i]];
}

// Sort one tile of a big node; writes the sorted (key, global position) pairs.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_tiles(const int32_t* __restrict__ tile_node,
                                                             const int32_t* __restrict__ tile_p0,
                                                             const int32_t* __restrict__ tile_p1,
                                                             const int8_t* __restrict__ axis_of,
                                                             const int32_t* __restrict__ order,
                                                             const double* __restrict__ mx,
                                                             const double* __restrict__ my,
                                                             const double* __restrict__ mz,
                                                             uint64_t* __restrict__ out_key,
                                                             uint32_t* __restrict__ out_pos) {
  __shared__ uint64_t key[TILE];
  __shared__ uint32_t pos[TILE];
  const int t = blockIdx.x;
  const int p0 = tile_p0[t], p1 = tile_p1[t];
  const int len = p1 - p0;
  int width = 1;
  while (width < len) width <<= 1;
  const int axis = axis_of[tile_node[t]];
  const double* coord = axis == 0 ? mx : (axis == 1 ? my : mz);
  for (int i = threadIdx.x; i < width; i += blockDim.x) {
    if (i < len) {
      key[i] = ordered_key(coord[order[p0 + i]]);
      pos[i] = (uint32_t)(p0 + i);
    } else {
      key[i] = ~0ull;
      pos[i] = 0xffffffffu;
    }
  }
  __syncthreads();
  lds_bitonic(key, pos, width);
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    out_key[p0 + i] = key[i];
    out_pos[p0 + i] = pos[i];
  }
}

// One merge pass: within each big node, runs of `width` elements (from the node start) are merged in
// pairs.  A workgroup produces one output tile; each thread finds its own start on the merge path by
// binary search and merges PER_THREAD outputs.
static constexpr int PER_THREAD = TILE / SORT_THREADS;

__global__ __launch_bounds__(SORT_THREADS) void k_merge_pass(const int32_t* __restrict__ tile_node,
                                                             const int32_t* __restrict__ tile_p0,
                                                             const int32_t* __restrict__ tile_p1,
                                                             const int32_t* __restrict__ clo,
                                                             const int32_t* __restrict__ chi, int width,
                                                             const uint64_t* __restrict__ in_key,
                                                             const uint32_t* __restrict__ in_pos,
                                                             uint64_t* __restrict__ out_key,
                                                             uint32_t* __restrict__ out_pos) {
  const int t = blockIdx.x;
  const int node = tile_node[t];
  const int a = clo[node], b = chi[node];
  const int p0 = tile_p0[t], p1 = tile_p1[t];
  const int pair_start = a + ((p0 - a) / (2 * width)) * (2 * width);
  const int a0 = pair_start, a1 = min(b, pair_start + width);
  const int b0 = a1, b1 = min(b, a1 + width);
  const int la = a1 - a0, lb = b1 - b0;
  const int o_begin = p0 + threadIdx.x * PER_THREAD;
  if (o_begin >= p1) return;
  const int o_end = min(p1, o_begin + PER_THREAD);
  const int diag = o_begin - pair_start;  // outputs before this thread's first within the pair
  // find i (taken from A) with i + j = diag such that A[i-1] < B[j] and B[j-1] < A[i]
  int lo = max(0, diag - lb), hi = min(diag, la);
  while (lo < hi) {
    const int i = (lo + hi) >> 1;
    const int j = diag - i - 1;  // compare A[i] with B[diag - i - 1]
    if (pair_less(in_key[b0 + j], in_pos[b0 + j], in_key[a0 + i], in_pos[a0 + i]))
      hi = i;
    else
      lo = i + 1;
  }
  int i = lo, j = diag - lo;
  for (int o = o_begin; o < o_end; ++o) {
    bool take_a;
    if (i >= la) take_a = false;
    else if (j >= lb) take_a = true;
    else take_a = !pair_less(in_key[b0 + j], in_pos[b0 + j], in_key[a0 + i], in_pos[a0 + i]);
    if (take_a) {
      out_key[o] = in_key[a0 + i];
      out_pos[o] = in_pos[a0 + i];
      ++i;
    } else {
      out_key[o] = in_key[b0 + j];
      out_pos[o] = in_pos[b0 + j];
      ++j;
    }
  }
}

__global__ __launch_bounds__(256) void k_apply_perm(const int32_t* __restrict__ tile_p0,
                                                    const int32_t* __restrict__ tile_p1,
                                                    const uint32_t* __restrict__ sorted_pos,
                                                    const int32_t* __restrict__ order, int32_t* __restrict__ tmp) {
  const int t = blockIdx.x;
  const int p0 = tile_p0[t], p1 = tile_p1[t];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) tmp[p] = order[sorted_pos[p]];
}

__global__ __launch_bounds__(256) void k_copy_tiles(const int32_t* __restrict__ tile_p0,
                                                    const int32_t* __restrict__ tile_p1,
                                                    const int32_t* __restrict__ tmp, int32_t* __restrict__ order) {
  const int t = blockIdx.x;
  const int p0 = tile_p0[t], p1 = tile_p1[t];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) order[p] = tmp[p];
}

// ------------------------------------------------------------------ host driver
namespace {

template <typename T>
T* carve(char*& base, size_t count) {
  T* p = reinterpret_cast<T*>(base);
  size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
  base += bytes;
  return p;
}

}  // namespace

int build_tree(Ctx& c, int64_t m64, const double* mx, const double* my, const double* mz, int64_t leaf64,
               int32_t* order, HostTree& t) {
  const int m = (int)m64;
  const int leaf = (int)std::min<int64_t>(leaf64, INT_MAX);
  t.m = m;
  t.lo.assign(1, 0);
  t.hi.assign(1, m);
  t.left.assign(1, -1);
  t.right.assign(1, -1);
  if (m == 0) {
    t.count = 1;
    return finish_tree(t);
  }
  hipLaunchKernelGGL(k_iota, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, order);
  OFW_HIP(hipGetLastError());
  // scratch: sort buffers (2 x (key, pos)), tmp order, per-level tables
  const size_t cap_nodes = (size_t)4 * (m / std::max(leaf, 1) + 1) + 1;
  const size_t max_tiles = (size_t)m / TILE + cap_nodes + 8;
  size_t bytes = 0;
  bytes += 2 * (((size_t)m * 8 + 255) & ~(size_t)255) + 2 * (((size_t)m * 4 + 255) & ~(size_t)255);
  bytes += (((size_t)m * 4 + 255) & ~(size_t)255);
  bytes += 8 * ((cap_nodes * 4 + 255) & ~(size_t)255) + 4 * ((max_tiles * 4 + 255) & ~(size_t)255);
  bytes += ((max_tiles * 6 * 8 + 255) & ~(size_t)255);
  OFW_HIP(c.tree[0].reserve(bytes));
  char* base = (char*)c.tree[0].ptr;
  uint64_t* keyA = carve<uint64_t>(base, m);
  uint64_t* keyB = carve<uint64_t>(base, m);
  uint32_t* posA = carve<uint32_t>(base, m);
  uint32_t* posB = carve<uint32_t>(base, m);
  int32_t* tmp = carve<int32_t>(base, m);
  int32_t* d_clo = carve<int32_t>(base, cap_nodes);
  int32_t* d_chi = carve<int32_t>(base, cap_nodes);
  int32_t* d_list = carve<int32_t>(base, cap_nodes);
  int32_t* d_first = carve<int32_t>(base, cap_nodes);
  int32_t* d_ntiles = carve<int32_t>(base, cap_nodes);
  int8_t* d_axis = carve<int8_t>(base, cap_nodes);
  int8_t* d_split = carve<int8_t>(base, cap_nodes);
  int32_t* d_slist = carve<int32_t>(base, cap_nodes);
  int32_t* d_tnode = carve<int32_t>(base, max_tiles);
  int32_t* d_tp0 = carve<int32_t>(base, max_tiles);
  int32_t* d_tp1 = carve<int32_t>(base, max_tiles);
  int32_t* d_unused = carve<int32_t>(base, max_tiles);
  (void)d_unused;
  double* d_partial = carve<double>(base, max_tiles * 6);

  std::vector<int32_t> level{0};
  std::vector<int32_t> cand, clo, chi, small, big, tnode, tp0, tp1, first, ntiles, slist_small, slist_big;
  std::vector<int8_t> axis, split;
  std::vector<int32_t> bt_node, bt_p0, bt_p1;  // alive until the next synchronisation (async uploads)
  while (!level.empty()) {
    cand.clear();
    for (int node : level)
      if (t.hi[node] - t.lo[node] > leaf) cand.push_back(node);
    if (cand.empty()) break;
    const int nc = (int)cand.size();
    clo.resize(nc);
    chi.resize(nc);
    small.clear();
    big.clear();
    tnode.clear();
    tp0.clear();
    tp1.clear();
    first.clear();
    ntiles.clear();
    for (int i = 0; i < nc; ++i) {
      clo[i] = t.lo[cand[i]];
      chi[i] = t.hi[cand[i]];
      if (chi[i] - clo[i] <= TILE) {
        small.push_back(i);
      } else {
        big.push_back(i);
        first.push_back((int)tnode.size());
        int k = 0;
        for (int p = clo[i]; p < chi[i]; p += TILE, ++k) {
          tnode.push_back(i);
          tp0.push_back(p);
          tp1.push_back(std::min(chi[i], p + TILE));
        }
        ntiles.push_back(k);
      }
    }
    // upload the level's tables
    OFW_HIP(hipMemcpyAsync(d_clo, clo.data(), nc * 4, hipMemcpyHostToDevice, c.stream));
    OFW_HIP(hipMemcpyAsync(d_chi, chi.data(), nc * 4, hipMemcpyHostToDevice, c.stream));
    if (!small.empty())
      OFW_HIP(hipMemcpyAsync(d_list, small.data(), small.size() * 4, hipMemcpyHostToDevice, c.stream));
    const int nt = (int)tnode.size();
    if (!big.empty()) {
      OFW_HIP(hipMemcpyAsync(d_slist, big.data(), big.size() * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_first, first.data(), big.size() * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_ntiles, ntiles.data(), big.size() * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_tnode, tnode.data(), nt * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_tp0, tp0.data(), nt * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_tp1, tp1.data(), nt * 4, hipMemcpyHostToDevice, c.stream));
    }
    // 1. boxes -> axis and split flag of every candidate
    if (!small.empty())
      hipLaunchKernelGGL(k_box_small, dim3(((int)small.size() + 3) / 4), dim3(256), 0, c.stream, (int)small.size(),
                         d_list, d_clo, d_chi, order, mx, my, mz, d_axis, d_split);
    if (!big.empty()) {
      hipLaunchKernelGGL(k_box_tiles, dim3(nt), dim3(256), 0, c.stream, d_tnode, d_tp0, d_tp1, order, mx, my, mz,
                         d_partial);
      hipLaunchKernelGGL(k_box_combine, dim3(((int)big.size() + 63) / 64), dim3(64), 0, c.stream, (int)big.size(),
                         d_slist, d_first, d_ntiles, d_partial, d_axis, d_split);
    }
    OFW_HIP(hipGetLastError());
    axis.resize(nc);
    split.resize(nc);
    OFW_HIP(hipMemcpyAsync(split.data(), d_split, nc, hipMemcpyDeviceToHost, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    // 2. sort the ranges of the nodes that split
    slist_small.clear();
    for (int i : small)
      if (split[i]) slist_small.push_back(i);
    bool any_big = false;
    for (int i : big) any_big = any_big || split[i];
    if (!slist_small.empty()) {
      OFW_HIP(hipMemcpyAsync(d_list, slist_small.data(), slist_small.size() * 4, hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_sort_small, dim3((unsigned)slist_small.size()), dim3(SORT_THREADS), 0, c.stream, d_list,
                         d_clo, d_chi, d_axis, order, mx, my, mz);
      OFW_HIP(hipGetLastError());
    }
    if (any_big) {
      // big nodes that do not split keep their order: restrict the tile list to splitting nodes
      bt_node.clear();
      bt_p0.clear();
      bt_p1.clear();
      int maxlen = 0;
      for (size_t k = 0; k < big.size(); ++k) {
        const int i = big[k];
        if (!split[i]) continue;
        maxlen = std::max(maxlen, chi[i] - clo[i]);
        for (int q = 0; q < ntiles[k]; ++q) {
          bt_node.push_back(i);
          bt_p0.push_back(tp0[first[k] + q]);
          bt_p1.push_back(tp1[first[k] + q]);
        }
      }
      const int nbt = (int)bt_node.size();
      OFW_HIP(hipMemcpyAsync(d_tnode, bt_node.data(), nbt * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_tp0, bt_p0.data(), nbt * 4, hipMemcpyHostToDevice, c.stream));
      OFW_HIP(hipMemcpyAsync(d_tp1, bt_p1.data(), nbt * 4, hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_sort_tiles, dim3(nbt), dim3(SORT_THREADS), 0, c.stream, d_tnode, d_tp0, d_tp1, d_axis,
                         order, mx, my, mz, keyA, posA);
      uint64_t *kin = keyA, *kout = keyB;
      uint32_t *pin = posA, *pout = posB;
      for (int width = TILE; width < maxlen; width <<= 1) {
        hipLaunchKernelGGL(k_merge_pass, dim3(nbt), dim3(SORT_THREADS), 0, c.stream, d_tnode, d_tp0, d_tp1, d_clo,
                           d_chi, width, kin, pin, kout, pout);
        std::swap(kin, kout);
        std::swap(pin, pout);
      }
      hipLaunchKernelGGL(k_apply_perm, dim3(nbt), dim3(256), 0, c.stream, d_tp0, d_tp1, pin, order, tmp);
      hipLaunchKernelGGL(k_copy_tiles, dim3(nbt), dim3(256), 0, c.stream, d_tp0, d_tp1, tmp, order);
      OFW_HIP(hipGetLastError());
    }
    // 3. children (vortex_tree.py:244-268)
    std::vector<int32_t> next;
    for (int i = 0; i < nc; ++i) {
      if (!split[i]) continue;
      const int node = cand[i];
      const int a = clo[i], b = chi[i], middle = (int)(((int64_t)a + b) / 2);
      const int first_child = (int)t.lo.size();
      t.lo.push_back(a);
      t.hi.push_back(middle);
      t.lo.push_back(middle);
      t.hi.push_back(b);
      t.left.push_back(-1);
      t.left.push_back(-1);
      t.right.push_back(-1);
      t.right.push_back(-1);
      t.left[node] = first_child;
      t.right[node] = first_child + 1;
      next.push_back(first_child);
      next.push_back(first_child + 1);
    }
    level.swap(next);
  }
  t.count = (int64_t)t.lo.size();
  return finish_tree(t);
}

// Reference numbering (vortex_tree.py:271-306, _renumber) and the pre-order layout used by the traversal.
int finish_tree(HostTree& t) {
  const int count = (int)t.count;
  std::vector<int32_t> number(count, 0), stack;
  t.ref_lo.assign(count, 0);
  t.ref_hi.assign(count, 0);
  t.ref_left.assign(count, -1);
  t.ref_right.assign(count, -1);
  t.ref_lo[0] = t.lo[0];
  t.ref_hi[0] = t.hi[0];
  int assigned = 1;
  stack.push_back(0);
  while (!stack.empty()) {
    const int node = stack.back();
    stack.pop_back();
    if (t.left[node] < 0) continue;
    const int f = assigned, s = assigned + 1;
    assigned += 2;
    number[t.left[node]] = f;
    number[t.right[node]] = s;
    t.ref_left[number[node]] = f;
    t.ref_right[number[node]] = s;
    t.ref_lo[f] = t.lo[t.left[node]];
    t.ref_hi[f] = t.hi[t.left[node]];
    t.ref_lo[s] = t.lo[t.right[node]];
    t.ref_hi[s] = t.hi[t.right[node]];
    stack.push_back(t.right[node]);
    stack.push_back(t.left[node]);
  }
  // pre-order (left first) of the reference-numbered tree: the visit order of its stack walk
  t.pre_ref.assign(count, 0);   // pre-order index -> reference node number
  t.pre_skip.assign(count, 0);  // pre-order index -> pre-order index after its subtree
  t.pre_parent.assign(count, -1);
  t.pre_depth.assign(count, 0);
  std::vector<int32_t> pstack;  // pre-order index of nodes on the stack
  std::vector<int32_t> size(count, 1);
  int next = 0;
  stack.clear();
  stack.push_back(0);
  pstack.push_back(-1);
  std::vector<int32_t> pre_of_ref(count, 0);
  while (!stack.empty()) {
    const int r = stack.back();
    stack.pop_back();
    const int parent = pstack.back();
    pstack.pop_back();
    const int i = next++;
    pre_of_ref[r] = i;
    t.pre_ref[i] = r;
    t.pre_parent[i] = parent;
    t.pre_depth[i] = parent < 0 ? 0 : t.pre_depth[parent] + 1;
    if (t.ref_left[r] >= 0) {
      stack.push_back(t.ref_right[r]);
      pstack.push_back(i);
      stack.push_back(t.ref_left[r]);
      pstack.push_back(i);
    }
  }
  for (int i = count - 1; i > 0; --i) size[t.pre_parent[i]] += size[i];
  for (int i = 0; i < count; ++i) t.pre_skip[i] = i + size[i];
  t.depth = 0;
  for (int i = 0; i < count; ++i) t.depth = std::max(t.depth, (int)t.pre_depth[i]);
  return OFW_OK;
}

}  // namespace ofw
