// Median-split binary tree over segment midpoints, built on the GPU element for element like the
// reference (vortex_tree.py:51-115 serial build, :170-306 level build + renumbering), for one tree or a
// forest of trees over disjoint segment ranges.  The whole build is enqueued without any host
// synchronisation (fixed launch sequence, counts kept on the device), so it can be captured in a graph.
//
// Every node with more than `leaf` segments: the box of its midpoints gives the split axis (longest extent,
// the lower axis on ties) and the split flag (extent > 0), with the reference's comparisons; its range of
// `order` is stably sorted by the midpoint coordinate on that axis; the range is cut at (a+b)//2.  A stable
// sort by key equals an ordinary sort by the unique pair (key, position), with the key the order-preserving
// 64-bit image of the double and -0.0 folded into +0.0 (np.argsort compares with <).
//
//  * Top phase, nodes of more than S_BOTTOM segments, level by level: partial boxes over tiles, the node's
//    box and sort of each 4,096-element tile (registers, then in-place LDS merge rounds), merge-path passes
//    between tiles, the last of which applies the sorted order into the other of two order buffers; a
//    one-workgroup kernel creates the children and plans the next level (tiles and deferred subtrees).
//  * Bottom phase: each node of at most S_BOTTOM segments is finished by one workgroup entirely in LDS,
//    level after level (box, each splitting node's (key, position) merge sort on its own thread group,
//    children), into a heap block.
//  * Finish: in a median-split tree the pre-order (the visit order of the reference's stack walk) is the
//    order of the nodes by (lo, depth), so pre(N) = #nodes with lo < N.lo + depth(N) - min depth at N.lo and
//    the escape index is #nodes with lo < N.hi: a histogram over lo and one scan give the whole layout.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "scan.h"
#include "tree.h"

namespace ofw {

static constexpr int TILE = 4096;           // elements per LDS tile of the top phase
static constexpr int TILE_THREADS = 1024;
static constexpr int TILE_E = TILE / TILE_THREADS;  // elements per thread of the tile sort
static constexpr int S_BOTTOM = 2048;       // largest subtree finished in LDS
static constexpr int BOTTOM_THREADS = 1024;
static constexpr int BOTTOM_E = S_BOTTOM / BOTTOM_THREADS;  // elements per thread of the subtree sort
static constexpr int HCAP = 2 * S_BOTTOM - 1;
static constexpr int PLAN_THREADS = 1024;
static constexpr int SUBS = 4;              // workgroups per tile of the elementwise top-phase kernels
// list build
static constexpr int RUN_MAX = 256;                  // longest run of equal coordinates the lists re-order
static constexpr int FIX_WAVES = PLAN_THREADS / 64;  // wavefronts (one run each) per workgroup of the re-order
static constexpr int FIX_BLOCKS = 64;                // workgroups of the re-order
static constexpr int32_t TIE_FLAG = INT32_MIN;       // list entry of a segment in a run of equal coordinates
static constexpr int32_t SEG_MASK = INT32_MAX;       // the segment id of a list entry

// Order-preserving 32-bit key of x rounded to float32 (round to nearest: a < b gives key(a) <= key(b)); -0.0 and
// +0.0 share the key of +0.0.
__device__ __forceinline__ uint32_t ordered_key32(double x) {
  const float f = (float)x;
  const uint32_t u = f == 0.0f ? 0u : __float_as_uint(f);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint64_t ordered_key(double x) {
  uint64_t u = (x == 0.0) ? 0ull : (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__host__ __device__ __forceinline__ int heap_slots(int size, int leaf) {
  int d = 0;
  while (size > leaf) {
    size = (size + 1) / 2;
    ++d;
  }
  return (2 << d) - 1;
}

// ------------------------------------------------------------------ init
__global__ void k_iota(int m, int32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) order[i] = i;
}

// The roots, the counters and the single-pass scans' tickets and status words, by one workgroup (every thread of
// it calls this).
__device__ void init_roots_block(const TreeArrays& t) {
  for (int k = threadIdx.x; k < t.K; k += blockDim.x) {
    // the root of an empty tree is no node (lo = -1): it takes no place in the pre-order layout
    const bool empty = t.roots_lo[k] == t.roots_lo[k + 1];
    t.lo[k] = empty ? -1 : t.roots_lo[k];
    t.hi[k] = empty ? -1 : t.roots_lo[k + 1];
    t.left[k] = -1;
    t.right[k] = -1;
    t.parent[k] = -1;
    t.depth[k] = 0;
    t.tree[k] = k;
    t.level_a[k] = k;
  }
  if (threadIdx.x < C_COUNT) t.counters[threadIdx.x] = 0;
  // the finish's single-pass scans: tickets and status words start empty in every call
  if (threadIdx.x < 2) t.scan_ticket[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < t.scan_words; i += blockDim.x) t.scan_status[i] = 0ull;
  // the list build's work-unit tickets and look-back status words start empty in every call
  if (t.lists) {
    for (int i = threadIdx.x; i < LEVEL_MAX; i += blockDim.x) t.tickets[i] = 0;
    for (int i = threadIdx.x; i < 3 * t.tile_cap; i += blockDim.x) t.lstat[i] = 0ull;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    t.counters[C_NODES] = t.K;
    t.counters[C_LEVEL] = t.K;
  }
}

__global__ void k_init_roots(TreeArrays t) { init_roots_block(t); }

template <typename T>
__global__ void k_fill_n(int64_t count, T* __restrict__ a, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = v;
}

__global__ void k_fill_i32(int count, int32_t* __restrict__ a, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = v;
}

struct I32Chunk {
  int count;
  int32_t v[240];
};

__global__ void k_store_chunk(I32Chunk ch, int32_t* __restrict__ dst) {
  if ((int)threadIdx.x < ch.count) dst[threadIdx.x] = ch.v[threadIdx.x];
}

int upload_i32(Ctx& c, const std::vector<int64_t>& values, int32_t* dst) {
  const int total = (int)values.size();
  for (int base = 0; base < total; base += 240) {
    I32Chunk ch{};
    ch.count = std::min(240, total - base);
    for (int q = 0; q < ch.count; ++q) ch.v[q] = (int32_t)values[base + q];
    hipLaunchKernelGGL(k_store_chunk, dim3(1), dim3(256), 0, c.stream, ch, dst + base);
  }
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

// ------------------------------------------------------------------ top phase
// Classes of a top-phase node: a big node (more than S_BOTTOM and `leaf` segments) is sorted by tiles, a node of
// leaf < size <= S_BOTTOM is deferred to the bottom phase, any other node is a leaf.
__host__ __device__ __forceinline__ bool is_big(int size, int leaf) { return size > S_BOTTOM && size > leaf; }
__host__ __device__ __forceinline__ bool is_small(int size, int leaf) { return size > leaf && size <= S_BOTTOM; }

// One workgroup: the plan of the first level (the roots): big roots get tiles, small ones go to the bottom.
__global__ __launch_bounds__(PLAN_THREADS) void k_level_plan(TreeArrays t, const int32_t* __restrict__ level) {
  const int count = live_count(t.counters, C_LEVEL);
  int nbig = 0, ntiles = 0, nsmall = 0;
  const int bottom0 = t.counters[C_BOTTOM];
  for (int base = 0; base < count; base += blockDim.x) {
    const int i = base + threadIdx.x;
    int node = -1, size = 0;
    if (i < count) {
      node = level[i];
      size = t.hi[node] - t.lo[node];
    }
    const bool big = node >= 0 && is_big(size, t.leaf);
    const bool small = node >= 0 && is_small(size, t.leaf);
    const int tiles = big ? (size + TILE - 1) / TILE : 0;
    const int sv[3] = {big ? 1 : 0, tiles, small ? 1 : 0};
    int sr[3], stot[3];
    block_exclusive_scan_n<3>(sv, sr, stot);
    const int rb = sr[0], rt = sr[1], rs = sr[2];
    const int tb = stot[0], tt = stot[1], ts = stot[2];
    if (big) {
      const int bi = nbig + rb;
      const int t0 = ntiles + rt;
      if (bi < t.cap && t0 + tiles <= t.tile_cap) {
        t.big[bi] = node;
        const int lo = t.lo[node], hi = t.hi[node];  // read once: the stores below may alias them
        for (int q = 0; q < tiles; ++q) {
          t.tile_node[t0 + q] = node;
          t.tile_p0[t0 + q] = lo + q * TILE;
          t.tile_p1[t0 + q] = min(hi, lo + (q + 1) * TILE);
        }
      } else {
        t.counters[C_OVERFLOW] = 1;
      }
    }
    if (small) {
      const int b = bottom0 + nsmall + rs;
      if (b < t.bottom_cap) t.bottom[b] = node;
      else t.counters[C_OVERFLOW] = 1;
    }
    nbig += tb;
    ntiles += tt;
    nsmall += ts;
  }
  if (threadIdx.x == 0) {
    t.counters[C_BIG] = nbig;
    t.counters[C_TILES] = ntiles;
    t.counters[C_BOTTOM] = bottom0 + nsmall;
  }
}

__device__ __forceinline__ void box_wave_reduce(double* v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double o = __shfl_xor(v[a], off, 64);
      v[a] = o < v[a] ? o : v[a];
      const double q = __shfl_xor(v[3 + a], off, 64);
      v[3 + a] = q > v[3 + a] ? q : v[3 + a];
    }
  }
}

// the reference's choice (vortex_tree.py:87-94): axis 0 first, a later axis only if strictly longer
__device__ __forceinline__ void choose_axis(const double* v, int8_t* axis_out, int8_t* split_out) {
  int axis = 0;
  double extent = v[3] - v[0];
  for (int c = 1; c < 3; ++c) {
    const double e = v[3 + c] - v[c];
    if (e > extent) {
      extent = e;
      axis = c;
    }
  }
  *axis_out = (int8_t)axis;
  *split_out = (int8_t)(extent > 0.0 ? 1 : 0);
}

// Part `sub` of SUBS equal parts of the tile range [p0, p1).
__device__ __forceinline__ void sub_range(int p0, int p1, int sub, int& a, int& b) {
  const int chunk = (p1 - p0 + SUBS - 1) / SUBS;
  a = min(p1, p0 + sub * chunk);
  b = min(p1, a + chunk);
}

static constexpr int BOX_THREADS = 256;
static constexpr int BOX_PER = TILE / SUBS / BOX_THREADS;  // elements per thread of a work unit
static_assert(TILE % (SUBS * BOX_THREADS) == 0, "a box work unit holds a whole number of elements per thread");

__global__ __launch_bounds__(BOX_THREADS) void k_box_tiles(TreeArrays t, const int32_t* __restrict__ order_in,
                                                   const double* __restrict__ mx, const double* __restrict__ my,
                                                   const double* __restrict__ mz) {
  __shared__ double red[4][6];
  const int ntiles = live_count(t.counters, C_TILES);
  for (int wu = blockIdx.x; wu < ntiles * SUBS; wu += gridDim.x) {
    const int tile = wu / SUBS;
    int p0, p1;
    sub_range(t.tile_p0[tile], t.tile_p1[tile], wu % SUBS, p0, p1);
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    // a work unit holds at most BOX_PER elements per thread: every order entry is read first, then the
    // midpoints, so that no load waits for another
    int sidx[BOX_PER];
#pragma unroll
    for (int e = 0; e < BOX_PER; ++e) {
      const int p = p0 + (int)threadIdx.x + e * BOX_THREADS;
      sidx[e] = p < p1 ? order_in[p] : -1;
    }
#pragma unroll
    for (int e = 0; e < BOX_PER; ++e)
      if (sidx[e] >= 0) {
        const int s = sidx[e];
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
      for (int q = 0; q < 6; ++q) t.box_partial[6 * (size_t)wu + q] = v[q];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t pa, uint64_t kb, uint32_t pb) {
  return ka < kb || (ka == kb && pa < pb);
}

// Workgroup sort of `width` (key, tag) elements in LDS (width a power of two, E <= width <= blockDim * E):
// each thread sorts E consecutive elements in registers (bitonic network), then runs of E, 2E, ... are
// merged in place: every thread finds its first output with a merge-path search in LDS, merges its E outputs
// into registers, and writes them back after a barrier.  Every thread of the workgroup must call it.
template <int E, typename Less>
__device__ void block_merge_sort(uint64_t* s_key, uint32_t* s_tag, int width, Less less) {
  const int tid = threadIdx.x;
  const bool active = tid * E < width;
  uint64_t k[E];
  uint32_t p[E];
  if (active) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      k[e] = s_key[tid * E + e];
      p[e] = s_tag[tid * E + e];
    }
#pragma unroll
    for (int size = 2; size <= E; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int j = i ^ stride;
          if (j > i) {
            const bool up = (i & size) == 0;
            const bool swap = up ? less(k[j], p[j], k[i], p[i]) : less(k[i], p[i], k[j], p[j]);
            if (swap) {
              const uint64_t tk = k[i];
              k[i] = k[j];
              k[j] = tk;
              const uint32_t tp = p[i];
              p[i] = p[j];
              p[j] = tp;
            }
          }
        }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      s_key[tid * E + e] = k[e];
      s_tag[tid * E + e] = p[e];
    }
  }
  for (int run = E; run < width; run <<= 1) {
    // a merge of two runs within one wavefront's 64 E elements reads and writes that wavefront's elements
    // only: its LDS order needs no workgroup barrier
    const bool local = 2 * run <= 64 * E;
    if (local) wave_lds_sync();
    else __syncthreads();  // the previous stores are visible
    if (active) {
      const int out0 = tid * E;
      const int pair0 = out0 & ~(2 * run - 1);
      const int diag = out0 - pair0;
      const uint64_t* ka = s_key + pair0;
      const uint32_t* pa = s_tag + pair0;
      const uint64_t* kb = ka + run;
      const uint32_t* pb = pa + run;
      int lo = max(0, diag - run), hi = min(diag, run);
      while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int j = diag - i - 1;
        if (less(kb[j], pb[j], ka[i], pa[i])) hi = i;
        else lo = i + 1;
      }
      int i = lo, j = diag - lo;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // outputs of one thread never pass the end of both runs: i + j < 2 * run
        const bool take_a = j >= run || (i < run && !less(kb[j], pb[j], ka[i], pa[i]));
        if (take_a) {
          k[e] = ka[i];
          p[e] = pa[i];
          ++i;
        } else {
          k[e] = kb[j];
          p[e] = pb[j];
          ++j;
        }
      }
    }
    if (local) wave_lds_sync();
    else __syncthreads();  // every read of this round is done
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        s_key[tid * E + e] = k[e];
        s_tag[tid * E + e] = p[e];
      }
    }
  }
  __syncthreads();
}

struct PairLess {
  __device__ bool operator()(uint64_t ka, uint32_t pa, uint64_t kb, uint32_t pb) const {
    return pair_less(ka, pa, kb, pb);
  }
};

// Box of a big node from the partial boxes of all its tiles, by every workgroup that sorts one of them (exact
// min / max, so every tile finds the same box); the reference's axis choice and split flag follow.
__device__ void node_axis(const TreeArrays& t, int tile, int node, int* axis_out, bool* split_out) {
  __shared__ double red[TILE_THREADS / 64][6];
  __shared__ int8_t s_axis, s_split;
  const int lo = t.lo[node], hi = t.hi[node];
  const int t0 = tile - (t.tile_p0[tile] - lo) / TILE;  // first tile of the node
  const int w1 = (t0 + (hi - lo + TILE - 1) / TILE) * SUBS;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int w = t0 * SUBS + (int)threadIdx.x; w < w1; w += blockDim.x)
    for (int c = 0; c < 3; ++c) {
      const double l = t.box_partial[6 * (size_t)w + c], h = t.box_partial[6 * (size_t)w + 3 + c];
      v[c] = l < v[c] ? l : v[c];
      v[3 + c] = h > v[3 + c] ? h : v[3 + c];
    }
  box_wave_reduce(v);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) red[threadIdx.x / 64][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / 64); ++k)
      for (int q = 0; q < 3; ++q) {
        v[q] = red[k][q] < v[q] ? red[k][q] : v[q];
        v[3 + q] = red[k][3 + q] > v[3 + q] ? red[k][3 + q] : v[3 + q];
      }
    choose_axis(v, &s_axis, &s_split);
    if (t0 == tile) {  // the node's first tile records it for the merge passes and the next plan
      t.axis[node] = s_axis;
      t.split[node] = s_split;
    }
  }
  __syncthreads();
  *axis_out = s_axis;
  *split_out = s_split != 0;
  __syncthreads();  // the next tile reuses red and the flags
}

// Sort of each tile of the level's big nodes by (key, position).  With no merge pass after it (every node fits
// one tile) the sorted positions are applied at once: order_out[p] = order_in[sorted position], else the sorted
// pairs go to keyA / posA.
__global__ __launch_bounds__(TILE_THREADS) void k_sort_tiles(TreeArrays t, const int32_t* __restrict__ order_in,
                                                             int32_t* __restrict__ order_out,
                                                             const double* __restrict__ mx,
                                                             const double* __restrict__ my,
                                                             const double* __restrict__ mz) {
  __shared__ uint64_t key[TILE];
  __shared__ uint32_t pos[TILE];
  const int ntiles = live_count(t.counters, C_TILES);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int node = t.tile_node[tile];
    int axis;
    bool split;
    node_axis(t, tile, node, &axis, &split);
    if (!split) continue;  // uniform over the workgroup
    const int p0 = t.tile_p0[tile], p1 = t.tile_p1[tile];
    const int len = p1 - p0;
    int width = TILE_E;
    while (width < len) width <<= 1;
    const double* coord = axis == 0 ? mx : (axis == 1 ? my : mz);
    // TILE_E elements per thread: the order entries first, then the coordinates
    int gidx[TILE_E];
#pragma unroll
    for (int e = 0; e < TILE_E; ++e) {
      const int i = (int)threadIdx.x + e * TILE_THREADS;
      gidx[e] = i < len ? order_in[p0 + i] : 0;
    }
#pragma unroll
    for (int e = 0; e < TILE_E; ++e) {
      const int i = (int)threadIdx.x + e * TILE_THREADS;
      if (i < len) {
        key[i] = ordered_key(coord[gidx[e]]);
        pos[i] = (uint32_t)(p0 + i);
      } else if (i < width) {
        key[i] = ~0ull;
        pos[i] = 0xffffffffu;
      }
    }
    __syncthreads();
    block_merge_sort<TILE_E>(key, pos, width, PairLess());
    if (order_out) {
      int g[TILE_E];
#pragma unroll
      for (int e = 0; e < TILE_E; ++e) {
        const int i = (int)threadIdx.x + e * TILE_THREADS;
        g[e] = i < len ? order_in[pos[i]] : 0;
      }
#pragma unroll
      for (int e = 0; e < TILE_E; ++e) {
        const int i = (int)threadIdx.x + e * TILE_THREADS;
        if (i < len) order_out[p0 + i] = g[e];
      }
    } else {
      for (int i = threadIdx.x; i < len; i += blockDim.x) {
        t.keyA[p0 + i] = key[i];
        t.posA[p0 + i] = pos[i];
      }
    }
    __syncthreads();
  }
}

// One merge pass: within each big split node, runs of `width` elements (from the node start) are merged in
// pairs.  A workgroup produces one slice of MERGE_OUT outputs: two merge-path searches in global memory (one
// wavefront each) give the slice's input ranges of both runs, which are staged in LDS; each thread finds its
// own start in LDS, merges MERGE_E outputs into registers, and the slice leaves through LDS with coalesced
// stores.
static constexpr int MERGE_THREADS = 1024;
static constexpr int MERGE_E = 4;                                // outputs per thread
static constexpr int MERGE_OUT = MERGE_THREADS * MERGE_E;        // outputs per work item
static constexpr int MERGE_SLICES = TILE / MERGE_OUT;            // work items per sort tile
// every merge pass must write its whole output (the host swaps the buffers after each pass), and a slice must
// not straddle a pair of runs: whole slices per tile
static_assert(TILE >= MERGE_OUT && TILE % MERGE_OUT == 0, "a sort tile holds a whole number of merge slices");
static_assert(TILE % TILE_THREADS == 0 && (TILE_E & (TILE_E - 1)) == 0, "a power-of-two tile sort per thread");

// number of elements taken from run A among the first `diag` outputs of merging A (la) and B (lb)
__device__ __forceinline__ int merge_split(const uint64_t* __restrict__ ka, const uint32_t* __restrict__ pa, int la,
                                           const uint64_t* __restrict__ kb, const uint32_t* __restrict__ pb, int lb,
                                           int diag) {
  int lo = max(0, diag - lb), hi = min(diag, la);
  while (lo < hi) {
    const int i = (lo + hi) >> 1;
    const int j = diag - i - 1;
    if (pair_less(kb[j], pb[j], ka[i], pa[i])) hi = i;
    else lo = i + 1;
  }
  return lo;
}

// merge_split by one wavefront: 64 probes per round trip narrow the range 65-fold (the predicate
// "B[diag - i - 1] < A[i]" is monotone in i), instead of one dependent load pair per halving
__device__ int wave_merge_split(const uint64_t* __restrict__ ka, const uint32_t* __restrict__ pa, int la,
                                const uint64_t* __restrict__ kb, const uint32_t* __restrict__ pb, int lb, int diag) {
  const int lane = threadIdx.x & 63;
  int lo = max(0, diag - lb), hi = min(diag, la);  // the answer (first i with the predicate, else hi) is in [lo, hi]
  while (true) {
    const int span = hi - lo;
    if (span <= 64) {
      int best = hi;
      const int q = lo + lane;
      if (q < hi) {
        const int j = diag - q - 1;
        if (pair_less(kb[j], pb[j], ka[q], pa[q])) best = q;
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) best = min(best, __shfl_xor(best, off, 64));
      return best;
    }
    // 64 distinct probes i_k = lo + floor((k + 1) * span / 65) < hi
    const int i = lo + (int)(((int64_t)(lane + 1) * span) / 65);
    const int j = diag - i - 1;
    const bool pred = pair_less(kb[j], pb[j], ka[i], pa[i]);
    const unsigned long long ballot = __builtin_amdgcn_ballot_w64(pred);
    if (ballot == 0) {
      lo = lo + (int)(((int64_t)64 * span) / 65) + 1;  // above the last probe
    } else {
      const int first = __builtin_ctzll(ballot);
      const int i_first = lo + (int)(((int64_t)(first + 1) * span) / 65);
      lo = first == 0 ? lo : lo + (int)(((int64_t)first * span) / 65) + 1;  // above the last false probe
      hi = i_first;
    }
  }
}

// The level's last pass (order_out not null) applies the sorted positions, order_out[q] = order_in[position],
// instead of storing the pairs.
__global__ __launch_bounds__(MERGE_THREADS) void k_merge_pass(TreeArrays t, int width,
                                                              const uint64_t* __restrict__ in_key,
                                                              const uint32_t* __restrict__ in_pos,
                                                              uint64_t* __restrict__ out_key,
                                                              uint32_t* __restrict__ out_pos,
                                                              const int32_t* __restrict__ order_in,
                                                              int32_t* __restrict__ order_out) {
  __shared__ uint64_t s_key[MERGE_OUT];
  __shared__ uint32_t s_pos[MERGE_OUT];
  __shared__ int s_split[2];
  const int ntiles = live_count(t.counters, C_TILES);
  // work item = one MERGE_OUT slice of a sort tile's output range
  for (int w = blockIdx.x; w < ntiles * MERGE_SLICES; w += gridDim.x) {
    const int tile = w / MERGE_SLICES, slice = w - tile * MERGE_SLICES;
    const int node = t.tile_node[tile];
    if (!t.split[node]) continue;  // uniform over the workgroup
    const int a = t.lo[node], b = t.hi[node];
    const int q0 = t.tile_p0[tile] + slice * MERGE_OUT;
    const int q1 = min(t.tile_p1[tile], q0 + MERGE_OUT);
    if (q0 >= q1) continue;
    const int pair_start = a + ((q0 - a) / (2 * width)) * (2 * width);
    const int a0 = pair_start, a1 = min(b, pair_start + width);
    const int b0 = a1, b1 = min(b, a1 + width);
    const int la = a1 - a0, lb = b1 - b0;
    const int d0 = q0 - pair_start, d1 = q1 - pair_start;
    if (threadIdx.x < 128) {
      // wavefront 0 finds the slice's first split, wavefront 1 its end
      const int wv = threadIdx.x >> 6;
      const int split = wave_merge_split(in_key + a0, in_pos + a0, la, in_key + b0, in_pos + b0, lb, wv ? d1 : d0);
      if ((threadIdx.x & 63) == 0) s_split[wv] = split;
    }
    __syncthreads();
    const int i0 = s_split[0], i1 = s_split[1];
    const int j0 = d0 - i0, j1 = d1 - i1;
    const int na = i1 - i0, nb = j1 - j0;  // na + nb = q1 - q0 <= MERGE_OUT
    {
      // MERGE_E inputs per thread, loaded together
      uint64_t kk[MERGE_E];
      uint32_t pp[MERGE_E];
#pragma unroll
      for (int r = 0; r < MERGE_E; ++r) {
        const int e = (int)threadIdx.x + r * MERGE_THREADS;
        if (e < na + nb) {
          const int src = e < na ? a0 + i0 + e : b0 + j0 + (e - na);
          kk[r] = in_key[src];
          pp[r] = in_pos[src];
        }
      }
#pragma unroll
      for (int r = 0; r < MERGE_E; ++r) {
        const int e = (int)threadIdx.x + r * MERGE_THREADS;
        if (e < na + nb) {
          s_key[e] = kk[r];
          s_pos[e] = pp[r];
        }
      }
    }
    __syncthreads();
    // each thread merges MERGE_E consecutive outputs into registers
    uint64_t k[MERGE_E];
    uint32_t p[MERGE_E];
    const int o_begin = threadIdx.x * MERGE_E;
    const int cnt = max(0, min(na + nb, o_begin + MERGE_E) - o_begin);
    if (cnt > 0) {
      int i = merge_split(s_key, s_pos, na, s_key + na, s_pos + na, nb, o_begin);
      int j = o_begin - i;
#pragma unroll
      for (int e = 0; e < MERGE_E; ++e) {
        if (e < cnt) {
          const bool take_a = j >= nb || (i < na && !pair_less(s_key[na + j], s_pos[na + j], s_key[i], s_pos[i]));
          if (take_a) {
            k[e] = s_key[i];
            p[e] = s_pos[i];
            ++i;
          } else {
            k[e] = s_key[na + j];
            p[e] = s_pos[na + j];
            ++j;
          }
        }
      }
    }
    __syncthreads();
    // back through LDS, so that the global stores are coalesced
#pragma unroll
    for (int e = 0; e < MERGE_E; ++e)
      if (e < cnt) {
        s_key[o_begin + e] = k[e];
        s_pos[o_begin + e] = p[e];
      }
    __syncthreads();
    if (order_out) {
      int g[MERGE_E];
#pragma unroll
      for (int r = 0; r < MERGE_E; ++r) {
        const int e = (int)threadIdx.x + r * MERGE_THREADS;
        g[r] = e < na + nb ? order_in[s_pos[e]] : 0;
      }
#pragma unroll
      for (int r = 0; r < MERGE_E; ++r) {
        const int e = (int)threadIdx.x + r * MERGE_THREADS;
        if (e < na + nb) order_out[q0 + e] = g[r];
      }
    } else {
      for (int e = threadIdx.x; e < na + nb; e += blockDim.x) {
        out_key[q0 + e] = s_key[e];
        out_pos[q0 + e] = s_pos[e];
      }
    }
    __syncthreads();  // the next work item reuses the LDS
  }
}

// After a level's partition (lists in lst): each run's members, in every node, in the order of their positions
// in `order`.  A run's members in one node's list are consecutive (equal keys), and nodes are ordered alike in
// `order` and in the lists, so the member with the k-th smallest position in `order` takes the k-th smallest of
// the run's list positions.  One wavefront per run (run `wave`, then every nwaves-th); positions are unique,
// so ranks are counts of smaller ones.
__device__ void tie_fix_runs(TreeArrays& t, int32_t* __restrict__ lst, int wave, int nwaves) {
  __shared__ int s_o[FIX_WAVES][RUN_MAX], s_p[FIX_WAVES][RUN_MAX], s_q[FIX_WAVES][RUN_MAX];
  const int runs = live_count(t.counters, C_RUNS);
  const int m = (int)t.m;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = wave; r < runs; r += nwaves) {
    const int off = t.run_off[r], len = t.run_len[r] >> 2, A = t.run_len[r] & 3;
    if (len > RUN_MAX || off < 0 || off + len > 3 * m) {  // cannot happen for a consistent scan; reported
      t.counters[C_OVERFLOW] = 1;
      continue;
    }
    const int32_t* __restrict__ seg = t.run_seg + off;
    int32_t* __restrict__ inv = t.invb + (size_t)A * m;
    for (int k = lane; k < len; k += 64) {
      const int s = seg[k];
      s_o[wv][k] = t.inv_o[s];
      s_p[wv][k] = inv[s];
    }
    wave_lds_sync();
    for (int k = lane; k < len; k += 64) {
      const int p = s_p[wv][k];
      int rank = 0;
      for (int j = 0; j < len; ++j) rank += s_p[wv][j] < p;
      s_q[wv][rank] = p;
    }
    wave_lds_sync();
    for (int k = lane; k < len; k += 64) {
      const int o = s_o[wv][k];
      int rank = 0;
      for (int j = 0; j < len; ++j) rank += s_o[wv][j] < o;
      const int dest = s_q[wv][rank];
      const int s = seg[k];
      if ((unsigned)dest < (unsigned)m) {
        lst[(size_t)A * m + dest] = s | TIE_FLAG;
        inv[s] = dest;
      } else {
        t.counters[C_OVERFLOW] = 1;
      }
    }
    wave_lds_sync();
  }
}

// One workgroup after the sort of level `lvl`: the children of its big nodes that split (vortex_tree.py:244-268)
// form level lvl + 1, which is planned at once (big children get tiles, small ones go to the bottom phase).
// A top-phase node at depth d holds its segments in the order buffer of d's parity (even: order, odd: tmp);
// the nodes that stop at an odd depth without going to the bottom phase (leaves, and big nodes that do not
// split) are listed for the final copy into `order`.
// With the list build (fix_lst: the lists just partitioned), workgroups 1, 2, ... put the runs of equal
// coordinates back in order (tie_fix_runs) while workgroup 0 plans.
__global__ __launch_bounds__(PLAN_THREADS) void k_level_next(TreeArrays t, int lvl, const int32_t* __restrict__ big_in,
                                                             int32_t* __restrict__ big_out,
                                                             int32_t* __restrict__ fix_lst) {
  if (blockIdx.x > 0) {
    if (fix_lst) tie_fix_runs(t, fix_lst, (blockIdx.x - 1) * FIX_WAVES + (threadIdx.x >> 6), (gridDim.x - 1) * FIX_WAVES);
    return;
  }
  const int nbig = live_count(t.counters, C_BIG);
  const int nodes0 = t.counters[C_NODES];
  const int bottom0 = t.counters[C_BOTTOM];
  const int strand0 = t.counters[C_STRAND];
  const bool odd = (lvl & 1) != 0;
  int made = 0, nnext = 0, ntiles = 0, nsmall = 0, nstrand = 0;
  for (int base = 0; base < nbig; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int node = i < nbig ? big_in[i] : -1;
    const bool sp = node >= 0 && t.split[node];
    int a = 0, b = 0, middle = 0;
    if (sp) {
      a = t.lo[node];
      b = t.hi[node];
      middle = (int)(((int64_t)a + b) / 2);
    }
    const int size[2] = {middle - a, b - middle};
    bool big[2], small[2], leaf[2];
    int tiles[2];
    for (int q = 0; q < 2; ++q) {
      big[q] = sp && is_big(size[q], t.leaf);
      small[q] = sp && is_small(size[q], t.leaf);
      leaf[q] = sp && !big[q] && !small[q];
      tiles[q] = big[q] ? (size[q] + TILE - 1) / TILE : 0;
    }
    const bool strand_self = node >= 0 && !sp && odd;           // stays where level lvl left it
    const int strands = (int)strand_self + (odd ? 0 : (int)leaf[0] + (int)leaf[1]);  // children: depth lvl + 1
    const int sv[5] = {sp ? 1 : 0, (int)big[0] + (int)big[1], tiles[0] + tiles[1], (int)small[0] + (int)small[1],
                       strands};
    int sr[5], stot[5];
    block_exclusive_scan_n<5>(sv, sr, stot);
    const int rm = sr[0], rb = sr[1], rt = sr[2], rs = sr[3], rst = sr[4];
    const int tm = stot[0], tb = stot[1], tt = stot[2], ts = stot[3], tst = stot[4];
    int st = strand0 + nstrand + rst;
    if (strand_self) {
      if (st < t.cap) t.strand[st] = node;
      else t.counters[C_OVERFLOW] = 1;
      ++st;
    }
    if (sp) {
      const int first = nodes0 + 2 * (made + rm);
      if (first + 1 < t.cap) {
        const int d = t.depth[node] + 1, tr = t.tree[node];
        int bi = nnext + rb, t0 = ntiles + rt, sb = bottom0 + nsmall + rs;
        for (int q = 0; q < 2; ++q) {
          const int child = first + q;
          const int clo = q ? middle : a, chi = q ? b : middle;
          t.lo[child] = clo;
          t.hi[child] = chi;
          t.left[child] = -1;
          t.right[child] = -1;
          t.parent[child] = node;
          t.depth[child] = d;
          t.tree[child] = tr;
          if (big[q]) {
            if (bi < t.cap && t0 + tiles[q] <= t.tile_cap) {
              big_out[bi] = child;
              for (int k = 0; k < tiles[q]; ++k) {
                t.tile_node[t0 + k] = child;
                t.tile_p0[t0 + k] = clo + k * TILE;
                t.tile_p1[t0 + k] = min(chi, clo + (k + 1) * TILE);
              }
            } else {
              t.counters[C_OVERFLOW] = 1;
            }
            ++bi;
            t0 += tiles[q];
          }
          if (small[q]) {
            if (sb < t.bottom_cap) t.bottom[sb] = child;
            else t.counters[C_OVERFLOW] = 1;
            ++sb;
          }
          if (leaf[q] && !odd) {  // a leaf at odd depth lvl + 1
            if (st < t.cap) t.strand[st] = child;
            else t.counters[C_OVERFLOW] = 1;
            ++st;
          }
        }
        t.left[node] = first;
        t.right[node] = first + 1;
      } else {
        t.counters[C_OVERFLOW] = 1;
      }
    }
    made += tm;
    nnext += tb;
    ntiles += tt;
    nsmall += ts;
    nstrand += tst;
  }
  __syncthreads();  // every thread has read the counters of this level
  if (threadIdx.x == 0) {
    t.counters[C_NODES] = nodes0 + 2 * made;
    t.counters[C_BIG] = nnext;
    t.counters[C_TILES] = ntiles;
    t.counters[C_BOTTOM] = bottom0 + nsmall;
    t.counters[C_STRAND] = strand0 + nstrand;
  }
}

// The listed nodes' segments, left in tmp (odd depth), move to `order`.
__global__ __launch_bounds__(256) void k_copy_strands(TreeArrays t) {
  const int count = live_count(t.counters, C_STRAND);
  for (int k = blockIdx.x; k < count; k += gridDim.x) {
    const int node = t.strand[k];
    for (int p = t.lo[node] + (int)threadIdx.x; p < t.hi[node]; p += blockDim.x) t.order[p] = t.tmp[p];
  }
}
// ------------------------------------------------------------------ list build (top phase)
// A stable sort of a node by one axis (vortex_tree.py:95-101) orders its segments by (coordinate, position in
// the node's order).  The top phase keeps, for every big node and axis, the node's segments in that order (the
// node's list of the axis), so that no level sorts anything:
//  * the node's box is the first and last entry of its lists (exact min / max), which give the axis and flag;
//  * its sorted order is its list of that axis: the order entries and each segment's side ((a+b)//2 cut) are
//    written;
//  * each list is partitioned, stably, into the children's ranges [a, mid) and [mid, b);
//  * a partitioned list is ordered by (coordinate, position in the parent's order), while the child's order is
//    the parent's list of the split axis: in a run of equal coordinates the entries are put in the order of
//    their new positions in `order`.  The segments of such runs carry TIE_FLAG in the lists, and their positions
//    in `order` (inv_o) and in each list (invb) are kept per segment.
// The roots' lists are each tree's segments sorted by (coordinate, index), its order at the root: an LSD radix
// sort of (tree id, 64-bit key) by 8-bit digits, three kernels per digit (the digit counts of every tile, their
// scan, the stable scatter of every tile).  Integer counts only: the result is the same for every schedule.  A
// run of equal coordinates longer than RUN_MAX sets C_DUP (and the overflow flag, so that nothing runs after it):
// the call is then made again with the per-level sorts.
static constexpr int RS_THREADS = 256;
static constexpr int RS_E = 8;                       // elements per thread
static constexpr int RS_TILE = RS_THREADS * RS_E;    // elements per radix tile
static constexpr int RS_WAVE = RS_TILE / (RS_THREADS / 64);  // elements per wavefront of a tile
static constexpr int RS_BINS = 256;
static constexpr int RS_KEY_PASSES = 4;
static constexpr int RS_MAX_PASSES = RS_KEY_PASSES + 3;
static constexpr int FIX_RUN = 64;  // longest run of equal 32-bit keys the presort orders by its doubles
static_assert(RS_WAVE % 64 == 0, "a wavefront ranks whole rounds of 64 elements");
static_assert(RS_THREADS == RS_BINS, "one thread per digit");

// tree of segment position s: the last k with roots_lo[k] <= s
__device__ __forceinline__ int tree_of(const int32_t* __restrict__ roots_lo, int K, int s) {
  int lo = 0, hi = K;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (roots_lo[mid] <= s) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int rs_digit(uint32_t key, int tree, int pass) {
  return pass < RS_KEY_PASSES ? (int)((key >> (8 * pass)) & 0xff) : ((tree >> (8 * (pass - RS_KEY_PASSES))) & 0xff);
}

// Keys and segment ids of the three axes (buffer 0); no segment is in a run yet.  The identity order, and (the
// first workgroup) the roots, counters and scan words: the list build needs no launch of its own for them.
__global__ void k_rs_init(TreeArrays t, const double* __restrict__ mx, const double* __restrict__ my,
                          const double* __restrict__ mz) {
  const int m = (int)t.m, A = blockIdx.y;
  if (blockIdx.x == 0 && blockIdx.y == 0) init_roots_block(t);
  const double* __restrict__ coord = A == 0 ? mx : (A == 1 ? my : mz);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const size_t g = (size_t)A * m + i;
    t.rkey[0][g] = ordered_key32(coord[i]);
    t.rval[0][g] = i;
    if (A == 0) {
      t.order[i] = i;
      t.tie[i] = 0;
      t.ord_inv[i] = i;
      // the finish's histogram over lo (m + 1 entries)
      t.lo_cnt[i] = 0;
      t.lo_mind[i] = INT_MAX;
      if (i == 0) {
        t.lo_cnt[m] = 0;
        t.lo_mind[m] = INT_MAX;
      }
    }
  }
}

// Digit counts of each tile of buffer src for digit `pass`, digit-major per axis: rcount[A][d][tile].
__global__ __launch_bounds__(RS_THREADS) void k_rs_count(TreeArrays t, int pass, int src) {
  __shared__ int h[RS_BINS];
  const int m = (int)t.m, A = blockIdx.y, tile = blockIdx.x;
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t* __restrict__ kin = t.rkey[src] + (size_t)A * m;
  const int32_t* __restrict__ vin = t.rval[src] + (size_t)A * m;
  const int i0 = tile * RS_TILE;
#pragma unroll 4
  for (int e = 0; e < RS_E; ++e) {
    const int i = i0 + e * RS_THREADS + (int)threadIdx.x;
    if (i < m) {
      const int tr = pass >= RS_KEY_PASSES ? tree_of(t.roots_lo, t.K, vin[i]) : 0;
      atomicAdd(&h[rs_digit(kin[i], tr, pass)], 1);
    }
  }
  __syncthreads();
  t.rcount[((size_t)A * RS_BINS + threadIdx.x) * t.rs_tiles + tile] = h[threadIdx.x];
}

// Per axis and digit (one wavefront each): the exclusive scan of the digit's tile counts, and its total.
__global__ __launch_bounds__(256) void k_rs_scan(TreeArrays t) {
  const int gw = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int A = gw / RS_BINS, d = gw % RS_BINS;  // the grid is 3 * RS_BINS / 4 workgroups
  const int T = t.rs_tiles;
  int32_t* __restrict__ x = t.rcount + ((size_t)A * RS_BINS + d) * T;
  int run = 0;
  for (int t0 = 0; t0 < T; t0 += 256) {
    int v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int q = t0 + c * 64 + lane;
      v[c] = q < T ? x[q] : 0;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int incl = v[c];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
      }
      const int q = t0 + c * 64 + lane;
      if (q < T) x[q] = run + incl - v[c];
      run += __shfl(incl, 63, 64);
    }
  }
  if (lane == 0) t.rtotal[A * RS_BINS + d] = run;
}

// One stable pass of the presort on digit `pass`, from buffer src to 1 - src: each tile ranks its elements
// (per-wavefront running counts, peers found by ballots on the digit's bits), orders them by digit in LDS and
// writes each digit's run to consecutive addresses from its base (the axis's block, the digits before it, the
// tiles before this one).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(TreeArrays t, int pass, int src) {
  __shared__ int s_cnt[RS_THREADS / 64][RS_BINS];  // running counts of each wavefront, then its digit bases
  __shared__ int s_base[RS_BINS];                  // destination of the digit's first element of the tile
  __shared__ int s_start[RS_BINS];                 // the digit's first place in the tile's digit order
  __shared__ uint32_t s_key[RS_TILE];
  __shared__ int32_t s_val[RS_TILE];
  __shared__ uint8_t s_dig[RS_TILE];
  const int m = (int)t.m, A = blockIdx.y, tile = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = threadIdx.x; q < (RS_THREADS / 64) * RS_BINS; q += blockDim.x) (&s_cnt[0][0])[q] = 0;
  {
    const int d = threadIdx.x;
    int total;
    const int below = block_exclusive_scan(t.rtotal[A * RS_BINS + d], &total);
    s_base[d] = (int)((int64_t)A * m) + below + t.rcount[((size_t)A * RS_BINS + d) * t.rs_tiles + tile];
  }
  __syncthreads();
  const uint32_t* __restrict__ kin = t.rkey[src] + (size_t)A * m;
  const int32_t* __restrict__ vin = t.rval[src] + (size_t)A * m;
  uint32_t key[RS_E];
  int32_t val[RS_E];
  int dig[RS_E], rank[RS_E];
  const int w0 = tile * RS_TILE + w * RS_WAVE;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < RS_E; ++r) {
    const int i = w0 + r * 64 + lane;
    const bool act = i < m;
    key[r] = act ? kin[i] : 0u;
    val[r] = act ? vin[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < RS_E; ++r) {
    const int i = w0 + r * 64 + lane;
    const bool act = i < m;
    const int tr = (pass >= RS_KEY_PASSES && act) ? tree_of(t.roots_lo, t.K, val[r]) : 0;
    const int d = rs_digit(key[r], tr, pass);
    // the lanes holding the same digit: AND over the digit's bits of the matching ballots
    unsigned long long peers = __builtin_amdgcn_ballot_w64(act);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long on = __builtin_amdgcn_ballot_w64(((d >> b) & 1) != 0);
      peers &= ((d >> b) & 1) ? on : ~on;
    }
    const int before = s_cnt[w][d];
    wave_lds_sync();
    dig[r] = d;
    rank[r] = before + __popcll(peers & below);
    const bool leader = act && (peers & below) == 0ull;
    if (leader) s_cnt[w][d] = before + __popcll(peers);
    wave_lds_sync();
  }
  __syncthreads();
  // digit d: the wavefronts' bases within the digit, and the digit's start in the tile's digit order
  {
    const int d = threadIdx.x;
    int tot = 0;
#pragma unroll
    for (int v = 0; v < RS_THREADS / 64; ++v) {
      const int c = s_cnt[v][d];
      s_cnt[v][d] = tot;
      tot += c;
    }
    int total;
    s_start[d] = block_exclusive_scan(tot, &total);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_E; ++r) {
    const int i = w0 + r * 64 + lane;
    if (i < m) {
      const int lp = s_start[dig[r]] + s_cnt[w][dig[r]] + rank[r];
      if ((unsigned)lp < (unsigned)RS_TILE) {
        s_key[lp] = key[r];
        s_val[lp] = val[r];
        s_dig[lp] = (uint8_t)dig[r];
      }
    }
  }
  __syncthreads();
  uint32_t* __restrict__ kout = t.rkey[1 - src];
  int32_t* __restrict__ vout = t.rval[1 - src];
  const int cnt = min(RS_TILE, m - tile * RS_TILE);
  for (int j = threadIdx.x; j < cnt; j += RS_THREADS) {
    const int d = s_dig[j];
    const int dest = s_base[d] + (j - s_start[d]);
    if ((unsigned)(dest - A * m) < (unsigned)m) {
      kout[dest] = s_key[j];
      vout[dest] = s_val[j];
    } else {
      t.counters[C_OVERFLOW] = 1;  // cannot happen for a consistent count; reported, never written past
    }
  }
}

// Positions i < j of a sorted list hold the same key of the same tree.
__device__ __forceinline__ bool same_tree(const TreeArrays& t, int i, int j) {
  return t.K == 1 || tree_of(t.roots_lo, t.K, i) == tree_of(t.roots_lo, t.K, j);
}

// The presort orders by the 32-bit keys of the coordinates rounded to float32; within a run of equal keys (one
// tree) the entries are put in the order of (coordinate, segment id), the order of a stable sort by the
// coordinate.  Values from buffer buf to buf ^ 1 (the keys stay: the run's keys are equal).  A run longer than
// FIX_RUN stops the list build.
__global__ void k_rs_fix(TreeArrays t, int buf, const double* __restrict__ mx, const double* __restrict__ my,
                         const double* __restrict__ mz) {
  const int m = (int)t.m, A = blockIdx.y;
  const double* __restrict__ coord = A == 0 ? mx : (A == 1 ? my : mz);
  const uint32_t* __restrict__ key = t.rkey[buf] + (size_t)A * m;
  const int32_t* __restrict__ in = t.rval[buf] + (size_t)A * m;
  int32_t* __restrict__ out = t.rval[buf ^ 1] + (size_t)A * m;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const uint32_t k = key[i];
    const int s = in[i];
    int rs = i, re = i + 1;
    while (rs > 0 && i - rs <= FIX_RUN && key[rs - 1] == k && same_tree(t, rs - 1, i)) --rs;
    while (re < m && re - i <= FIX_RUN && key[re] == k && same_tree(t, re, i)) ++re;
    if (re - rs == 1) {
      out[i] = s;
      continue;
    }
    if (re - rs > FIX_RUN) {
      out[i] = s;
      t.counters[C_DUP] = 1;
      t.counters[C_OVERFLOW] = 1;
      continue;
    }
    const double c = coord[s];
    int rank = 0;
    for (int j = rs; j < re; ++j) {
      const int sj = in[j];
      const double cj = coord[sj];
      rank += (cj < c || (cj == c && sj < s)) ? 1 : 0;
    }
    out[rs + rank] = s;
  }
}

// Positions i < j of a sorted list hold the same coordinate of the same tree.
__device__ __forceinline__ bool same_run(const TreeArrays& t, const uint32_t* __restrict__ key,
                                         const int32_t* __restrict__ val, const double* __restrict__ coord, int i,
                                         int j) {
  return key[i] == key[j] && coord[val[i]] == coord[val[j]] && same_tree(t, i, j);
}

// The runs of equal coordinates of the roots' sorted lists (keys in kbuf, lists in buf): each run's members, axis
// and length are
// recorded (in any order: runs are disjoint, so the re-order does not depend on it), its segments marked, and
// their positions in `order` (the identity at the roots) and in the axis's list set.  A run longer than RUN_MAX
// stops the list build.
__global__ void k_tie_scan(TreeArrays t, int kbuf, int buf, const double* __restrict__ mx,
                           const double* __restrict__ my, const double* __restrict__ mz) {
  const int m = (int)t.m, A = blockIdx.y;
  const double* __restrict__ coord = A == 0 ? mx : (A == 1 ? my : mz);
  const uint32_t* __restrict__ key = t.rkey[kbuf] + (size_t)A * m;
  const int32_t* __restrict__ val = t.rval[buf] + (size_t)A * m;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < m; i += gridDim.x * blockDim.x) {
    if (!same_run(t, key, val, coord, i, i + 1) || (i > 0 && same_run(t, key, val, coord, i - 1, i)))
      continue;  // first entries only
    int len = 2;
    while (i + len < m && len <= RUN_MAX && same_run(t, key, val, coord, i, i + len)) ++len;
    if (len > RUN_MAX) {
      t.counters[C_DUP] = 1;
      t.counters[C_OVERFLOW] = 1;
      continue;
    }
    const int r = atomicAdd(&t.counters[C_RUNS], 1);
    const int off = atomicAdd(&t.counters[C_RUNMEM], len);
    t.run_off[r] = off;
    t.run_len[r] = (len << 2) | A;
    for (int k = 0; k < len; ++k) {
      const int s = val[i + k];
      t.run_seg[off + k] = s;
      t.tie[s] = 1;
      t.inv_o[s] = s;
      t.invb[(size_t)A * m + s] = i + k;
    }
  }
}

// The flag of every list entry whose segment is in a run of any axis.
__global__ void k_tie_mark(TreeArrays t, int buf) {
  if (live_count(t.counters, C_RUNS) == 0) return;
  const int m = (int)t.m, A = blockIdx.y;
  int32_t* __restrict__ val = t.rval[buf] + (size_t)A * m;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int s = val[i];
    if (t.tie[s]) val[i] = s | TIE_FLAG;
  }
}

// Box, axis and split flag of a big node from its lists' first and last entries.
__device__ __forceinline__ void list_axis(const TreeArrays& t, const int32_t* __restrict__ lst, int lo, int hi,
                                          const double* __restrict__ mx, const double* __restrict__ my,
                                          const double* __restrict__ mz, int8_t* axis, int8_t* split) {
  const int m = (int)t.m;
  double v[6];
  const double* cs[3] = {mx, my, mz};
#pragma unroll
  for (int A = 0; A < 3; ++A) {
    const int a = lst[(size_t)A * m + lo] & SEG_MASK, b = lst[(size_t)A * m + hi - 1] & SEG_MASK;
    v[A] = cs[A][(unsigned)a < (unsigned)m ? a : 0];
    v[3 + A] = cs[A][(unsigned)b < (unsigned)m ? b : 0];
  }
  choose_axis(v, axis, split);
}

// One top level of the list build in one launch: per work unit (a 4,096-entry tile of one node and one axis),
// the side of each of the tile's list entries, its rank among the tile's left entries, the left entries of the
// node's earlier tiles (single-pass look-back over the node's units of that axis), and the stable partition of
// the list into the children's ranges [lo, mid) and [mid, hi); the units of the split axis write the node's
// sorted order.  The node's sorted order (vortex_tree.py:95-101) is its list of the split axis, cut at
// (a+b)//2: an entry goes left when its coordinate on that axis is below the median entry's, or equal to it
// and before it in that list (then it is in a run of equal coordinates, whose list positions are kept).  The
// node's first unit records its axis and split flag for the next plan.
static constexpr int LL_THREADS = 256;
static constexpr int LL_E = TILE / LL_THREADS;  // entries per thread (strided by LL_THREADS)
static_assert(LL_E * (LL_THREADS / 64) == 64, "the per-(row, wavefront) left counts are scanned by one wavefront");
__global__ __launch_bounds__(LL_THREADS) void k_list_level(TreeArrays t, int lvl, const int32_t* __restrict__ lst_in,
                                                           int32_t* __restrict__ lst_out,
                                                           const double* __restrict__ mx,
                                                           const double* __restrict__ my,
                                                           const double* __restrict__ mz) {
  __shared__ int s_unit;
  __shared__ int s_cell[64];  // left entries of each (row, wavefront), then their exclusive scan
  __shared__ int s_prefix;
  const int ntiles = live_count(t.counters, C_TILES);
  const int units = 3 * ntiles;
  const int m = (int)t.m;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned tag = (unsigned)lvl + 1u;
  const double* cs[3] = {mx, my, mz};
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(&t.tickets[lvl], 1);
    __syncthreads();
    const int u = s_unit;
    __syncthreads();  // every thread has read the ticket before the next one is drawn
    if (u >= units) break;
    const int tile = u / 3, A = u - 3 * tile;
    const int node = t.tile_node[tile];
    const int p0 = t.tile_p0[tile], p1 = t.tile_p1[tile];
    const int32_t* __restrict__ src = lst_in + (size_t)A * m;
    int sv[LL_E];  // the tile's entries (loaded beside the node's)
#pragma unroll
    for (int e = 0; e < LL_E; ++e) {
      const int p = p0 + (int)threadIdx.x + e * LL_THREADS;
      sv[e] = p < p1 ? src[p] : 0;
    }
    const int lo = t.lo[node], hi = t.hi[node];
    const int mid = (int)(((int64_t)lo + hi) / 2);
    int8_t axis, split;
    list_axis(t, lst_in, lo, hi, mx, my, mz, &axis, &split);
    if (p0 == lo && A == 0 && threadIdx.x == 0) {
      t.axis[node] = axis;
      t.split[node] = split;
    }
    if (!split) continue;  // uniform over the workgroup
    const int s_med = lst_in[(size_t)axis * m + mid] & SEG_MASK;
    const double* __restrict__ ca = cs[axis];
    const double c_med = ca[s_med < m ? s_med : 0];
    const int32_t* __restrict__ run_pos = t.invb + (size_t)axis * m;
    int f[LL_E];
    if (A == axis) {
#pragma unroll
      for (int e = 0; e < LL_E; ++e) {
        const int p = p0 + (int)threadIdx.x + e * LL_THREADS;
        f[e] = p < p1 && p < mid ? 1 : 0;
      }
    } else {
      double x[LL_E];
#pragma unroll
      for (int e = 0; e < LL_E; ++e) {
        const int q = sv[e] & SEG_MASK;
        const int p = p0 + (int)threadIdx.x + e * LL_THREADS;
        x[e] = p < p1 && q < m ? ca[q] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < LL_E; ++e) {
        const int q = sv[e] & SEG_MASK;
        const int p = p0 + (int)threadIdx.x + e * LL_THREADS;
        f[e] = (p < p1 && q < m &&
                (x[e] < c_med || (x[e] == c_med && q != s_med && run_pos[q] < mid))) ? 1 : 0;
      }
    }
    // ranks among the tile's left entries: row e of wavefront w covers positions p0 + 256 e + 64 w + lane, so
    // the (row, wavefront) cells in row-major order are in position order
    int rank[LL_E];
#pragma unroll
    for (int e = 0; e < LL_E; ++e) {
      const unsigned long long mask = __builtin_amdgcn_ballot_w64(f[e] != 0);
      rank[e] = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      if (lane == 0) s_cell[e * (LL_THREADS / 64) + wv] = __popcll(mask);
    }
    __syncthreads();
    if (wv == 0) {
      const int v = s_cell[lane];
      int incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      s_cell[lane] = incl - v;
      const int total = __builtin_amdgcn_readlane(incl, 63);
      // the node's tiles are consecutive; its first publishes its inclusive count at once
      const int first = tile - (p0 - lo) / TILE;
      unsigned long long* st = t.lstat + A;
      int prefix = 0;
      if (tile == first) {
        if (lane == 0) lb_publish(st + 3 * (size_t)tile, tag, true, total);
      } else {
        if (lane == 0) lb_publish(st + 3 * (size_t)tile, tag, false, total);
        bool ok = true;
        prefix = lb_lookback(st, 3, first, tile, tag, &ok);
        if (lane == 0) {
          if (!ok) t.counters[C_OVERFLOW] = 1;
          lb_publish(st + 3 * (size_t)tile, tag, true, prefix + total);
        }
      }
      if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const int prefix = s_prefix;
    const bool small_l = mid - lo <= S_BOTTOM, small_r = hi - mid <= S_BOTTOM;  // children that leave the lists
    int32_t* __restrict__ dst = lst_out + (size_t)A * m;
    int32_t* __restrict__ inv = t.invb + (size_t)A * m;
#pragma unroll
    for (int e = 0; e < LL_E; ++e) {
      const int p = p0 + (int)threadIdx.x + e * LL_THREADS;
      if (p < p1) {
        const int left = prefix + s_cell[e * (LL_THREADS / 64) + wv] + rank[e];
        const int dest = f[e] ? lo + left : mid + (p - lo - left);
        const int q = sv[e] & SEG_MASK;
        if (dest >= lo && dest < hi && q < m) {
          dst[dest] = sv[e];
          // the split axis's list keeps every position (dest == p), so its run positions, which the other
          // axes' units read for their sides, are not written
          if (sv[e] < 0 && A != axis) inv[q] = dest;
          if (A == axis) {  // the node's sorted order
            t.order[p] = q;
            if (sv[e] < 0) t.inv_o[q] = p;
            if (p < mid ? small_l : small_r) t.ord_inv[q] = p;
          }
        } else {
          t.counters[C_OVERFLOW] = 1;
        }
      }
    }
    __syncthreads();  // s_cell and s_prefix are read before the next unit writes them
  }
}

// ------------------------------------------------------------------ bottom phase
__global__ __launch_bounds__(PLAN_THREADS) void k_bottom_bases(TreeArrays t) {
  const int count = live_count(t.counters, C_BOTTOM);
  const int nodes0 = t.counters[C_NODES];
  int acc = 0;
  for (int base = 0; base < count; base += blockDim.x) {
    const int b = base + threadIdx.x;
    int slots = 0;
    if (b < count) {
      const int node = t.bottom[b];
      slots = heap_slots(t.hi[node] - t.lo[node], t.leaf) - 1;  // the heap root is the node itself
    }
    int total;
    const int r = block_exclusive_scan(slots, &total);
    if (b < count) t.heap_base[b] = nodes0 + acc + r;
    acc += total;
  }
  if (threadIdx.x == 0) {
    if ((int64_t)nodes0 + acc > t.cap) t.counters[C_OVERFLOW] = 1;
    t.counters[C_HEAP] = acc;
    t.counters[C_NODES] = nodes0 + acc;
  }
}

// Sort of every splitting node of heap level k (slots [first, last)) by (key, position), in place in LDS: node
// first + g on group g of blockDim >> k threads, E elements per thread (a sorting network in registers), then
// merge rounds of runs E, 2E, ... within the node (a merge-path search per thread) while a run is shorter than
// `maxlen`, the largest node of the level.  Other nodes' ranges are left as they are.  Every thread calls it.
template <int E>
__device__ void sort_heap_level(uint64_t* key, uint32_t* tag, const uint16_t* hlo, const uint16_t* hhi,
                                const uint8_t* hstate, int k, int first, int last, int maxlen) {
  const int T = max(1, (int)blockDim.x >> k);  // a splitting node holds at most E * T elements
  const int g = threadIdx.x / T, r = threadIdx.x - g * T;
  const int j = first + g;
  int lo = 0, len = 0;
  if (j < last && hstate[j] == 2) {
    lo = hlo[j];
    len = hhi[j] - lo;
  }
  uint64_t* kn = key + lo;
  uint32_t* tn = tag + lo;
  const int o = r * E;
  const int cnt = max(0, min(E, len - o));
  uint64_t kv[E];
  uint32_t tv[E];
  if (cnt > 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      kv[e] = e < cnt ? kn[o + e] : ~0ull;
      tv[e] = e < cnt ? tn[o + e] : 0xffffffffu;
    }
#pragma unroll
    for (int size = 2; size <= E; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int q = i ^ stride;
          if (q > i) {
            const bool up = (i & size) == 0;
            const bool swap = up ? pair_less(kv[q], tv[q], kv[i], tv[i]) : pair_less(kv[i], tv[i], kv[q], tv[q]);
            if (swap) {
              const uint64_t a = kv[i];
              kv[i] = kv[q];
              kv[q] = a;
              const uint32_t b = tv[i];
              tv[i] = tv[q];
              tv[q] = b;
            }
          }
        }
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (e < cnt) {
        kn[o + e] = kv[e];
        tn[o + e] = tv[e];
      }
  }
  for (int w = E; w < maxlen; w <<= 1) {
    // the threads of a merge (a node's group, or the 2w / E threads of two runs) inside one wavefront read
    // and write that wavefront's elements only: their LDS order needs no workgroup barrier
    const bool local = T <= 64 || 2 * w <= 64 * E;
    if (local) wave_lds_sync();
    else __syncthreads();
    if (cnt > 0) {
      const int pair0 = o & ~(2 * w - 1);
      const int a1 = min(pair0 + w, len), b1 = min(pair0 + 2 * w, len);
      const int la = a1 - pair0, lb = b1 - a1, diag = o - pair0;
      const uint64_t* ka = kn + pair0;
      const uint32_t* ta = tn + pair0;
      const uint64_t* kb = kn + a1;
      const uint32_t* tb = tn + a1;
      int lo_s = max(0, diag - lb), hi_s = min(diag, la);
      while (lo_s < hi_s) {
        const int mid = (lo_s + hi_s) >> 1;
        const int q = diag - mid - 1;
        if (pair_less(kb[q], tb[q], ka[mid], ta[mid])) hi_s = mid;
        else lo_s = mid + 1;
      }
      int ia = lo_s, ib = diag - lo_s;
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (e < cnt) {
          const bool take_a = ib >= lb || (ia < la && !pair_less(kb[ib], tb[ib], ka[ia], ta[ia]));
          if (take_a) {
            kv[e] = ka[ia];
            tv[e] = ta[ia];
            ++ia;
          } else {
            kv[e] = kb[ib];
            tv[e] = tb[ib];
            ++ib;
          }
        }
    }
    if (local) wave_lds_sync();
    else __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (e < cnt) {
        kn[o + e] = kv[e];
        tn[o + e] = tv[e];
      }
  }
  __syncthreads();
}

// Node-table entries of heap slot j of bottom subtree b (root `root` at segment a, heap nodes from hb on).
__device__ __forceinline__ void heap_node_out(TreeArrays& t, int root, int hb, int a, int d0, int tr, int j, bool exists,
                                              bool sp, int lo_j, int hi_j) {
  const int gid = j == 0 ? root : hb + j - 1;
  if (j > 0) {
    t.lo[gid] = exists ? a + lo_j : -1;
    t.hi[gid] = exists ? a + hi_j : -1;
    const int pj = (j - 1) / 2;
    t.parent[gid] = pj == 0 ? root : hb + pj - 1;
    t.tree[gid] = tr;
    int d = 0;
    for (int q = j; q > 0; q = (q - 1) / 2) ++d;
    t.depth[gid] = d0 + d;
  }
  t.left[gid] = sp ? hb + 2 * j : -1;       // heap slot 2j+1 -> hb + 2j
  t.right[gid] = sp ? hb + 2 * j + 1 : -1;  // heap slot 2j+2 -> hb + 2j+1
}

// One workgroup per subtree of at most S_BOTTOM segments: every remaining level in LDS (the sort build).
__global__ __launch_bounds__(BOTTOM_THREADS) void k_bottom(TreeArrays t, const double* __restrict__ mx,
                                                           const double* __restrict__ my,
                                                           const double* __restrict__ mz) {
  __shared__ double cx[S_BOTTOM], cy[S_BOTTOM], cz[S_BOTTOM];
  __shared__ int32_t seg[S_BOTTOM], itmp[S_BOTTOM];
  __shared__ uint64_t key[S_BOTTOM];
  __shared__ uint32_t tag[S_BOTTOM];
  __shared__ uint16_t nodeof[S_BOTTOM];
  __shared__ uint16_t hlo[HCAP], hhi[HCAP];
  __shared__ uint8_t hstate[HCAP], haxis[HCAP];  // state: 0 absent, 1 leaf so far, 2 split
  __shared__ int any_split;
  const int count = live_count(t.counters, C_BOTTOM);
  const int nwaves = blockDim.x / 64, wv = threadIdx.x / 64, lane = threadIdx.x & 63;
  for (int b = blockIdx.x; b < count; b += gridDim.x) {
    const int root = t.bottom[b];
    const int a = t.lo[root];
    const int s = t.hi[root] - a;
    const int slots = heap_slots(s, t.leaf);
    // the buffer of the root's depth parity
    const int32_t* src = (t.depth[root] & 1) ? t.tmp : t.order;
    for (int i = threadIdx.x; i < s; i += blockDim.x) {
      const int g = src[a + i];
      seg[i] = g;
      cx[i] = mx[g];
      cy[i] = my[g];
      cz[i] = mz[g];
      nodeof[i] = 0;
    }
    for (int j = threadIdx.x; j < slots; j += blockDim.x) {
      hstate[j] = j == 0 ? 1 : 0;
      hlo[j] = 0;
      hhi[j] = j == 0 ? (uint16_t)s : 0;
      haxis[j] = 0;
    }
    __syncthreads();
    for (int k = 0;; ++k) {
      const int first = (1 << k) - 1;
      if (first >= slots) break;
      const int last = min(slots, (2 << k) - 1);
      if (threadIdx.x == 0) any_split = 0;
      __syncthreads();
      // boxes: one wavefront per node of this level
      for (int j = first + wv; j < last; j += nwaves) {
        if (hstate[j] == 0) continue;
        const int lo = hlo[j], hi = hhi[j];
        if (hi - lo <= t.leaf) continue;
        double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int i = lo + lane; i < hi; i += 64) {
          const double c[3] = {cx[i], cy[i], cz[i]};
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            v[q] = c[q] < v[q] ? c[q] : v[q];
            v[3 + q] = c[q] > v[3 + q] ? c[q] : v[3 + q];
          }
        }
        box_wave_reduce(v);
        if (lane == 0) {
          int8_t ax, sp;
          choose_axis(v, &ax, &sp);
          haxis[j] = (uint8_t)ax;
          if (sp) {
            hstate[j] = 2;
            any_split = 1;
          }
        }
      }
      __syncthreads();
      if (!any_split) break;
      // sort every splitting node's range by (key, position); other nodes keep their order
      for (int i = threadIdx.x; i < s; i += blockDim.x) {
        const int j = nodeof[i];
        uint64_t kk = 0;
        if (hstate[j] == 2) {
          const int ax = haxis[j];
          kk = ordered_key(ax == 0 ? cx[i] : (ax == 1 ? cy[i] : cz[i]));
        }
        key[i] = kk;
        tag[i] = (uint32_t)i;
      }
      __syncthreads();
      int maxlen = s;  // the largest node of heap level k
      for (int q = 0; q < k; ++q) maxlen = (maxlen + 1) / 2;
      sort_heap_level<BOTTOM_E>(key, tag, hlo, hhi, hstate, k, first, last, maxlen);
      // permute segments and midpoints by the sorted positions (one array at a time through a scratch)
      double* dtmp = reinterpret_cast<double*>(key);
      for (int i = threadIdx.x; i < s; i += blockDim.x) itmp[i] = seg[tag[i]];
      __syncthreads();
      for (int i = threadIdx.x; i < s; i += blockDim.x) seg[i] = itmp[i];
      __syncthreads();
      double* arrays[3] = {cx, cy, cz};
      for (int q = 0; q < 3; ++q) {
        double* arr = arrays[q];
        for (int i = threadIdx.x; i < s; i += blockDim.x) dtmp[i] = arr[tag[i]];
        __syncthreads();
        for (int i = threadIdx.x; i < s; i += blockDim.x) arr[i] = dtmp[i];
        __syncthreads();
      }
      // children of the splitting nodes, then every element's node for the next level
      for (int j = first + threadIdx.x; j < last; j += blockDim.x) {
        if (hstate[j] != 2) continue;
        const int lo = hlo[j], hi = hhi[j], middle = (lo + hi) / 2;
        hstate[2 * j + 1] = 1;
        hlo[2 * j + 1] = (uint16_t)lo;
        hhi[2 * j + 1] = (uint16_t)middle;
        hstate[2 * j + 2] = 1;
        hlo[2 * j + 2] = (uint16_t)middle;
        hhi[2 * j + 2] = (uint16_t)hi;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < s; i += blockDim.x) {
        const int j = nodeof[i];
        if (hstate[j] == 2) nodeof[i] = (uint16_t)(i < hhi[2 * j + 1] ? 2 * j + 1 : 2 * j + 2);
      }
      __syncthreads();
    }
    // results
    for (int i = threadIdx.x; i < s; i += blockDim.x) t.order[a + i] = seg[i];
    const int hb = t.heap_base[b];
    const int d0 = t.depth[root], tr = t.tree[root];
    for (int j = threadIdx.x; j < slots; j += blockDim.x)
      heap_node_out(t, root, hb, a, d0, tr, j, hstate[j] != 0, hstate[j] == 2, hlo[j], hhi[j]);
    __syncthreads();
  }
}

// The list build's bottom phase: one workgroup per subtree of at most S_BOTTOM segments.  Local ids are the
// positions in the root's order; every level in LDS splits each node's three lists by a stable partition (no
// sorting): the node's box is the first and last entry of its lists, its sides and its children's order come from
// its list of the split axis, and the order of a leaf is its parent's list of that axis restricted to the leaf.
// When the subtree holds segments of runs of equal coordinates, each partitioned run is put back in the order of
// the new positions (ranks counted within the run; runs are at most RUN_MAX long).
static constexpr int BL_THREADS = 512;
static constexpr int BL_E = S_BOTTOM / BL_THREADS;   // consecutive positions per thread
static constexpr int BL_LEVEL_NODES = S_BOTTOM / 2;  // splitting nodes of one heap level (two segments or more)
static constexpr int RID_NONE = 2 * S_BOTTOM;
__global__ __launch_bounds__(BL_THREADS) void k_bottom_lists(TreeArrays t, int lbuf0, const double* __restrict__ mx,
                                                             const double* __restrict__ my,
                                                             const double* __restrict__ mz) {
  __shared__ uint16_t L[2][3][S_BOTTOM];  // the lists (local ids), two buffers
  __shared__ uint16_t rid[3][S_BOTTOM];   // local id -> first position of its run of equal coordinates in the
                                          // root's list, or RID_NONE + position when it is in no run of that axis
  __shared__ uint16_t pos_new[S_BOTTOM];  // local id -> its position in the order after this level's split
  __shared__ uint8_t side[S_BOTTOM];      // local id -> 1 when it goes left at this level
  __shared__ uint16_t nodeof[S_BOTTOM];
  __shared__ uint16_t hlo[HCAP], hhi[HCAP];
  __shared__ uint8_t hinfo[HCAP];         // state (0 absent, 1 leaf so far, 2 split) | axis << 2 | parent's axis << 4
  __shared__ uint16_t nbase[BL_LEVEL_NODES][3];
  __shared__ int any_split, any_tie;
  const int count = live_count(t.counters, C_BOTTOM);
  const int m = (int)t.m;
  const double* cs[3] = {mx, my, mz};
  for (int b = blockIdx.x; b < count; b += gridDim.x) {
    const int root = t.bottom[b];
    const int a = t.lo[root];
    const int s = t.hi[root] - a;
    const int slots = heap_slots(s, t.leaf);
    const int32_t* __restrict__ lst = t.rval[(lbuf0 + t.depth[root]) & 1];
    const int32_t* __restrict__ seg = t.order + a;  // local id -> segment
    __syncthreads();  // every thread has read the previous subtree's flags
    if (threadIdx.x == 0) any_tie = 0;
    __syncthreads();
    for (int A = 0; A < 3; ++A)
      for (int i = threadIdx.x; i < s; i += blockDim.x) {
        const int v = lst[(size_t)A * m + a + i];
        const int q = v & SEG_MASK;
        const int loc = q < m ? t.ord_inv[q] - a : -1;
        if (v < 0) any_tie = 1;
        if ((unsigned)loc < (unsigned)s) {
          L[0][A][i] = (uint16_t)loc;
          if (A == 0) side[loc] = v < 0 ? 1 : 0;  // the segment is in a run of some axis
        } else {
          L[0][A][i] = 0;
          t.counters[C_OVERFLOW] = 1;  // cannot happen for a consistent build; reported
        }
      }
    for (int i = threadIdx.x; i < s; i += blockDim.x) nodeof[i] = 0;
    for (int j = threadIdx.x; j < slots; j += blockDim.x) {
      hinfo[j] = j == 0 ? (uint8_t)(1 | (3 << 4)) : 0;
      hlo[j] = 0;
      hhi[j] = j == 0 ? (uint16_t)s : 0;
    }
    __syncthreads();
    const bool ties = any_tie != 0;
    if (ties) {  // run ids from the root's lists (coordinates compared for flagged neighbours only)
      // the three axes together (the second list buffer, free until the first split, holds the equal flags)
      for (int idx = threadIdx.x; idx < 3 * s; idx += blockDim.x) {
        const int A = idx / s, i = idx - A * s;
        const int x = L[0][A][i];
        bool eq = false;
        if (i > 0 && side[x]) {
          const int y = L[0][A][i - 1];
          eq = side[y] && cs[A][seg[x]] == cs[A][seg[y]];
        }
        L[1][A][i] = eq ? 1 : 0;  // equal to the entry before it
      }
      __syncthreads();
      for (int idx = threadIdx.x; idx < 3 * s; idx += blockDim.x) {
        const int A = idx / s, i = idx - A * s;
        const uint16_t* __restrict__ eq = L[1][A];
        const bool member = eq[i] || (i + 1 < s && eq[i + 1]);
        int r = i;
        while (r > 0 && eq[r]) --r;
        rid[A][L[0][A][i]] = (uint16_t)(member ? r : RID_NONE + i);
      }
      __syncthreads();
    }
    int cur = 0;
    for (int k = 0;; ++k) {
      const int first = (1 << k) - 1;
      if (first >= slots) break;
      const int last = min(slots, (2 << k) - 1);
      if (threadIdx.x == 0) any_split = 0;
      __syncthreads();
      // axis and split flag of each node of this level from its lists' ends
      for (int j = first + (int)threadIdx.x; j < last; j += blockDim.x) {
        if ((hinfo[j] & 3) == 0) continue;
        const int lo = hlo[j], hi = hhi[j];
        if (hi - lo <= t.leaf) continue;
        double v[6];
#pragma unroll
        for (int A = 0; A < 3; ++A) {
          v[A] = cs[A][seg[L[cur][A][lo]]];
          v[3 + A] = cs[A][seg[L[cur][A][hi - 1]]];
        }
        int8_t ax, sp;
        choose_axis(v, &ax, &sp);
        if (sp) {
          hinfo[j] = (uint8_t)((hinfo[j] & 0xF0) | 2 | (ax << 2));
          any_split = 1;
        }
      }
      __syncthreads();
      if (!any_split) break;
      // sides and new positions from each splitting node's list of its axis
      for (int i = threadIdx.x; i < s; i += blockDim.x) {
        const int j = nodeof[i];
        if ((hinfo[j] & 3) == 2) {
          const int mid = (hlo[j] + hhi[j]) / 2;
          const int x = L[cur][(hinfo[j] >> 2) & 3][i];
          side[x] = i < mid ? 1 : 0;
          pos_new[x] = (uint16_t)i;
        }
      }
      __syncthreads();
      // stable partition of the three lists within every splitting node (one scan for the three)
      const int i0 = (int)threadIdx.x * BL_E;
      int f[3][BL_E], cnt[3], excl[3], tot[3];
#pragma unroll
      for (int A = 0; A < 3; ++A) {
        cnt[A] = 0;
#pragma unroll
        for (int e = 0; e < BL_E; ++e) {
          const int i = i0 + e;
          f[A][e] = (i < s && (hinfo[nodeof[i]] & 3) == 2) ? side[L[cur][A][i]] : 0;
          cnt[A] += f[A][e];
        }
      }
      block_exclusive_scan_n<3>(cnt, excl, tot);
#pragma unroll
      for (int A = 0; A < 3; ++A) {
        int run = excl[A];
#pragma unroll
        for (int e = 0; e < BL_E; ++e) {
          const int i = i0 + e;
          if (i < s) {
            const int j = nodeof[i];
            if ((hinfo[j] & 3) == 2 && hlo[j] == i) nbase[j - first][A] = (uint16_t)run;
          }
          run += f[A][e];
        }
      }
      __syncthreads();
#pragma unroll
      for (int A = 0; A < 3; ++A) {
        int run = excl[A];
#pragma unroll
        for (int e = 0; e < BL_E; ++e) {
          const int i = i0 + e;
          if (i < s) {
            const int j = nodeof[i];
            int dest = i;
            if ((hinfo[j] & 3) == 2) {
              const int lo = hlo[j], hi = hhi[j], mid = (lo + hi) / 2;
              const int left = run - nbase[j - first][A];
              dest = f[A][e] ? lo + left : mid + (i - lo - left);
            }
            if ((unsigned)dest < (unsigned)s) L[cur ^ 1][A][dest] = L[cur][A][i];
            else t.counters[C_OVERFLOW] = 1;
          }
          run += f[A][e];
        }
      }
      __syncthreads();
      cur ^= 1;
      if (ties) {
        // each run of equal coordinates within a child, in the order of the new positions
        for (int A = 0; A < 3; ++A)
          for (int i = threadIdx.x; i < s; i += blockDim.x) {
            const int j = nodeof[i];
            const int x = L[cur][A][i];
            int dest = i;
            const int r = rid[A][x];
            // the split axis's lists are in the new order already
            if ((hinfo[j] & 3) == 2 && r < RID_NONE && ((hinfo[j] >> 2) & 3) != A) {
              const int lo = hlo[j], hi = hhi[j], mid = (lo + hi) / 2;
              const int cl = i < mid ? lo : mid, ch = i < mid ? mid : hi;
              const uint16_t* __restrict__ Lc = L[cur][A];
              const uint16_t* __restrict__ rA = rid[A];
              // the run's ends within the child, four entries a step (their loads issued together)
              int rs = i, re = i + 1;
              for (;;) {
                const int r1 = rA[Lc[max(rs - 1, cl)]], r2 = rA[Lc[max(rs - 2, cl)]];
                const int r3 = rA[Lc[max(rs - 3, cl)]], r4 = rA[Lc[max(rs - 4, cl)]];
                if (rs - 1 < cl || r1 != r) break;
                --rs;
                if (rs - 1 < cl || r2 != r) break;
                --rs;
                if (rs - 1 < cl || r3 != r) break;
                --rs;
                if (rs - 1 < cl || r4 != r) break;
                --rs;
              }
              for (;;) {
                const int r1 = rA[Lc[min(re, ch - 1)]], r2 = rA[Lc[min(re + 1, ch - 1)]];
                const int r3 = rA[Lc[min(re + 2, ch - 1)]], r4 = rA[Lc[min(re + 3, ch - 1)]];
                if (re >= ch || r1 != r) break;
                ++re;
                if (re >= ch || r2 != r) break;
                ++re;
                if (re >= ch || r3 != r) break;
                ++re;
                if (re >= ch || r4 != r) break;
                ++re;
              }
              if (re - rs > 1) {
                const int p = pos_new[x];
                int rank = 0;
#pragma unroll 4
                for (int q = rs; q < re; ++q) rank += pos_new[Lc[q]] < p;
                dest = rs + rank;
              }
            }
            L[cur ^ 1][A][dest] = (uint16_t)x;
          }
        __syncthreads();
        cur ^= 1;
      }
      // children of the splitting nodes, then every position's node for the next level
      for (int j = first + (int)threadIdx.x; j < last; j += blockDim.x) {
        if ((hinfo[j] & 3) != 2) continue;
        const int ax = (hinfo[j] >> 2) & 3;
        const int lo = hlo[j], hi = hhi[j], middle = (lo + hi) / 2;
        hinfo[2 * j + 1] = (uint8_t)(1 | (ax << 4));
        hlo[2 * j + 1] = (uint16_t)lo;
        hhi[2 * j + 1] = (uint16_t)middle;
        hinfo[2 * j + 2] = (uint8_t)(1 | (ax << 4));
        hlo[2 * j + 2] = (uint16_t)middle;
        hhi[2 * j + 2] = (uint16_t)hi;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < s; i += blockDim.x) {
        const int j = nodeof[i];
        if ((hinfo[j] & 3) == 2) nodeof[i] = (uint16_t)(i < hhi[2 * j + 1] ? 2 * j + 1 : 2 * j + 2);
      }
      __syncthreads();
    }
    // results: each leaf's order is its parent's list of the split axis (a root that did not split keeps its own)
    int out[BL_E];
    for (int e = 0; e < BL_E; ++e) {
      const int i = (int)threadIdx.x + e * BL_THREADS;
      out[e] = -1;
      if (i < s) {
        const int pa = (hinfo[nodeof[i]] >> 4) & 3;
        if (pa != 3) out[e] = seg[L[cur][pa][i]];
      }
    }
    __syncthreads();  // every read of the root's order is done before it is overwritten
    for (int e = 0; e < BL_E; ++e) {
      const int i = (int)threadIdx.x + e * BL_THREADS;
      if (out[e] >= 0) t.order[a + i] = out[e];
    }
    const int hb = t.heap_base[b];
    const int d0 = t.depth[root], tr = t.tree[root];
    for (int j = threadIdx.x; j < slots; j += blockDim.x)
      heap_node_out(t, root, hb, a, d0, tr, j, (hinfo[j] & 3) != 0, (hinfo[j] & 3) == 2, hlo[j], hhi[j]);
  }
}

// ------------------------------------------------------------------ finish: pre-order layout
__global__ void k_node_hist(TreeArrays t) {
  const int count = live_count(t.counters, C_NODES);
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < count; id += gridDim.x * blockDim.x) {
    const int lo = t.lo[id];
    if (lo < 0) continue;
    atomicAdd(&t.lo_cnt[lo], 1);
    atomicMin(&t.lo_mind[lo], t.depth[id]);
  }
}

// exclusive scan of n int32 values in blocks of SCAN_ITEMS (three kernels)
static constexpr int SCAN_THREADS = 256, SCAN_PER = 16, SCAN_ITEMS = SCAN_THREADS * SCAN_PER;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(int n, const int32_t* __restrict__ in,
                                                              int32_t* __restrict__ part) {
  const int base = blockIdx.x * SCAN_ITEMS;
  int s = 0;
  for (int q = 0; q < SCAN_PER; ++q) {
    const int i = base + threadIdx.x * SCAN_PER + q;
    if (i < n) s += in[i];
  }
  int total;
  (void)block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ __launch_bounds__(PLAN_THREADS) void k_scan_parts(int blocks, int32_t* __restrict__ part,
                                                            int32_t* __restrict__ total_out) {
  int acc = 0;
  for (int base = 0; base < blocks; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int v = i < blocks ? part[i] : 0;
    int total;
    const int r = block_exclusive_scan(v, &total);
    if (i < blocks) part[i] = acc + r;
    acc += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = acc;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(int n, const int32_t* __restrict__ in,
                                                             const int32_t* __restrict__ part,
                                                             int32_t* __restrict__ out) {
  const int base = blockIdx.x * SCAN_ITEMS;
  int v[SCAN_PER];
  int s = 0;
  for (int q = 0; q < SCAN_PER; ++q) {
    const int i = base + threadIdx.x * SCAN_PER + q;
    v[q] = i < n ? in[i] : 0;
    s += v[q];
  }
  int total;
  int r = block_exclusive_scan(s, &total) + part[blockIdx.x];
  for (int q = 0; q < SCAN_PER; ++q) {
    const int i = base + threadIdx.x * SCAN_PER + q;
    if (i < n) out[i] = r;
    r += v[q];
  }
}

void scan_exclusive(Ctx& c, int n, const int32_t* in, int32_t* out, int32_t* part, int32_t* total_out) {
  const int blocks = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
  hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(SCAN_THREADS), 0, c.stream, n, in, part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(PLAN_THREADS), 0, c.stream, blocks, part, total_out);
  hipLaunchKernelGGL(k_scan_apply, dim3(blocks), dim3(SCAN_THREADS), 0, c.stream, n, in, part, out);
}

// Exclusive scan of n int32 values in one launch: tiles of SCAN_ITEMS taken by ticket, each tile's prefix by a
// single-pass look-back over the tiles before it.  The ticket and the status words (scan_lb_words(n) of them)
// must be zero when the kernel starts; the last tile writes the total to *total_out when it is not null.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_lb(int n, const int32_t* __restrict__ in,
                                                          int32_t* __restrict__ out, int32_t* __restrict__ ticket,
                                                          unsigned long long* __restrict__ status,
                                                          int32_t* __restrict__ total_out,
                                                          int32_t* __restrict__ counters) {
  __shared__ int s_tile, s_prefix;
  const int tiles = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1);
  __syncthreads();
  const int tile = s_tile;
  if (tile >= tiles) return;
  const int base = tile * SCAN_ITEMS;
  int v[SCAN_PER];
  int sum = 0;
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    const int i = base + (int)threadIdx.x * SCAN_PER + q;
    v[q] = i < n ? in[i] : 0;
    sum += v[q];
  }
  int total;
  int r = block_exclusive_scan(sum, &total);
  if (threadIdx.x < 64) {
    int prefix = 0;
    if (tile == 0) {
      if (threadIdx.x == 0) lb_publish(status, 1u, true, total);
    } else {
      if (threadIdx.x == 0) lb_publish(status + tile, 1u, false, total);
      bool ok = true;
      prefix = lb_lookback(status, 1, 0, tile, 1u, &ok);
      if (threadIdx.x == 0) {
        lb_publish(status + tile, 1u, true, prefix + total);
        if (!ok) counters[C_OVERFLOW] = 1;
      }
    }
    if (threadIdx.x == 0) {
      s_prefix = prefix;
      if (tile == tiles - 1 && total_out) *total_out = prefix + total;
    }
  }
  __syncthreads();
  r += s_prefix;
#pragma unroll
  for (int q = 0; q < SCAN_PER; ++q) {
    const int i = base + (int)threadIdx.x * SCAN_PER + q;
    if (i < n) out[i] = r;
    r += v[q];
  }
}

int scan_lb_words(int64_t n) { return (int)((n + SCAN_ITEMS - 1) / SCAN_ITEMS) + 1; }

void scan_exclusive_lb(Ctx& c, int n, const int32_t* in, int32_t* out, int32_t* ticket, unsigned long long* status,
                       int32_t* total_out, int32_t* counters) {
  const int tiles = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
  if (tiles == 0) return;
  hipLaunchKernelGGL(k_scan_lb, dim3(tiles), dim3(SCAN_THREADS), 0, c.stream, n, in, out, ticket, status, total_out,
                     counters);
}

__global__ void k_node_pre(TreeArrays t) {
  const int count = live_count(t.counters, C_NODES);
  int deepest = 0;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < count; id += gridDim.x * blockDim.x) {
    const int lo = t.lo[id];
    if (lo < 0) {
      t.pre[id] = -1;
      continue;
    }
    // read before the stores, which the compiler cannot tell apart from these arrays
    const int hi = t.hi[id], d = t.depth[id], tr = t.tree[id];
    const int p = t.lo_base[lo] + d - t.lo_mind[lo];
    const int skip = t.lo_base[hi];
    t.pre[id] = p;
    t.p_lo[p] = lo;
    t.p_hi[p] = hi;
    t.p_skip[p] = skip;
    t.p_depth[p] = d;
    t.p_tree[p] = tr;
    t.p_node[p] = id;
    deepest = max(deepest, d);
  }
  // the largest depth: one atomic per wavefront
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) deepest = max(deepest, __shfl_xor(deepest, off, 64));
  if ((threadIdx.x & 63) == 0 && deepest > 0) atomicMax(&t.counters[C_DEPTH], deepest);
}

__global__ void k_node_pre_links(TreeArrays t) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < t.cap; p += gridDim.x * blockDim.x) {
    int internal = 0;
    if (p < count) {
      const int id = t.p_node[p];
      const int par = t.parent[id];
      t.p_parent[p] = par >= 0 ? t.pre[par] : -1;
      internal = t.left[id] >= 0 ? 1 : 0;
    }
    t.p_flag[p] = internal;  // internal flags in pre-order, zero beyond the node count
  }
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= t.K; k += gridDim.x * blockDim.x)
    t.tree_first[k] = t.lo_base[t.roots_lo[k]];
}

// ------------------------------------------------------------------ host driver
int tree_layout(Ctx& c, int buf, int64_t m64, int K, int leaf, TreeArrays& t) {
  const int64_t m = m64;
  t.K = K;
  t.m = m;
  t.leaf = leaf;
  const int64_t cap = 8 * (m / std::max(leaf, 1) + 1) + 16 * (int64_t)K + 1024;
  if (cap > INT_MAX / 8 || m > (1 << 28)) {
    set_error("tree too large for one build");
    return OFW_ERR_ARG;
  }
  t.cap = (int)cap;
  t.tile_cap = (int)(m / TILE + 2 * (m / (S_BOTTOM / 2) + 1) + K + 64);
  t.bottom_cap = (int)(m / (std::max(leaf, 1) + 1) + K + 64);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t M = (size_t)std::max<int64_t>(m, 1);
  const size_t C = (size_t)t.cap;
  size_t need = 0;
  need += al((K + 1) * 4) + al(M * 4);                                    // roots_lo, order
  need += 7 * al(C * 4) + 2 * al(C);                                      // node table, axis, split
  need += al(C_COUNT * 4) + 5 * al(C * 4);                                // counters, roots, big lists, strands, flags
  need += 3 * al((size_t)t.tile_cap * 4) + al((size_t)t.tile_cap * SUBS * 48);   // tiles, box partials
  need += 2 * al(M * 8) + 2 * al(M * 4) + al(std::max(M, C) * 4);         // sort buffers, tmp
  need += 2 * al((size_t)t.bottom_cap * 4);                               // bottom, heap base
  need += 3 * al((M + 1 + C) * 4) + al(((M + C) / 4096 + 8) * 4);         // lo_cnt, lo_mind, lo_base, parts
  need += al(C * 4) + 8 * al(C * 4) + al((K + 1) * 4);                    // pre, p_*, tree_first
  // list build: keys and ids of the three axes (two buffers), sides, partition counts, presort bookkeeping
  t.rs_tiles = (int)((M + RS_TILE - 1) / RS_TILE);
  t.rs_passes = RS_KEY_PASSES + (K <= 1 ? 0 : (K <= 256 ? 1 : (K <= 65536 ? 2 : 3)));
  const size_t RUNS = 3 * M / 2 + 1;
  need += 2 * al(3 * M * 4) + 2 * al(3 * M * 4) + al((size_t)t.tile_cap * 3 * 8) + al(LEVEL_MAX * 4);
  need += al(2 * 4) + al(((size_t)scan_lb_words(m + 1) + scan_lb_words(t.cap)) * 8);  // single-pass scans
  need += al(3 * (size_t)t.rs_tiles * RS_BINS * 4) + al(3 * RS_BINS * 4);
  need += al(3 * M * 4) + 2 * al(RUNS * 4) + al(M) + al(M * 4) + al(3 * M * 4);  // runs
  need += al(M * 4);                                                            // bottom phase on the lists
  OFW_HIP(c.tree[buf].reserve(need));
  char* base = (char*)c.tree[buf].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return p;
  };
  t.roots_lo = (int32_t*)carve((K + 1) * 4);
  t.order = (int32_t*)carve(M * 4);
  t.lo = (int32_t*)carve(C * 4);
  t.hi = (int32_t*)carve(C * 4);
  t.left = (int32_t*)carve(C * 4);
  t.right = (int32_t*)carve(C * 4);
  t.parent = (int32_t*)carve(C * 4);
  t.depth = (int32_t*)carve(C * 4);
  t.tree = (int32_t*)carve(C * 4);
  t.axis = (int8_t*)carve(C);
  t.split = (int8_t*)carve(C);
  t.counters = (int32_t*)carve(C_COUNT * 4);
  t.level_a = (int32_t*)carve(C * 4);
  t.big = (int32_t*)carve(C * 4);
  t.big_alt = (int32_t*)carve(C * 4);
  t.strand = (int32_t*)carve(C * 4);
  t.p_flag = (int32_t*)carve(C * 4);
  t.tile_node = (int32_t*)carve((size_t)t.tile_cap * 4);
  t.tile_p0 = (int32_t*)carve((size_t)t.tile_cap * 4);
  t.tile_p1 = (int32_t*)carve((size_t)t.tile_cap * 4);
  t.box_partial = (double*)carve((size_t)t.tile_cap * SUBS * 48);
  t.keyA = (uint64_t*)carve(M * 8);
  t.keyB = (uint64_t*)carve(M * 8);
  t.posA = (uint32_t*)carve(M * 4);
  t.posB = (uint32_t*)carve(M * 4);
  t.tmp = (int32_t*)carve(std::max(M, C) * 4);
  t.bottom = (int32_t*)carve((size_t)t.bottom_cap * 4);
  t.heap_base = (int32_t*)carve((size_t)t.bottom_cap * 4);
  t.lo_cnt = (int32_t*)carve((M + 1 + C) * 4);
  t.lo_mind = (int32_t*)carve((M + 1 + C) * 4);
  t.lo_base = (int32_t*)carve((M + 1 + C) * 4);
  t.scan_part = (int32_t*)carve(((M + C) / 4096 + 8) * 4);
  t.pre = (int32_t*)carve(C * 4);
  t.p_lo = (int32_t*)carve(C * 4);
  t.p_hi = (int32_t*)carve(C * 4);
  t.p_skip = (int32_t*)carve(C * 4);
  t.p_parent = (int32_t*)carve(C * 4);
  t.p_depth = (int32_t*)carve(C * 4);
  t.p_tree = (int32_t*)carve(C * 4);
  t.p_node = (int32_t*)carve(C * 4);
  t.p_irank = (int32_t*)carve(C * 4);
  t.tree_first = (int32_t*)carve((K + 1) * 4);
  for (int b = 0; b < 2; ++b) t.rkey[b] = (uint32_t*)carve(3 * M * 4);
  for (int b = 0; b < 2; ++b) t.rval[b] = (int32_t*)carve(3 * M * 4);
  t.lstat = (unsigned long long*)carve((size_t)t.tile_cap * 3 * 8);
  t.scan_ticket = (int32_t*)carve(2 * 4);
  t.scan_words = scan_lb_words(m + 1) + scan_lb_words(t.cap);
  t.scan_status = (unsigned long long*)carve((size_t)t.scan_words * 8);
  t.tickets = (int32_t*)carve(LEVEL_MAX * 4);
  t.rcount = (int32_t*)carve(3 * (size_t)t.rs_tiles * RS_BINS * 4);
  t.rtotal = (int32_t*)carve(3 * RS_BINS * 4);
  t.run_seg = (int32_t*)carve(3 * M * 4);
  t.run_off = (int32_t*)carve(RUNS * 4);
  t.run_len = (int32_t*)carve(RUNS * 4);
  t.tie = (uint8_t*)carve(M);
  t.inv_o = (int32_t*)carve(M * 4);
  t.ord_inv = (int32_t*)carve(M * 4);
  t.invb = (int32_t*)carve(3 * M * 4);
  t.lists = c.sort_build ? 0 : 1;
  return OFW_OK;
}

int tree_build(Ctx& c, TreeArrays& t, const double* mx, const double* my, const double* mz,
               const std::vector<int64_t>& offsets) {
  const int m = (int)t.m;
  t.mid[0] = mx;
  t.mid[1] = my;
  t.mid[2] = mz;
  int64_t max_root = 0;
  for (int k = 0; k < t.K; ++k) max_root = std::max(max_root, offsets[k + 1] - offsets[k]);
  const int grid = c.cu_count * 2;
  const bool list_init = t.lists && m > 0;  // the presort's first kernel takes the order and the roots
  if (!list_init) {
    if (m > 0) hipLaunchKernelGGL(k_iota, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, t.order);
    hipLaunchKernelGGL(k_init_roots, dim3(1), dim3(256), 0, c.stream, t);
  }
  OFW_HIP(hipGetLastError());
  // top levels: while the largest node of a level can hold more than S_BOTTOM segments.  Level lvl reads the
  // order buffer of its parity and leaves its sorted nodes in the other one.
  int lbuf = 0;  // the list buffer of the current level
  int lbuf0 = 0;  // the buffer of the roots' lists (a node at depth d holds its lists in lbuf0 ^ (d & 1))
  if (t.lists && m > 0) {
    // presort of the three axes' lists and the runs of equal coordinates in them
    const unsigned ge = (unsigned)std::min<int64_t>((m + 255) / 256, 4 * (int64_t)grid);
    hipLaunchKernelGGL(k_rs_init, dim3(ge, 3), dim3(256), 0, c.stream, t, mx, my, mz);
    for (int p = 0; p < t.rs_passes; ++p) {
      hipLaunchKernelGGL(k_rs_count, dim3((unsigned)t.rs_tiles, 3), dim3(RS_THREADS), 0, c.stream, t, p, lbuf);
      hipLaunchKernelGGL(k_rs_scan, dim3(3 * RS_BINS / 4), dim3(256), 0, c.stream, t);
      hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)t.rs_tiles, 3), dim3(RS_THREADS), 0, c.stream, t, p, lbuf);
      lbuf ^= 1;
    }
    hipLaunchKernelGGL(k_rs_fix, dim3(ge, 3), dim3(256), 0, c.stream, t, lbuf, mx, my, mz);
    hipLaunchKernelGGL(k_tie_scan, dim3(ge, 3), dim3(256), 0, c.stream, t, lbuf, lbuf ^ 1, mx, my, mz);
    lbuf ^= 1;
    lbuf0 = lbuf;
    hipLaunchKernelGGL(k_tie_mark, dim3(ge, 3), dim3(256), 0, c.stream, t, lbuf);
    OFW_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_level_plan, dim3(1), dim3(PLAN_THREADS), 0, c.stream, t, t.level_a);
  int32_t* const bigs[2] = {t.big, t.big_alt};
  int32_t* const orders[2] = {t.order, t.tmp};
  for (int lvl = 0;; ++lvl) {
    const int64_t biggest = (max_root + ((int64_t)1 << lvl) - 1) >> lvl;  // ceil(max_root / 2^lvl)
    if (!is_big((int)std::min<int64_t>(biggest, INT_MAX), t.leaf)) break;  // planned: all to the bottom
    // no more tiles than m / TILE plus one partial tile per node of the level: grids beyond that only idle
    const int64_t tile_bound = (int64_t)m / TILE + ((int64_t)t.K << std::min(lvl, 30)) + 1;
    if (t.lists) {
      if (lvl >= LEVEL_MAX) {
        set_error("tree too deep for the list build");
        return OFW_ERR_ARG;
      }
      const unsigned gl = (unsigned)std::min<int64_t>(4 * (int64_t)c.cu_count, 3 * tile_bound);
      hipLaunchKernelGGL(k_list_level, dim3(gl), dim3(LL_THREADS), 0, c.stream, t, lvl, t.rval[lbuf],
                         t.rval[lbuf ^ 1], mx, my, mz);
      lbuf ^= 1;
    } else {
      const int32_t* oin = orders[lvl & 1];
      int32_t* oout = orders[(lvl + 1) & 1];
      int passes = 0;
      for (int64_t width = TILE; width < biggest; width <<= 1) ++passes;
      const unsigned g1 = (unsigned)std::min<int64_t>(grid, tile_bound);
      const unsigned g2 = (unsigned)std::min<int64_t>(2 * (int64_t)grid, tile_bound);
      const unsigned g4 = (unsigned)std::min<int64_t>(4 * (int64_t)grid, tile_bound * SUBS);
      hipLaunchKernelGGL(k_box_tiles, dim3(g4), dim3(BOX_THREADS), 0, c.stream, t, oin, mx, my, mz);
      hipLaunchKernelGGL(k_sort_tiles, dim3(g1), dim3(TILE_THREADS), 0, c.stream, t, oin,
                         passes == 0 ? oout : nullptr, mx, my, mz);
      uint64_t *kin = t.keyA, *kout = t.keyB;
      uint32_t *pin = t.posA, *pout = t.posB;
      for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        hipLaunchKernelGGL(k_merge_pass, dim3(g2), dim3(MERGE_THREADS), 0, c.stream, t, (int)(TILE << p), kin,
                           pin, kout, pout, last ? oin : nullptr, last ? oout : nullptr);
        std::swap(kin, kout);
        std::swap(pin, pout);
      }
    }
    hipLaunchKernelGGL(k_level_next, dim3(t.lists ? 1 + FIX_BLOCKS : 1), dim3(PLAN_THREADS), 0, c.stream, t, lvl,
                       bigs[lvl & 1], bigs[(lvl + 1) & 1], t.lists ? t.rval[lbuf] : nullptr);
    OFW_HIP(hipGetLastError());
  }
  if (!t.lists) hipLaunchKernelGGL(k_copy_strands, dim3(grid), dim3(256), 0, c.stream, t);
  // bottom phase
  hipLaunchKernelGGL(k_bottom_bases, dim3(1), dim3(PLAN_THREADS), 0, c.stream, t);
  if (t.lists) hipLaunchKernelGGL(k_bottom_lists, dim3(grid), dim3(BL_THREADS), 0, c.stream, t, lbuf0, mx, my, mz);
  else hipLaunchKernelGGL(k_bottom, dim3(grid), dim3(BOTTOM_THREADS), 0, c.stream, t, mx, my, mz);
  OFW_HIP(hipGetLastError());
  // finish: pre-order layout from (lo, depth)
  const int n1 = m + 1;
  if (!(t.lists && m > 0)) {  // the list build's presort clears them
    hipLaunchKernelGGL(k_fill_i32, dim3((n1 + 255) / 256), dim3(256), 0, c.stream, n1, t.lo_cnt, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3((n1 + 255) / 256), dim3(256), 0, c.stream, n1, t.lo_mind, INT_MAX);
  }
  hipLaunchKernelGGL(k_node_hist, dim3(grid), dim3(256), 0, c.stream, t);
  scan_exclusive_lb(c, n1, t.lo_cnt, t.lo_base, t.scan_ticket, t.scan_status, t.counters + C_FINAL, t.counters);
  hipLaunchKernelGGL(k_node_pre, dim3(grid), dim3(256), 0, c.stream, t);
  hipLaunchKernelGGL(k_node_pre_links, dim3(grid), dim3(256), 0, c.stream, t);
  // internal ranks in pre-order (for the reference numbering)
  scan_exclusive_lb(c, t.cap, t.p_flag, t.p_irank, t.scan_ticket + 1, t.scan_status + scan_lb_words(t.m + 1), nullptr,
                    t.counters);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

// Reference numbering (vortex_tree.py:271-306): root 0; the children of the split node with internal rank
// j in pre-order are 2j+1 and 2j+2.  Host arrays for tests and ofw_build_tree (single tree).
int tree_reference_arrays(Ctx& c, const TreeArrays& t, std::vector<int64_t>& lo, std::vector<int64_t>& hi,
                          std::vector<int64_t>& left, std::vector<int64_t>& right, std::vector<int32_t>& ref_of_pre,
                          int64_t& count) {
  int32_t counters[C_COUNT];
  OFW_HIP(hipMemcpyAsync(counters, t.counters, sizeof(counters), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (counters[C_OVERFLOW]) {
    set_error("tree build exceeded its node capacity");
    return OFW_ERR_OVERFLOW;
  }
  const int n = counters[C_FINAL];
  std::vector<int32_t> plo(n), phi(n), pnode(n), pparent(n), irank(n), left_id(t.cap);
  OFW_HIP(hipMemcpyAsync(plo.data(), t.p_lo, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(phi.data(), t.p_hi, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(pnode.data(), t.p_node, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(pparent.data(), t.p_parent, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(irank.data(), t.p_irank, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(left_id.data(), t.left, (size_t)t.cap * 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  count = n;
  ref_of_pre.assign(n, 0);
  lo.assign(n, 0);
  hi.assign(n, 0);
  left.assign(n, -1);
  right.assign(n, -1);
  for (int p = 0; p < n; ++p) {
    const int par = pparent[p];
    int r;
    if (par < 0) {
      r = 0;
    } else {
      // this node is the left child of its parent when it directly follows it in pre-order
      r = 2 * irank[par] + (par + 1 == p ? 1 : 2);
    }
    ref_of_pre[p] = r;
  }
  for (int p = 0; p < n; ++p) {
    const int r = ref_of_pre[p];
    lo[r] = plo[p];
    hi[r] = phi[p];
    const int id = pnode[p];
    if (left_id[id] >= 0) {
      left[r] = 2 * irank[p] + 1;
      right[r] = 2 * irank[p] + 2;
    }
  }
  return OFW_OK;
}

}  // namespace ofw
