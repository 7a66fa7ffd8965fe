// Incident-exclusion pairs of a wake network on the GPU (network_numba.py:121-253, incident_pairs): for the
// free points nodes[first:], every (point, segment) pair in which the point coincides bit for bit with an end
// of the segment, in segment order, the start end before the end end, and for one end the coinciding points
// in increasing order.  A node with a NaN coordinate never matches.  Point indices count from `first`.
//
// Integer work, bound by memory latency and L2 atomics, not by arithmetic:
//   1. k_group_points: every free point joins the group of its coordinate bit pattern in an open-addressing
//      hash table of H = 2^k >= 2 F slots (linear probing, one CAS per new pattern); the slot is the group id
//      and count[slot] its size.
//   2. A stable LSD radix sort of the point indices by slot (8-bit digits, per-tile histograms, wavefront
//      ballots for the stable rank inside a tile) leaves every group's members contiguous and in increasing
//      order at start[slot] = the exclusive scan of count, whatever order the threads won the table in.
//   3. every node's group: the free points' from step 1, the others' by one probe (k_fixed_groups); each segment end
//      reads its node's group (k_end_groups) and the pair counts are scanned.
//   4. k_write_pairs: one thread per output pair finds its segment end by binary search over the offsets.
#include <climits>

#include "common.h"
#include "internal.h"
#include "scan.h"
#include "tree.h"

namespace ofw {

namespace {

constexpr int RADIX_BITS = 8, RADIX = 1 << RADIX_BITS;
constexpr int SORT_THREADS = 256, SORT_PER = 8, SORT_TILE = SORT_THREADS * SORT_PER;
constexpr int SORT_WAVES = SORT_THREADS / 64;
static_assert(SORT_THREADS == RADIX, "one histogram bin per thread");

struct NodeBits {
  uint64_t x, y, z;
};

__device__ __forceinline__ NodeBits node_bits(const double* __restrict__ nodes, int i) {
  return NodeBits{(uint64_t)__double_as_longlong(nodes[3 * (size_t)i]),
                  (uint64_t)__double_as_longlong(nodes[3 * (size_t)i + 1]),
                  (uint64_t)__double_as_longlong(nodes[3 * (size_t)i + 2])};
}

__device__ __forceinline__ bool has_nan(const double* __restrict__ nodes, int i) {
  return isnan(nodes[3 * (size_t)i]) || isnan(nodes[3 * (size_t)i + 1]) || isnan(nodes[3 * (size_t)i + 2]);
}

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

__device__ __forceinline__ bool same_bits(const NodeBits& a, const NodeBits& b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}

// splitmix64 finaliser (Steele, Lea and Flood, "Fast splittable pseudorandom number generators", 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t v) {
  v ^= v >> 30;
  v *= 0xbf58476d1ce4e5b9ull;
  v ^= v >> 27;
  v *= 0x94d049bb133111ebull;
  return v ^ (v >> 31);
}

__device__ __forceinline__ uint32_t slot_of(const NodeBits& k, uint32_t mask) {
  return (uint32_t)mix64(k.x ^ mix64(k.y ^ mix64(k.z))) & mask;
}

}  // namespace

__global__ void k_net_fill(int count, int32_t* __restrict__ a, int32_t v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) a[i] = v;
}

// table[slot] holds the node index of the group's first inserted point (-1 = empty).  key[i] = the point's
// slot, or H (sorted past every group) for a point with a NaN coordinate; val[i] = i.
__global__ void k_group_points(int nfree, int first, const double* __restrict__ nodes, uint32_t H,
                               int32_t* __restrict__ table, int32_t* __restrict__ count, int32_t* __restrict__ key,
                               int32_t* __restrict__ val, int32_t* __restrict__ node_group) {
  const uint32_t mask = H - 1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nfree; i += gridDim.x * blockDim.x) {
    const int node = first + i;
    val[i] = i;
    if (has_nan(nodes, node)) {
      key[i] = (int32_t)H;
      node_group[node] = -1;
      continue;
    }
    const NodeBits bits = node_bits(nodes, node);
    uint32_t h = slot_of(bits, mask);
    // the table has H >= 2 nfree slots, so an empty slot ends every probe sequence
    while (true) {
      int cur = table[h];
      if (cur < 0) {
        const int prev = atomicCAS(&table[h], -1, node);
        if (prev < 0) break;  // this point opens the group
        cur = prev;
      }
      if (same_bits(node_bits(nodes, cur), bits)) break;
      h = (h + 1) & mask;
    }
    key[i] = (int32_t)h;
    node_group[node] = (int32_t)h;
    atomicAdd(&count[h], 1);
  }
}

// The group of each node before `first` (not inserted): one probe; -1 when none or a NaN coordinate.
__global__ void k_fixed_groups(int first, const double* __restrict__ nodes, uint32_t H,
                               const int32_t* __restrict__ table, int32_t* __restrict__ node_group) {
  const uint32_t mask = H - 1;
  for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < first; node += gridDim.x * blockDim.x) {
    int g = -1;
    if (!has_nan(nodes, node)) {
      const NodeBits bits = node_bits(nodes, node);
      uint32_t h = slot_of(bits, mask);
      while (true) {
        const int cur = table[h];
        if (cur < 0) break;
        if (same_bits(node_bits(nodes, cur), bits)) {
          g = (int)h;
          break;
        }
        h = (h + 1) & mask;
      }
    }
    node_group[node] = g;
  }
}

// Digit histogram of one tile, digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(int n, int shift, const int32_t* __restrict__ key,
                                                             int32_t* __restrict__ hist) {
  __shared__ int s_hist[RADIX];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE;
  for (int q = 0; q < SORT_PER; ++q) {
    const int i = base + q * SORT_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&s_hist[(key[i] >> shift) & (RADIX - 1)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = s_hist[threadIdx.x];
}

// Stable scatter of one tile: the tile's items go in rounds of 256 in index order; inside a round a wavefront
// finds the lanes sharing its digit from 8 ballots, the wavefronts before it add their counts, and the rounds
// before it add a running count per digit.
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(int n, int shift, const int32_t* __restrict__ key,
                                                                const int32_t* __restrict__ val,
                                                                const int32_t* __restrict__ offset,
                                                                int32_t* __restrict__ key_out,
                                                                int32_t* __restrict__ val_out) {
  __shared__ int s_base[RADIX];
  __shared__ int s_wave[SORT_WAVES][RADIX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  s_base[threadIdx.x] = offset[threadIdx.x * gridDim.x + blockIdx.x];
  const uint64_t below = (1ull << lane) - 1;
  const int base = blockIdx.x * SORT_TILE;
  for (int q = 0; q < SORT_PER; ++q) {
    for (int w = 0; w < SORT_WAVES; ++w) s_wave[w][threadIdx.x] = 0;
    __syncthreads();
    const int i = base + q * SORT_THREADS + threadIdx.x;
    const bool live = i < n;
    const int k = live ? key[i] : 0;
    const int d = (k >> shift) & (RADIX - 1);
    uint64_t peers = __builtin_amdgcn_ballot_w64(live);
    for (int b = 0; b < RADIX_BITS; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t set = __builtin_amdgcn_ballot_w64(bit);
      peers &= bit ? set : ~set;
    }
    const int rank = __popcll(peers & below);
    if (live && rank == 0) s_wave[wave][d] = __popcll(peers);
    __syncthreads();
    if (live) {
      int pos = s_base[d] + rank;
      for (int w = 0; w < wave; ++w) pos += s_wave[w][d];
      key_out[pos] = k;
      val_out[pos] = val[i];
    }
    __syncthreads();
    int add = 0;
    for (int w = 0; w < SORT_WAVES; ++w) add += s_wave[w][threadIdx.x];
    s_base[threadIdx.x] += add;
  }
}

// Per segment end q = 2 s + e: its node's group (-1 when none) and the number of pairs it gives.  status[0] sums
// the pairs over all ends; status[1] is set when an edge names a node outside [0, n).
__global__ void k_end_groups(int m, int n, const int64_t* __restrict__ edges, const int32_t* __restrict__ node_group,
                             const int32_t* __restrict__ count, int32_t* __restrict__ end_group,
                             int32_t* __restrict__ end_count, unsigned long long* __restrict__ status) {
  unsigned long long mine = 0;
  bool bad = false;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * m; q += gridDim.x * blockDim.x) {
    const int64_t e = edges[q];
    int g = -1;
    if (e < 0 || e >= n)
      bad = true;
    else
      g = node_group[e];
    end_group[q] = g;
    const int c = g >= 0 ? count[g] : 0;
    end_count[q] = c;
    mine += (unsigned long long)c;
  }
  // one atomic per workgroup on the shared total: wavefront sums, then the workgroup's in LDS
  __shared__ unsigned long long s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
  const bool lead = (threadIdx.x & 63) == 0;
  if (lead && mine) atomicAdd(&s_sum, mine);
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && lead) atomicOr(&status[1], 1ull);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) atomicAdd(&status[0], s_sum);
}

// One thread per pair p < min(total, capacity): its end q is the last one with offset[q] <= p (the ends after q
// start beyond p, so end_count[q] > 0 and p lies inside q's range).  Nothing is written after a bad edge or when the
// total exceeds 2^31 - 1 (the int32 offsets would not hold it); the range checks keep every access in bounds.
// Each thread writes PAIR_RUN consecutive pairs: one binary search for the first, then a walk along the ends.
constexpr int PAIR_RUN = 8;

__global__ void k_write_pairs(int64_t capacity, int ends, const unsigned long long* __restrict__ status,
                              const int32_t* __restrict__ end_group, const int32_t* __restrict__ offset,
                              const int32_t* __restrict__ start, const int32_t* __restrict__ count,
                              const int32_t* __restrict__ members, int64_t* __restrict__ pairs) {
  const unsigned long long all = status[0];
  if (status[1] || all > (unsigned long long)INT_MAX) return;
  const int64_t total = (int64_t)all < capacity ? (int64_t)all : capacity;
  const int64_t runs = (total + PAIR_RUN - 1) / PAIR_RUN;
  for (int64_t run = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; run < runs;
       run += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p0 = run * PAIR_RUN, p1 = p0 + PAIR_RUN < total ? p0 + PAIR_RUN : total;
    int q = 0, hi = ends - 1;
    while (q < hi) {
      const int mid = (q + hi + 1) >> 1;
      if (offset[mid] <= p0)
        q = mid;
      else
        hi = mid - 1;
    }
    for (int64_t p = p0; p < p1; ++p) {
      // the last end starting at or before p: a few steps forward, then a binary search over the rest (runs of
      // ends without pairs can be long)
      int steps = 0;
      while (q + 1 < ends && offset[q + 1] <= p && steps < 4) {
        ++q;
        ++steps;
      }
      if (q + 1 < ends && offset[q + 1] <= p) {
        int lo = q + 1, top = ends - 1;
        while (lo < top) {
          const int mid = (lo + top + 1) >> 1;
          if (offset[mid] <= p)
            lo = mid;
          else
            top = mid - 1;
        }
        q = lo;
      }
      const int g = end_group[q];
      const int64_t r = p - offset[q];
      if (g < 0 || r < 0 || r >= count[g]) continue;
      pairs[2 * p] = members[start[g] + r];
      pairs[2 * p + 1] = q >> 1;
    }
  }
}

int radix_hist_size(int n) { return RADIX * ((n + SORT_TILE - 1) / SORT_TILE); }

// Stable LSD radix sort of (key, val) by the low 8 * passes bits of key; returns the buffer index (0 or 1)
// that holds the result.  hist and hist_off hold RADIX * tiles entries, part the scan's partials.
int radix_sort(Ctx& c, int n, int passes, int32_t* key[2], int32_t* val[2], int32_t* hist, int32_t* hist_off,
                      int32_t* part) {
  const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(SORT_THREADS), 0, c.stream, n, p * RADIX_BITS, key[cur],
                       hist);
    scan_exclusive(c, RADIX * tiles, hist, hist_off, part, nullptr);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(SORT_THREADS), 0, c.stream, n, p * RADIX_BITS, key[cur],
                       val[cur], hist_off, key[cur ^ 1], val[cur ^ 1]);
    cur ^= 1;
  }
  return cur;
}

// Number of 8-bit passes for keys in [0, top].
int radix_passes(int64_t top) {
  int bits = 1;
  while ((1ll << bits) <= top) ++bits;
  return (bits + RADIX_BITS - 1) / RADIX_BITS;
}

// Scratch of one incident-pair call (buffer 6 of the context).
struct PairScratch {
  uint32_t H = 0;
  int F = 0, E = 0, passes = 0;
  int32_t *table, *cnt, *start, *key[2], *val[2], *end_group, *end_count, *offset, *hist, *hist_off, *part;
  int32_t* node_group;
  unsigned long long* status;  // [0] the pair total, [1] a bad edge
};

static int pair_scratch(Ctx& c, int64_t n, int64_t m, int64_t first, PairScratch& s) {
  const int64_t nfree = n - first;
  s.H = 2;
  while ((int64_t)s.H < 2 * nfree) s.H <<= 1;
  s.passes = radix_passes(s.H);  // keys run to H inclusive
  s.F = (int)nfree;
  s.E = (int)(2 * m);
  const int tiles = (s.F + SORT_TILE - 1) / SORT_TILE;
  const int hist_n = RADIX * tiles;
  const int scan_max = std::max<int>(std::max<int>((int)s.H, s.E), hist_n);
  const size_t spart = ((size_t)(scan_max + 4095) / 4096 + 8) * 4;
  const size_t H4 = (size_t)s.H * 4, F4 = (size_t)s.F * 4, E4 = (size_t)s.E * 4;
  const size_t need = 3 * al(H4) + 4 * al(F4) + 3 * al(E4) + 2 * al((size_t)hist_n * 4) + al(spart) + al(16) +
                      al((size_t)n * 4);
  OFW_HIP(c.tree[6].reserve(need));
  char* base = (char*)c.tree[6].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (int32_t*)p;
  };
  s.table = carve(H4);
  s.cnt = carve(H4);
  s.start = carve(H4);
  s.key[0] = carve(F4);
  s.key[1] = carve(F4);
  s.val[0] = carve(F4);
  s.val[1] = carve(F4);
  s.end_group = carve(E4);
  s.end_count = carve(E4);
  s.offset = carve(E4);
  s.hist = carve((size_t)hist_n * 4);
  s.hist_off = carve((size_t)hist_n * 4);
  s.part = carve(spart);
  s.status = (unsigned long long*)carve(16);
  s.node_group = carve((size_t)n * 4);
  return OFW_OK;
}

// Phase 1: the groups, each end's group and count, the 64-bit total (status[0]).  Returns the buffer index of the
// sorted members.
static int pairs_count(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, int64_t first,
                       PairScratch& s) {
  const int grid = c.cu_count * 8;
  auto blocks = [&](int64_t items) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (items + 255) / 256)));
  };
  hipLaunchKernelGGL(k_net_fill, blocks(s.H), dim3(256), 0, c.stream, (int)s.H, s.table, -1);
  hipLaunchKernelGGL(k_net_fill, blocks(s.H), dim3(256), 0, c.stream, (int)s.H, s.cnt, 0);
  hipLaunchKernelGGL(k_net_fill, dim3(1), dim3(64), 0, c.stream, 4, (int32_t*)s.status, 0);
  hipLaunchKernelGGL(k_group_points, blocks(s.F), dim3(256), 0, c.stream, s.F, (int)first, nodes, s.H, s.table, s.cnt,
                     s.key[0], s.val[0], s.node_group);
  if (first > 0)
    hipLaunchKernelGGL(k_fixed_groups, blocks(first), dim3(256), 0, c.stream, (int)first, nodes, s.H, s.table,
                       s.node_group);
  scan_exclusive(c, (int)s.H, s.cnt, s.start, s.part, nullptr);
  const int cur = radix_sort(c, s.F, s.passes, s.key, s.val, s.hist, s.hist_off, s.part);
  hipLaunchKernelGGL(k_end_groups, blocks(s.E), dim3(256), 0, c.stream, (int)m, (int)n, edges, s.node_group, s.cnt,
                     s.end_group, s.end_count, s.status);
  OFW_HIP(hipGetLastError());
  return cur;
}

// Phase 2: the offsets and the pairs, at most `capacity` of them.
static int pairs_write(Ctx& c, const PairScratch& s, int members_buf, int64_t capacity, int64_t* pairs) {
  if (capacity <= 0) return OFW_OK;
  scan_exclusive(c, s.E, s.end_count, s.offset, s.part, nullptr);
  const int64_t runs = (capacity + PAIR_RUN - 1) / PAIR_RUN;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(c.cu_count * 8, (runs + 255) / 256));
  hipLaunchKernelGGL(k_write_pairs, dim3(grid), dim3(256), 0, c.stream, capacity, s.E, s.status, s.end_group,
                     s.offset, s.start, s.cnt, s.val[members_buf], pairs);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

static int pairs_args(int64_t n, int64_t m) {
  if (n > INT_MAX / 4 || m > INT_MAX / 4) {
    set_error("ofw_incident_pairs: more than 2^29 nodes or segments");
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

int incident_pairs(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, int64_t first,
                   int64_t* pairs, int64_t capacity, int64_t* count_out) {
  *count_out = 0;
  if (n - first <= 0 || m <= 0) return OFW_OK;
  int st = pairs_args(n, m);
  if (st) return st;
  PairScratch s;
  st = pair_scratch(c, n, m, first, s);
  if (st) return st;
  const int cur = pairs_count(c, n, nodes, m, edges, first, s);
  unsigned long long host[2] = {0, 0};
  OFW_HIP(hipMemcpyAsync(host, s.status, 16, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (host[1]) {
    set_error("ofw_incident_pairs: an edge names a node outside [0, n)");
    return OFW_ERR_ARG;
  }
  *count_out = (int64_t)host[0];
  if (host[0] > (unsigned long long)INT_MAX) {
    set_error("ofw_incident_pairs: more than 2^31 - 1 pairs");
    return OFW_ERR_ARG;
  }
  if ((int64_t)host[0] > capacity) {
    set_error("ofw_incident_pairs: the pair buffer is too small (count_out holds the number needed)");
    return OFW_ERR_OVERFLOW;
  }
  if (host[0] == 0) return OFW_OK;
  return pairs_write(c, s, cur, (int64_t)host[0], pairs);
}

int incident_pairs_enqueue(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, int64_t first,
                           int64_t* pairs, int64_t capacity, const unsigned long long** status_out) {
  *status_out = nullptr;
  if (n - first <= 0 || m <= 0) return OFW_OK;
  int st = pairs_args(n, m);
  if (st) return st;
  PairScratch s;
  st = pair_scratch(c, n, m, first, s);
  if (st) return st;
  const int cur = pairs_count(c, n, nodes, m, edges, first, s);
  *status_out = s.status;
  return pairs_write(c, s, cur, capacity, pairs);
}

// ---------------------------------------------------------------------------------------------------------------
// Ring-edge aggregation (network_numba.py:36-117, pack_ring_edges; the loop of wake.py:300-307).  Ring (i, j) of the
// (depth - 1) x panels lattice runs (i, left) -> (i, right) -> (i+1, right) -> (i+1, left) -> (i, left) over the
// node ids row * width + column; contribution c = 4 (i panels + j) + k is its k-th edge, stored undirected as
// (lo, hi) with +strength when it runs lo -> hi and -strength otherwise.  Contributions sharing (lo, hi) are summed
// from 0.0 in increasing c; the edges come in the order of their first contribution, exact zeros omitted.
//
// Same machinery as the incident pairs: groups by a hash of the 64-bit key lo * nodes + hi, members in increasing c
// by the stable radix sort, one thread per group for the ordered sum, and an exclusive scan over the first
// contributions for the output positions.

struct RingEdge {
  int64_t lo, hi;
  double value;
};

__device__ __forceinline__ RingEdge ring_contribution(int64_t c, int64_t width, int panels,
                                                      const int64_t* __restrict__ panel_edges,
                                                      const double* __restrict__ strengths) {
  const int64_t ring = c >> 2;
  const int k = (int)(c & 3);
  const int64_t i = ring / panels;
  const int j = (int)(ring - i * panels);
  const int64_t left = panel_edges[2 * j], right = panel_edges[2 * j + 1];
  const int64_t row = i * width, next = (i + 1) * width;
  const int64_t a = k == 0 ? row + left : k == 1 ? row + right : k == 2 ? next + right : next + left;
  const int64_t b = k == 0 ? row + right : k == 1 ? next + right : k == 2 ? next + left : row + left;
  const double g = strengths[ring];
  return a < b ? RingEdge{a, b, g} : RingEdge{b, a, -g};
}

constexpr unsigned long long EMPTY_KEY = ~0ull;

// status[1] is set when a panel names a column outside [0, width) or joins a column to itself.
__global__ void k_ring_group(int64_t total, int64_t width, int64_t nodes, int panels,
                             const int64_t* __restrict__ panel_edges, const double* __restrict__ strengths,
                             uint32_t H, unsigned long long* __restrict__ table, int32_t* __restrict__ count,
                             int32_t* __restrict__ key, int32_t* __restrict__ val,
                             unsigned long long* __restrict__ status) {
  const uint32_t mask = H - 1;
  bool bad = false;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < total; c += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)((c >> 2) % panels);
    const int64_t left = panel_edges[2 * j], right = panel_edges[2 * j + 1];
    val[c] = (int32_t)c;
    if (left < 0 || left >= width || right < 0 || right >= width || left == right) {
      bad = true;
      key[c] = (int32_t)H;
      continue;
    }
    const RingEdge e = ring_contribution(c, width, panels, panel_edges, strengths);
    const unsigned long long k = (unsigned long long)e.lo * (unsigned long long)nodes + (unsigned long long)e.hi;
    uint32_t h = (uint32_t)mix64(k) & mask;
    while (true) {
      unsigned long long cur = table[h];
      if (cur == EMPTY_KEY) {
        const unsigned long long prev = atomicCAS(&table[h], EMPTY_KEY, k);
        if (prev == EMPTY_KEY || prev == k) break;
        cur = prev;
      }
      if (cur == k) break;
      h = (h + 1) & mask;
    }
    key[c] = (int32_t)h;
    atomicAdd(&count[h], 1);
  }
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(&status[1], 1ull);
}

// One thread per table slot: the group's sum in member order from 0.0; a nonzero sum marks the group's first
// contribution.
__global__ void k_ring_sums(uint32_t H, const int32_t* __restrict__ count, const int32_t* __restrict__ start,
                            const int32_t* __restrict__ members, int64_t width, int panels,
                            const int64_t* __restrict__ panel_edges, const double* __restrict__ strengths,
                            int32_t* __restrict__ mark, double* __restrict__ sums) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < H; h += gridDim.x * blockDim.x) {
    const int size = count[h];
    if (size == 0) continue;
    const int a = start[h];
    double total = 0.0;
    for (int q = a; q < a + size; ++q) total = total + ring_contribution(members[q], width, panels, panel_edges,
                                                                           strengths).value;
    if (total != 0.0) {
      const int first = members[a];
      mark[first] = 1;
      sums[first] = total;
    }
  }
}

__global__ void k_ring_emit(int64_t total, const int32_t* __restrict__ mark, const int32_t* __restrict__ pos,
                            const double* __restrict__ sums, int64_t width, int panels,
                            const int64_t* __restrict__ panel_edges, const double* __restrict__ strengths,
                            int64_t* __restrict__ edges, double* __restrict__ gamma) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < total; c += (int64_t)gridDim.x * blockDim.x) {
    if (!mark[c]) continue;
    const RingEdge e = ring_contribution(c, width, panels, panel_edges, strengths);
    const int64_t o = pos[c];
    edges[2 * o] = e.lo;
    edges[2 * o + 1] = e.hi;
    gamma[o] = sums[c];
  }
}

int pack_ring_edges(Ctx& c, int64_t depth, int64_t width, int64_t panels, const int64_t* panel_edges,
                    const double* strengths, int64_t* edges, double* gamma, int64_t capacity, int64_t* count_out) {
  *count_out = 0;
  if (depth < 2 || width < 2 || panels == 0) return OFW_OK;
  if (depth > INT_MAX || width > INT_MAX || depth * width > INT_MAX || panels > (1 << 28) ||
      4 * (depth - 1) * panels > (1 << 28)) {
    set_error("ofw_pack_ring_edges: more than 2^31 - 1 nodes or 2^28 ring-edge contributions");
    return OFW_ERR_ARG;
  }
  const int T = (int)(4 * (depth - 1) * panels);
  uint32_t H = 2;
  while ((int64_t)H < 2 * (int64_t)T) H <<= 1;
  const int passes = radix_passes(H);  // keys run to H inclusive (H marks a bad panel)
  const int tiles = (T + SORT_TILE - 1) / SORT_TILE;
  const int hist_n = RADIX * tiles;
  const int scan_max = std::max<int>(std::max<int>((int)H, T), hist_n);
  const size_t spart = ((size_t)(scan_max + 4095) / 4096 + 8) * 4;
  const size_t need = al((size_t)H * 8) + 2 * al((size_t)H * 4) + 4 * al((size_t)T * 4) + 2 * al((size_t)T * 4) +
                      al((size_t)T * 8) + 2 * al((size_t)hist_n * 4) + al(spart) + al(16) + al(4);
  OFW_HIP(c.tree[7].reserve(need));
  char* base = (char*)c.tree[7].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  unsigned long long* table = (unsigned long long*)carve((size_t)H * 8);
  int32_t* cnt = (int32_t*)carve((size_t)H * 4);
  int32_t* start = (int32_t*)carve((size_t)H * 4);
  int32_t* key[2] = {(int32_t*)carve((size_t)T * 4), (int32_t*)carve((size_t)T * 4)};
  int32_t* val[2] = {(int32_t*)carve((size_t)T * 4), (int32_t*)carve((size_t)T * 4)};
  int32_t* mark = (int32_t*)carve((size_t)T * 4);
  int32_t* pos = (int32_t*)carve((size_t)T * 4);
  double* sums = (double*)carve((size_t)T * 8);
  int32_t* hist = (int32_t*)carve((size_t)hist_n * 4);
  int32_t* hist_off = (int32_t*)carve((size_t)hist_n * 4);
  int32_t* part = (int32_t*)carve(spart);
  unsigned long long* status = (unsigned long long*)carve(16);
  int32_t* kept = (int32_t*)carve(4);

  const int grid = c.cu_count * 8;
  auto blocks = [&](int64_t items) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (items + 255) / 256)));
  };
  hipLaunchKernelGGL(k_net_fill, blocks(2 * (int64_t)H), dim3(256), 0, c.stream, (int)(2 * H), (int32_t*)table, -1);
  hipLaunchKernelGGL(k_net_fill, blocks(H), dim3(256), 0, c.stream, (int)H, cnt, 0);
  hipLaunchKernelGGL(k_net_fill, blocks(T), dim3(256), 0, c.stream, T, mark, 0);
  hipLaunchKernelGGL(k_net_fill, dim3(1), dim3(64), 0, c.stream, 4, (int32_t*)status, 0);
  hipLaunchKernelGGL(k_ring_group, blocks(T), dim3(256), 0, c.stream, (int64_t)T, width, depth * width, (int)panels,
                     panel_edges, strengths, H, table, cnt, key[0], val[0], status);
  scan_exclusive(c, (int)H, cnt, start, part, nullptr);
  const int32_t* members = val[radix_sort(c, T, passes, key, val, hist, hist_off, part)];
  hipLaunchKernelGGL(k_ring_sums, blocks(H), dim3(256), 0, c.stream, H, cnt, start, members, width, (int)panels,
                     panel_edges, strengths, mark, sums);
  scan_exclusive(c, T, mark, pos, part, kept);
  OFW_HIP(hipGetLastError());
  unsigned long long host[2] = {0, 0};
  int32_t host_kept = 0;
  OFW_HIP(hipMemcpyAsync(host, status, 16, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipMemcpyAsync(&host_kept, kept, 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (host[1]) {
    set_error("ofw_pack_ring_edges: invalid spanwise panel connectivity");
    return OFW_ERR_ARG;
  }
  *count_out = host_kept;
  if (host_kept > capacity) {
    set_error("ofw_pack_ring_edges: the edge buffer is too small (count_out holds the number needed)");
    return OFW_ERR_OVERFLOW;
  }
  if (host_kept == 0) return OFW_OK;
  hipLaunchKernelGGL(k_ring_emit, blocks(T), dim3(256), 0, c.stream, (int64_t)T, mark, pos, sums, width, (int)panels,
                     panel_edges, strengths, edges, gamma);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Material core of every wake edge (cores.py:129-149, edge_radii).  Node id = row * width + column over the rows
// (bound, trailing edge, free rows...).  A spanwise edge (same row r) of a free row (r >= 2) takes
// span[r - 2, panel], the panel whose columns are the edge's (the last such panel, as the reference's dictionary
// keeps the last); a streamwise edge (same column c, rows r and r + 1) with r >= 1 takes stream[r - 1, c]; the
// bound and trailing-edge edges keep `initial`.  status bit 0: an edge that is neither; bit 1: a spanwise edge
// without a panel; bit 2: a row beyond the core arrays.

__global__ void k_panel_table(int panels, int64_t width, const int64_t* __restrict__ panel_edges,
                              int32_t* __restrict__ table) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < panels; j += gridDim.x * blockDim.x) {
    const int64_t a = panel_edges[2 * j], b = panel_edges[2 * j + 1];
    if (a < 0 || b < 0 || a >= width || b >= width) continue;
    atomicMax(&table[(a < b ? a : b) * width + (a < b ? b : a)], j);
  }
}

__global__ void k_edge_radii(int m, const int64_t* __restrict__ edges, int64_t offset, int64_t width, int panels,
                             const int32_t* __restrict__ table, int64_t span_rows, const double* __restrict__ span,
                             int64_t stream_rows, const double* __restrict__ stream, double initial,
                             double* __restrict__ out, int32_t* __restrict__ status) {
  int bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int64_t e0 = edges[2 * (size_t)i] + offset, e1 = edges[2 * (size_t)i + 1] + offset;
    const int64_t ra = e0 / width, a = e0 % width, rb = e1 / width, b = e1 % width;
    double r = initial;
    if (e0 < 0 || e1 < 0) {
      bad |= 1;
    } else if (ra == rb) {
      if (ra >= 2) {
        const int j = table[(a < b ? a : b) * width + (a < b ? b : a)];
        if (j < 0) {
          bad |= 2;
        } else if (ra - 2 >= span_rows) {
          bad |= 4;
        } else {
          r = span[(ra - 2) * panels + j];
        }
      }
    } else if (a == b && (rb - ra == 1 || ra - rb == 1)) {
      const int64_t row = ra < rb ? ra : rb;
      if (row >= 1) {
        if (row - 1 >= stream_rows)
          bad |= 4;
        else
          r = stream[(row - 1) * width + a];
      }
    } else {
      bad |= 1;
    }
    out[i] = r;
  }
  if (__builtin_amdgcn_ballot_w64(bad != 0) != 0) {
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o);
    if ((threadIdx.x & 63) == 0) atomicOr(status, bad);
  }
}

}  // namespace ofw

extern "C" int ofw_edge_radii(ofw_ctx* x, int64_t m, const int64_t* edges, int64_t offset, int64_t width,
                              int64_t panels, const int64_t* panel_edges, int64_t span_rows, const double* span,
                              int64_t stream_rows, const double* stream, double initial, double* out) {
  if (!x || m < 0 || m > ((int64_t)1 << 30) || width < 1 || width > 46340 || panels < 0 || panels > INT_MAX ||
      span_rows < 0 || stream_rows < 0 || (m > 0 && (!edges || !out)) || (panels > 0 && !panel_edges) ||
      (span_rows > 0 && !span) || (stream_rows > 0 && !stream)) {
    ofw::set_error("ofw_edge_radii: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (m == 0) return OFW_OK;
  const size_t cells = (size_t)(width * width);
  OFW_HIP(c.tree[9].reserve(ofw::al(cells * 4) + ofw::al(4)));
  int32_t* table = (int32_t*)c.tree[9].ptr;
  int32_t* status = (int32_t*)((char*)c.tree[9].ptr + ofw::al(cells * 4));
  const int grid = c.cu_count * 8;
  auto blocks = [&](int64_t items) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (items + 255) / 256)));
  };
  hipLaunchKernelGGL(ofw::k_net_fill, blocks((int64_t)cells), dim3(256), 0, c.stream, (int)cells, table, -1);
  hipLaunchKernelGGL(ofw::k_net_fill, dim3(1), dim3(64), 0, c.stream, 1, status, 0);
  if (panels > 0)
    hipLaunchKernelGGL(ofw::k_panel_table, blocks(panels), dim3(256), 0, c.stream, (int)panels, width, panel_edges,
                       table);
  hipLaunchKernelGGL(ofw::k_edge_radii, blocks(m), dim3(256), 0, c.stream, (int)m, edges, offset, width,
                     (int)panels, table, span_rows, span, stream_rows, stream, initial, out, status);
  OFW_HIP(hipGetLastError());
  int32_t host = 0;
  OFW_HIP(hipMemcpyAsync(&host, status, 4, hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (host & 1) {
    ofw::set_error("edge is not a spanwise or streamwise wake segment");
    return OFW_ERR_ARG;
  }
  if (host & 2) {
    ofw::set_error("spanwise wake edge without a panel");
    return OFW_ERR_ARG;
  }
  if (host & 4) {
    ofw::set_error("wake edge beyond the core history");
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

extern "C" int ofw_pack_ring_edges(ofw_ctx* x, int64_t depth, int64_t width, int64_t panels,
                                   const int64_t* panel_edges, const double* strengths, int64_t* edges,
                                   double* gamma, int64_t capacity, int64_t* count_out) {
  if (count_out) *count_out = 0;
  if (!x || !count_out || depth < 0 || width < 0 || panels < 0 || capacity < 0 ||
      (depth >= 2 && width >= 2 && panels > 0 && (!panel_edges || !strengths)) ||
      (capacity > 0 && (!edges || !gamma))) {
    ofw::set_error("ofw_pack_ring_edges: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  OFW_HIP(c.mark());
  const int st = ofw::pack_ring_edges(c, depth, width, panels, panel_edges, strengths, edges, gamma, capacity,
                                      count_out);
  if (st) return st;
  OFW_HIP(c.mark());
  return OFW_OK;
}

extern "C" int ofw_incident_pairs(ofw_ctx* x, int64_t n, const double* nodes, int64_t m, const int64_t* edges,
                                  int64_t first, int64_t* pairs, int64_t capacity, int64_t* count_out) {
  if (count_out) *count_out = 0;
  if (!x || !count_out || n < 0 || m < 0 || first < 0 || first > n || capacity < 0 || (n > 0 && !nodes) ||
      (m > 0 && !edges) || (capacity > 0 && !pairs)) {
    ofw::set_error("ofw_incident_pairs: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  OFW_HIP(c.mark());
  const int st = ofw::incident_pairs(c, n, nodes, m, edges, first, pairs, capacity, count_out);
  if (st) return st;
  OFW_HIP(c.mark());
  return OFW_OK;
}
