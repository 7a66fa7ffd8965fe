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
//   3. k_end_groups: each segment end finds its group with one probe; the pair counts are scanned.
//   4. k_write_pairs: one thread per output pair finds its segment end by binary search over the offsets.
#include <climits>

#include "common.h"
#include "internal.h"
#include "scan.h"
#include "tree.h"

namespace ofw {

namespace {

constexpr int RADIX_BITS = 8, RADIX = 1 << RADIX_BITS;
constexpr int SORT_THREADS = 256, SORT_PER = 16, SORT_TILE = SORT_THREADS * SORT_PER;
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
                               int32_t* __restrict__ val) {
  const uint32_t mask = H - 1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nfree; i += gridDim.x * blockDim.x) {
    const int node = first + i;
    val[i] = i;
    if (has_nan(nodes, node)) {
      key[i] = (int32_t)H;
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
    atomicAdd(&count[h], 1);
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

// Per segment end q = 2 s + e: its group (-1 when none) and the number of pairs it gives.  status[0] sums the
// pairs over all ends; status[1] is set when an edge names a node outside [0, n).
__global__ void k_end_groups(int m, int n, const int64_t* __restrict__ edges, const double* __restrict__ nodes,
                             uint32_t H, const int32_t* __restrict__ table, const int32_t* __restrict__ count,
                             int32_t* __restrict__ end_group, int32_t* __restrict__ end_count,
                             unsigned long long* __restrict__ status) {
  const uint32_t mask = H - 1;
  unsigned long long mine = 0;
  bool bad = false;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * m; q += gridDim.x * blockDim.x) {
    const int64_t e = edges[q];
    int g = -1;
    if (e < 0 || e >= n) {
      bad = true;
    } else if (!has_nan(nodes, (int)e)) {
      const NodeBits bits = node_bits(nodes, (int)e);
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
    end_group[q] = g;
    const int c = g >= 0 ? count[g] : 0;
    end_count[q] = c;
    mine += (unsigned long long)c;
  }
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
  const bool lead = (threadIdx.x & 63) == 0;
  if (lead && mine) atomicAdd(&status[0], mine);
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && lead) atomicOr(&status[1], 1ull);
}

// One thread per pair p: its end q is the last one with offset[q] <= p (the ends after q start beyond p, so
// end_count[q] > 0 and p lies inside q's range).
__global__ void k_write_pairs(int total, int ends, const int32_t* __restrict__ end_group,
                              const int32_t* __restrict__ offset, const int32_t* __restrict__ start,
                              const int32_t* __restrict__ members, int64_t* __restrict__ pairs) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
    int lo = 0, hi = ends - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offset[mid] <= p)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int g = end_group[lo];
    pairs[2 * (int64_t)p] = members[start[g] + (p - offset[lo])];
    pairs[2 * (int64_t)p + 1] = lo >> 1;
  }
}

int incident_pairs(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, int64_t first,
                   int64_t* pairs, int64_t capacity, int64_t* count_out) {
  *count_out = 0;
  const int64_t nfree = n - first;
  if (nfree <= 0 || m <= 0) return OFW_OK;
  if (n > INT_MAX / 4 || m > INT_MAX / 4) {
    set_error("ofw_incident_pairs: more than 2^29 nodes or segments");
    return OFW_ERR_ARG;
  }
  uint32_t H = 2;
  while ((int64_t)H < 2 * nfree) H <<= 1;
  int key_bits = 1;  // keys run to H inclusive
  while ((1ll << key_bits) <= (int64_t)H) ++key_bits;
  const int passes = (key_bits + RADIX_BITS - 1) / RADIX_BITS;
  const int F = (int)nfree, E = (int)(2 * m);
  const int tiles = (F + SORT_TILE - 1) / SORT_TILE;
  const int hist_n = RADIX * tiles;
  const int scan_max = std::max<int>(std::max<int>((int)H, E), hist_n);
  const size_t spart = ((size_t)(scan_max + 4095) / 4096 + 8) * 4;
  const size_t need = 3 * al((size_t)H * 4) + 4 * al((size_t)F * 4) + 3 * al((size_t)E * 4) +
                      2 * al((size_t)hist_n * 4) + al(spart) + al(16);
  OFW_HIP(c.tree[6].reserve(need));
  char* base = (char*)c.tree[6].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  int32_t* table = (int32_t*)carve((size_t)H * 4);
  int32_t* cnt = (int32_t*)carve((size_t)H * 4);
  int32_t* start = (int32_t*)carve((size_t)H * 4);
  int32_t* key[2] = {(int32_t*)carve((size_t)F * 4), (int32_t*)carve((size_t)F * 4)};
  int32_t* val[2] = {(int32_t*)carve((size_t)F * 4), (int32_t*)carve((size_t)F * 4)};
  int32_t* end_group = (int32_t*)carve((size_t)E * 4);
  int32_t* end_count = (int32_t*)carve((size_t)E * 4);
  int32_t* offset = (int32_t*)carve((size_t)E * 4);
  int32_t* hist = (int32_t*)carve((size_t)hist_n * 4);
  int32_t* hist_off = (int32_t*)carve((size_t)hist_n * 4);
  int32_t* part = (int32_t*)carve(spart);
  unsigned long long* status = (unsigned long long*)carve(16);

  const int grid = c.cu_count * 8;
  auto blocks = [&](int64_t items) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (items + 255) / 256))); };
  hipLaunchKernelGGL(k_net_fill, blocks(H), dim3(256), 0, c.stream, (int)H, table, -1);
  hipLaunchKernelGGL(k_net_fill, blocks(H), dim3(256), 0, c.stream, (int)H, cnt, 0);
  hipLaunchKernelGGL(k_net_fill, dim3(1), dim3(64), 0, c.stream, 4, (int32_t*)status, 0);
  hipLaunchKernelGGL(k_group_points, blocks(F), dim3(256), 0, c.stream, F, (int)first, nodes, H, table, cnt, key[0],
                     val[0]);
  scan_exclusive(c, (int)H, cnt, start, part, nullptr);
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(SORT_THREADS), 0, c.stream, F, p * RADIX_BITS, key[cur],
                       hist);
    scan_exclusive(c, hist_n, hist, hist_off, part, nullptr);
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(SORT_THREADS), 0, c.stream, F, p * RADIX_BITS, key[cur],
                       val[cur], hist_off, key[cur ^ 1], val[cur ^ 1]);
    cur ^= 1;
  }
  const int32_t* members = val[cur];
  hipLaunchKernelGGL(k_end_groups, blocks(E), dim3(256), 0, c.stream, (int)m, (int)n, edges, nodes, H, table, cnt,
                     end_group, end_count, status);
  OFW_HIP(hipGetLastError());
  unsigned long long host[2] = {0, 0};
  OFW_HIP(hipMemcpyAsync(host, status, 16, hipMemcpyDeviceToHost, c.stream));
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
  scan_exclusive(c, E, end_count, offset, part, nullptr);
  hipLaunchKernelGGL(k_write_pairs, blocks((int64_t)host[0]), dim3(256), 0, c.stream, (int)host[0], E, end_group,
                     offset, start, members, pairs);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

}  // namespace ofw

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
