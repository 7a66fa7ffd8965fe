// Node quantities of the tree on the GPU: centre, bounding radius, largest core, moments m0, m1, m2 and
// their contractions (vortex_tree.py:118-167 node_moments, :425-447 node_contractions).
//
// Segments are first permuted into tree order, so that every node's segments are one contiguous range and
// every pass reads HBM with coalesced loads.  Each node's range is cut into items of at most ITEM segments
// (planned on the device with one scan); one wavefront reduces one item (lanes stride, then a fixed
// butterfly), and one wavefront per node combines the node's items (lanes stride, fixed butterfly).  Box,
// centre and core are exact (min / max), so they equal the reference's bit for bit; the bounding radius is
// the square root of an exact maximum of squared distances computed operation by operation as the reference
// does, so it is bit-exact too.  The moments are sums in another order: equal to the reference to rounding.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree.h"
#include "tree_eval.h"
#include "tree_records.h"

namespace ofw {

static constexpr int ITEM = 1024;
static constexpr int MPARTS = 31;  // r2, m0 (3), m1 (9), m2 (18 entries j <= l)

// vortex_tree.py:608-611, elementwise in FP64 without contraction: the midpoints the tree is built on
__global__ __launch_bounds__(256) void k_tree_prep(int m, const double* __restrict__ starts,
                                                   const double* __restrict__ directions,
                                                   const double* __restrict__ lengths, double* __restrict__ mx,
                                                   double* __restrict__ my, double* __restrict__ mz) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const double L = lengths[s];
  double mid[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double st = starts[3 * (size_t)s + c];
    const double lv = directions[3 * (size_t)s + c] * L;
    const double en = st + lv;
    mid[c] = 0.5 * (st + en);
  }
  mx[s] = mid[0];
  my[s] = mid[1];
  mz[s] = mid[2];
}

// Tree-order copies: the moment inputs (start, end, alpha, length vector, radius, gamma) and the pair
// records of the traversal's near field.
__global__ __launch_bounds__(256) void k_tree_permute(int m, const int32_t* __restrict__ order,
                                                      const double* __restrict__ starts,
                                                      const double* __restrict__ directions,
                                                      const double* __restrict__ lengths,
                                                      const double* __restrict__ gamma,
                                                      const double* __restrict__ radius, int model,
                                                      MomentInputs mi, double* __restrict__ rec64,
                                                      float* __restrict__ rec32) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int s = order[k];
  const double L = lengths[s], g = gamma[s];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double st = starts[3 * (size_t)s + c];
    const double lv = directions[3 * (size_t)s + c] * L;
    mi.s[c][k] = st;
    mi.e[c][k] = st + lv;
    mi.a[c][k] = g * lv;
    mi.l[c][k] = lv;
  }
  mi.rad[k] = radius[s];
  mi.gam[k] = g;
  pack_record(s, starts, directions, lengths, gamma, radius, model, rec64 + (size_t)k * REC64,
              rec32 + (size_t)k * REC32);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// ---------------------------------------------------------------- item plan
__global__ void k_items_count(TreeArrays t, int32_t* __restrict__ cnt) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < t.cap; p += gridDim.x * blockDim.x) {
    int c = 0;
    if (p < count) c = max(1, (t.p_hi[p] - t.p_lo[p] + ITEM - 1) / ITEM);
    cnt[p] = c;
  }
}

__global__ void k_items_fill(TreeArrays t, const int32_t* __restrict__ first, int item_cap,
                             int32_t* __restrict__ inode, int32_t* __restrict__ ip0, int32_t* __restrict__ ip1) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += gridDim.x * blockDim.x) {
    const int a = t.p_lo[p], b = t.p_hi[p];
    const int n = max(1, (b - a + ITEM - 1) / ITEM);
    const int f = first[p];
    if (f + n > item_cap) {
      t.counters[C_OVERFLOW] = 1;
      continue;
    }
    for (int q = 0; q < n; ++q) {
      inode[f + q] = p;
      ip0[f + q] = a + q * ITEM;
      ip1[f + q] = min(b, a + (q + 1) * ITEM);
    }
    if (p == count - 1) t.counters[C_ITEMS] = f + n;
  }
}

// ---------------------------------------------------------------- passes
// pass 1: box of start and end points, largest core (7 values per item)
__global__ __launch_bounds__(256) void k_moments_box(const int32_t* __restrict__ counters,
                                                     const int32_t* __restrict__ item_p0,
                                                     const int32_t* __restrict__ item_p1, MomentInputs mi,
                                                     double* __restrict__ part) {
  const int items = live_count(counters, C_ITEMS);
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * 4 + threadIdx.x / 64; w < items; w += gridDim.x * 4) {
    const int p0 = item_p0[w], p1 = item_p1[w];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double big = 0.0;
    for (int k = p0 + lane; k < p1; k += 64) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double a = mi.s[c][k], b = mi.e[c][k];
        lo[c] = a < lo[c] ? a : lo[c];
        hi[c] = a > hi[c] ? a : hi[c];
        lo[c] = b < lo[c] ? b : lo[c];
        hi[c] = b > hi[c] ? b : hi[c];
      }
      const double r = mi.rad[k];
      big = r > big ? r : big;
    }
    double out[7];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      out[c] = wave_min(lo[c]);
      out[3 + c] = wave_max(hi[c]);
    }
    out[6] = wave_max(big);
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 7; ++q) part[7 * (size_t)w + q] = out[q];
  }
}

// one wavefront per node: combine its item boxes -> centre and largest core
__global__ __launch_bounds__(256) void k_moments_centre(const int32_t* __restrict__ counters,
                                                        const int32_t* __restrict__ first,
                                                        const int32_t* __restrict__ ncount,
                                                        const double* __restrict__ part, double* __restrict__ centre,
                                                        double* __restrict__ core) {
#pragma clang fp contract(off)
  const int count = live_count(counters, C_FINAL);
  const int lane = threadIdx.x & 63;
  for (int node = blockIdx.x * 4 + threadIdx.x / 64; node < count; node += gridDim.x * 4) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double big = 0.0;
    for (int it = first[node] + lane; it < first[node] + ncount[node]; it += 64) {
      const double* p = part + 7 * (size_t)it;
      for (int c = 0; c < 3; ++c) {
        lo[c] = p[c] < lo[c] ? p[c] : lo[c];
        hi[c] = p[3 + c] > hi[c] ? p[3 + c] : hi[c];
      }
      big = p[6] > big ? p[6] : big;
    }
    for (int c = 0; c < 3; ++c) {
      lo[c] = wave_min(lo[c]);
      hi[c] = wave_max(hi[c]);
    }
    big = wave_max(big);
    if (lane == 0) {
      for (int c = 0; c < 3; ++c) centre[3 * (size_t)node + c] = 0.5 * (lo[c] + hi[c]);
      core[node] = big;
    }
  }
}

// pass 2: squared bounding radius and the moments about the node centre
__global__ __launch_bounds__(256) void k_moments_sum(const int32_t* __restrict__ counters,
                                                     const int32_t* __restrict__ item_node,
                                                     const int32_t* __restrict__ item_p0,
                                                     const int32_t* __restrict__ item_p1, MomentInputs mi,
                                                     const double* __restrict__ centre, double* __restrict__ part) {
  const int items = live_count(counters, C_ITEMS);
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * 4 + threadIdx.x / 64; w < items; w += gridDim.x * 4) {
    const int node = item_node[w];
    const int p0 = item_p0[w], p1 = item_p1[w];
    const double c0 = centre[3 * (size_t)node], c1 = centre[3 * (size_t)node + 1], c2 = centre[3 * (size_t)node + 2];
    double r2 = 0.0;
    double acc[30];
#pragma unroll
    for (int q = 0; q < 30; ++q) acc[q] = 0.0;
    for (int k = p0 + lane; k < p1; k += 64) {
      {
#pragma clang fp contract(off)
        const double s0 = mi.s[0][k] - c0, s1 = mi.s[1][k] - c1, s2 = mi.s[2][k] - c2;
        const double qs = s0 * s0 + s1 * s1 + s2 * s2;
        const double e0 = mi.e[0][k] - c0, e1 = mi.e[1][k] - c1, e2 = mi.e[2][k] - c2;
        const double qe = e0 * e0 + e1 * e1 + e2 * e2;
        r2 = qs > r2 ? qs : r2;
        r2 = qe > r2 ? qe : r2;
      }
      if (mi.gam[k] == 0.0) continue;
      const double d[3] = {0.5 * (mi.s[0][k] + mi.e[0][k]) - c0, 0.5 * (mi.s[1][k] + mi.e[1][k]) - c1,
                           0.5 * (mi.s[2][k] + mi.e[2][k]) - c2};
      const double al[3] = {mi.a[0][k], mi.a[1][k], mi.a[2][k]};
      const double lv[3] = {mi.l[0][k], mi.l[1][k], mi.l[2][k]};
      double dd[6];
      int q = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int l = j; l < 3; ++l) dd[q++] = d[j] * d[l] + lv[j] * lv[l] / 12.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        acc[i] += al[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[3 + 3 * i + j] += al[i] * d[j];
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[12 + 6 * i + u] += al[i] * dd[u];
      }
    }
    r2 = wave_max(r2);
    double* out = part + MPARTS * (size_t)w;
    if (lane == 0) out[0] = r2;
#pragma unroll
    for (int q = 0; q < 30; ++q) {
      const double v = wave_sum(acc[q]);
      if (lane == 0) out[1 + q] = v;
    }
  }
}

__device__ __forceinline__ int sym6(int j, int l) {
  const int a = j < l ? j : l, b = j < l ? l : j;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// One wavefront per node: combine its item partials (lanes stride, fixed butterfly); contractions
// (vortex_tree.py:425-447); full outputs (optional) and the packed traversal records.
__global__ __launch_bounds__(256) void k_moments_final(TreeArrays t, const int32_t* __restrict__ first,
                                                       const int32_t* __restrict__ ncount,
                                                       const double* __restrict__ part,
                                                       const double* __restrict__ centre,
                                                       const double* __restrict__ core, double clearance,
                                                       NodeOutputs no, double* __restrict__ acc_rec,
                                                       double* __restrict__ far64, float* __restrict__ far32) {
  const int count = live_count(t.counters, C_FINAL);
  const int lane = threadIdx.x & 63;
  for (int node = blockIdx.x * 4 + threadIdx.x / 64; node < count; node += gridDim.x * 4) {
    double r2 = 0.0, v[30];
    for (int q = 0; q < 30; ++q) v[q] = 0.0;
    for (int it = first[node] + lane; it < first[node] + ncount[node]; it += 64) {
      const double* p = part + MPARTS * (size_t)it;
      r2 = p[0] > r2 ? p[0] : r2;
      for (int q = 0; q < 30; ++q) v[q] += p[1 + q];
    }
    r2 = wave_max(r2);
    for (int q = 0; q < 30; ++q) v[q] = wave_sum(v[q]);
    if (lane != 0) continue;
    double bound;
    {
#pragma clang fp contract(off)
      bound = __builtin_sqrt(r2) * (1.0 + 1e-12);
    }
    const double* m0 = v;
    const double* m1 = v + 3;
    const double* m2s = v + 12;
    double m2[27];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) m2[9 * i + 3 * j + l] = m2s[6 * i + sym6(j, l)];
    double w1[3], t2[3], w2[9];
    w1[0] = m1[3 * 1 + 2] - m1[3 * 2 + 1];
    w1[1] = m1[3 * 2 + 0] - m1[3 * 0 + 2];
    w1[2] = m1[3 * 0 + 1] - m1[3 * 1 + 0];
    for (int i = 0; i < 3; ++i) t2[i] = m2[9 * i + 0] + m2[9 * i + 4] + m2[9 * i + 8];
    for (int k = 0; k < 3; ++k) {
      w2[0 * 3 + k] = m2[9 * 1 + 3 * 2 + k] - m2[9 * 2 + 3 * 1 + k];
      w2[1 * 3 + k] = m2[9 * 2 + 3 * 0 + k] - m2[9 * 0 + 3 * 2 + k];
      w2[2 * 3 + k] = m2[9 * 0 + 3 * 1 + k] - m2[9 * 1 + 3 * 0 + k];
    }
    if (no.bound) {
      no.bound[node] = bound;
      for (int q = 0; q < 3; ++q) no.m0[3 * (size_t)node + q] = m0[q];
      for (int q = 0; q < 9; ++q) no.m1[9 * (size_t)node + q] = m1[q];
      for (int q = 0; q < 27; ++q) no.m2[27 * (size_t)node + q] = m2[q];
      for (int q = 0; q < 3; ++q) no.w1[3 * (size_t)node + q] = w1[q];
      for (int q = 0; q < 3; ++q) no.t2[3 * (size_t)node + q] = t2[q];
      for (int q = 0; q < 9; ++q) no.w2[9 * (size_t)node + q] = w2[q];
    }
    double* a = acc_rec + (size_t)node * ACC_REC;
    a[A_CX] = centre[3 * (size_t)node];
    a[A_CY] = centre[3 * (size_t)node + 1];
    a[A_CZ] = centre[3 * (size_t)node + 2];
    a[A_BOUND] = bound;
    a[A_CLEAR] = clearance * core[node];
    int32_t* ai = reinterpret_cast<int32_t*>(a + A_INTS);
    ai[0] = t.p_lo[node];
    ai[1] = t.p_hi[node];
    ai[2] = t.p_skip[node];
    ai[3] = t.p_tree[node];
    a[A_PAD] = 0.0;
    double* f = far64 + (size_t)node * FAR_REC;
    float* g = far32 + (size_t)node * FAR_REC;
    double rec[FAR_REC];
    for (int q = 0; q < 3; ++q) rec[FR_M0 + q] = m0[q];
    for (int q = 0; q < 9; ++q) rec[FR_M1 + q] = m1[q];
    for (int q = 0; q < 3; ++q) rec[FR_W1 + q] = w1[q];
    // symmetric second moments: the diagonal (j == l) once, the off-diagonal doubled
    for (int i = 0; i < 3; ++i)
      for (int u = 0; u < 6; ++u) {
        const bool diag = (u == 0 || u == 3 || u == 5);
        rec[FR_M2 + 6 * i + u] = diag ? m2s[6 * i + u] : 2.0 * m2s[6 * i + u];
      }
    for (int q = 0; q < 3; ++q) rec[FR_T2 + q] = t2[q];
    for (int q = 0; q < 9; ++q) rec[FR_W2 + q] = w2[q];
    rec[FR_PAD0] = rec[FR_PAD1] = rec[FR_PAD2] = 0.0;
    for (int q = 0; q < FAR_REC; ++q) {
      f[q] = rec[q];
      g[q] = (float)rec[q];
    }
  }
}

int tree_moments(Ctx& c, int buf, const TreeArrays& t, const double* starts, const double* directions,
                 const double* lengths, const double* gamma, const double* radius, int model, double clearance,
                 TreeDevice& td, bool want_outputs) {
  const int m = (int)t.m, cap = t.cap;
  int depth_bound = 2;
  for (int64_t s = std::max<int64_t>(m, 1); s > 1; s = (s + 1) / 2) ++depth_bound;
  const int64_t item_cap64 = (int64_t)m * depth_bound / ITEM + 2 * (int64_t)cap + 64;
  const int item_cap = (int)std::min<int64_t>(item_cap64, INT_MAX / 64);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t M = (size_t)std::max(m, 1);
  size_t need = 14 * al(M * 8) + al(M * REC64 * 8) + al(M * REC32 * 4);
  need += 3 * al((size_t)item_cap * 4) + 2 * al((size_t)cap * 4);
  need += al((size_t)item_cap * MPARTS * 8) + al((size_t)cap * 3 * 8) + al((size_t)cap * 8);
  need += al((size_t)cap * ACC_REC * 8) + al((size_t)cap * FAR_REC * 8) + al((size_t)cap * FAR_REC * 4);
  OFW_HIP(c.tree[buf].reserve(need));
  char* base = (char*)c.tree[buf].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  MomentInputs mi;
  for (int q = 0; q < 3; ++q) mi.s[q] = (double*)carve(M * 8);
  for (int q = 0; q < 3; ++q) mi.e[q] = (double*)carve(M * 8);
  for (int q = 0; q < 3; ++q) mi.a[q] = (double*)carve(M * 8);
  for (int q = 0; q < 3; ++q) mi.l[q] = (double*)carve(M * 8);
  mi.rad = (double*)carve(M * 8);
  mi.gam = (double*)carve(M * 8);
  td.rec64 = (double*)carve(M * REC64 * 8);
  td.rec32 = (float*)carve(M * REC32 * 4);
  int32_t* d_inode = (int32_t*)carve((size_t)item_cap * 4);
  int32_t* d_ip0 = (int32_t*)carve((size_t)item_cap * 4);
  int32_t* d_ip1 = (int32_t*)carve((size_t)item_cap * 4);
  int32_t* d_first = (int32_t*)carve((size_t)cap * 4);
  int32_t* d_ncount = (int32_t*)carve((size_t)cap * 4);
  double* d_part = (double*)carve((size_t)item_cap * MPARTS * 8);
  td.centre = (double*)carve((size_t)cap * 3 * 8);
  td.core = (double*)carve((size_t)cap * 8);
  td.acc = (double*)carve((size_t)cap * ACC_REC * 8);
  td.far64 = (double*)carve((size_t)cap * FAR_REC * 8);
  td.far32 = (float*)carve((size_t)cap * FAR_REC * 4);
  td.count = cap;
  const int grid = c.cu_count * 4;
  if (m > 0)
    hipLaunchKernelGGL(k_tree_permute, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, t.order, starts, directions,
                       lengths, gamma, radius, model, mi, td.rec64, td.rec32);
  hipLaunchKernelGGL(k_items_count, dim3(grid), dim3(256), 0, c.stream, t, d_ncount);
  scan_exclusive(c, cap, d_ncount, d_first, t.scan_part, nullptr);
  hipLaunchKernelGGL(k_items_fill, dim3(grid), dim3(256), 0, c.stream, t, d_first, item_cap, d_inode, d_ip0, d_ip1);
  hipLaunchKernelGGL(k_moments_box, dim3(grid), dim3(256), 0, c.stream, t.counters, d_ip0, d_ip1, mi, d_part);
  hipLaunchKernelGGL(k_moments_centre, dim3(grid), dim3(256), 0, c.stream, t.counters, d_first, d_ncount, d_part,
                     td.centre, td.core);
  hipLaunchKernelGGL(k_moments_sum, dim3(grid), dim3(256), 0, c.stream, t.counters, d_inode, d_ip0, d_ip1, mi,
                     td.centre, d_part);
  NodeOutputs no = want_outputs ? td.out : NodeOutputs{};
  hipLaunchKernelGGL(k_moments_final, dim3(grid), dim3(256), 0, c.stream, t, d_first, d_ncount, d_part, td.centre,
                     td.core, clearance, no, td.acc, td.far64, td.far32);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

int tree_prep(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths, double* mx,
              double* my, double* mz) {
  if (m == 0) return OFW_OK;
  hipLaunchKernelGGL(k_tree_prep, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, starts,
                     directions, lengths, mx, my, mz);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

}  // namespace ofw
