// Node quantities of the tree on the GPU: centre, bounding radius, largest core, moments m0, m1, m2 and
// their contractions (vortex_tree.py:118-167 node_moments, :425-447 node_contractions).
//
// Segments are first permuted into tree order, so that every node's segments are one contiguous range.
//  * Leaves: one thread per leaf computes its box, centre, core, squared bound and moments from its
//    segments, in the reference's operation order (so a leaf's quantities equal the reference's bit for bit).
//  * Internal nodes, bottom-up, one launch per depth: a thread computes a node from its two children -- box
//    and core by exact min / max (bit-exact), moments shifted to the node's centre by the parallel-axis
//    identity (equal to the reference's direct sums to rounding).
//  * The bounding radius of an internal node needs its own centre and every end point below it: its range is
//    cut into items of at most ITEM segments (planned on the device with one scan), one wavefront takes the
//    exact maximum over an item, and one wavefront per node combines the items.  The square root of an exact
//    maximum of squared distances computed operation by operation as the reference does: bit-exact.
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

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// ---------------------------------------------------------------- item plan (internal nodes)
__device__ __forceinline__ bool is_leaf(const TreeArrays& t, int p) { return t.p_skip[p] == p + 1; }

__global__ void k_items_count(TreeArrays t, int32_t* __restrict__ cnt) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < t.cap; p += gridDim.x * blockDim.x) {
    int c = 0;
    if (p < count && !is_leaf(t, p)) c = (t.p_hi[p] - t.p_lo[p] + ITEM - 1) / ITEM;
    cnt[p] = c;
  }
}

__global__ void k_items_fill(TreeArrays t, const int32_t* __restrict__ first, int item_cap,
                             int32_t* __restrict__ inode, int32_t* __restrict__ ip0, int32_t* __restrict__ ip1) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += gridDim.x * blockDim.x) {
    const int a = t.p_lo[p], b = t.p_hi[p];
    const int n = is_leaf(t, p) ? 0 : (b - a + ITEM - 1) / ITEM;
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

// Per-node scratch of the bottom-up pass (pre-order index): box of the end points (low xyz, high xyz),
// moments (m0 3, m1 9, m2 as 3 x 6 entries j <= l), squared bound of a leaf.
struct NodeScratch {
  double* box;
  double* mom;
  double* r2;
};

// The parent's moments from its two children's, about its own centre: with delta = c_child - c_parent,
// m1 += m0 (x) delta, m2_ijl += m1_ij delta_l + m1_il delta_j + m0_i delta_j delta_l (the length terms do
// not depend on the centre).
__device__ __forceinline__ void shift_add(const double* __restrict__ cm, const double* dl, double* pm) {
#pragma clang fp contract(off)
  const int jj[6] = {0, 0, 0, 1, 1, 2}, ll[6] = {0, 1, 2, 1, 2, 2};
  for (int i = 0; i < 3; ++i) {
    const double a0 = cm[i];
    pm[i] += a0;
    for (int j = 0; j < 3; ++j) pm[3 + 3 * i + j] += cm[3 + 3 * i + j] + a0 * dl[j];
    for (int u = 0; u < 6; ++u) {
      const int j = jj[u], l = ll[u];
      pm[12 + 6 * i + u] += cm[12 + 6 * i + u] + cm[3 + 3 * i + j] * dl[l] + cm[3 + 3 * i + l] * dl[j] +
                            a0 * (dl[j] * dl[l]);
    }
  }
}

// One thread per leaf: its box, centre, largest core, squared bound and moments from its segments in tree
// order, operation by operation as node_moments does (vortex_tree.py:118-167).
__global__ __launch_bounds__(256) void k_leaves(TreeArrays t, MomentInputs mi, NodeScratch ns,
                                                double* __restrict__ centre, double* __restrict__ core) {
#pragma clang fp contract(off)
  const int count = live_count(t.counters, C_FINAL);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count || !is_leaf(t, p)) return;
  const int a = t.p_lo[p], b = t.p_hi[p];
  double low[3] = {INFINITY, INFINITY, INFINITY}, high[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = a; k < b; ++k)
    for (int axis = 0; axis < 3; ++axis) {
      const double v0 = mi.s[axis][k], v1 = mi.e[axis][k];
      if (v0 < low[axis]) low[axis] = v0;
      if (v0 > high[axis]) high[axis] = v0;
      if (v1 < low[axis]) low[axis] = v1;
      if (v1 > high[axis]) high[axis] = v1;
    }
  double c[3];
  for (int axis = 0; axis < 3; ++axis) c[axis] = 0.5 * (low[axis] + high[axis]);
  double m[30];
  for (int q = 0; q < 30; ++q) m[q] = 0.0;
  double r2 = 0.0, largest = 0.0;
  for (int k = a; k < b; ++k) {
    const double s0 = mi.s[0][k] - c[0], s1 = mi.s[1][k] - c[1], s2 = mi.s[2][k] - c[2];
    const double qs = s0 * s0 + s1 * s1 + s2 * s2;
    if (qs > r2) r2 = qs;
    const double e0 = mi.e[0][k] - c[0], e1 = mi.e[1][k] - c[1], e2 = mi.e[2][k] - c[2];
    const double qe = e0 * e0 + e1 * e1 + e2 * e2;
    if (qe > r2) r2 = qe;
    const double rad = mi.rad[k];
    if (rad > largest) largest = rad;
    if (mi.gam[k] == 0.0) continue;
    double d[3], al[3], lv[3];
    for (int axis = 0; axis < 3; ++axis) {
      d[axis] = 0.5 * (mi.s[axis][k] + mi.e[axis][k]) - c[axis];
      al[axis] = mi.a[axis][k];
      lv[axis] = mi.l[axis][k];
    }
    for (int i = 0; i < 3; ++i) {
      m[i] += al[i];
      for (int j = 0; j < 3; ++j) m[3 + 3 * i + j] += al[i] * d[j];
      int u = 0;
      for (int j = 0; j < 3; ++j)
        for (int l = j; l < 3; ++l) m[12 + 6 * i + u++] += al[i] * (d[j] * d[l] + lv[j] * lv[l] / 12.0);
    }
  }
  for (int q = 0; q < 3; ++q) {
    ns.box[6 * (size_t)p + q] = low[q];
    ns.box[6 * (size_t)p + 3 + q] = high[q];
    centre[3 * (size_t)p + q] = c[q];
  }
  for (int q = 0; q < 30; ++q) ns.mom[30 * (size_t)p + q] = m[q];
  core[p] = largest;
  ns.r2[p] = r2;
}

// The internal nodes of one depth from their two children (computed by the previous launch or by k_leaves):
// box and core exactly by min / max, moments by shift_add, in the fixed order left then right.
__global__ __launch_bounds__(256) void k_internal_level(TreeArrays t, int depth, NodeScratch ns,
                                                        double* __restrict__ centre, double* __restrict__ core) {
#pragma clang fp contract(off)
  const int count = live_count(t.counters, C_FINAL);
  for (int par = blockIdx.x * blockDim.x + threadIdx.x; par < count; par += gridDim.x * blockDim.x) {
    if (t.p_depth[par] != depth || is_leaf(t, par)) continue;
    const int left = par + 1, right = t.p_skip[left];
    double pl[3], ph[3], pc[3];
    for (int q = 0; q < 3; ++q) {
      const double l0 = ns.box[6 * (size_t)left + q], r0 = ns.box[6 * (size_t)right + q];
      const double l1 = ns.box[6 * (size_t)left + 3 + q], r1 = ns.box[6 * (size_t)right + 3 + q];
      // the left child's segments come first: on equal values its bound is kept, as a sequential scan keeps it
      pl[q] = r0 < l0 ? r0 : l0;
      ph[q] = r1 > l1 ? r1 : l1;
      pc[q] = 0.5 * (pl[q] + ph[q]);
    }
    double pm[30];
    for (int q = 0; q < 30; ++q) pm[q] = 0.0;
    double dl[3];
    for (int q = 0; q < 3; ++q) dl[q] = centre[3 * (size_t)left + q] - pc[q];
    shift_add(ns.mom + 30 * (size_t)left, dl, pm);
    for (int q = 0; q < 3; ++q) dl[q] = centre[3 * (size_t)right + q] - pc[q];
    shift_add(ns.mom + 30 * (size_t)right, dl, pm);
    const double cl = core[left], cr = core[right];
    for (int q = 0; q < 3; ++q) {
      ns.box[6 * (size_t)par + q] = pl[q];
      ns.box[6 * (size_t)par + 3 + q] = ph[q];
      centre[3 * (size_t)par + q] = pc[q];
    }
    for (int q = 0; q < 30; ++q) ns.mom[30 * (size_t)par + q] = pm[q];
    core[par] = cr > cl ? cr : cl;
  }
}

// Squared bounding radius of internal nodes: one wavefront per item (lanes stride, fixed butterfly),
// the same operations as node_moments.
__global__ __launch_bounds__(256) void k_bound_items(const int32_t* __restrict__ counters,
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
    for (int k = p0 + lane; k < p1; k += 64) {
#pragma clang fp contract(off)
      const double s0 = mi.s[0][k] - c0, s1 = mi.s[1][k] - c1, s2 = mi.s[2][k] - c2;
      const double qs = s0 * s0 + s1 * s1 + s2 * s2;
      const double e0 = mi.e[0][k] - c0, e1 = mi.e[1][k] - c1, e2 = mi.e[2][k] - c2;
      const double qe = e0 * e0 + e1 * e1 + e2 * e2;
      r2 = qs > r2 ? qs : r2;
      r2 = qe > r2 ? qe : r2;
    }
    r2 = wave_max(r2);
    if (lane == 0) part[w] = r2;
  }
}

// index of the symmetric pair (j, l) among 00 01 02 11 12 22
__device__ __forceinline__ int sym6(int j, int l) {
  const int a = j < l ? j : l, b = j < l ? l : j;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// One wavefront per node: the squared bound (a leaf's own, or the maximum over the node's items), the
// contractions (vortex_tree.py:425-447), full outputs (optional) and the packed traversal records.
__global__ __launch_bounds__(256) void k_moments_final(TreeArrays t, const int32_t* __restrict__ first,
                                                       const int32_t* __restrict__ ncount,
                                                       const double* __restrict__ part, NodeScratch ns,
                                                       const double* __restrict__ centre,
                                                       const double* __restrict__ core, double clearance,
                                                       NodeOutputs no, double* __restrict__ acc_rec,
                                                       double* __restrict__ far64, float* __restrict__ far32) {
  const int count = live_count(t.counters, C_FINAL);
  const int lane = threadIdx.x & 63;
  for (int node = blockIdx.x * 4 + threadIdx.x / 64; node < count; node += gridDim.x * 4) {
    double r2 = 0.0;
    for (int it = first[node] + lane; it < first[node] + ncount[node]; it += 64) r2 = part[it] > r2 ? part[it] : r2;
    r2 = wave_max(r2);
    if (lane != 0) continue;
    if (ncount[node] == 0) r2 = ns.r2[node];  // a leaf
    double bound;
    {
#pragma clang fp contract(off)
      bound = __builtin_sqrt(r2) * (1.0 + 1e-12);
    }
    const double* v = ns.mom + 30 * (size_t)node;
    double m0[3], m1[9], m2s[18];
    for (int q = 0; q < 3; ++q) m0[q] = v[q];
    for (int q = 0; q < 9; ++q) m1[q] = v[3 + q];
    for (int q = 0; q < 18; ++q) m2s[q] = v[12 + q];
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
  const size_t C = (size_t)cap;
  size_t need = 14 * al(M * 8) + al(M * REC64 * 8) + al(M * REC32 * 4);
  need += 3 * al((size_t)item_cap * 4) + 2 * al(C * 4);
  need += al((size_t)item_cap * 8) + al(C * 3 * 8) + al(C * 8);
  need += al(C * ACC_REC * 8) + al(C * FAR_REC * 8) + al(C * FAR_REC * 4);
  need += al(C * 6 * 8) + al(C * 30 * 8) + al(C * 8);
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
  int32_t* d_first = (int32_t*)carve(C * 4);
  int32_t* d_ncount = (int32_t*)carve(C * 4);
  double* d_part = (double*)carve((size_t)item_cap * 8);
  td.centre = (double*)carve(C * 3 * 8);
  td.core = (double*)carve(C * 8);
  td.acc = (double*)carve(C * ACC_REC * 8);
  td.far64 = (double*)carve(C * FAR_REC * 8);
  td.far32 = (float*)carve(C * FAR_REC * 4);
  NodeScratch ns;
  ns.box = (double*)carve(C * 6 * 8);
  ns.mom = (double*)carve(C * 30 * 8);
  ns.r2 = (double*)carve(C * 8);
  td.count = cap;
  const int grid = c.cu_count * 4;
  if (m > 0)
    hipLaunchKernelGGL(k_tree_permute, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, t.order, starts, directions,
                       lengths, gamma, radius, model, mi, td.rec64, td.rec32);
  hipLaunchKernelGGL(k_items_count, dim3(grid), dim3(256), 0, c.stream, t, d_ncount);
  scan_exclusive(c, cap, d_ncount, d_first, t.scan_part, nullptr);
  hipLaunchKernelGGL(k_items_fill, dim3(grid), dim3(256), 0, c.stream, t, d_first, item_cap, d_inode, d_ip0, d_ip1);
  hipLaunchKernelGGL(k_leaves, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, c.stream, t, mi, ns, td.centre,
                     td.core);
  // internal nodes level by level, deepest first (the depth of a median-split tree of s segments is at most
  // ceil(log2 s))
  for (int d = depth_bound; d >= 0; --d)
    hipLaunchKernelGGL(k_internal_level, dim3((unsigned)std::max(1, std::min(grid, (cap + 255) / 256))), dim3(256), 0,
                       c.stream, t, d, ns, td.centre, td.core);
  hipLaunchKernelGGL(k_bound_items, dim3(grid), dim3(256), 0, c.stream, t.counters, d_inode, d_ip0, d_ip1, mi,
                     td.centre, d_part);
  NodeOutputs no = want_outputs ? td.out : NodeOutputs{};
  hipLaunchKernelGGL(k_moments_final, dim3(grid), dim3(256), 0, c.stream, t, d_first, d_ncount, d_part, ns, td.centre,
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
