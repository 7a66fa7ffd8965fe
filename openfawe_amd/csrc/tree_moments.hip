// Node quantities of the tree on the GPU: centre, bounding radius, largest core, moments m0, m1, m2 and
// their contractions (vortex_tree.py:118-167 node_moments, :425-447 node_contractions).
//
// Segments are first permuted into tree order, so that every node's segments are one contiguous range.
//  * Below a cut depth, one workgroup finishes a whole subtree in LDS: each leaf on half a wavefront (box,
//    centre, core, squared bound and moments in the reference's operation order, so a leaf's quantities equal
//    the reference's bit for bit); the internal nodes level by level from their two children (box and core by
//    exact min / max, bit-exact; moments shifted to the node's centre by the parallel-axis identity, equal to
//    the reference's direct sums to rounding); the exact squared bound of each internal node in one sweep
//    over the subtree's end points.
//  * Above the cut, one workgroup per tree finishes the internal nodes level by level, and one pass over all
//    segments gives their exact squared bounds by integer atomic maxima on the bits.
//  * A last pass takes the square roots, the contractions and packs the traversal records.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "scan.h"
#include "tree.h"
#include "tree_eval.h"
#include "tree_records.h"

namespace ofw {

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

// Tree-order copies: the moment inputs (start, length vector, radius, gamma) and the pair records of the
// traversal's near field: the FP32 near-field records, and the FP64 records only when rec64 is given (the FP32
// near field rebuilds the few it needs from the engine inputs).
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
    mi.l[c][k] = lv;
  }
  mi.rad[k] = radius[s];
  mi.gam[k] = g;
  if (rec64) pack_record(s, starts, directions, lengths, gamma, radius, model, rec64 + (size_t)k * REC64, nullptr);
  pack_near32(s, directions, lengths, gamma, radius, model, rec32 + (size_t)k * NEAR32);
}

// One step of a wavefront maximum of 64-bit patterns: each lane against the lane the DPP control names (rows
// outside ROW_MASK keep their value).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long max_dpp(unsigned long long v) {
  const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROW_MASK, 0xf, false);
  const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROW_MASK, 0xf, false);
  const unsigned long long o = ((unsigned long long)ohi << 32) | olo;
  return o > v ? o : v;
}

// The maximum of a 64-bit pattern over the wavefront (every lane active), by DPP moves: quad swaps, the half-row
// and row mirrors, then the row broadcasts that carry each row's maximum into the next rows; lane 63 ends with the
// whole wavefront's.
__device__ __forceinline__ unsigned long long wave_max_bits(unsigned long long v) {
  v = max_dpp<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  v = max_dpp<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  v = max_dpp<0x141, 0xF>(v);  // row_half_mirror
  v = max_dpp<0x140, 0xF>(v);  // row_mirror
  v = max_dpp<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = max_dpp<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// ---------------------------------------------------------------- bottom-up node quantities
__device__ __forceinline__ bool is_leaf(const TreeArrays& t, int p) { return t.p_skip[p] == p + 1; }

// Each tree is cut at depth (its depth bound - MB_DEPTH): one workgroup finishes each subtree below the cut
// (k_moments_bottom, at most 2^(MB_DEPTH+1) - 1 nodes, held in LDS); one workgroup per tree finishes the
// internal nodes of the MB_MAX_CUT levels under the root (k_moments_top), and a launch per level the
// internal nodes between those and the cut (trees deeper than MB_MAX_CUT + MB_DEPTH).
static constexpr int MB_DEPTH = 5;
static constexpr int SUB_CAP = 64;
static constexpr int MB_MAX_CUT = 8;
static constexpr int LEAF_LANES = 32;                  // lanes of one leaf job (half a wavefront)
static constexpr int BOTTOM_MT = 256;                  // threads of k_moments_bottom
static constexpr int HALVES = BOTTOM_MT / LEAF_LANES;  // leaf jobs at a time
static constexpr int STAGE_SEGS = 16;                  // segments of a leaf staged in LDS at a time

// Per-tree tables and lists of the moment pass (device, int32).
struct MomentForest {
  int K;
  const int32_t* cut;  // K: the cut depth of each tree
  int32_t* top_count;  // K x MB_MAX_CUT: internal nodes listed per tree and level (cleared by the table upload)
  int32_t* top_list;   // K x 2^MB_MAX_CUT: level d of tree k at k * 2^MB_MAX_CUT + 2^d - 1
  int32_t* roots;      // subtree roots, in any order (their count is C_SUBTREES)
};

// the last k with table[k] <= v (table[0] = 0 <= v < table[K])
__device__ __forceinline__ int last_le(const int32_t* __restrict__ table, int K, int v) {
  int lo = 0, hi = K;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (table[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Per-node scratch of the bottom-up pass (pre-order index): box of the end points (low xyz, high xyz),
// moments (m0 3, m1 9, m2 as 3 x 6 entries j <= l), squared bound.
struct NodeScratch {
  double* box;
  double* mom;
  double* r2;
  int32_t* slot;  // top nodes: the index in their level's list
};

// Node quantities addressed by a node index, in global memory (pre-order index) or in LDS (index within a
// subtree).
struct NodeRef {
  double* box;     // 6 per node
  double* centre;  // 3 per node
  double* mom;     // 30 per node
  double* core;    // 1 per node
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

// One node's quantities (box 6, centre 3, moments 30, core 1), in global memory or in LDS.
struct NodeP {
  double *box, *centre, *mom, *core;
};
__device__ __forceinline__ NodeP node_at(NodeRef v, int i) {
  return NodeP{v.box + 6 * (size_t)i, v.centre + 3 * (size_t)i, v.mom + 30 * (size_t)i, v.core + i};
}

// An internal node from its two children: box and core exactly by min / max, the centre from the box,
// moments by shift_add in the fixed order left then right; written to out and, when copy.box is set, to copy.
__device__ __forceinline__ void combine_nodes(NodeP L, NodeP R, NodeP out, NodeP copy) {
#pragma clang fp contract(off)
  double pl[3], ph[3], pc[3];
  for (int q = 0; q < 3; ++q) {
    const double l0 = L.box[q], r0 = R.box[q];
    const double l1 = L.box[3 + q], r1 = R.box[3 + q];
    // the left child's segments come first: on equal values its bound is kept, as a sequential scan keeps it
    pl[q] = r0 < l0 ? r0 : l0;
    ph[q] = r1 > l1 ? r1 : l1;
    pc[q] = 0.5 * (pl[q] + ph[q]);
  }
  double pm[30];
  for (int q = 0; q < 30; ++q) pm[q] = 0.0;
  double dl[3];
  for (int q = 0; q < 3; ++q) dl[q] = L.centre[q] - pc[q];
  shift_add(L.mom, dl, pm);
  for (int q = 0; q < 3; ++q) dl[q] = R.centre[q] - pc[q];
  shift_add(R.mom, dl, pm);
  const double cl = L.core[0], cr = R.core[0];
  const double pcore = cr > cl ? cr : cl;
  for (int q = 0; q < 3; ++q) {
    out.box[q] = pl[q];
    out.box[3 + q] = ph[q];
    out.centre[q] = pc[q];
  }
  for (int q = 0; q < 30; ++q) out.mom[q] = pm[q];
  out.core[0] = pcore;
  if (copy.box) {
    for (int q = 0; q < 3; ++q) {
      copy.box[q] = pl[q];
      copy.box[3 + q] = ph[q];
      copy.centre[q] = pc[q];
    }
    for (int q = 0; q < 30; ++q) copy.mom[q] = pm[q];
    copy.core[0] = pcore;
  }
}

__device__ __forceinline__ void combine_children(NodeRef v, int par, int left, int right) {
  combine_nodes(node_at(v, left), node_at(v, right), node_at(v, par), NodeP{});
}

__device__ __forceinline__ double half_max(double v) {
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }

// One leaf [a, b) on 32 lanes (hl = lane within the half wavefront), operation by operation as node_moments
// does (vortex_tree.py:118-167): every accumulator is updated in the order of the leaf's segments, each by
// its own lane, so the leaf's quantities equal the sequential loop's bit for bit.  The segments are staged
// in LDS 32 at a time by coalesced loads.
//  * box: lanes 0-2 the low bound of axis hl, lanes 3-5 the high bound of axis hl - 3 (start, then end);
//  * squared bound and largest core: one segment per lane, then the maximum over the lanes (exact);
//  * moments: lane q < 30 accumulates m[q] (m0 3, m1 9, m2 as 3 x 6).
// The results go to node i of v; the squared bound to r2b[i] as its bits.
__device__ void leaf_job(const MomentInputs& mi, NodeRef v, unsigned long long* r2b, int i, int a, int b,
                         double (*sh)[STAGE_SEGS], int hl) {
#pragma clang fp contract(off)
  const int axis = hl < 3 ? hl : (hl < 6 ? hl - 3 : 0);
  const bool low = hl < 3;
  double bnd = low ? INFINITY : -INFINITY;
  for (int base = a; base < b; base += STAGE_SEGS) {
    const int n = min(STAGE_SEGS, b - base);
    if (hl < n) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sh[c][hl] = mi.s[c][base + hl];
        sh[3 + c][hl] = mi.end(c, base + hl);
      }
    }
    wave_lds_sync();
    if (hl < 6) {
#pragma unroll 4
      for (int j = 0; j < n; ++j) {
        const double v0 = sh[axis][j], v1 = sh[3 + axis][j];
        if (low) {
          if (v0 < bnd) bnd = v0;
          if (v1 < bnd) bnd = v1;
        } else {
          if (v0 > bnd) bnd = v0;
          if (v1 > bnd) bnd = v1;
        }
      }
    }
    wave_lds_sync();
  }
  const int h0 = threadIdx.x & 32;  // first lane of this half wavefront
  double c[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) c[q] = 0.5 * (__shfl(bnd, h0 + q, 64) + __shfl(bnd, h0 + 3 + q, 64));
  // accumulator of this lane: kind 0 = m0_i, 1 = m1_ij, 2 = m2_ijl (j <= l)
  const int kind = hl < 3 ? 0 : (hl < 12 ? 1 : 2);
  const int ai = hl < 3 ? hl : (hl < 12 ? (hl - 3) / 3 : min((hl - 12) / 6, 2));
  const int u = hl < 12 ? 0 : (hl - 12) % 6;
  const int aj = kind == 1 ? (hl - 3) % 3 : (u < 3 ? 0 : (u < 5 ? 1 : 2));
  const int al = kind == 2 ? (u == 0 ? 0 : (u == 1 || u == 3) ? 1 : 2) : 0;
  double acc = 0.0, r2 = 0.0, largest = 0.0;
  // staged per segment: d (0-2), alpha (3-5), length vector (6-8), gamma (9)
  for (int base = a; base < b; base += STAGE_SEGS) {
    const int n = min(STAGE_SEGS, b - base);
    if (hl < n) {
      const int k = base + hl;
      double sv[3], ev[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        sv[q] = mi.s[q][k];
        ev[q] = mi.end(q, k);
      }
      const double s0 = sv[0] - c[0], s1 = sv[1] - c[1], s2 = sv[2] - c[2];
      const double qs = s0 * s0 + s1 * s1 + s2 * s2;
      if (qs > r2) r2 = qs;
      const double e0 = ev[0] - c[0], e1 = ev[1] - c[1], e2 = ev[2] - c[2];
      const double qe = e0 * e0 + e1 * e1 + e2 * e2;
      if (qe > r2) r2 = qe;
      const double rad = mi.rad[k];
      if (rad > largest) largest = rad;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        sh[q][hl] = 0.5 * (sv[q] + ev[q]) - c[q];
        sh[3 + q][hl] = mi.alpha(q, k);
        sh[6 + q][hl] = mi.l[q][k];
      }
      sh[9][hl] = mi.gam[k];
    }
    wave_lds_sync();
    if (hl < 30) {
#pragma unroll 4
      for (int j = 0; j < n; ++j) {
        const double alpha = sh[3 + ai][j];
        const double dj = sh[aj][j], dl = sh[al][j];
        const double lj = sh[6 + aj][j], ll = sh[6 + al][j];
        const double x2 = dj * dl + lj * ll / 12.0;
        // alpha * 1.0 == alpha exactly; a segment without circulation leaves the sum as it is
        const double x = kind == 0 ? 1.0 : (kind == 1 ? dj : x2);
        const double next = acc + alpha * x;
        acc = sh[9][j] == 0.0 ? acc : next;
      }
    }
    wave_lds_sync();
  }
  r2 = half_max(r2);
  largest = half_max(largest);
  if (hl < 30) v.mom[30 * i + hl] = acc;
  if (hl < 6) v.box[6 * i + hl] = bnd;
  if (hl < 3) v.centre[3 * i + hl] = c[hl == 0 ? 0 : (hl == 1 ? 1 : 2)];
  if (hl == 0) {
    v.core[i] = largest;
    r2b[i] = dbits(r2);
  }
}

// Steps of halving a node of s segments until it holds at most `leaf`: a bound on the depth below it.
__device__ __host__ __forceinline__ int halving_depth(int64_t s, int leaf) {
  int d = 0;
  while (s > leaf) {
    s = (s + 1) / 2;
    ++d;
  }
  return d;
}

// Ranges of k_bound_top: BT_SPAN consecutive positions each, and its table of the internal nodes above the cut
// that meet a range (per range and depth, slot 0: the node holding the range's first position, slot 1: a node
// starting inside the range).  The table is not cleared: k_bound_top checks every entry against the tree.
static constexpr int BT_LEVELS = 32;  // the host caps the cut depth at 30
static constexpr int BT_SPAN = 1024;

// One pass over the nodes: the subtree roots of k_moments_bottom (the nodes at their tree's cut depth and
// the leaves above it), per tree and level the internal nodes above min(cut, MB_MAX_CUT) for k_moments_top,
// and the range table of k_bound_top.  The order of the lists does not matter: every entry is finished on its own.
__global__ __launch_bounds__(256) void k_moment_lists(TreeArrays t, MomentForest mf, int32_t* __restrict__ btab) {
  const int count = live_count(t.counters, C_FINAL);
  const int lane = threadIdx.x & 63;
  // whole wavefronts step together: the range table is written by all lanes of a wavefront at once
  for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    const int p = base + (int)threadIdx.x;
    const bool valid = p < count;
    const int k = valid ? t.p_tree[p] : 0, d = valid ? t.p_depth[p] : 0, cut = valid ? mf.cut[k] : 0;
    const bool leaf = valid && is_leaf(t, p);
    // subtree roots appended with one atomic per wavefront
    const bool root = valid && (d == cut || (d < cut && leaf));
    const unsigned long long ballot = __builtin_amdgcn_ballot_w64(root);
    if (ballot) {
      const int leader = __builtin_ctzll(ballot);
      int rbase = 0;
      if (lane == leader) rbase = atomicAdd(t.counters + C_SUBTREES, __popcll(ballot));
      rbase = __shfl(rbase, leader, 64);
      if (root) mf.roots[rbase + __popcll(ballot & ((1ull << lane) - 1ull))] = p;
    }
    if (valid && d < min(cut, MB_MAX_CUT) && !leaf) {
      const int slot = atomicAdd(mf.top_count + (size_t)k * MB_MAX_CUT + d, 1);
      mf.top_list[((size_t)k << MB_MAX_CUT) + (1 << d) - 1 + slot] = p;
    }
    unsigned long long tops = __builtin_amdgcn_ballot_w64(valid && d < min(cut, BT_LEVELS) && !leaf);
    const int lo_p = valid ? t.p_lo[p] : 0, hi_p = valid ? t.p_hi[p] : 0;
    while (tops) {
      const int l = __builtin_ctzll(tops);
      tops &= tops - 1ull;
      const int q = __builtin_amdgcn_readlane(p, l), dq = __builtin_amdgcn_readlane(d, l);
      const int lo = __builtin_amdgcn_readlane(lo_p, l), hi = __builtin_amdgcn_readlane(hi_p, l);
      for (int b = lo / BT_SPAN + lane; b <= (hi - 1) / BT_SPAN; b += 64)
        btab[((size_t)b * BT_LEVELS + dq) * 2 + (lo <= b * BT_SPAN ? 0 : 1)] = q;
    }
  }
}

// index of the symmetric pair (j, l) among 00 01 02 11 12 22
__device__ __forceinline__ int sym6(int j, int l) {
  const int a = j < l ? j : l, b = j < l ? l : j;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// Field q of a node's far-field record from its moments v (m0 3, m1 9, m2 as 3 x 6 entries j <= l): the
// contractions of vortex_tree.py:425-447 and the symmetric second moments (the diagonal once, the off-diagonal
// doubled), by the operations of k_moments_final.
__device__ __forceinline__ double far_field(const double* v, int q) {
  const double* m1 = v + 3;
  const double* m2s = v + 12;
  if (q < FR_M1) return v[q];
  if (q < FR_W1) return m1[q - FR_M1];
  if (q < FR_M2) {
    const int i = q - FR_W1;  // w1[i] = m1[a][b] - m1[b][a] for (a, b) = (1, 2), (2, 0), (0, 1)
    const int a = (i + 1) % 3, b = (i + 2) % 3;
    return m1[3 * a + b] - m1[3 * b + a];
  }
  if (q < FR_T2) {
    const int x = q - FR_M2, u = x % 6;
    const bool diag = (u == 0 || u == 3 || u == 5);
    return diag ? m2s[x] : 2.0 * m2s[x];
  }
  if (q < FR_W2) {
    const int i = q - FR_T2;
    return m2s[6 * i + 0] + m2s[6 * i + 3] + m2s[6 * i + 5];
  }
  if (q < FR_PAD0) {
    const int x = q - FR_W2, i = x / 3, k = x % 3;  // w2[i][k] = m2[a][b][k] - m2[b][a][k]
    const int a = (i + 1) % 3, b = (i + 2) % 3;
    return m2s[6 * a + sym6(b, k)] - m2s[6 * b + sym6(a, k)];
  }
  return 0.0;
}

// Where k_moments_bottom writes the final records of its subtree's nodes (on = 0: k_moments_final writes them).
struct FinalOut {
  int on;
  double clearance;
  double* acc;
  double* far64;
  float* far32;
};

// One workgroup per subtree below the cut depth, all of it in LDS: its leaves (a half wavefront each), its
// internal nodes level by level from the deepest, then the exact squared bound of each internal node in one
// sweep over the subtree's segments (each segment against the centres of the internal nodes on its path;
// the maximum over a wavefront's lanes that share a node, then an LDS atomic maximum on the bits, which is
// exact and independent of the order: non-negative doubles order as their bit patterns).
__global__ __launch_bounds__(BOTTOM_MT, 4) void k_moments_bottom(TreeArrays t, MomentForest mf, MomentInputs mi,
                                                              NodeScratch ns, double* __restrict__ centre,
                                                              double* __restrict__ core, FinalOut fo) {
#pragma clang fp contract(off)
  __shared__ double s_box[SUB_CAP * 6], s_cen[SUB_CAP * 3], s_mom[SUB_CAP * 30], s_core[SUB_CAP];
  __shared__ unsigned long long s_r2[SUB_CAP];
  __shared__ int s_lo[SUB_CAP], s_hi[SUB_CAP], s_skip[SUB_CAP], s_depth[SUB_CAP], s_leaves[SUB_CAP];
  __shared__ double stage[HALVES][10][STAGE_SEGS];
  if (t.counters[C_OVERFLOW] || (int)blockIdx.x >= t.counters[C_SUBTREES]) return;
  const int r = mf.roots[blockIdx.x];
  const int nn = t.p_skip[r] - r;  // at most 2^(MB_DEPTH+1) - 1 <= SUB_CAP
  const int tid = threadIdx.x;
  int leaf_flag = 0;
  if (tid < nn) {
    const int p = r + tid;
    const int skip = t.p_skip[p];
    s_lo[tid] = t.p_lo[p];
    s_hi[tid] = t.p_hi[p];
    s_skip[tid] = skip - r;
    s_depth[tid] = t.p_depth[p];
    s_r2[tid] = 0ull;
    leaf_flag = skip == p + 1 ? 1 : 0;
  }
  int nleaves;
  const int rank = block_exclusive_scan(leaf_flag, &nleaves);
  if (leaf_flag) s_leaves[rank] = tid;
  __syncthreads();
  const NodeRef lds{s_box, s_cen, s_mom, s_core};
  const int hw = tid >> 5, hl = tid & 31;
  for (int q = hw; q < nleaves; q += HALVES) {
    const int i = s_leaves[q];
    leaf_job(mi, lds, s_r2, i, s_lo[i], s_hi[i], stage[hw], hl);
  }
  __syncthreads();
  if (nn > 1) {
    const int d0 = s_depth[0];
    const int dmax = d0 + halving_depth(s_hi[0] - s_lo[0], t.leaf);
    for (int lev = dmax - 1; lev >= d0; --lev) {
      if (tid < nn && s_depth[tid] == lev && s_skip[tid] != tid + 1) combine_children(lds, tid, tid + 1, s_skip[tid + 1]);
      __syncthreads();
    }
    const int lane = tid & 63;
    const int a = s_lo[0], b = s_hi[0];
    for (int base = a; base < b; base += BOTTOM_MT) {
      const int kk = base + tid;
      bool live = kk < b;
      double sv[3] = {0.0, 0.0, 0.0}, ev[3] = {0.0, 0.0, 0.0};
      if (live)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          sv[q] = mi.s[q][kk];
          ev[q] = mi.end(q, kk);
        }
      int i = 0;
      while (wave_any(live)) {
        if (live && s_skip[i] == i + 1) live = false;
        double v = 0.0;
        int id = -1;
        if (live) {
          const double c0 = s_cen[3 * i], c1 = s_cen[3 * i + 1], c2 = s_cen[3 * i + 2];
          const double s0 = sv[0] - c0, s1 = sv[1] - c1, s2 = sv[2] - c2;
          const double qs = s0 * s0 + s1 * s1 + s2 * s2;
          const double e0 = ev[0] - c0, e1 = ev[1] - c1, e2 = ev[2] - c2;
          const double qe = e0 * e0 + e1 * e1 + e2 * e2;
          v = qs > qe ? qs : qe;
          id = i;
        }
        const int first = __shfl(id, 0, 64);
        if (wave_all(id == first || id < 0) && first >= 0) {
          v = wave_max(v);
          if (lane == 0) atomicMax(&s_r2[first], dbits(v));
        } else if (id >= 0) {
          atomicMax(&s_r2[id], dbits(v));
        }
        if (live) i = kk < s_hi[i + 1] ? i + 1 : s_skip[i + 1];
      }
    }
    __syncthreads();
  }
  // the subtree's nodes are the pre-order indices r .. r + nn - 1: contiguous rows of every array.  With fo.on
  // the nodes' final records are written here, field by field over the workgroup, and the levels above read the
  // root's quantities only.
  const int keep = fo.on ? 1 : nn;
  for (int x = tid; x < 30 * keep; x += BOTTOM_MT) ns.mom[30 * (size_t)r + x] = s_mom[x];
  for (int x = tid; x < 6 * keep; x += BOTTOM_MT) ns.box[6 * (size_t)r + x] = s_box[x];
  for (int x = tid; x < 3 * nn; x += BOTTOM_MT) centre[3 * (size_t)r + x] = s_cen[x];
  if (tid < nn) {
    core[r + tid] = s_core[tid];
    if (!fo.on || tid == 0) ns.r2[r + tid] = __longlong_as_double((long long)s_r2[tid]);
  }
  if (fo.on) {
    for (int x = tid; x < FAR_REC * nn; x += BOTTOM_MT) {
      const int i = x / FAR_REC, q = x - i * FAR_REC;
      const double v = far_field(s_mom + 30 * i, q);
      fo.far64[FAR_REC * (size_t)r + x] = v;
      fo.far32[FAR_REC * (size_t)r + x] = (float)v;
    }
    const int tr = t.p_tree[r];
    for (int x = tid; x < ACC_REC * nn; x += BOTTOM_MT) {
      const int i = x / ACC_REC, q = x - i * ACC_REC;
      double v;
      if (q < A_BOUND) {
        v = s_cen[3 * i + q];
      } else if (q == A_BOUND) {
        v = __builtin_sqrt(__longlong_as_double((long long)s_r2[i])) * (1.0 + 1e-12);
      } else if (q == A_CLEAR) {
        v = fo.clearance * s_core[i];
      } else {
        // p_lo, p_hi | p_skip, tree | right child (-1 for a leaf), 0
        int32_t w[2];
        const int node = r + i;
        if (q == A_INTS) {
          w[0] = s_lo[i];
          w[1] = s_hi[i];
        } else if (q == A_INTS2) {
          w[0] = r + s_skip[i];
          w[1] = tr;
        } else {
          w[0] = s_skip[i] != i + 1 ? r + s_skip[i + 1] : -1;
          w[1] = 0;
        }
        (void)node;
        v = __hiloint2double(w[1], w[0]);
      }
      fo.acc[ACC_REC * (size_t)r + x] = v;
    }
  }
}

// The internal nodes of one depth between MB_MAX_CUT and their tree's cut (trees deeper than
// MB_MAX_CUT + MB_DEPTH), from their children; their squared bounds are cleared for k_bound_top.
__global__ __launch_bounds__(256) void k_internal_level(TreeArrays t, MomentForest mf, int depth, NodeScratch ns,
                                                        double* __restrict__ centre, double* __restrict__ core) {
  const int count = live_count(t.counters, C_FINAL);
  const NodeRef g{ns.box, centre, ns.mom, core};
  for (int par = blockIdx.x * blockDim.x + threadIdx.x; par < count; par += gridDim.x * blockDim.x) {
    if (t.p_depth[par] != depth || is_leaf(t, par) || depth >= mf.cut[t.p_tree[par]]) continue;
    combine_children(g, par, par + 1, t.p_skip[par + 1]);
    ns.r2[par] = 0.0;
  }
}

// One workgroup per tree: the internal nodes of the first min(cut, MB_MAX_CUT) levels (listed by
// k_moment_lists), level by level from the deepest; their children were finished by k_moments_bottom or
// k_internal_level.  Their squared bounds are cleared for k_bound_top.
// The levels of a tree's top at most TOP_LDS_LEVEL deep (level d holds at most 2^d nodes) keep their quantities
// in LDS for the level above, besides writing them to global memory: the even levels in the slots from 0, the
// odd ones after the deepest even level's 2^TOP_EVEN slots.  (Wider levels, whose children's quantities come
// from global memory, are spread over the device by k_internal_level: one workgroup reads them at about one
// cache line per clock.)
static constexpr int TOP_LDS_LEVEL = MB_MAX_CUT - 1;
static constexpr int TOP_EVEN = TOP_LDS_LEVEL & ~1;                                      // deepest even kept level
static constexpr int TOP_ODD = (TOP_LDS_LEVEL & 1) ? TOP_LDS_LEVEL : TOP_LDS_LEVEL - 1;  // deepest odd kept level
static constexpr int TOP_QN = (1 << TOP_EVEN) + (1 << TOP_ODD);                          // slots of both parities
__global__ __launch_bounds__(512) void k_moments_top(TreeArrays t, MomentForest mf, NodeScratch ns,
                                                     double* __restrict__ centre, double* __restrict__ core) {
  __shared__ int s_node[1 << MB_MAX_CUT], s_right[1 << MB_MAX_CUT];
  __shared__ int16_t s_lq[1 << MB_MAX_CUT], s_rq[1 << MB_MAX_CUT];  // a child's LDS slot, or -1: global memory
  __shared__ int s_first[MB_MAX_CUT + 1];                          // listed nodes of the levels before each
  __shared__ double q_box[TOP_QN][6], q_cen[TOP_QN][3], q_mom[TOP_QN][30], q_core[TOP_QN];
  if (t.counters[C_OVERFLOW]) return;
  const int k = blockIdx.x;
  const int D = min(mf.cut[k], MB_MAX_CUT);
  if (D == 0 || t.tree_first[k] == t.tree_first[k + 1]) return;
  const int32_t* list = mf.top_list + ((size_t)k << MB_MAX_CUT);
  const int32_t* cnt = mf.top_count + (size_t)k * MB_MAX_CUT;
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int d = 0; d < D; ++d) {
      s_first[d] = acc;
      acc += cnt[d];
    }
    s_first[D] = acc;
  }
  __syncthreads();
  // every listed node (all levels at once) and its right child; a node's slot in its level's list
  auto level_of = [&](int f) {
    int d = 0;
    while (s_first[d + 1] <= f) ++d;
    return d;
  };
  const int total = s_first[D];
  for (int f = (int)threadIdx.x; f < total; f += blockDim.x) {
    const int d = level_of(f), i = f - s_first[d];
    const int p = list[(1 << d) - 1 + i];
    s_node[(1 << d) - 1 + i] = p;
    s_right[(1 << d) - 1 + i] = t.p_skip[p + 1];
    ns.slot[p] = i;
  }
  __syncthreads();
  auto qbase = [](int d) { return (d & 1) ? (1 << TOP_EVEN) : 0; };
  // where each node's children are: in LDS when this kernel makes them at a level kept there, else global
  for (int f = (int)threadIdx.x; f < total; f += blockDim.x) {
    const int d = level_of(f), i = f - s_first[d], dc = d + 1;
    const int j = (1 << d) - 1 + i;
    const int c[2] = {s_node[j] + 1, s_right[j]};
    int16_t q[2];
    for (int h = 0; h < 2; ++h) {
      int sl = -1;
      if (dc <= TOP_LDS_LEVEL && dc < D && !is_leaf(t, c[h])) {
        sl = ns.slot[c[h]] + qbase(dc);
        if ((unsigned)sl >= (unsigned)TOP_QN) {  // cannot happen (level dc holds at most 2^dc nodes); reported
          t.counters[C_OVERFLOW] = 1;
          sl = -1;
        }
      }
      q[h] = (int16_t)sl;
    }
    s_lq[j] = q[0];
    s_rq[j] = q[1];
  }
  __syncthreads();
  const NodeRef g{ns.box, centre, ns.mom, core};
  auto child = [&](int c, int sl) -> NodeP {
    return sl >= 0 ? NodeP{q_box[sl], q_cen[sl], q_mom[sl], &q_core[sl]} : node_at(g, c);
  };
  for (int d = D - 1; d >= 0; --d) {
    for (int i = (int)threadIdx.x; i < s_first[d + 1] - s_first[d]; i += blockDim.x) {
      const int j = (1 << d) - 1 + i;
      const int p = s_node[j];
      const int sl = i + qbase(d);
      const NodeP copy = (d <= TOP_LDS_LEVEL && sl < TOP_QN) ? NodeP{q_box[sl], q_cen[sl], q_mom[sl], &q_core[sl]}
                                                             : NodeP{};
      combine_nodes(child(p + 1, s_lq[j]), child(s_right[j], s_rq[j]), node_at(g, p), copy);
      ns.r2[p] = 0.0;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void max_bits(unsigned long long* a, double v) {
  const unsigned long long b = dbits(v);
  // the stored value only grows: a stale read can only cause an unneeded atomic, and skipping the others
  // keeps thousands of wavefronts from queueing on the root's bound
  if (b > *reinterpret_cast<volatile unsigned long long*>(a)) atomicMax(a, b);
}

// Exact squared bounds of the internal nodes above the cut depth in one pass over the segments (four per
// lane): each segment's end points against the centre of each such node on its root path, maxima combined over
// the workgroup, then an integer atomic maximum on the bits (exact in any order: non-negative doubles order as
// their bits).  A workgroup's range of BT_SPAN consecutive positions meets at most two nodes of a depth when the
// nodes above the cut are larger than the range: then the range's nodes come from the table of k_moment_lists and
// every depth takes two centres from LDS and two atomics.  Otherwise (a range over several trees, a leaf above the
// cut, small top nodes) each lane walks its segments' paths and the maxima are combined over the wavefronts that
// share a node.
__global__ __launch_bounds__(256) void k_bound_top(TreeArrays t, MomentForest mf, MomentInputs mi,
                                                   const double* __restrict__ centre, double* __restrict__ r2out,
                                                   const int32_t* __restrict__ btab) {
#pragma clang fp contract(off)
  constexpr int J = 4, WAVES = 4;
  static_assert(J * WAVES * 64 == BT_SPAN, "a workgroup's positions are one range of the table");
  __shared__ int s_node[WAVES];  // per wavefront: its one node, -1 nothing live, -2 several nodes
  __shared__ unsigned long long s_bits[WAVES];
  __shared__ unsigned long long s_lev_bits[BT_LEVELS][2][WAVES];  // per depth and node: the wavefronts' maxima
  __shared__ int s_path[2][BT_LEVELS];     // the internal nodes on the root paths of the range's end positions
  __shared__ int s_lo[2][BT_LEVELS], s_hi[2][BT_LEVELS];
  __shared__ double s_cen[2][BT_LEVELS][3];
  __shared__ int s_depth[1], s_tree[2];    // depths to do (the cut), the trees of the range's ends
  __shared__ int s_fast;
  if (t.counters[C_OVERFLOW]) return;
  const int m = (int)t.m;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(r2out);
  const int stride = gridDim.x * WAVES * 64 * J;
  // the wavefronts of a workgroup run the same iterations and levels (they combine their maxima)
  for (int bb = blockIdx.x * WAVES * 64 * J; bb < m; bb += stride) {
    const int base = bb + wv * 64 * J;
    int kk[J];
    double s[J][3], e[J][3];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      kk[j] = base + 64 * j + lane;
      const bool valid = kk[j] < m;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        s[j][q] = valid ? mi.s[q][kk[j]] : 0.0;
        e[j][q] = valid ? mi.end(q, kk[j]) : 0.0;
      }
    }
    // the range's nodes from the table, every entry checked against the tree (a stale or missing one sends the
    // range to the per-lane walk); a leaf above the cut or a range over two trees does too
    const int b0 = bb, b1 = min(bb + WAVES * 64 * J, m) - 1;
    const int count = live_count(t.counters, C_FINAL);
    if (threadIdx.x == 0) {
      const int tr = last_le(t.roots_lo, t.K, b0);
      s_tree[0] = tr;
      s_tree[1] = last_le(t.roots_lo, t.K, b1);
      s_depth[0] = min(mf.cut[tr], BT_LEVELS);
      s_fast = 1;
    }
    __syncthreads();
    if (threadIdx.x < BT_LEVELS) {
      // one lane per depth: the node of the range's first position, and the node of its last position (the same
      // node when it holds both, else the table's node starting inside the range)
      const int lev = threadIdx.x, tr = s_tree[0];
      if (lev < s_depth[0]) {
        const int32_t* e = btab + ((size_t)(bb / BT_SPAN) * BT_LEVELS + lev) * 2;
        const int pa = e[0], pb = e[1];
        int lo[2] = {0, 0}, hi[2] = {0, 0}, pp[2] = {pa, pb};
        bool ok[2];
        double c[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = pp[h];
          ok[h] = q >= 0 && q < count;
          if (ok[h]) {
            lo[h] = t.p_lo[q];
            hi[h] = t.p_hi[q];
            ok[h] = t.p_depth[q] == lev && t.p_tree[q] == tr && t.p_skip[q] != q + 1;
#pragma unroll
            for (int x = 0; x < 3; ++x) c[h][x] = centre[3 * (size_t)q + x];
          }
        }
        ok[0] = ok[0] && lo[0] <= b0 && b0 < hi[0];
        const bool one = ok[0] && b1 < hi[0];
        ok[1] = one || (ok[1] && lo[1] <= b1 && b1 < hi[1]);
        // selects, not indexing: the arrays stay in registers
        s_path[0][lev] = pa;
        s_lo[0][lev] = lo[0];
        s_hi[0][lev] = hi[0];
        s_path[1][lev] = one ? pa : pb;
        s_lo[1][lev] = one ? lo[0] : lo[1];
        s_hi[1][lev] = one ? hi[0] : hi[1];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          s_cen[0][lev][x] = c[0][x];
          s_cen[1][lev][x] = one ? c[0][x] : c[1][x];
        }
        if (!(ok[0] && ok[1])) s_fast = 0;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_fast) {
      bool fast = s_tree[0] == s_tree[1];
      for (int lev = 0; fast && lev < s_depth[0]; ++lev)
        fast = s_path[0][lev] == s_path[1][lev] || s_hi[0][lev] == s_lo[1][lev];
      s_fast = fast ? 1 : 0;
    }
    __syncthreads();
    if (s_fast) {
      // every depth's two wavefront maxima into LDS, then one lane per (depth, node) combines the wavefronts and
      // takes the atomic: a single round trip to memory for all depths
      const int D = s_depth[0];
      for (int lev = 0; lev < D; ++lev) {
        const int split = s_path[0][lev] == s_path[1][lev] ? INT_MAX : s_hi[0][lev];  // below it: the first node
        double v0 = 0.0, v1 = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          if (kk[j] >= m) continue;
          const int h = kk[j] < split ? 0 : 1;
          const double c0 = s_cen[h][lev][0], c1 = s_cen[h][lev][1], c2 = s_cen[h][lev][2];
          const double s0 = s[j][0] - c0, s1 = s[j][1] - c1, s2 = s[j][2] - c2;
          const double qs = s0 * s0 + s1 * s1 + s2 * s2;
          const double e0 = e[j][0] - c0, e1 = e[j][1] - c1, e2 = e[j][2] - c2;
          const double qe = e0 * e0 + e1 * e1 + e2 * e2;
          const double v = qs > qe ? qs : qe;
          if (h == 0) v0 = v > v0 ? v : v0;
          else v1 = v > v1 ? v : v1;
        }
        // a wavefront's 256 positions lie in one node at most depths: one reduction then
        unsigned long long u0 = 0ull, u1 = 0ull;
        if (split <= base) {
          u1 = wave_max_bits(dbits(v1));
        } else if (split >= base + 64 * J) {
          u0 = wave_max_bits(dbits(v0));
        } else {
          u0 = wave_max_bits(dbits(v0));
          u1 = wave_max_bits(dbits(v1));
        }
        if (lane == 0) {
          s_lev_bits[lev][0][wv] = u0;
          s_lev_bits[lev][1][wv] = u1;
        }
      }
      __syncthreads();
      if ((int)threadIdx.x < 2 * D) {
        const int lev = threadIdx.x >> 1, h = threadIdx.x & 1;
        const int n = s_path[h][lev];
        if (h == 0 || n != s_path[0][lev]) {
          unsigned long long mb = 0ull;
          for (int w = 0; w < WAVES; ++w) mb = s_lev_bits[lev][h][w] > mb ? s_lev_bits[lev][h][w] : mb;
          unsigned long long* a = bits + n;
          if (mb > *reinterpret_cast<volatile unsigned long long*>(a)) atomicMax(a, mb);
        }
      }
      __syncthreads();  // the LDS tables are written again by the next range
      continue;
    }
    int p[J], D[J];
    bool live[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const bool valid = kk[j] < m;
      const int tr = last_le(t.roots_lo, t.K, valid ? kk[j] : min(base, m - 1));
      D[j] = valid ? mf.cut[tr] : 0;
      p[j] = valid ? t.tree_first[tr] : -1;
      live[j] = valid && D[j] > 0;
    }
    for (int lev = 0;; ++lev) {
      bool any = false;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if (live[j] && (lev >= D[j] || is_leaf(t, p[j]))) live[j] = false;
        any = any || live[j];
      }
      if (!__syncthreads_or(any ? 1 : 0)) break;
      double v[J];
      int id[J];
      bool mine = false;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        v[j] = 0.0;
        id[j] = -1;
        if (live[j]) {
          const double c0 = centre[3 * (size_t)p[j]], c1 = centre[3 * (size_t)p[j] + 1],
                       c2 = centre[3 * (size_t)p[j] + 2];
          const double s0 = s[j][0] - c0, s1 = s[j][1] - c1, s2 = s[j][2] - c2;
          const double qs = s0 * s0 + s1 * s1 + s2 * s2;
          const double e0 = e[j][0] - c0, e1 = e[j][1] - c1, e2 = e[j][2] - c2;
          const double qe = e0 * e0 + e1 * e1 + e2 * e2;
          v[j] = qs > qe ? qs : qe;
          id[j] = p[j];
          mine = true;
        }
      }
      const int first = __shfl(id[0], 0, 64);
      bool same = first >= 0;
#pragma unroll
      for (int j = 0; j < J; ++j) same = same && (id[j] == first || id[j] < 0);
      const bool wsame = wave_all(same);
      const bool wlive = wave_any(mine);
      double vm = v[0];
#pragma unroll
      for (int j = 1; j < J; ++j) vm = v[j] > vm ? v[j] : vm;
      vm = wave_max(vm);
      if (lane == 0) {
        s_node[wv] = wsame ? first : (wlive ? -2 : -1);
        s_bits[wv] = dbits(vm);
      }
      __syncthreads();
      int node = -1;
      bool one = true;
      unsigned long long mb = 0ull;
      for (int w = 0; w < WAVES; ++w) {
        const int st = s_node[w];
        if (st == -1) continue;
        if (st == -2 || (node >= 0 && st != node)) {
          one = false;
        } else {
          node = st;
          mb = s_bits[w] > mb ? s_bits[w] : mb;
        }
      }
      if (one && node >= 0) {
        if (threadIdx.x == 0 && mb > *reinterpret_cast<volatile unsigned long long*>(bits + node))
          atomicMax(bits + node, mb);
      } else if (wsame) {
        if (lane == 0) max_bits(bits + first, vm);
      } else {
#pragma unroll
        for (int j = 0; j < J; ++j)
          if (id[j] >= 0) max_bits(bits + id[j], v[j]);
      }
      __syncthreads();  // s_node and s_bits are written again at the next level
#pragma unroll
      for (int j = 0; j < J; ++j)
        if (live[j]) {
          const int left = p[j] + 1;
          p[j] = kk[j] < t.p_hi[left] ? left : t.p_skip[left];
        }
    }
  }
}

// One thread per node: the bounding radius from the exact squared bound, the contractions
// (vortex_tree.py:425-447), full outputs (optional) and the packed traversal records.
__global__ __launch_bounds__(256) void k_moments_final(TreeArrays t, NodeScratch ns, const double* __restrict__ centre,
                                                       const double* __restrict__ core, double clearance,
                                                       NodeOutputs no, double* __restrict__ acc_rec,
                                                       double* __restrict__ far64, float* __restrict__ far32,
                                                       const int32_t* __restrict__ top_cut) {
  const int count = live_count(t.counters, C_FINAL);
  for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < count; node += gridDim.x * blockDim.x) {
    // top_cut: only the internal nodes above their tree's cut (k_moments_bottom finished the others)
    if (top_cut && !(t.p_depth[node] < top_cut[t.p_tree[node]] && !is_leaf(t, node))) continue;
    const double r2 = ns.r2[node];
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
    // the right child (the left one is node + 1), -1 for a leaf: a breadth-first walk queues both children
    const int skip = t.p_skip[node];
    ai[4] = skip != node + 1 ? t.p_skip[node + 1] : -1;
    ai[5] = 0;
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

int tree_moments(Ctx& c, int buf, const TreeArrays& t, const std::vector<int64_t>& offsets, const double* starts,
                 const double* directions, const double* lengths, const double* gamma, const double* radius, int model,
                 double clearance, TreeDevice& td, bool want_outputs, bool want_rec64) {
  const int m = (int)t.m, cap = t.cap, K = t.K;
  // per tree: the cut depth and the subtrees below it (at most one per path of that depth); the table also
  // clears the per-level counters of the top lists
  std::vector<int64_t> table((size_t)K * (1 + MB_MAX_CUT), 0);
  int64_t wgs = 0;
  int cut_max = 0;
  for (int k = 0; k < K; ++k) {
    const int64_t mk = offsets[k + 1] - offsets[k];
    const int cut = std::max(0, halving_depth(mk, t.leaf) - MB_DEPTH);
    if (cut > 30) {
      set_error("tree too deep for the moment pass");
      return OFW_ERR_ARG;
    }
    cut_max = std::max(cut_max, cut);
    table[k] = cut;
    wgs += mk > 0 ? ((int64_t)1 << cut) : 0;
  }
  if (wgs > INT_MAX / 2) {
    set_error("tree too large for one moment pass");
    return OFW_ERR_ARG;
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t M = (size_t)std::max(m, 1);
  const size_t C = (size_t)cap;
  size_t need = 8 * al(M * 8) + (want_rec64 ? al(M * REC64 * 8) : 0) + al(M * NEAR32 * 4) + al(table.size() * 4);
  need += al(((size_t)K << MB_MAX_CUT) * 4) + al((size_t)std::max<int64_t>(wgs, 1) * 4);
  need += al(C * 3 * 8) + al(C * 8);
  need += al(C * ACC_REC * 8) + al(C * FAR_REC * 8) + al(C * FAR_REC * 4);
  need += al(C * 6 * 8) + al(C * 30 * 8) + al(C * 8) + al(C * 4);
  const size_t nblk = (size_t)std::max<int64_t>(1, (m + BT_SPAN - 1) / BT_SPAN);
  need += al(nblk * BT_LEVELS * 2 * 4);
  OFW_HIP(c.tree[buf].reserve(need));
  char* base = (char*)c.tree[buf].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  MomentInputs mi;
  for (int q = 0; q < 3; ++q) mi.s[q] = (double*)carve(M * 8);
  for (int q = 0; q < 3; ++q) mi.l[q] = (double*)carve(M * 8);
  mi.rad = (double*)carve(M * 8);
  mi.gam = (double*)carve(M * 8);
  td.rec64 = want_rec64 ? (double*)carve(M * REC64 * 8) : nullptr;
  for (int q = 0; q < 3; ++q) {
    td.start[q] = mi.s[q];
    td.lenv[q] = mi.l[q];
  }
  td.rec32 = (float*)carve(M * NEAR32 * 4);
  int32_t* d_table = (int32_t*)carve(table.size() * 4);
  int32_t* d_top = (int32_t*)carve(((size_t)K << MB_MAX_CUT) * 4);
  int32_t* d_roots = (int32_t*)carve((size_t)std::max<int64_t>(wgs, 1) * 4);
  td.centre = (double*)carve(C * 3 * 8);
  td.core = (double*)carve(C * 8);
  td.acc = (double*)carve(C * ACC_REC * 8);
  td.far64 = (double*)carve(C * FAR_REC * 8);
  td.far32 = (float*)carve(C * FAR_REC * 4);
  NodeScratch ns;
  ns.box = (double*)carve(C * 6 * 8);
  ns.mom = (double*)carve(C * 30 * 8);
  ns.r2 = (double*)carve(C * 8);
  ns.slot = (int32_t*)carve(C * 4);
  int32_t* btab = (int32_t*)carve(nblk * BT_LEVELS * 2 * 4);
  td.count = cap;
  td.src = SegSource{t.order, starts, directions, lengths, gamma, radius, model};
  int st = upload_i32(c, table, d_table);
  if (st) return st;
  MomentForest mf{K, d_table, d_table + K, d_top, d_roots};
  const int grid = c.cu_count * 4;
  if (m > 0) {
    hipLaunchKernelGGL(k_tree_permute, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, t.order, starts, directions,
                       lengths, gamma, radius, model, mi, td.rec64, td.rec32);
    hipLaunchKernelGGL(k_moment_lists, dim3((unsigned)std::max(1, std::min(grid, (cap + 255) / 256))), dim3(256), 0,
                       c.stream, t, mf, btab);
    if (c.fork_pending) {  // a tree call's traversal plan may start on its second stream from here
      c.fork_pending = false;
      OFW_HIP(hipEventRecord(c.fork_ev, c.stream));
    }
    const FinalOut fo{want_outputs ? 0 : 1, clearance, td.acc, td.far64, td.far32};
    hipLaunchKernelGGL(k_moments_bottom, dim3((unsigned)wgs), dim3(BOTTOM_MT), 0, c.stream, t, mf, mi, ns,
                       td.centre, td.core, fo);
    for (int d = cut_max - 1; d >= MB_MAX_CUT; --d)
      hipLaunchKernelGGL(k_internal_level, dim3((unsigned)std::max(1, std::min(grid, (cap + 255) / 256))), dim3(256),
                         0, c.stream, t, mf, d, ns, td.centre, td.core);
    hipLaunchKernelGGL(k_moments_top, dim3(K), dim3(512), 0, c.stream, t, mf, ns, td.centre, td.core);
    // one range of BT_SPAN positions per workgroup
    hipLaunchKernelGGL(k_bound_top, dim3((unsigned)nblk), dim3(256), 0, c.stream, t, mf, mi, td.centre, ns.r2, btab);
  }
  NodeOutputs no = want_outputs ? td.out : NodeOutputs{};
  const int32_t* top_cut = (m > 0 && !want_outputs) ? d_table : nullptr;
  hipLaunchKernelGGL(k_moments_final, dim3((unsigned)std::max(1, std::min(grid, (cap + 255) / 256))), dim3(256), 0,
                     c.stream, t, ns, td.centre, td.core, clearance, no, td.acc, td.far64, td.far32, top_cut);
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
