// Tree-code evaluation on gfx950: a wavefront-cooperative, stackless walk of the node tree.
//
// A wavefront owns 64 targets (one per lane).  All lanes walk the tree together in pre-order (left child
// first, the order in which the reference's per-target stack visits nodes, vortex_tree.py:508-561), so the
// node index is wave-uniform: every node record and every near-field segment record is read with scalar
// loads and shared by the 64 lanes.  Each lane keeps its own decisions: a lane that accepts a node (far
// field) or finishes a leaf (near field) skips that node's subtree by waiting until the wave reaches the
// node's escape index (pre-order index + subtree size); the wave descends into a node's first child when
// at least one lane opens it, and jumps to the escape index otherwise.  Each lane therefore visits exactly
// the nodes its own walk in the reference visits, in the same order, and accumulates in that order.
//
// Occupancy: 20,570 targets are only 322 wavefronts.  The tree is cut at a frontier (the nodes at a
// fixed depth, and shallower leaves); a wavefront walks one frontier subtree for its 64 targets, after
// testing the frontier node's ancestors: an ancestor accepted by a lane is added by the subtree that
// holds the ancestor's leftmost segment, which is where the reference's walk adds it.  The per-subtree
// partial sums are added in pre-order by a second kernel, so the result is the reference's ordered sum
// re-associated at subtree boundaries, and the accepted nodes are exactly the reference's.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree.h"
#include "tree_eval.h"
#include "scan.h"
#include "tree_records.h"

namespace ofw {

static constexpr double INV_FOUR_PI_D = 1.0 / (4.0 * 3.141592653589793);  // vortex_tree.py:47

// The reference's acceptance test, operation by operation (vortex_tree.py:526-531): FP64 in every precision.
__device__ __forceinline__ bool accept_test(double x0, double x1, double x2, const double* __restrict__ a,
                                            double theta, double& rx, double& ry, double& rz, double& dist) {
#pragma clang fp contract(off)
  rx = x0 - a[A_CX];
  ry = x1 - a[A_CY];
  rz = x2 - a[A_CZ];
  dist = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
  const double bound = a[A_BOUND];
  return bound < theta * dist && dist - bound >= a[A_CLEAR];
}

// The acceptance test decided from the squared distance d2 where its answer is certain.  Both conditions are
// monotone in dist = sqrt(d2), and so is their floating-point evaluation (rounding is monotone): with
// b2 = (bound / theta)^2 and c2 = (bound + clear)^2, d2 above both by a relative 1e-9 (far more than the
// rounding of any of these steps) accepts as the reference's operations do, and d2 below either by 1e-9
// rejects.  Returns 1 (accept), 0 (reject) or -1: d2 within 1e-9 of a threshold or outside the range where
// these bounds hold, to be decided by accept_test.
__device__ __forceinline__ int accept_sure(double x0, double x1, double x2, const double* __restrict__ a,
                                           double inv_theta2) {
#pragma clang fp contract(off)
  const double rx = x0 - a[A_CX], ry = x1 - a[A_CY], rz = x2 - a[A_CZ];
  const double d2 = rx * rx + ry * ry + rz * rz;
  const double bound = a[A_BOUND];
  const double b2 = (bound * bound) * inv_theta2;
  const double sum = bound + a[A_CLEAR];
  const double c2 = sum * sum;
  const bool sane = d2 >= 1e-200 && d2 <= 1e200 && b2 <= 1e200 && c2 <= 1e200;
  const bool yes = d2 > b2 * (1.0 + 1e-9) && d2 > c2 * (1.0 + 1e-9);
  const bool no = d2 < b2 * (1.0 - 1e-9) || d2 < c2 * (1.0 - 1e-9);
  return !sane ? -1 : (yes ? 1 : (no ? 0 : -1));
}

// accept_sure split in two for a node tested by many lanes: the node's thresholds once (above t_yes the test
// accepts, below t_no it rejects; infinite when the node's b2 or c2 is out of range), then per lane its d2 against
// them.  Both answers equal accept_sure's: d2 > b2(1 + 1e-9) && d2 > c2(1 + 1e-9) is d2 > their maximum, and
// d2 < b2(1 - 1e-9) || d2 < c2(1 - 1e-9) is d2 < theirs.
__device__ __forceinline__ void accept_thresholds(double bound, double clear, double inv_theta2, double& t_yes,
                                                  double& t_no) {
#pragma clang fp contract(off)
  const double b2 = (bound * bound) * inv_theta2;
  const double sum = bound + clear;
  const double c2 = sum * sum;
  const bool sane = b2 <= 1e200 && c2 <= 1e200;
  t_yes = sane ? fmax(b2 * (1.0 + 1e-9), c2 * (1.0 + 1e-9)) : __builtin_inf();
  t_no = sane ? fmax(b2 * (1.0 - 1e-9), c2 * (1.0 - 1e-9)) : -__builtin_inf();
}
__device__ __forceinline__ int accept_sure_t(double x0, double x1, double x2, double cx, double cy, double cz,
                                             double t_yes, double t_no) {
#pragma clang fp contract(off)
  const double rx = x0 - cx, ry = x1 - cy, rz = x2 - cz;
  const double d2 = rx * rx + ry * ry + rz * rz;
  const bool sane = d2 >= 1e-200 && d2 <= 1e200;
  return sane && d2 > t_yes ? 1 : (sane && d2 < t_no ? 0 : -1);
}

// The same test on a record given as values (centre, bound, clear).
__device__ __forceinline__ bool accept_test5(double x0, double x1, double x2, const double* rec, double theta,
                                             double& rx, double& ry, double& rz, double& dist) {
#pragma clang fp contract(off)
  rx = x0 - rec[0];
  ry = x1 - rec[1];
  rz = x2 - rec[2];
  dist = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
  return rec[3] < theta * dist && dist - rec[3] >= rec[4];
}

// far_velocity (vortex_tree.py:450-497) in FP64, from the contracted moments of one node record
__device__ __forceinline__ void far64(double rx, double ry, double rz, double dist, const double* __restrict__ f,
                                      int order, double& ux, double& uy, double& uz) {
  const double ir = 1.0 / dist;
  const double ir2 = ir * ir;
  const double inv3 = ir2 * ir;
  const double inv5 = inv3 * ir2;
  const double a0 = f[FR_M0], a1 = f[FR_M0 + 1], a2 = f[FR_M0 + 2];
  double u0 = inv3 * (a1 * rz - a2 * ry);
  double u1 = inv3 * (a2 * rx - a0 * rz);
  double u2 = inv3 * (a0 * ry - a1 * rx);
  if (order >= 1) {
    const double v0 = f[FR_M1 + 0] * rx + f[FR_M1 + 1] * ry + f[FR_M1 + 2] * rz;
    const double v1 = f[FR_M1 + 3] * rx + f[FR_M1 + 4] * ry + f[FR_M1 + 5] * rz;
    const double v2 = f[FR_M1 + 6] * rx + f[FR_M1 + 7] * ry + f[FR_M1 + 8] * rz;
    const double k3 = 3.0 * inv5;
    u0 += k3 * (v1 * rz - v2 * ry) - inv3 * f[FR_W1];
    u1 += k3 * (v2 * rx - v0 * rz) - inv3 * f[FR_W1 + 1];
    u2 += k3 * (v0 * ry - v1 * rx) - inv3 * f[FR_W1 + 2];
    if (order >= 2) {
      const double inv7 = inv5 * ir2;
      const double xx = rx * rx, xy = rx * ry, xz = rx * rz, yy = ry * ry, yz = ry * rz, zz = rz * rz;
      double q[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double* m = f + FR_M2 + 6 * i;
        q[i] = m[0] * xx + m[1] * xy + m[2] * xz + m[3] * yy + m[4] * yz + m[5] * zz;
      }
      const double k75 = 7.5 * inv7, k15 = 1.5 * inv5;
      const double b0 = k75 * q[0] - k15 * f[FR_T2], b1 = k75 * q[1] - k15 * f[FR_T2 + 1],
                   b2 = k75 * q[2] - k15 * f[FR_T2 + 2];
      const double s0 = f[FR_W2 + 0] * rx + f[FR_W2 + 1] * ry + f[FR_W2 + 2] * rz;
      const double s1 = f[FR_W2 + 3] * rx + f[FR_W2 + 4] * ry + f[FR_W2 + 5] * rz;
      const double s2 = f[FR_W2 + 6] * rx + f[FR_W2 + 7] * ry + f[FR_W2 + 8] * rz;
      u0 += (b1 * rz - b2 * ry) - k3 * s0;
      u1 += (b2 * rx - b0 * rz) - k3 * s1;
      u2 += (b0 * ry - b1 * rx) - k3 * s2;
    }
  }
  ux += INV_FOUR_PI_D * u0;
  uy += INV_FOUR_PI_D * u1;
  uz += INV_FOUR_PI_D * u2;
}

// The same expansion in FP32 (moments rounded to FP32, R = x - c formed in FP64 and rounded).
__device__ __forceinline__ void far32(float rx, float ry, float rz, const float* __restrict__ f, int order,
                                      float& ux, float& uy, float& uz) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ir2 = ir * ir;
  const float inv3 = ir2 * ir;
  const float inv5 = inv3 * ir2;
  const float a0 = f[FR_M0], a1 = f[FR_M0 + 1], a2 = f[FR_M0 + 2];
  float u0 = inv3 * (a1 * rz - a2 * ry);
  float u1 = inv3 * (a2 * rx - a0 * rz);
  float u2 = inv3 * (a0 * ry - a1 * rx);
  if (order >= 1) {
    const float v0 = f[FR_M1 + 0] * rx + f[FR_M1 + 1] * ry + f[FR_M1 + 2] * rz;
    const float v1 = f[FR_M1 + 3] * rx + f[FR_M1 + 4] * ry + f[FR_M1 + 5] * rz;
    const float v2 = f[FR_M1 + 6] * rx + f[FR_M1 + 7] * ry + f[FR_M1 + 8] * rz;
    const float k3 = 3.0f * inv5;
    u0 += k3 * (v1 * rz - v2 * ry) - inv3 * f[FR_W1];
    u1 += k3 * (v2 * rx - v0 * rz) - inv3 * f[FR_W1 + 1];
    u2 += k3 * (v0 * ry - v1 * rx) - inv3 * f[FR_W1 + 2];
    if (order >= 2) {
      const float inv7 = inv5 * ir2;
      const float xx = rx * rx, xy = rx * ry, xz = rx * rz, yy = ry * ry, yz = ry * rz, zz = rz * rz;
      float q[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float* m = f + FR_M2 + 6 * i;
        q[i] = m[0] * xx + m[1] * xy + m[2] * xz + m[3] * yy + m[4] * yz + m[5] * zz;
      }
      const float k75 = 7.5f * inv7, k15 = 1.5f * inv5;
      const float b0 = k75 * q[0] - k15 * f[FR_T2], b1 = k75 * q[1] - k15 * f[FR_T2 + 1],
                  b2 = k75 * q[2] - k15 * f[FR_T2 + 2];
      const float s0 = f[FR_W2 + 0] * rx + f[FR_W2 + 1] * ry + f[FR_W2 + 2] * rz;
      const float s1 = f[FR_W2 + 3] * rx + f[FR_W2 + 4] * ry + f[FR_W2 + 5] * rz;
      const float s2 = f[FR_W2 + 6] * rx + f[FR_W2 + 7] * ry + f[FR_W2 + 8] * rz;
      u0 += (b1 * rz - b2 * ry) - k3 * s0;
      u1 += (b2 * rx - b0 * rz) - k3 * s1;
      u2 += (b0 * ry - b1 * rx) - k3 * s2;
    }
  }
  const float k = (float)INV_FOUR_PI_D;
  ux = k * u0;
  uy = k * u1;
  uz = k * u2;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// value of lane `src` (wave-uniform index) for every lane
__device__ __forceinline__ float bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double bcast(double v, int src) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Per-tree tables of a forest evaluation (K+1 entries each, int32, device).
struct ForestTables {
  int K;
  int node_cap;               // entries of the node record arrays
  int seg_count;              // entries of the segment record arrays
  const int32_t* toff;        // first target of each tree
  const int32_t* task_first;  // first traversal task of each tree
  const int32_t* wave_off;    // first work item (wavefront) of each tree
  const int32_t* part_off;    // first partial-sum slot of each tree
  const EvalTask* tasks;      // the traversal tasks (frontier subtrees) of all trees, tree after tree
  const int32_t* counters;
};

// A near-field work entry: a leaf (its segment range) and the lanes of the wavefront that opened it.
// An interaction-list entry of a walk wavefront: an opened leaf (its segment range lo < hi) or an accepted
// node (lo = -1 - node), and the lanes of the wavefront that opened / accepted it.
struct NearEntry {
  int32_t lo, hi;
  uint32_t mask_lo, mask_hi;
};
static constexpr int CHUNK = 8;  // list entries per work item of the list kernel
static_assert(CHUNK <= 64, "a work item's entries are loaded one per lane");
// cost classes of the work items (dispatched heaviest first; see k_item_waves)
static constexpr int ITEM_CLASSES = 8;
static constexpr int ITEM_CLASS_W = 32;  // segments per class step
static constexpr int ITEM_FAR = 2;       // cost of an accepted node's far field, in segments
static_assert(CHUNK == 8, "a work item's cost is summed over eight consecutive lanes");

// The cost class of the work item of entry b0 + lane of a list of `count` entries (b0 a multiple of 64): every
// lane of the item's eight holds it, and *leader is set on the first lane of each item that has entries.  With
// ATOMIC the entries are read past the vector cache (the calling wavefront has just written them).
template <bool ATOMIC>
__device__ __forceinline__ int item_class(const NearEntry* __restrict__ list, size_t lbase, int count, int b0,
                                          int lane, bool* leader) {
  const int e = b0 + lane;
  int cost = 0;
  if (e < count) {
    int lo, hi;
    if (ATOMIC) {
      const unsigned long long w0 = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(list + lbase + e),
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      lo = (int32_t)(uint32_t)w0;
      hi = (int32_t)(w0 >> 32);
    } else {
      lo = list[lbase + e].lo;
      hi = list[lbase + e].hi;
    }
    cost = lo < 0 ? ITEM_FAR : hi - lo;
  }
  cost += __shfl_xor(cost, 1, 64);
  cost += __shfl_xor(cost, 2, 64);
  cost += __shfl_xor(cost, 4, 64);
  *leader = (lane & 7) == 0 && e < count;
  return min(cost / ITEM_CLASS_W, ITEM_CLASSES - 1);
}

// the tree k of a wavefront / target: the last k with table[k] <= v (table[0] = 0 <= v < table[K])
__device__ __forceinline__ int find_tree(const int32_t* __restrict__ table, int K, int v) {
  int lo = 0, hi = K;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (table[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The interaction list of walk wavefront `local` of tree k (groups of 64 targets per task): its first entry, and
// its capacity in *cap.
__device__ __forceinline__ size_t wave_list(const ForestTables& ft, int k, int local, int groups, int* cap) {
  const int q = local / groups, g = local - q * groups;
  const EvalTask* T = ft.tasks + ft.task_first[k] + q;
  *cap = uniform(T->list_cap);
  return (size_t)uniform(T->list_off) + (size_t)g * (size_t)*cap;
}

// Acceptance records are read through a per-wavefront window of WIN consecutive pre-order nodes in LDS: the
// walk only moves forward in pre-order, so one batch of coalesced loads serves many visits.  The walk only
// decides: accepted nodes and opened leaves go to the wavefront's interaction list, whose far fields and
// pair sums the list kernel evaluates spread over the whole GPU.  Segments of the near field are staged
// one per lane in a window of 64 consecutive tree positions and broadcast with readlane.
static constexpr int WIN = 32;
static_assert(WIN * ACC_REC % 128 == 0, "the node window is loaded in 16-byte chunks by 64 lanes");
// The breadth-first walk: a FIFO of nodes in LDS (children queued in pairs that share one lane mask), the
// records of 64 queued nodes loaded together, and a bitmap of the list entries over the subtree's pre-order
// indices (two bits per node) that ranks every entry into pre-order.  Subtrees of more than BFS_NODES nodes,
// and walks whose queue or list outgrows its bound, take the stackless walk.
static constexpr int QCAP = 512;
static constexpr int BFS_NODES = 8192;
static_assert((QCAP & (QCAP - 1)) == 0, "the queue is a power-of-two ring");
static constexpr int EVAL_WAVES = 4;
// The walk and the list kernel run one wavefront per workgroup: their work items differ widely in cost, and a
// workgroup keeps its slot (and its LDS) until its slowest wavefront ends, so single-wavefront workgroups let the
// dispatcher start the next item wherever one has finished.
static constexpr int NEAR_WAVES = 1;

template <int MODEL, int PREC>
__global__ __launch_bounds__(64 * NEAR_WAVES) void k_tree_eval(ForestTables ft, const double* __restrict__ pts,
                                                   const int32_t* __restrict__ ex_ptr,
                                                   const int32_t* __restrict__ ex_pos,
                                                   const EvalTask* __restrict__ tasks,
                                                   const double* __restrict__ accr, const double* __restrict__ fr64,
                                                   const float* __restrict__ fr32, const double* __restrict__ rec64,
                                                   const float* __restrict__ rec32, double theta, int order,
                                                   int raise_singular, double* __restrict__ part,
                                                   int8_t* __restrict__ pflag, int32_t* __restrict__ pfar,
                                                   int32_t* __restrict__ pnear, int8_t* __restrict__ pnz,
                                                   NearEntry* __restrict__ list, int32_t* __restrict__ list_count,
                                                   int32_t* __restrict__ cls_wcnt) {
  typedef typename std::conditional<PREC == OFW_FP32, float, double>::type acc_t;
  // the stackless walk's node window, or the breadth-first walk's test records of one batch (8 doubles per node)
  __shared__ double s_acc[NEAR_WAVES][64 * 8];
  __shared__ int s_q[NEAR_WAVES][QCAP];
  __shared__ unsigned long long s_qm[NEAR_WAVES][QCAP / 2];
  __shared__ uint32_t s_bits[NEAR_WAVES][BFS_NODES / 16];
  __shared__ int s_cp[NEAR_WAVES][64];
  const int wv = threadIdx.x >> 6;
  // Workgroups go to the XCDs round-robin: the long walks, which belong to neighbouring groups of one task, are
  // spread over the XCDs and their compute units instead of sharing SIMDs.
  const int wave = uniform((int)blockIdx.x * NEAR_WAVES + wv);
  // whole wavefront beyond the work, or a plan that overflowed its capacity (reported by the host)
  if (wave >= ft.counters[C_WORK] || ft.counters[C_OVERFLOW]) return;
  const int lane = threadIdx.x & 63;
  // work item -> (tree, task of the tree, group of 64 targets): the groups of one task are adjacent
  const int k = uniform(find_tree(ft.wave_off, ft.K, wave));
  const int t0 = uniform(ft.toff[k]);
  const int nt = uniform(ft.toff[k + 1]) - t0;
  const int groups = (nt + 63) >> 6;
  const int local = wave - uniform(ft.wave_off[k]);
  const int q = local / groups;
  const int g = local - q * groups;
  const int task = uniform(ft.task_first[k]) + q;
  const int jl = g * 64 + lane;
  const bool valid = jl < nt;
  const int j = t0 + jl;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  int cur = 0, cend = 0;
  if (valid) {
    x0 = pts[3 * (size_t)j];
    x1 = pts[3 * (size_t)j + 1];
    x2 = pts[3 * (size_t)j + 2];
    if (ex_ptr) {
      cur = ex_ptr[j];
      cend = ex_ptr[j + 1];
    }
  }
  // The lane's excluded tree positions (sorted): the first eight in registers and their range, so that the
  // test of a node's segment range [lo, hi) needs no memory access (and no wait on the wavefront's stores) unless
  // the lane has more than eight and the range reaches past the eighth.
  int xs[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) xs[u] = cur + u < cend ? ex_pos[cur + u] : INT_MAX;
  const int xmin = xs[0];
  const int xmax = cend > cur ? ex_pos[cend - 1] : INT_MIN;
  auto has_ex = [&](int lo, int hi) -> bool {
    if (hi <= xmin || lo > xmax) return false;
    bool h = false;
#pragma unroll
    for (int u = 0; u < 8; ++u) h = h || (xs[u] >= lo && xs[u] < hi);
    if (!h && cend - cur > 8 && xs[7] < hi) {
      for (int c = cur + 8; c < cend; ++c) {
        const int v = ex_pos[c];
        if (v >= hi) break;
        if (v >= lo) {
          h = true;
          break;
        }
      }
    }
    return h;
  };
  acc_t ux = 0, uy = 0, uz = 0;
  int flag = OFW_FLAG_OK, far = 0, near = 0;
  bool alive = valid;
  const EvalTask& T = tasks[task];
  const int F = uniform(T.node);
  const int nanc = uniform(T.nanc);
  const double inv_theta2 = 1.0 / (theta * theta);

  auto add_far = [&](double rx, double ry, double rz, double dist, int node, const float* f32) {
    if (PREC == OFW_FP64) {
      double vx = 0.0, vy = 0.0, vz = 0.0;
      far64(rx, ry, rz, dist, fr64 + (size_t)node * FAR_REC, order, vx, vy, vz);
      ux += vx;
      uy += vy;
      uz += vz;
    } else {
      float vx, vy, vz;
      far32((float)rx, (float)ry, (float)rz, f32, order, vx, vy, vz);
      ux += (acc_t)vx;
      uy += (acc_t)vy;
      uz += (acc_t)vz;
    }
    ++far;
  };

  // ---- ancestors of the frontier node: their acceptance records loaded together, one per lane
  {
    double acx = 0.0, acy = 0.0, acz = 0.0, abound = 0.0, aclear = 0.0;
    int alo_l = 0, ahi_l = 0;
    if (lane < nanc) {
      const double* a = accr + (size_t)T.anc[lane] * ACC_REC;
      acx = a[A_CX];
      acy = a[A_CY];
      acz = a[A_CZ];
      abound = a[A_BOUND];
      aclear = a[A_CLEAR];
      const int32_t* ai = reinterpret_cast<const int32_t*>(a + A_INTS);
      alo_l = ai[0];
      ahi_l = ai[1];
    }
    const int flo = uniform(reinterpret_cast<const int32_t*>(accr + (size_t)F * ACC_REC + A_INTS)[0]);
    for (int a_i = 0; a_i < nanc; ++a_i) {
      const double rec[5] = {bcast(acx, a_i), bcast(acy, a_i), bcast(acz, a_i), bcast(abound, a_i),
                             bcast(aclear, a_i)};
      const int alo = __builtin_amdgcn_readlane(alo_l, a_i), ahi = __builtin_amdgcn_readlane(ahi_l, a_i);
      double rx, ry, rz, dist;
      bool acc = alive && accept_test5(x0, x1, x2, rec, theta, rx, ry, rz, dist);
      if (acc) acc = !has_ex(alo, ahi);
      if (acc) {
        alive = false;
        // the subtree that holds the ancestor's leftmost segment adds it, where the reference's walk does
        if (alo == flo) {
          const int A = uniform(T.anc[a_i]);
          add_far(rx, ry, rz, dist, A, fr32 + (size_t)A * FAR_REC);
        }
      }
    }
  }

  const int end = uniform(reinterpret_cast<const int32_t*>(accr + (size_t)F * ACC_REC + A_INTS)[2]);
  int wbase = INT_MIN / 2;
  auto load_window = [&](int base) {
    wbase = base;
    // the window is one contiguous block of each record array: each lane copies 16 bytes per chunk,
    // stored with the element type the window is read with
    const double* src = accr + (size_t)base * ACC_REC;
    constexpr int CHUNKS = WIN * ACC_REC / 128;  // 16-byte chunks per lane
    double v[2 * CHUNKS];
#pragma unroll
    for (int c = 0; c < CHUNKS; ++c) {
      const int e = 2 * lane + 128 * c;  // first double of this lane's chunk
      const bool in = base + e / ACC_REC < ft.node_cap;
      const double2 d = in ? *reinterpret_cast<const double2*>(src + e) : make_double2(0.0, 0.0);
      v[2 * c] = d.x;
      v[2 * c + 1] = d.y;
    }
#pragma unroll
    for (int c = 0; c < CHUNKS; ++c) {
      s_acc[wv][2 * lane + 128 * c] = v[2 * c];
      s_acc[wv][2 * lane + 128 * c + 1] = v[2 * c + 1];
    }
  };
  // near-field list of this wavefront
  const int lcap = uniform(T.list_cap);
  const size_t lbase = (size_t)uniform(T.list_off) + (size_t)g * lcap;
  int nlist = 0;
  bool walked = false;
  if (end - F <= BFS_NODES) {
    // ---- breadth-first walk of the frontier subtree.  A batch of up to 64 queued nodes is taken in two
    // phases: (A) every node's acceptance test for every lane, four nodes at a time so that their arithmetic
    // overlaps, the node's accept and open lane masks written into the lane that holds the node; (B) every lane's
    // node emits its list entries and queues its children at once, at offsets from one scan over the lanes.
    const int far0 = far;
    const int words = (2 * (end - F) + 31) >> 5;
    for (int w = lane; w < words; w += 64) s_bits[wv][w] = 0u;
    const unsigned long long alive_mask = __builtin_amdgcn_ballot_w64(alive);
    int head = 0, tail = 0;
    if (alive_mask) {
      if (lane == 0) {
        s_q[wv][0] = F;
        s_q[wv][1] = -1;
        s_qm[wv][0] = alive_mask;
      }
      tail = 2;
    }
    wave_lds_sync();
    int nb = 0;
    bool ok = true;
    while (head < tail) {
      const int nbatch = min(64, tail - head);
      int my_node = -1;
      unsigned long long my_mask = 0ull;
      if (lane < nbatch) {
        const int idx = (head + lane) & (QCAP - 1);
        my_node = s_q[wv][idx];
        my_mask = s_qm[wv][idx >> 1];
      }
      head += nbatch;
      // The batch's records, one per lane (queued nodes lie in [F, end): children have larger pre-order indices
      // within the frontier subtree, so the walk ends).  The lane that holds a node writes its test record to LDS
      // for broadcast reads: the centre, the two squared-distance thresholds of accept_sure (above t_yes it
      // accepts, below t_no it rejects, in between the reference's operations decide), (lo, hi), the lane mask
      // and the node index.
      const bool mine = my_node >= F && my_node < end;
      int r_skip = 0, r_right = -1;
      {
        double cx = 0.0, cy = 0.0, cz = 0.0, t_yes = 0.0, t_no = 0.0;
        long long lohi = 0;
        if (mine) {
          const double2* src = reinterpret_cast<const double2*>(accr + (size_t)my_node * ACC_REC);
          const double2 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
          cx = v0.x;
          cy = v0.y;
          cz = v1.x;
          const double bound = v1.y, clear = v2.x;
          lohi = __double_as_longlong(v2.y);
          const long long i1 = __double_as_longlong(v3.x), i2 = __double_as_longlong(v3.y);
          r_skip = (int)(i1 & 0xffffffffll);
          r_right = (int)(i2 & 0xffffffffll);
          accept_thresholds(bound, clear, inv_theta2, t_yes, t_no);
        }
        double* t = &s_acc[wv][lane * 8];
        t[0] = cx;
        t[1] = cy;
        t[2] = cz;
        t[3] = t_yes;
        t[4] = t_no;
        t[5] = __longlong_as_double(lohi);
        t[6] = __longlong_as_double(mine ? (long long)my_mask : 0ll);
        t[7] = __longlong_as_double((long long)my_node);
      }
      // Wait here, once per batch, for the loads above: a wait on a load inside the loops below would also wait
      // for every list store issued before it (loads and stores share one counter).
      __builtin_amdgcn_s_waitcnt(0);
      wave_lds_sync();
      // (A) lane e collects node e's masks
      int am_lo = 0, am_hi = 0, om_lo = 0, om_hi = 0;
      for (int e0 = 0; e0 < nbatch; e0 += 4) {
        bool act[4], acc[4], unc[4];
        int lo[4], hi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double* t = &s_acc[wv][min(e0 + u, 63) * 8];
          const double cx = t[0], cy = t[1], cz = t[2], t_yes = t[3], t_no = t[4];
          const long long lohi = __double_as_longlong(t[5]);
          const unsigned long long am = (unsigned long long)__double_as_longlong(t[6]);
          lo[u] = (int)(lohi & 0xffffffffll);
          hi[u] = (int)(lohi >> 32);
          act[u] = e0 + u < nbatch && ((am >> lane) & 1ull);
          const int sure = accept_sure_t(x0, x1, x2, cx, cy, cz, t_yes, t_no);
          acc[u] = act[u] && sure > 0;
          unc[u] = act[u] && sure < 0;
        }
        if (wave_any(unc[0] || unc[1] || unc[2] || unc[3])) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (unc[u]) {
              const double* t = &s_acc[wv][min(e0 + u, 63) * 8];
              const int node = (int)__double_as_longlong(t[7]);
              const double* g = accr + (size_t)node * ACC_REC;
              double rx, ry, rz, dist;
              acc[u] = accept_test(x0, x1, x2, g, theta, rx, ry, rz, dist);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (acc[u]) acc[u] = !has_ex(lo[u], hi[u]);
          far += acc[u] ? 1 : 0;
          const unsigned long long amask = __builtin_amdgcn_ballot_w64(acc[u]);
          const unsigned long long omask = __builtin_amdgcn_ballot_w64(act[u] && !acc[u]);
          if (lane == e0 + u) {
            am_lo = (int)(uint32_t)amask;
            am_hi = (int)(uint32_t)(amask >> 32);
            om_lo = (int)(uint32_t)omask;
            om_hi = (int)(uint32_t)(omask >> 32);
          }
        }
      }
      // (B) list entries (an accepted node, then an opened leaf) staged from the top of the wavefront's list,
      // their bits marking their pre-order places, and both children of an opened node queued with the lanes
      // that opened it (vortex_tree.py:526-561 for each lane)
      const unsigned long long my_am = ((unsigned long long)(uint32_t)am_hi << 32) | (uint32_t)am_lo;
      const unsigned long long my_om = ((unsigned long long)(uint32_t)om_hi << 32) | (uint32_t)om_lo;
      const bool leaf = mine && r_skip == my_node + 1;
      const int n_ent = mine ? (my_am ? 1 : 0) + (leaf && my_om ? 1 : 0) : 0;
      const int n_push = mine && !leaf && my_om ? 2 : 0;
      const int packed = n_ent | (n_push << 16);  // at most 2 each per lane: 16-bit fields suffice
      int incl = packed;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      const int total = __builtin_amdgcn_readlane(incl, 63);
      const int tot_ent = total & 0xffff, tot_push = total >> 16;
      if (2 * (nb + tot_ent) > lcap || tail + tot_push - head > QCAP) {
        ok = false;
        break;
      }
      if (mine) {
        const int before = incl - packed;
        const int code = 2 * (my_node - F);
        int b = nb + (before & 0xffff);
        if (my_am) {
          list[lbase + lcap - 1 - b] = NearEntry{code, 0, (uint32_t)my_am, (uint32_t)(my_am >> 32)};
          atomicOr(&s_bits[wv][code >> 5], 1u << (code & 31));
          ++b;
        }
        if (leaf && my_om) {
          list[lbase + lcap - 1 - b] = NearEntry{code + 1, 0, (uint32_t)my_om, (uint32_t)(my_om >> 32)};
          atomicOr(&s_bits[wv][(code + 1) >> 5], 1u << ((code + 1) & 31));
        }
        if (n_push) {
          const int t = tail + (before >> 16);
          s_q[wv][t & (QCAP - 1)] = my_node + 1;
          s_q[wv][(t + 1) & (QCAP - 1)] = r_right;
          s_qm[wv][(t & (QCAP - 1)) >> 1] = my_om;
        }
      }
      nb += tot_ent;
      tail += tot_push;
      wave_lds_sync();  // the queue and bitmap writes come before the next batch's reads
    }
    if (ok) {
      // rank of each entry: its bit's place among the set bits (a scan of per-lane runs of words)
      const int per = (words + 63) >> 6;
      int cnt = 0;
      for (int w = lane * per; w < min(words, lane * per + per); ++w) cnt += __popc(s_bits[wv][w]);
      int incl = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      s_cp[wv][lane] = incl - cnt;
      wave_lds_sync();
      // the staged entries were written by other lanes: their stores complete first, and they are read from L2
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0);
      for (int b0 = 0; b0 < nb; b0 += 64) {
        const int b = b0 + lane;
        if (b < nb) {
          unsigned long long* sp = reinterpret_cast<unsigned long long*>(list + lbase + lcap - 1 - b);
          const unsigned long long w0 = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned long long w1 = __hip_atomic_load(sp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const NearEntry st{(int32_t)(uint32_t)w0, (int32_t)(w0 >> 32), (uint32_t)w1, (uint32_t)(w1 >> 32)};
          const int code = st.lo;
          if (code >= 0 && code < 2 * (end - F)) {
            const int w = code >> 5;
            const int c = w / per;
            int r = s_cp[wv][c];
            for (int u = c * per; u < w; ++u) r += __popc(s_bits[wv][u]);
            r += __popc(s_bits[wv][w] & ((1u << (code & 31)) - 1u));
            const int node = F + (code >> 1);
            NearEntry out{-1 - node, 0, st.mask_lo, st.mask_hi};
            if (code & 1) {
              const int32_t* ai = reinterpret_cast<const int32_t*>(accr + (size_t)node * ACC_REC + A_INTS);
              out.lo = ai[0];
              out.hi = ai[1];
            }
            if (r < nb) list[lbase + r] = out;
          }
        }
      }
      nlist = nb;
      walked = true;
    } else {
      far = far0;
    }
  }
  if (!walked) {
    // ---- stackless walk of the frontier subtree through the node window
    int i = F;
    int skip_until = F;
    while (i < end) {
      const bool act = alive && i >= skip_until;
      if (!wave_any(act)) break;
      if (i >= wbase + WIN) load_window(i);
      const double* a = &s_acc[wv][(i - wbase) * ACC_REC];
      // (lo, hi) and (skip, tree) are int32 pairs in two record slots
      const long long ints0 = __double_as_longlong(a[A_INTS]), ints1 = __double_as_longlong(a[A_INTS2]);
      const int lo = uniform((int)(ints0 & 0xffffffffll)), hi = uniform((int)(ints0 >> 32));
      const int skip = uniform((int)(ints1 & 0xffffffffll));
      int sure = 0;
      if (act) sure = accept_sure(x0, x1, x2, a, inv_theta2);
      bool acc = sure > 0;
      if (wave_any(act && sure < 0)) {
        if (act && sure < 0) {
          double rx, ry, rz, dist;
          acc = accept_test(x0, x1, x2, a, theta, rx, ry, rz, dist);
        }
      }
      if (acc) acc = !has_ex(lo, hi);
      if (wave_any(acc)) {
        // the accepted node goes to the wavefront's interaction list with the lanes that accepted it; the list
        // kernel adds its far field (vortex_tree.py:526-537 for each of them)
        const unsigned long long amask = __builtin_amdgcn_ballot_w64(acc);
        if (nlist < lcap) {
          if (lane == 0) list[lbase + nlist] = NearEntry{-1 - i, 0, (uint32_t)amask, (uint32_t)(amask >> 32)};
          ++nlist;
        } else if (lane == 0) {
          atomicOr(const_cast<int32_t*>(ft.counters) + C_OVERFLOW, 1);
        }
        if (acc) ++far;
      }
      const bool open = act && !acc;
      if (skip == i + 1) {
        // leaf: the near field of the lanes that did not accept it (vortex_tree.py:538-557)
        if (wave_any(open)) {
          // the leaf goes to the wavefront's near-field list with the lanes that opened it; the near-field
          // kernel evaluates the lists in short chunks spread over the whole GPU
          const unsigned long long mask = __builtin_amdgcn_ballot_w64(open);
          if (nlist < lcap) {
            if (lane == 0) list[lbase + nlist] = NearEntry{lo, hi, (uint32_t)mask, (uint32_t)(mask >> 32)};
            ++nlist;
          } else if (lane == 0) {
            atomicOr(const_cast<int32_t*>(ft.counters) + C_OVERFLOW, 1);
          }
        }
        i = i + 1;
      } else {
        if (acc) skip_until = skip;
        i = wave_any(open) ? i + 1 : skip;
      }
    }
  }
  // partial sums only from wavefronts that added anything (the reduction reads the flag)
  const bool any = wave_any(far > 0 || near > 0);
  if (lane == 0) {
    pnz[wave] = any ? 1 : 0;
    list_count[wave] = nlist;
  }
  // the list's work items per cost class (lane c counts class c), for the heaviest-first dispatch of the list
  // kernel; the entries were written by other lanes, whose stores complete first
  {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    int cnt = 0;
    for (int b0 = 0; b0 < nlist; b0 += 64) {
      bool lead;
      const int cls = item_class<true>(list, lbase, nlist, b0, lane, &lead);
#pragma unroll
      for (int cc = 0; cc < ITEM_CLASSES; ++cc) {
        const int nc = __popcll(__builtin_amdgcn_ballot_w64(lead && cls == cc));
        if (lane == cc) cnt += nc;
      }
    }
    if (lane < ITEM_CLASSES) cls_wcnt[(size_t)wave * ITEM_CLASSES + lane] = cnt;
  }
  if (any && valid) {
    const size_t o = (size_t)uniform(ft.part_off[k]) + (size_t)q * nt + jl;
    part[3 * o] = (double)ux;
    part[3 * o + 1] = (double)uy;
    part[3 * o + 2] = (double)uz;
    pflag[o] = (int8_t)flag;
    if (pfar) {
      pfar[o] = far;
      pnear[o] = near;
    }
  }
}

// Near field of the traversal: one wavefront per work item, CHUNK consecutive entries of one walk
// wavefront's list, for the same 64 targets.  Below FP64 the leaf's segments are staged one per lane in a
// window of 64 consecutive tree positions and broadcast with readlane; in FP64 the 128-byte record is read
// with scalar loads.  Every lane evaluates the pair and the lanes that opened the leaf select the result
// (vortex_tree.py:538-557 for each of them).  A grid-stride
// loop covers the work items, whose count is known on the device only.
// The near field's segments in tree order: FP64 records (FP64 evaluations), or the starts and, for the few
// pairs that take the FP64 path, the engine inputs their records are packed from.
struct NearSegs {
  const double* rec64;
  const double* start[3];
  const double* lenv[3];
  SegSource src;
};

template <int MODEL, int PREC>
__global__ __launch_bounds__(64 * NEAR_WAVES) __attribute__((amdgpu_waves_per_eu(4))) void k_tree_near(ForestTables ft, const double* __restrict__ pts,
                                                   const int32_t* __restrict__ ex_ptr,
                                                   const int32_t* __restrict__ ex_pos,
                                                   NearSegs segs, const float* __restrict__ rec32,
                                                   const double* __restrict__ accr, const double* __restrict__ fr64,
                                                   const float* __restrict__ fr32, int order,
                                                   int raise_singular, const NearEntry* __restrict__ list,
                                                   const int32_t* __restrict__ list_count,
                                                   const int32_t* __restrict__ chunk_off,
                                                   const int32_t* __restrict__ item_wave, int chunk_slots,
                                                   const int32_t* __restrict__ cls_count,
                                                   const int32_t* __restrict__ cls_list,
                                                   double* __restrict__ cpart, int8_t* __restrict__ cflag,
                                                   int32_t* __restrict__ cnear) {
  typedef typename std::conditional<PREC == OFW_FP32, float, double>::type acc_t;
  typedef typename std::conditional<PREC == OFW_FP64, double, float>::type far_t;
  // far-field records (and centres) of the accepted nodes of one work item, per wavefront
  __shared__ far_t s_far[NEAR_WAVES][CHUNK][FAR_REC];
  __shared__ double s_cen[NEAR_WAVES][CHUNK][4];
  // below FP64, a window of 64 consecutive segments: FP32 records and FP64 starts
  constexpr int SEGW = PREC == OFW_FP64 ? 1 : 64;
  __shared__ float s_seg[NEAR_WAVES][SEGW][NEAR32];
  __shared__ double s_start[NEAR_WAVES][SEGW][4];
  __shared__ double s_lenv[NEAR_WAVES][SEGW][4];   // length vectors: the pair's cross product in FP64
  // below FP64, the FP64 record of the segment whose pairs take the FP64 path (the same for every lane), and the
  // engine index of each window segment
  __shared__ double s_rec[NEAR_WAVES][REC64];
  __shared__ int s_orig[NEAR_WAVES][SEGW];
  const far_t* __restrict__ frec = PREC == OFW_FP64 ? (const far_t*)fr64 : (const far_t*)fr32;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int works = min(live_count(ft.counters, C_CHUNKS), chunk_slots);
  const int nwaves = live_count(ft.counters, C_WORK);
  // one dispatch slot per wavefront (the grid covers every slot): the loop runs once, so nothing stays live
  // from one item to the next
  for (int slot = uniform(blockIdx.x * NEAR_WAVES + (int)(threadIdx.x >> 6)); slot < works; slot = works) {
    // the item of this dispatch slot: the classes heaviest first
    int item = -1;
    {
      int before = 0;
      for (int cc = ITEM_CLASSES - 1; cc >= 0; --cc) {
        const int cnt = uniform(cls_count[cc]);
        if (slot < before + cnt) {
          item = uniform(cls_list[(size_t)cc * chunk_slots + (slot - before)]);
          break;
        }
        before += cnt;
      }
    }
    if (item < 0) continue;  // cannot happen: the classes list every item below the slots
    // the walk wavefront of this item
    const int w = uniform(item_wave[item]);
    const int c = item - uniform(chunk_off[w]);
    const int k = uniform(find_tree(ft.wave_off, ft.K, w));
    const int t0 = uniform(ft.toff[k]);
    const int nt = uniform(ft.toff[k + 1]) - t0;
    const int groups = (nt + 63) >> 6;
    const int local = w - uniform(ft.wave_off[k]);
    const int g = local % groups;
    const int jl = g * 64 + lane;
    const bool valid = jl < nt;
    const int j = t0 + jl;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (valid) {
      x0 = pts[3 * (size_t)j];
      x1 = pts[3 * (size_t)j + 1];
      x2 = pts[3 * (size_t)j + 2];
    }
    acc_t ux = 0, uy = 0, uz = 0;
    int flag = OFW_FLAG_OK, near = 0;
    int sbase = INT_MIN / 2;
    int lcap_unused;
    const size_t lbase = wave_list(ft, k, local, groups, &lcap_unused);
    const int e0 = c * CHUNK;
    const int ne = min(uniform(list_count[w]), e0 + CHUNK) - e0;
    // the item's entries, one per lane, and the records of its accepted nodes staged in LDS together (lane
    // 8 q' + u copies part u of the record of entry q' + 8 r in round r), so that no entry waits for its own loads
    NearEntry mine{0, 0, 0u, 0u};
    if (lane < ne) mine = list[lbase + e0 + lane];
    wave_lds_sync();  // the previous item's reads of the stage come first
#pragma unroll
    for (int r = 0; r < (CHUNK + 7) / 8; ++r) {
      const int q = (lane >> 3) + 8 * r, u = lane & 7;
      const int lo_q = __shfl(mine.lo, q, 64);
      if (q < ne && lo_q < 0) {
        const int node = -1 - lo_q;
        const far_t* src = frec + (size_t)node * FAR_REC;
#pragma unroll
        for (int v = u; v < FAR_REC; v += 8) s_far[wv][q][v] = src[v];
        if (u < 3) s_cen[wv][q][u] = accr[(size_t)node * ACC_REC + A_CX + u];
      }
    }
    wave_lds_sync();
    // the accepted nodes first, then the opened leaves: the far field's registers and the near field's (the
    // exclusion cursor, the segment window) are never live together
    for (int e = 0; e < ne; ++e) {
      const int lo = __builtin_amdgcn_readlane(mine.lo, e);
      if (lo >= 0) continue;
      const unsigned long long mask =
          ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)mine.mask_hi, e) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)mine.mask_lo, e);
      const bool open = valid && ((mask >> lane) & 1ull);
      // an accepted node: its far field for the lanes that accepted it, as the walk's acceptance test formed
      // R and its length (vortex_tree.py:526-537)
      if (open) {
#pragma clang fp contract(off)
        const double rx = x0 - s_cen[wv][e][0], ry = x1 - s_cen[wv][e][1], rz = x2 - s_cen[wv][e][2];
        if (PREC == OFW_FP64) {
          const double dist = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
          double vx = 0.0, vy = 0.0, vz = 0.0;
          far64(rx, ry, rz, dist, (const double*)&s_far[wv][e][0], order, vx, vy, vz);
          ux += vx;
          uy += vy;
          uz += vz;
        } else {
          float vx, vy, vz;
          far32((float)rx, (float)ry, (float)rz, (const float*)&s_far[wv][e][0], order, vx, vy, vz);
          ux += (acc_t)vx;
          uy += (acc_t)vy;
          uz += (acc_t)vz;
        }
      }
    }
    int cur = 0, cend = 0;
    if (valid && ex_ptr) {
      cur = ex_ptr[j];
      cend = ex_ptr[j + 1];
    }
    // the next four excluded positions in registers (see k_tree_eval)
    int nextex = INT_MAX, e1 = INT_MAX, e2 = INT_MAX, e3 = INT_MAX;
    auto refill = [&]() {
      nextex = cur < cend ? ex_pos[cur] : INT_MAX;
      e1 = cur + 1 < cend ? ex_pos[cur + 1] : INT_MAX;
      e2 = cur + 2 < cend ? ex_pos[cur + 2] : INT_MAX;
      e3 = cur + 3 < cend ? ex_pos[cur + 3] : INT_MAX;
    };
    refill();
    auto advance_to = [&](int pos) {
      while (nextex < pos) {
        ++cur;
        nextex = e1;
        e1 = e2;
        e2 = e3;
        e3 = INT_MAX;
        if (nextex == INT_MAX && cur < cend) refill();
      }
    };
    for (int e = 0; e < ne; ++e) {
      const int lo = __builtin_amdgcn_readlane(mine.lo, e), hi = __builtin_amdgcn_readlane(mine.hi, e);
      if (lo < 0) continue;
      const unsigned long long mask =
          ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)mine.mask_hi, e) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)mine.mask_lo, e);
      const bool open = valid && ((mask >> lane) & 1ull);
      if (open) advance_to(lo);
      if (PREC == OFW_FP64) {
        // FP64: the segment record is wave-uniform (scalar loads); an excluded pair is evaluated, not added
        for (int s = lo; s < hi; ++s) {
          const bool hit = open && nextex == s;
          if (wave_any(hit)) {
            if (hit) advance_to(s + 1);
          }
          const bool use = open && !hit;
          double vx = 0.0, vy = 0.0, vz = 0.0;
          const int fl = pair_fast64<MODEL>(x0, x1, x2, segs.rec64 + (size_t)s * REC64, raise_singular, use, vx, vy,
                                            vz);
          ux = use ? ux + (acc_t)vx : ux;
          uy = use ? uy + (acc_t)vy : uy;
          uz = use ? uz + (acc_t)vz : uz;
          near += use ? 1 : 0;
          if (fl != OFW_FLAG_OK && flag == OFW_FLAG_OK) flag = fl;
        }
        continue;
      }
      for (int base = lo; base < hi;) {
        if (base < sbase || base >= sbase + 64) {
          sbase = base;
          const int s = base + lane;
          wave_lds_sync();  // the previous window's reads come first
          if (s < ft.seg_count) {
            const float4* f4 = reinterpret_cast<const float4*>(rec32 + (size_t)s * NEAR32);
            const float4 fa = f4[0], fb = f4[1];
            const double qx = segs.start[0][s], qy = segs.start[1][s], qz = segs.start[2][s];
            const double lx = segs.lenv[0][s], ly = segs.lenv[1][s], lz = segs.lenv[2][s];
            s_orig[wv][lane] = segs.src.order[s];
            s_lenv[wv][lane][0] = lx;
            s_lenv[wv][lane][1] = ly;
            s_lenv[wv][lane][2] = lz;
            float4* d4 = reinterpret_cast<float4*>(&s_seg[wv][lane][0]);
            d4[0] = fa;
            d4[1] = fb;
            s_start[wv][lane][0] = qx;
            s_start[wv][lane][1] = qy;
            s_start[wv][lane][2] = qz;
          }
          wave_lds_sync();
        }
        const int stop = min(hi, sbase + 64);
        // the window's pairs summed in FP32 and added to the target's sum once; excluded segments are looked
        // for pair by pair only in a window where an open lane has one
        const bool any_ex = wave_any(open && nextex < stop);
        float lx = 0.0f, ly = 0.0f, lz = 0.0f;
        int hits = 0;
        for (int s = base; s < stop; ++s) {
          // the pair's record and start: the same LDS address on every lane (a broadcast read)
          const int tl = s - sbase;
          const Near32 fc = load_near32(&s_seg[wv][tl][0]);
          const double cx0 = s_start[wv][tl][0], cy0 = s_start[wv][tl][1], cz0 = s_start[wv][tl][2];
          // every lane evaluates the FP32 pair; the lanes that use it select the result, and only the rare
          // excluded segment and the FP64 pairs branch (on the wavefront's ballot)
          bool use = open;
          if (any_ex) {
            const bool hit = open && nextex == s;
            if (wave_any(hit)) {
              if (hit) advance_to(s + 1);
            }
            use = open && !hit;
            hits += hit ? 1 : 0;
          }
          const double sx = x0 - cx0, sy = x1 - cy0, sz = x2 - cz0;
          // C = l x r formed in FP64: h keeps its FP32 accuracy next to the line
          const double gx = s_lenv[wv][tl][0], gy = s_lenv[wv][tl][1], gz = s_lenv[wv][tl][2];
          const float ccx = (float)(gy * sz - gz * sy), ccy = (float)(gz * sx - gx * sz),
                      ccz = (float)(gx * sy - gy * sx);
          float coef;
          const bool need64 = pair_near32<MODEL>((float)sx, (float)sy, (float)sz, ccx, ccy, ccz, fc, coef);
          const bool add = use && !need64;
          const float cw = add ? coef : 0.0f;
          lx = __builtin_fmaf(cw, ccx, lx);
          ly = __builtin_fmaf(cw, ccy, ly);
          lz = __builtin_fmaf(cw, ccz, lz);
          const bool slow = use && need64;
          const unsigned long long slow_lanes = __builtin_amdgcn_ballot_w64(slow);
          if (slow_lanes) {
            // the segment's FP64 record, packed from the engine inputs by one lane
            double* r = &s_rec[wv][0];
            if (lane == (int)__builtin_ctzll(slow_lanes)) {
              const SegSource& q = segs.src;
              pack_record(s_orig[wv][tl], q.starts, q.directions, q.lengths, q.gamma, q.radius, q.model, r, nullptr);
            }
            wave_lds_sync();
            if (slow) {
              double wx = 0.0, wy = 0.0, wz = 0.0;
              const int fl = pair_fast64<MODEL>(x0, x1, x2, r, raise_singular, wx, wy, wz);
              if (fl != OFW_FLAG_OK) {
                if (flag == OFW_FLAG_OK) flag = fl;
              } else {
                ux += (acc_t)wx;
                uy += (acc_t)wy;
                uz += (acc_t)wz;
              }
            }
            wave_lds_sync();  // every lane has read the record before the next one is packed
          }
        }
        ux += (acc_t)lx;
        uy += (acc_t)ly;
        uz += (acc_t)lz;
        near += open ? (stop - base) - hits : 0;
        base = stop;
      }
    }
    if (valid) {
      const size_t o = (size_t)item * 64 + lane;
      cpart[3 * o] = (double)ux;
      cpart[3 * o + 1] = (double)uy;
      cpart[3 * o + 2] = (double)uz;
      cflag[o] = (int8_t)flag;
      if (cnear) cnear[o] = near;
    }
  }
}

// The work items of the list kernel are dispatched heaviest first: the kernel's end is then set by the light
// items, not by a heavy one started late.  An item's cost is its leaf entries' segments plus ITEM_FAR per accepted
// node; class c = min(cost / ITEM_CLASS_W, ITEM_CLASSES - 1) lists the item in cls_list[c] (order within a class
// from atomics: the dispatch order does not change any result, every item writes its own partial slot).

// One wavefront per walk wavefront: the walk wavefront of each of its work items (within the `slots` partial
// slots), and the items listed by cost class (cls_list[c], `slots` entries per class) at the wavefront's offsets
// within each class (cls_woff), in item order; more work items than the fixed number `limit` of slots is reported
// as an overflow.
__global__ __launch_bounds__(256) void k_item_waves(ForestTables ft, const int32_t* __restrict__ chunk_off, int slots,
                                                    int limit, const NearEntry* __restrict__ list,
                                                    const int32_t* __restrict__ list_count,
                                                    const int32_t* __restrict__ cls_woff,
                                                    int32_t* __restrict__ item_wave, int32_t* __restrict__ cls_list) {
  int32_t* counters = const_cast<int32_t*>(ft.counters);
  if (blockIdx.x == 0 && threadIdx.x == 0 && counters[C_CHUNKS] > limit) counters[C_OVERFLOW] = 1;
  const int nwaves = live_count(ft.counters, C_WORK);
  const int w = uniform((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (w >= nwaves) return;
  const int lane = threadIdx.x & 63;
  const int k = uniform(find_tree(ft.wave_off, ft.K, w));
  const int local = w - uniform(ft.wave_off[k]);
  const int nt = uniform(ft.toff[k + 1]) - uniform(ft.toff[k]);
  int lcap_unused;
  const size_t lbase = wave_list(ft, k, local, (nt + 63) >> 6, &lcap_unused);
  const int count = uniform(list_count[w]);
  const int it0 = uniform(chunk_off[w]), it1 = min(uniform(chunk_off[w + 1]), slots);
  const int woff = lane < ITEM_CLASSES ? cls_woff[(size_t)w * ITEM_CLASSES + lane] : 0;
  int run = 0;  // lane c: items of class c placed so far
  for (int b0 = 0; b0 < count; b0 += 64) {
    bool lead;
    const int cls = item_class<false>(list, lbase, count, b0, lane, &lead);
    const int it = it0 + ((b0 + lane) >> 3);
    if (lead && it < it1) item_wave[it] = w;
#pragma unroll
    for (int cc = 0; cc < ITEM_CLASSES; ++cc) {
      const bool mine = lead && cls == cc;
      const unsigned long long mask = __builtin_amdgcn_ballot_w64(mine);
      if (mask == 0ull) continue;
      const int base = __builtin_amdgcn_readlane(woff + run, cc);
      const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      const int pos = base + rank;
      if (mine && it < it1 && pos < slots) cls_list[(size_t)cc * slots + pos] = it;
      if (lane == cc) run += __popcll(mask);
    }
  }
}

// The near-field work items of each walk wavefront, ceil(list entries / CHUNK) (zero past the planned waves),
// scanned into chunk_off (W1 entries), and the wavefronts' items per cost class (cls_wcnt, counted by the walk)
// scanned into each wavefront's offset within the class's list (cls_woff): 256 wavefronts per workgroup, a
// single-pass look-back across the workgroups (tiles taken by ticket; the status words and the ticket were
// cleared by k_work_plan).  The last tile writes the totals: counters[C_CHUNKS] and cls_count.
static constexpr int PLAN_VALUES = 1 + ITEM_CLASSES;
__global__ __launch_bounds__(256) void k_chunk_plan(ForestTables ft, int W1, const int32_t* __restrict__ list_count,
                                                    const int32_t* __restrict__ cls_wcnt,
                                                    int32_t* __restrict__ chunk_off, int32_t* __restrict__ cls_woff,
                                                    int32_t* __restrict__ counters, int32_t* __restrict__ cls_count,
                                                    int32_t* __restrict__ ticket, unsigned long long* __restrict__ status) {
  __shared__ int s_tile;
  __shared__ int s_prefix[PLAN_VALUES];
  const int nwaves = live_count(ft.counters, C_WORK);
  const int ntiles = (W1 + 255) / 256;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1);
  __syncthreads();
  const int tile = s_tile;
  if (tile >= ntiles) return;
  const int w = tile * 256 + (int)threadIdx.x;
  const bool live = w < nwaves;
  int v[PLAN_VALUES];
  v[0] = live ? (list_count[w] + CHUNK - 1) / CHUNK : 0;
#pragma unroll
  for (int cc = 0; cc < ITEM_CLASSES; ++cc) v[1 + cc] = live ? cls_wcnt[(size_t)w * ITEM_CLASSES + cc] : 0;
  int r[PLAN_VALUES], tot[PLAN_VALUES];
  block_exclusive_scan_n<PLAN_VALUES>(v, r, tot);
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    if (lane < PLAN_VALUES) {
      int mine = 0;
#pragma unroll
      for (int q = 0; q < PLAN_VALUES; ++q) mine = lane == q ? tot[q] : mine;
      lb_publish(status + (size_t)lane * ntiles + tile, 1u, tile == 0, mine);
    }
    int pre[PLAN_VALUES];
    bool ok = true;
#pragma unroll
    for (int q = 0; q < PLAN_VALUES; ++q) pre[q] = tile == 0 ? 0 : lb_lookback(status + (size_t)q * ntiles, 1, 0, tile, 1u, &ok);
    if (lane < PLAN_VALUES) {
      int p = 0, t = 0;
#pragma unroll
      for (int q = 0; q < PLAN_VALUES; ++q) {
        p = lane == q ? pre[q] : p;
        t = lane == q ? tot[q] : t;
      }
      if (tile > 0) lb_publish(status + (size_t)lane * ntiles + tile, 1u, true, p + t);
      s_prefix[lane] = p;
      if (tile == ntiles - 1) {
        if (lane == 0) counters[C_CHUNKS] = p + t;
        else cls_count[lane - 1] = p + t;
      }
    }
    if (!ok && lane == 0) counters[C_OVERFLOW] = 1;
  }
  __syncthreads();
  if (w < W1) {
    chunk_off[w] = s_prefix[0] + r[0];
#pragma unroll
    for (int cc = 0; cc < ITEM_CLASSES; ++cc) cls_woff[(size_t)w * ITEM_CLASSES + cc] = s_prefix[1 + cc] + r[1 + cc];
  }
}

// Per target: its tree's subtree partials and their near-field partials.  A target's items are, subtree after
// subtree in pre-order, the subtree partial (when the subtree's wavefront added anything) and the near-field
// partials of that wavefront's list, in list order.  A workgroup serves one group of 64 targets of one tree
// (groups numbered tree after tree from goff) with eight wavefronts: 64 subtrees at a time, the batch's items are
// cut into eight equal ranges, wavefront p adds the p-th range in order (four items' loads in flight together),
// and the eight sums are added in that order, a fixed association.  The flag kept is the first in item order.
static constexpr int REDUCE_PARTS = 8;
static constexpr int REDUCE_UNROLL = 4;
template <bool WIDE>
__global__ __launch_bounds__(64 * REDUCE_PARTS) void k_tree_reduce(
    ForestTables ft, const int32_t* __restrict__ goff, const double* __restrict__ part,
    const int8_t* __restrict__ pflag, const int32_t* __restrict__ pfar, const int32_t* __restrict__ pnear,
    const int8_t* __restrict__ pnz, const int32_t* __restrict__ chunk_off, const double* __restrict__ cpart,
    const int8_t* __restrict__ cflag, const int32_t* __restrict__ cnear, double* __restrict__ out,
    int8_t* __restrict__ flags, int64_t* __restrict__ far_counts, int64_t* __restrict__ near_counts) {
  typedef typename std::conditional<WIDE, double, float>::type acc_t;
  constexpr int PARTS = REDUCE_PARTS;
  __shared__ acc_t s_u[PARTS][64][3];
  __shared__ long long s_key[PARTS][64], s_far[PARTS][64], s_near[PARTS][64];
  __shared__ int s_flag[PARTS][64];
  const int lane = threadIdx.x & 63, pid = (int)(threadIdx.x >> 6);
  const int k = find_tree(goff, ft.K, (int)blockIdx.x);
  const int g = (int)blockIdx.x - goff[k];
  const int t0 = ft.toff[k], nt = ft.toff[k + 1] - t0;
  const int tl = g * 64 + lane;  // the target's index in its tree
  const bool valid = tl < nt;
  const int ntasks = ft.counters[C_OVERFLOW] ? 0 : ft.task_first[k + 1] - ft.task_first[k];
  const int groups = (nt + 63) >> 6;
  const int w0 = ft.wave_off[k] + g;
  const size_t base = (size_t)ft.part_off[k] + tl;
  acc_t ux = 0, uy = 0, uz = 0;
  int flag = 0;
  long long fkey = LLONG_MAX;  // (task << 32) | (0 for the subtree partial, 2 i + 1 for near-field partial i)
  long long far = 0, near = 0;
  for (int qb = 0; qb < ntasks; qb += 64) {
    // lane l describes subtree qb + l: its first chunk, its items, and whether its partial is there
    const int qn = min(64, ntasks - qb);
    int c0 = 0, cnt = 0, nz = 0;
    if (lane < qn) {
      const int w = w0 + (qb + lane) * groups;
      nz = pnz[w] ? 1 : 0;
      c0 = chunk_off ? chunk_off[w] : 0;
      cnt = (chunk_off ? chunk_off[w + 1] - c0 : 0) + nz;
    }
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);
    const int i0 = (int)((long long)total * pid / PARTS), i1 = (int)((long long)total * (pid + 1) / PARTS);
    for (int i = i0; i < i1; i += REDUCE_UNROLL) {
      double vx[REDUCE_UNROLL], vy[REDUCE_UNROLL], vz[REDUCE_UNROLL];
      int vf[REDUCE_UNROLL], vfar[REDUCE_UNROLL], vnear[REDUCE_UNROLL];
      long long vkey[REDUCE_UNROLL];
#pragma unroll
      for (int u = 0; u < REDUCE_UNROLL; ++u) {
        vx[u] = vy[u] = vz[u] = 0.0;
        vf[u] = vfar[u] = vnear[u] = 0;
        vkey[u] = LLONG_MAX;
        const int ii = i + u;
        if (ii >= i1) continue;
        // the subtree holding item ii: the lanes whose items end at or before it
        const int sl = __popcll(__ballot(incl <= ii));
        const int first = __builtin_amdgcn_readlane(incl - cnt, sl);
        const int fz = __builtin_amdgcn_readlane(nz, sl), fc = __builtin_amdgcn_readlane(c0, sl);
        const int r = ii - first;
        const long long q = qb + sl;
        if (!valid) continue;
        if (fz && r == 0) {
          const size_t o = base + (size_t)q * nt;
          vx[u] = part[3 * o];
          vy[u] = part[3 * o + 1];
          vz[u] = part[3 * o + 2];
          vf[u] = pflag[o];
          vkey[u] = q << 32;
          if (pfar) {
            vfar[u] = pfar[o];
            vnear[u] = pnear[o];
          }
        } else {
          const int ci = r - fz;
          const size_t o = (size_t)(fc + ci) * 64 + (tl & 63);
          vx[u] = cpart[3 * o];
          vy[u] = cpart[3 * o + 1];
          vz[u] = cpart[3 * o + 2];
          vf[u] = cflag[o];
          vkey[u] = (q << 32) | (long long)(2 * ci + 1);
          if (cnear) vnear[u] = cnear[o];
        }
      }
#pragma unroll
      for (int u = 0; u < REDUCE_UNROLL; ++u) {
        if (i + u >= i1) break;
        ux += (acc_t)vx[u];
        uy += (acc_t)vy[u];
        uz += (acc_t)vz[u];
        if (flag == 0 && vf[u] != 0) {
          flag = vf[u];
          fkey = vkey[u];
        }
        far += vfar[u];
        near += vnear[u];
      }
    }
  }
  s_u[pid][lane][0] = ux;
  s_u[pid][lane][1] = uy;
  s_u[pid][lane][2] = uz;
  s_key[pid][lane] = fkey;
  s_flag[pid][lane] = flag;
  s_far[pid][lane] = far;
  s_near[pid][lane] = near;
  __syncthreads();
  if (pid != 0 || !valid) return;
  acc_t vx = s_u[0][lane][0], vy = s_u[0][lane][1], vz = s_u[0][lane][2];
  int fl = s_flag[0][lane];
  long long best = fl != 0 ? s_key[0][lane] : LLONG_MAX;
  for (int q = 1; q < PARTS; ++q) {
    vx += s_u[q][lane][0];
    vy += s_u[q][lane][1];
    vz += s_u[q][lane][2];
    far += s_far[q][lane];
    near += s_near[q][lane];
    if (s_flag[q][lane] != 0 && s_key[q][lane] < best) {
      best = s_key[q][lane];
      fl = s_flag[q][lane];
    }
  }
  if (!(isfinite((double)vx) && isfinite((double)vy) && isfinite((double)vz)) && fl == 0) fl = OFW_FLAG_NONFINITE;
  const size_t j = (size_t)t0 + tl;
  out[3 * j] = (double)vx;
  out[3 * j + 1] = (double)vy;
  out[3 * j + 2] = (double)vz;
  flags[j] = (int8_t)fl;
  if (far_counts) far_counts[j] = far;
  if (near_counts) near_counts[j] = near;
}

// ------------------------------------------------------------------ exclusions in tree positions
__global__ void k_inverse(int m, const int32_t* __restrict__ order, int32_t* __restrict__ inverse) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < m) inverse[order[s]] = s;
}

// _positions (vortex_tree.py:571-580): per target, the sorted tree positions of its excluded segments
__global__ void k_ex_positions(int n, int m, const int32_t* __restrict__ ex_ptr, const int32_t* __restrict__ ex_seg,
                               const int32_t* __restrict__ inverse, int32_t* __restrict__ pos,
                               const int32_t* __restrict__ counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || counters[C_OVERFLOW]) return;  // an exclusion count beyond the call's capacity converts nothing
  const int a = ex_ptr[p], b = ex_ptr[p + 1];
  for (int s = a; s < b; ++s) {
    const int seg = ex_seg[s];
    pos[s] = (seg >= 0 && seg < m) ? inverse[seg] : INT_MAX;  // an id out of range excludes nothing
  }
  for (int s = a + 1; s < b; ++s) {
    const int v = pos[s];
    int q = s - 1;
    while (q >= a && pos[q] > v) {
      pos[q + 1] = pos[q];
      --q;
    }
    pos[q + 1] = v;
  }
}

// ------------------------------------------------------------------ traversal plan on the device
// Boxes for the frontier, B workgroups per box, each over a B-th of its range (partial boxes at box_part[(j B + b)
// 6 ...]): slot j < K 2^D is the j-th node at depth D of tree j >> D (the bits of j, first bit first, choose left or
// right from the root; a slot whose path meets a leaf has no node), boxed over its segments' midpoints, its slot
// recorded at its pre-order index; slot K 2^D + k is tree k's targets.
__global__ __launch_bounds__(256) void k_task_boxes(TreeArrays t, int D, int B, const double* __restrict__ pts,
                                                    const int32_t* __restrict__ toff, int32_t* __restrict__ node_slot,
                                                    double* __restrict__ box_part) {
  __shared__ double s_v[6][4];
  const int slots = t.K << D;
  const int j = blockIdx.x / B, b = blockIdx.x - j * B;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (t.counters[C_OVERFLOW]) return;
  if (j < slots) {
    const int k = j >> D;
    int p = t.tree_first[k];
    if (p >= t.tree_first[k + 1]) return;  // an empty tree
    for (int d = D - 1; d >= 0 && p >= 0; --d) {
      if (t.p_skip[p] == p + 1) p = -1;     // a leaf above depth D
      else p = ((j >> d) & 1) ? t.p_skip[p + 1] : p + 1;
    }
    if (p < 0) return;
    if (b == 0 && threadIdx.x == 0) node_slot[p] = j;
    const int64_t lo = t.p_lo[p], len = t.p_hi[p] - lo;
    const int q0 = (int)(lo + len * b / B), q1 = (int)(lo + len * (b + 1) / B);
    for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
      const int s = t.order[q];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double x = t.mid[c][s];
        v[c] = fmin(v[c], x);
        v[3 + c] = fmax(v[3 + c], x);
      }
    }
  } else {
    const int k = j - slots;
    const int64_t lo = toff[k], len = toff[k + 1] - lo;
    const int i0 = (int)(lo + len * b / B), i1 = (int)(lo + len * (b + 1) / B);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double x = pts[3 * (size_t)i + c];
        v[c] = fmin(v[c], x);
        v[3 + c] = fmax(v[3 + c], x);
      }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] = fmin(v[c], __shfl_xor(v[c], off, 64));
      v[3 + c] = fmax(v[3 + c], __shfl_xor(v[3 + c], off, 64));
    }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 6; ++c) s_v[c][wv] = v[c];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    double r = s_v[c][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = c < 3 ? fmin(r, s_v[c][w]) : fmax(r, s_v[c][w]);
    box_part[6 * ((size_t)j * B + b) + c] = r;
  }
}

// the box of slot j from its B partial boxes
__device__ __forceinline__ void slot_box(const double* __restrict__ box_part, int j, int B, double* box) {
  for (int c = 0; c < 6; ++c) box[c] = c < 3 ? INFINITY : -INFINITY;
  for (int b = 0; b < B; ++b) {
    const double* q = box_part + 6 * ((size_t)j * B + b);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      box[c] = fmin(box[c], q[c]);
      box[3 + c] = fmax(box[3 + c], q[3 + c]);
    }
  }
}

// A frontier subtree at depth D is split deeper when the box of its midpoints meets the box of its tree's targets:
// the walks of the groups among them open most of its nodes.  (A choice of the schedule only: every target's
// accepted nodes and near pairs are the reference's whatever the frontier.)
__device__ __forceinline__ bool near_targets(const double* __restrict__ b, const double* __restrict__ tb) {
  bool meet = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) meet = meet && tb[c] <= b[3 + c] && b[c] <= tb[3 + c];
  return meet;
}

// Frontier nodes (pre-order flags; zero past the node count): depth D, or a leaf shallower than D; in a subtree of
// depth D near its tree's targets, depth D + extra instead (or a leaf shallower than that).  The single-pass scan
// of the flags that follows: its ticket and status words are cleared here.
__global__ void k_task_flags(TreeArrays t, int D, int extra, int B, const int32_t* __restrict__ node_slot,
                             const double* __restrict__ box_part, int32_t* __restrict__ flag,
                             int32_t* __restrict__ scan_ticket, unsigned long long* __restrict__ scan_status,
                             int scan_words) {
  const int count = live_count(t.counters, C_FINAL);
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  if (gt == 0) *scan_ticket = 0;
  for (int i = gt; i < scan_words; i += gridDim.x * blockDim.x) scan_status[i] = 0ull;
  for (int p = gt; p <= t.cap; p += gridDim.x * blockDim.x) {
    int f = 0;
    if (p < count) {
      const int d = t.p_depth[p];
      const bool leaf = t.p_skip[p] == p + 1;
      if (d < D) {
        f = leaf ? 1 : 0;
      } else if (d == D && (leaf || extra == 0)) {
        f = 1;
      } else if (d <= D + extra) {
        int a = p;
        for (int q = d; q > D; --q) a = t.p_parent[a];  // the ancestor at depth D
        double nb[6], tb[6];
        slot_box(box_part, node_slot[a], B, nb);
        slot_box(box_part, (t.K << D) + t.p_tree[p], B, tb);
        const bool split = near_targets(nb, tb);
        f = split ? ((d == D + extra || leaf) ? 1 : 0) : (d == D ? 1 : 0);
      }
    }
    flag[p] = f;
  }
}

__global__ void k_task_fill(TreeArrays t, const int32_t* __restrict__ flag, const int32_t* __restrict__ index,
                            int task_cap, EvalTask* __restrict__ tasks) {
  const int count = t.counters[C_OVERFLOW] ? 0 : t.counters[C_FINAL];
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += gridDim.x * blockDim.x) {
    if (!flag[p]) continue;
    const int idx = index[p];
    if (idx >= task_cap) {
      t.counters[C_OVERFLOW] = 1;
      continue;
    }
    EvalTask task;
    const int d = t.p_depth[p];
    task.node = p;
    task.nanc = d;
    int a = t.p_parent[p];
    for (int q = d - 1; q >= 0; --q) {
      task.anc[q] = a;
      a = t.p_parent[a];
    }
    for (int q = d; q < MAX_ANC; ++q) task.anc[q] = -1;
    tasks[idx] = task;
  }
}

// One workgroup: per tree, its tasks (from the flag scan at the tree's first node), work items and partials.
__global__ __launch_bounds__(1024) void k_work_plan(TreeArrays t, const int32_t* __restrict__ index,
                                                    const int32_t* __restrict__ toff, int32_t* __restrict__ task_first,
                                                    int32_t* __restrict__ wave_off, int32_t* __restrict__ part_off,
                                                    int wave_cap, int part_cap, int32_t* __restrict__ plan_ticket,
                                                    unsigned long long* __restrict__ plan_status, int plan_words,
                                                    EvalTask* __restrict__ tasks, int min_leaf, int64_t entry_cap) {
  // the list plan's tile ticket and look-back status words start empty in every call
  if (plan_ticket) {
    if (threadIdx.x == 0) *plan_ticket = 0;
    for (int i = threadIdx.x; i < plan_words; i += blockDim.x) plan_status[i] = 0ull;
  }
  int waves = 0, parts = 0;
  for (int base = 0; base <= t.K; base += blockDim.x) {
    const int k = base + threadIdx.x;
    int w = 0, pp = 0;
    if (k < t.K) {
      const int first = index[t.tree_first[k]];
      const int ntasks = index[t.tree_first[k + 1]] - first;
      const int nt = toff[k + 1] - toff[k];
      task_first[k] = first;
      w = ntasks * ((nt + 63) >> 6);
      pp = ntasks * nt;
    } else if (k == t.K) {
      task_first[k] = index[t.tree_first[k]];
    }
    const int sv[2] = {w, pp};
    int sr[2], stot[2];
    block_exclusive_scan_n<2>(sv, sr, stot);
    const int rw = sr[0], rp = sr[1], tw = stot[0], tp = stot[1];
    if (k <= t.K) {
      wave_off[k] = waves + rw;
      part_off[k] = parts + rp;
    }
    waves += tw;
    parts += tp;
  }
  // each task's interaction lists: groups x (3 entries per leaf its subtree can hold, plus 6), in task order
  const int ntasks = index[t.tree_first[t.K]];
  int64_t entries = 0;
  for (int base = 0; base < ntasks; base += blockDim.x) {
    const int i = base + threadIdx.x;
    int cap = 0, sz = 0;
    if (i < ntasks) {
      const int p = tasks[i].node;
      const int k = t.p_tree[p];
      const int groups = (toff[k + 1] - toff[k] + 63) >> 6;
      cap = 3 * ((t.p_hi[p] - t.p_lo[p]) / min_leaf + 2);
      sz = groups * cap;
    }
    int excl, tot;
    excl = block_exclusive_scan(sz, &tot);
    if (i < ntasks) {
      const int64_t off = entries + excl;
      tasks[i].list_off = off <= (int64_t)INT_MAX ? (int32_t)off : 0;
      tasks[i].list_cap = cap;
    }
    entries += tot;
  }
  if (threadIdx.x == 0) {
    if (waves > wave_cap || parts > part_cap || entries > entry_cap) {
      t.counters[C_OVERFLOW] = 1;
      waves = 0;
    }
    t.counters[C_WORK] = waves;
    t.counters[C_TASKS] = ntasks;
  }
}

// ------------------------------------------------------------------ host driver
template <int MODEL>
static void launch_eval(int precision, dim3 grid, hipStream_t st, const ForestTables& ft, const double* pts,
                        const int32_t* ex_ptr, const int32_t* ex_pos, const EvalTask* tasks, const TreeDevice& td,
                        double theta, int order, int raise, double* part, int8_t* pflag, int32_t* pfar,
                        int32_t* pnear, int8_t* pnz, NearEntry* list, int32_t* list_count, int32_t* cls_wcnt) {
  if (precision == OFW_FP64)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP64>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear, pnz,
                       list, list_count, cls_wcnt);
  else if (precision == OFW_MIXED)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_MIXED>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, tasks,
                       td.acc, td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear,
                       pnz, list, list_count, cls_wcnt);
  else
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP32>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear, pnz,
                       list, list_count, cls_wcnt);
}

template <int MODEL>
static void launch_near(int precision, dim3 grid, hipStream_t st, const ForestTables& ft, const double* pts,
                        const int32_t* ex_ptr, const int32_t* ex_pos, const TreeDevice& td, int order, int raise,
                        const NearEntry* list, const int32_t* list_count, const int32_t* chunk_off,
                        const int32_t* item_wave, int slots, const int32_t* cls_count, const int32_t* cls_list,
                        double* cpart, int8_t* cflag, int32_t* cnear) {
  const NearSegs segs{td.rec64, {td.start[0], td.start[1], td.start[2]}, {td.lenv[0], td.lenv[1], td.lenv[2]}, td.src};
  if (precision == OFW_FP64)
    hipLaunchKernelGGL((k_tree_near<MODEL, OFW_FP64>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, segs,
                       td.rec32, td.acc, td.far64, td.far32, order, raise, list, list_count, chunk_off, item_wave, slots, cls_count,
                       cls_list, cpart, cflag, cnear);
  else if (precision == OFW_MIXED)
    hipLaunchKernelGGL((k_tree_near<MODEL, OFW_MIXED>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, segs,
                       td.rec32, td.acc, td.far64, td.far32, order, raise, list, list_count, chunk_off, item_wave, slots, cls_count,
                       cls_list, cpart, cflag, cnear);
  else
    hipLaunchKernelGGL((k_tree_near<MODEL, OFW_FP32>), grid, dim3(64 * NEAR_WAVES), 0, st, ft, pts, ex_ptr, ex_pos, segs,
                       td.rec32, td.acc, td.far64, td.far32, order, raise, list, list_count, chunk_off, item_wave, slots, cls_count,
                       cls_list, cpart, cflag, cnear);
}

int choose_split_depth(const Ctx& c, int64_t groups, int K) {
  (void)K;
  const int64_t want = (int64_t)c.cu_count * 4 * 8;
  int depth = 0;
  while ((groups << depth) < want && depth < 8) ++depth;
  return std::min(depth, MAX_ANC);
}

int tree_evaluate(Ctx& c, int buf, const TreeArrays& t, const TreeDevice& td, const EvalShape& shape, int precision,
                  const double* points, int model, int raise_singular, const int32_t* ex_ptr, const int32_t* ex_seg,
                  int64_t n_ex, double theta, int expansion_order, double* out, int8_t* flags, int64_t* far_counts,
                  int64_t* near_counts) {
  const int K = shape.K;
  const int64_t n = shape.n, m = t.m;
  const int D = std::max(0, std::min(shape.split_depth, MAX_ANC));
  // deeper subtrees near the targets only where the frontier depth was raised to fill the GPU with walks: a call of
  // many target groups (D = 0) has walks enough
  const int64_t slots64 = ((int64_t)K << D) + K;
  const int extra = D > 0 && slots64 <= (1 << 20) ? std::max(0, std::min(EXTRA_DEPTH, MAX_ANC - D)) : 0;
  // the frontier boxes (K 2^D subtrees, K target sets): B partial boxes each, about two workgroups per CU in all
  const int box_slots = extra > 0 ? (int)slots64 : 0;
  const int B = extra > 0 ? (int)std::min<int64_t>(16, (2 * (int64_t)c.cu_count + slots64 - 1) / slots64) : 1;
  const bool lists = true;
  // host bounds: a tree's frontier is an antichain of at most min(2^(D + extra), m_k) subtrees; a task's lists
  // hold 3 entries per leaf its subtree can have (per visited node one entry, accepted or opened; a leaf two)
  // plus 6, and a subtree of S segments has at most S / floor((leaf + 1) / 2) + 1 leaves
  const int64_t min_leaf = std::max(1, (t.leaf + 1) / 2);
  int64_t task_cap = 0, wave_cap = 0, part_cap = 0, entry_cap = 0;
  for (int k = 0; k < K; ++k) {
    const int64_t mk = shape.seg_off[k + 1] - shape.seg_off[k];
    const int64_t nk = shape.target_off[k + 1] - shape.target_off[k];
    const int64_t tasks_k = mk > 0 ? std::min<int64_t>((int64_t)1 << (D + extra), mk) : 0;
    const int64_t groups_k = (nk + 63) / 64;
    task_cap += tasks_k;
    wave_cap += tasks_k * groups_k;
    part_cap += tasks_k * nk;
    entry_cap += lists ? groups_k * 3 * (mk / min_leaf + 3 * tasks_k) : 0;
  }
  if (part_cap > INT_MAX || wave_cap > INT_MAX / 4) {
    set_error("tree traversal: split depth too large for the target count");
    return OFW_ERR_ARG;
  }
  if (entry_cap > INT_MAX) {
    set_error("tree traversal: too many targets and segments for one call (near-field list bound)");
    return OFW_ERR_ARG;
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t P = (size_t)std::max<int64_t>(part_cap, 1);
  const bool counts = far_counts || near_counts;
  const size_t W1 = (size_t)wave_cap + 1;
  const size_t E = (size_t)std::max<int64_t>(entry_cap, 1);
  size_t need = al((size_t)std::max<int64_t>(m, 1) * 4) + al((size_t)std::max<int64_t>(n_ex, 1) * 4);
  const int task_words = scan_lb_words((int64_t)t.cap + 1);
  need += 2 * al(((size_t)t.cap + 1) * 4) + al((size_t)std::max<int64_t>(task_cap, 1) * sizeof(EvalTask));
  need += al(4) + al((size_t)task_words * 8);
  need += al((size_t)(K + 1) * 8) + 3 * al((size_t)(K + 1) * 4) + al(P * 24) + al(P) + (counts ? 2 * al(P * 4) : 0);
  need += al(((size_t)t.cap + 1) * 4) + al((size_t)std::max(box_slots, 1) * B * 6 * 8);  // frontier boxes
  need += al(W1);
  const int plan_words = PLAN_VALUES * (int)((W1 + 255) / 256);
  if (lists) need += al(E * sizeof(NearEntry)) + 2 * al(W1 * 4) + al(ITEM_CLASSES * 4) + 2 * al(W1 * ITEM_CLASSES * 4) +
                     al(4) + al((size_t)plan_words * 8);
  OFW_HIP(c.tree[buf].reserve(need));
  char* base = (char*)c.tree[buf].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  int32_t* inverse = (int32_t*)carve((size_t)std::max<int64_t>(m, 1) * 4);
  int32_t* ex_pos = (int32_t*)carve((size_t)std::max<int64_t>(n_ex, 1) * 4);
  int32_t* flag = (int32_t*)carve(((size_t)t.cap + 1) * 4);
  int32_t* index = (int32_t*)carve(((size_t)t.cap + 1) * 4);
  int32_t* task_ticket = (int32_t*)carve(4);
  unsigned long long* task_status = (unsigned long long*)carve((size_t)task_words * 8);
  EvalTask* tasks = (EvalTask*)carve((size_t)std::max<int64_t>(task_cap, 1) * sizeof(EvalTask));
  // the host tables side by side, one upload: first target and first group of 64 targets of each tree
  int32_t* toff = (int32_t*)carve((size_t)(K + 1) * 2 * 4);
  int32_t* goff = toff + (K + 1);
  int32_t* node_slot = (int32_t*)carve(((size_t)t.cap + 1) * 4);  // written at the depth-D nodes only
  double* box_part = (double*)carve((size_t)std::max(box_slots, 1) * B * 6 * 8);
  int32_t* task_first = (int32_t*)carve((size_t)(K + 1) * 4);
  int32_t* wave_off = (int32_t*)carve((size_t)(K + 1) * 4);
  int32_t* part_off = (int32_t*)carve((size_t)(K + 1) * 4);
  double* part = (double*)carve(P * 24);
  int8_t* pflag = (int8_t*)carve(P);
  int32_t* pfar = counts ? (int32_t*)carve(P * 4) : nullptr;
  int32_t* pnear = counts ? (int32_t*)carve(P * 4) : nullptr;
  int8_t* pnz = (int8_t*)carve(W1);
  NearEntry* list = nullptr;
  int32_t *list_count = nullptr, *chunk_off = nullptr, *cls_count = nullptr, *cls_wcnt = nullptr, *cls_woff = nullptr;
  int32_t* plan_ticket = nullptr;
  unsigned long long* plan_status = nullptr;
  double* cpart = nullptr;
  int8_t* cflag = nullptr;
  int32_t* cnear = nullptr;
  if (lists) {
    list = (NearEntry*)carve(E * sizeof(NearEntry));
    list_count = (int32_t*)carve(W1 * 4);
    chunk_off = (int32_t*)carve(W1 * 4);
    cls_count = (int32_t*)carve(ITEM_CLASSES * 4);
    cls_wcnt = (int32_t*)carve(W1 * ITEM_CLASSES * 4);
    cls_woff = (int32_t*)carve(W1 * ITEM_CLASSES * 4);
    plan_ticket = (int32_t*)carve(4);
    plan_status = (unsigned long long*)carve((size_t)plan_words * 8);
  }

  std::vector<int64_t> tables(shape.target_off);
  int64_t groups_total = 0;
  for (int k = 0; k <= K; ++k) {
    tables.push_back(groups_total);
    if (k < K) groups_total += (shape.target_off[k + 1] - shape.target_off[k] + 63) / 64;
  }
  int st = upload_i32(c, tables, toff);
  if (st) return st;
  const int grid = c.cu_count * 2;
  if (n_ex > 0 && ex_ptr) {
    hipLaunchKernelGGL(k_inverse, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, t.order,
                       inverse);
    hipLaunchKernelGGL(k_ex_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, (int)m,
                       ex_ptr, ex_seg, inverse, ex_pos, t.counters);
  }
  const int32_t* ex_ptr_used = (n_ex > 0) ? ex_ptr : nullptr;
  if (extra > 0)
    hipLaunchKernelGGL(k_task_boxes, dim3((unsigned)(box_slots * B)), dim3(256), 0, c.stream, t, D, B, points, toff,
                       node_slot, box_part);
  hipLaunchKernelGGL(k_task_flags, dim3(grid), dim3(256), 0, c.stream, t, D, extra, B, node_slot, box_part, flag,
                     task_ticket, task_status, task_words);
  scan_exclusive_lb(c, t.cap + 1, flag, index, task_ticket, task_status, nullptr, t.counters);
  hipLaunchKernelGGL(k_task_fill, dim3(grid), dim3(256), 0, c.stream, t, flag, index, (int)task_cap, tasks);
  hipLaunchKernelGGL(k_work_plan, dim3(1), dim3(1024), 0, c.stream, t, index, toff, task_first, wave_off, part_off,
                     (int)wave_cap, (int)part_cap, plan_ticket, plan_status, plan_words, tasks, (int)min_leaf,
                     (int64_t)entry_cap);
  OFW_HIP(hipGetLastError());
  if (c.plan_forked) {
    // the plan ran on the second stream beside the node quantities: the call's stream waits for it here
    OFW_HIP(hipEventRecord(c.join_ev, c.stream));
    c.stream = c.plan_main;
    c.plan_forked = false;
    OFW_HIP(hipStreamWaitEvent(c.stream, c.join_ev, 0));
  }
  OFW_HIP(c.mark());
  ForestTables ft{K, td.count, (int)m, toff, task_first, wave_off, part_off, tasks, t.counters};
  if (wave_cap > 0) {
    dim3 egrid((unsigned)((wave_cap + NEAR_WAVES - 1) / NEAR_WAVES));
    switch (model) {
      case 0: launch_eval<0>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear, pnz, list, list_count, cls_wcnt); break;
      case 1: launch_eval<1>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear, pnz, list, list_count, cls_wcnt); break;
      default: launch_eval<2>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear, pnz, list, list_count, cls_wcnt); break;
    }
  }
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  if (lists && wave_cap > 0) {
    // near-field work items: CHUNK list entries each, numbered in walk-wavefront order
    hipLaunchKernelGGL(k_chunk_plan, dim3((unsigned)((W1 + 255) / 256)), dim3(256), 0, c.stream, ft, (int)W1, list_count,
                       cls_wcnt, chunk_off, cls_woff, t.counters, cls_count, plan_ticket, plan_status);
    // partial sums of the near-field work items (scratch buffer 5 of the context)
    int64_t slots = shape.chunk_slots;
    if (slots < 0) {
      int32_t count = 0;
      OFW_HIP(hipMemcpyAsync(&count, t.counters + C_CHUNKS, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
      OFW_HIP(hipStreamSynchronize(c.stream));
      slots = count;
      c.stats_i[6] = count;
    }
    const size_t CH = (size_t)std::max<int64_t>(slots, 1);
    OFW_HIP(c.tree[5].reserve(al(CH * 64 * 24) + al(CH * 64) + al(CH * 4) + al(CH * ITEM_CLASSES * 4) +
                              (counts ? al(CH * 64 * 4) : 0)));
    int32_t *item_wave, *cls_list;
    {
      char* cb = (char*)c.tree[5].ptr;
      cpart = (double*)cb;
      cb += al(CH * 64 * 24);
      cflag = (int8_t*)cb;
      cb += al(CH * 64);
      item_wave = (int32_t*)cb;
      cb += al(CH * 4);
      cls_list = (int32_t*)cb;
      cb += al(CH * ITEM_CLASSES * 4);
      cnear = counts ? (int32_t*)cb : nullptr;
    }
    hipLaunchKernelGGL(k_item_waves, dim3((unsigned)((W1 + 3) / 4)), dim3(256), 0, c.stream, ft, chunk_off, (int)CH,
                       (int)std::min<int64_t>(slots, INT_MAX), list, list_count, cls_woff, item_wave, cls_list);
    OFW_HIP(c.mark());
    // one wavefront per work item: the dispatcher starts the next item wherever a wavefront has finished
    const dim3 ngrid((unsigned)std::max<int64_t>(1, ((int64_t)CH + NEAR_WAVES - 1) / NEAR_WAVES));
    switch (model) {
      case 0: launch_near<0>(precision, ngrid, c.stream, ft, points, ex_ptr_used, ex_pos, td, expansion_order, raise_singular, list, list_count, chunk_off, item_wave, (int)CH, cls_count, cls_list, cpart, cflag, cnear); break;
      case 1: launch_near<1>(precision, ngrid, c.stream, ft, points, ex_ptr_used, ex_pos, td, expansion_order, raise_singular, list, list_count, chunk_off, item_wave, (int)CH, cls_count, cls_list, cpart, cflag, cnear); break;
      default: launch_near<2>(precision, ngrid, c.stream, ft, points, ex_ptr_used, ex_pos, td, expansion_order, raise_singular, list, list_count, chunk_off, item_wave, (int)CH, cls_count, cls_list, cpart, cflag, cnear); break;
    }
    OFW_HIP(hipGetLastError());
  }
  OFW_HIP(c.mark());
  if (groups_total > 0) {
    if (precision == OFW_FP32)
      hipLaunchKernelGGL(k_tree_reduce<false>, dim3((unsigned)groups_total), dim3(64 * REDUCE_PARTS), 0, c.stream, ft,
                         goff, part, pflag, pfar, pnear, pnz, (const int32_t*)chunk_off, cpart, cflag, cnear, out,
                         flags, far_counts, near_counts);
    else
      hipLaunchKernelGGL(k_tree_reduce<true>, dim3((unsigned)groups_total), dim3(64 * REDUCE_PARTS), 0, c.stream, ft,
                         goff, part, pflag, pfar, pnear, pnz, (const int32_t*)chunk_off, cpart, cflag, cnear, out,
                         flags, far_counts, near_counts);
  }
  OFW_HIP(hipGetLastError());
  c.stats_i[3] = D;
  return OFW_OK;
}

}  // namespace ofw
