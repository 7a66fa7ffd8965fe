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

__device__ __forceinline__ bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Per-tree tables of a forest evaluation (K+1 entries each, int32, device).
struct ForestTables {
  int K;
  const int32_t* toff;        // first target of each tree
  const int32_t* task_first;  // first traversal task of each tree
  const int32_t* wave_off;    // first work item (wavefront) of each tree
  const int32_t* part_off;    // first partial-sum slot of each tree
  const int32_t* counters;
};

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

template <int MODEL, int PREC>
__global__ __launch_bounds__(256) void k_tree_eval(ForestTables ft, const double* __restrict__ pts,
                                                   const int32_t* __restrict__ ex_ptr,
                                                   const int32_t* __restrict__ ex_pos,
                                                   const EvalTask* __restrict__ tasks,
                                                   const double* __restrict__ accr, const double* __restrict__ fr64,
                                                   const float* __restrict__ fr32, const double* __restrict__ rec64,
                                                   const float* __restrict__ rec32, double theta, int order,
                                                   int raise_singular, double* __restrict__ part,
                                                   int8_t* __restrict__ pflag, int32_t* __restrict__ pfar,
                                                   int32_t* __restrict__ pnear) {
  typedef typename std::conditional<PREC == OFW_FP32, float, double>::type acc_t;
  const int wave = uniform(xcd_block(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
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
  acc_t ux = 0, uy = 0, uz = 0;
  int flag = OFW_FLAG_OK, far = 0, near = 0;
  bool alive = valid;
  const EvalTask& T = tasks[task];
  const int F = uniform(T.node);
  const int nanc = uniform(T.nanc);
  const int32_t* fi = reinterpret_cast<const int32_t*>(accr + (size_t)F * ACC_REC + A_INTS);
  const int flo = uniform(fi[0]);

  // ---- ancestors of the frontier node
  for (int a_i = 0; a_i < nanc; ++a_i) {
    const int A = uniform(T.anc[a_i]);
    const double* a = accr + (size_t)A * ACC_REC;
    const int32_t* ai = reinterpret_cast<const int32_t*>(a + A_INTS);
    const int alo = ai[0], ahi = ai[1];
    double rx, ry, rz, dist;
    bool acc = alive && accept_test(x0, x1, x2, a, theta, rx, ry, rz, dist);
    if (acc) {
      while (cur < cend && ex_pos[cur] < alo) ++cur;
      acc = !(cur < cend && ex_pos[cur] < ahi);
    }
    if (acc) {
      alive = false;
      if (alo == flo) {
        if (PREC == OFW_FP64) {
          double vx = 0.0, vy = 0.0, vz = 0.0;
          far64(rx, ry, rz, dist, fr64 + (size_t)A * FAR_REC, order, vx, vy, vz);
          ux += vx;
          uy += vy;
          uz += vz;
        } else {
          float vx, vy, vz;
          far32((float)rx, (float)ry, (float)rz, fr32 + (size_t)A * FAR_REC, order, vx, vy, vz);
          ux += (acc_t)vx;
          uy += (acc_t)vy;
          uz += (acc_t)vz;
        }
        ++far;
      }
    }
  }

  // ---- stackless walk of the frontier subtree
  int i = F;
  const int end = uniform(fi[2]);
  int skip_until = F;
  while (i < end) {
    const bool act = alive && i >= skip_until;
    if (!wave_any(act)) break;
    const double* a = accr + (size_t)i * ACC_REC;
    const int32_t* ai = reinterpret_cast<const int32_t*>(a + A_INTS);
    const int lo = uniform(ai[0]), hi = uniform(ai[1]), skip = uniform(ai[2]);
    double rx = 0.0, ry = 0.0, rz = 0.0, dist = 1.0;
    bool acc = false;
    if (act) {
      acc = accept_test(x0, x1, x2, a, theta, rx, ry, rz, dist);
      while (cur < cend && ex_pos[cur] < lo) ++cur;
      if (acc) acc = !(cur < cend && ex_pos[cur] < hi);
    }
    if (wave_any(acc)) {
      if (acc) {
        if (PREC == OFW_FP64) {
          double vx = 0.0, vy = 0.0, vz = 0.0;
          far64(rx, ry, rz, dist, fr64 + (size_t)i * FAR_REC, order, vx, vy, vz);
          ux += vx;
          uy += vy;
          uz += vz;
        } else {
          float vx, vy, vz;
          far32((float)rx, (float)ry, (float)rz, fr32 + (size_t)i * FAR_REC, order, vx, vy, vz);
          ux += (acc_t)vx;
          uy += (acc_t)vy;
          uz += (acc_t)vz;
        }
        ++far;
      }
    }
    const bool open = act && !acc;
    if (skip == i + 1) {
      // leaf: the near field of the lanes that did not accept it (vortex_tree.py:538-557)
      if (wave_any(open)) {
        for (int s = lo; s < hi; ++s) {
          bool use = open;
          if (use && cur < cend && ex_pos[cur] == s) {
            ++cur;
            use = false;
          }
          if (!wave_any(use)) continue;
          if (use) {
            ++near;
            if (PREC == OFW_FP64) {
              double vx = 0.0, vy = 0.0, vz = 0.0;
              const int f = pair_fast64<MODEL>(x0, x1, x2, rec64 + (size_t)s * REC64, raise_singular, vx, vy, vz);
              if (f != OFW_FLAG_OK) {
                if (flag == OFW_FLAG_OK) flag = f;
              } else {
                ux += vx;
                uy += vy;
                uz += vz;
              }
            } else {
              const double* r = rec64 + (size_t)s * REC64;
              const float sx = (float)(x0 - r[R_SX]), sy = (float)(x1 - r[R_SY]), sz = (float)(x2 - r[R_SZ]);
              float vx = 0.0f, vy = 0.0f, vz = 0.0f;
              if (pair_f32<MODEL>(sx, sy, sz, rec32 + (size_t)s * REC32, vx, vy, vz)) {
                double wx = 0.0, wy = 0.0, wz = 0.0;
                const int f = pair_fast64<MODEL>(x0, x1, x2, r, raise_singular, wx, wy, wz);
                if (f != OFW_FLAG_OK) {
                  if (flag == OFW_FLAG_OK) flag = f;
                } else {
                  ux += (acc_t)wx;
                  uy += (acc_t)wy;
                  uz += (acc_t)wz;
                }
              } else {
                ux += (acc_t)vx;
                uy += (acc_t)vy;
                uz += (acc_t)vz;
              }
            }
          }
        }
      }
      i = i + 1;
    } else {
      if (acc) skip_until = skip;
      i = wave_any(open) ? i + 1 : skip;
    }
  }
  if (valid) {
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

// Per target: its tree's subtree partials added in pre-order of the subtrees.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_tree_reduce(ForestTables ft, int n, const double* __restrict__ part,
                                                     const int8_t* __restrict__ pflag,
                                                     const int32_t* __restrict__ pfar,
                                                     const int32_t* __restrict__ pnear, double* __restrict__ out,
                                                     int8_t* __restrict__ flags, int64_t* __restrict__ far_counts,
                                                     int64_t* __restrict__ near_counts) {
  typedef typename std::conditional<WIDE, double, float>::type acc_t;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = find_tree(ft.toff, ft.K, j);
  const int t0 = ft.toff[k], nt = ft.toff[k + 1] - t0;
  const int ntasks = ft.counters[C_OVERFLOW] ? 0 : ft.task_first[k + 1] - ft.task_first[k];
  const size_t base = (size_t)ft.part_off[k] + (j - t0);
  acc_t ux = 0, uy = 0, uz = 0;
  int flag = 0;
  int64_t far = 0, near = 0;
  for (int q = 0; q < ntasks; ++q) {
    const size_t o = base + (size_t)q * nt;
    ux += (acc_t)part[3 * o];
    uy += (acc_t)part[3 * o + 1];
    uz += (acc_t)part[3 * o + 2];
    if (flag == 0) flag = pflag[o];
    if (pfar) {
      far += pfar[o];
      near += pnear[o];
    }
  }
  if (!(isfinite((double)ux) && isfinite((double)uy) && isfinite((double)uz)) && flag == 0) flag = OFW_FLAG_NONFINITE;
  out[3 * (size_t)j] = (double)ux;
  out[3 * (size_t)j + 1] = (double)uy;
  out[3 * (size_t)j + 2] = (double)uz;
  flags[j] = (int8_t)flag;
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
                               const int32_t* __restrict__ inverse, int32_t* __restrict__ pos) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
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
// frontier nodes: depth == D, or leaves shallower than D (pre-order flags; zero past the node count)
__global__ void k_task_flags(TreeArrays t, int D, int32_t* __restrict__ flag) {
  const int count = live_count(t.counters, C_FINAL);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= t.cap; p += gridDim.x * blockDim.x) {
    int f = 0;
    if (p < count) {
      const int d = t.p_depth[p];
      f = (d == D || (d < D && t.p_skip[p] == p + 1)) ? 1 : 0;
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
                                                    int wave_cap, int part_cap) {
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
    int tw, tp;
    const int rw = block_exclusive_scan(w, &tw);
    const int rp = block_exclusive_scan(pp, &tp);
    if (k <= t.K) {
      wave_off[k] = waves + rw;
      part_off[k] = parts + rp;
    }
    waves += tw;
    parts += tp;
  }
  if (threadIdx.x == 0) {
    if (waves > wave_cap || parts > part_cap) {
      t.counters[C_OVERFLOW] = 1;
      waves = 0;
    }
    t.counters[C_WORK] = waves;
    t.counters[C_TASKS] = index[t.tree_first[t.K]];
  }
}

// ------------------------------------------------------------------ host driver
template <int MODEL>
static void launch_eval(int precision, dim3 grid, hipStream_t st, const ForestTables& ft, const double* pts,
                        const int32_t* ex_ptr, const int32_t* ex_pos, const EvalTask* tasks, const TreeDevice& td,
                        double theta, int order, int raise, double* part, int8_t* pflag, int32_t* pfar,
                        int32_t* pnear) {
  if (precision == OFW_FP64)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP64>), grid, dim3(256), 0, st, ft, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
  else if (precision == OFW_MIXED)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_MIXED>), grid, dim3(256), 0, st, ft, pts, ex_ptr, ex_pos, tasks,
                       td.acc, td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
  else
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP32>), grid, dim3(256), 0, st, ft, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
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
  // host bounds: a tree has at most min(2^D, m_k) frontier subtrees
  int64_t task_cap = 0, wave_cap = 0, part_cap = 0;
  for (int k = 0; k < K; ++k) {
    const int64_t mk = shape.seg_off[k + 1] - shape.seg_off[k];
    const int64_t nk = shape.target_off[k + 1] - shape.target_off[k];
    const int64_t tasks_k = mk > 0 ? std::min<int64_t>((int64_t)1 << D, mk) : 0;
    task_cap += tasks_k;
    wave_cap += tasks_k * ((nk + 63) / 64);
    part_cap += tasks_k * nk;
  }
  if (part_cap > INT_MAX || wave_cap > INT_MAX / 4) {
    set_error("tree traversal: split depth too large for the target count");
    return OFW_ERR_ARG;
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t P = (size_t)std::max<int64_t>(part_cap, 1);
  const bool counts = far_counts || near_counts;
  size_t need = al((size_t)std::max<int64_t>(m, 1) * 4) + al((size_t)std::max<int64_t>(n_ex, 1) * 4);
  need += 2 * al(((size_t)t.cap + 1) * 4) + al((size_t)std::max<int64_t>(task_cap, 1) * sizeof(EvalTask));
  need += 4 * al((size_t)(K + 1) * 4) + al(P * 24) + al(P) + (counts ? 2 * al(P * 4) : 0);
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
  EvalTask* tasks = (EvalTask*)carve((size_t)std::max<int64_t>(task_cap, 1) * sizeof(EvalTask));
  int32_t* toff = (int32_t*)carve((size_t)(K + 1) * 4);
  int32_t* task_first = (int32_t*)carve((size_t)(K + 1) * 4);
  int32_t* wave_off = (int32_t*)carve((size_t)(K + 1) * 4);
  int32_t* part_off = (int32_t*)carve((size_t)(K + 1) * 4);
  double* part = (double*)carve(P * 24);
  int8_t* pflag = (int8_t*)carve(P);
  int32_t* pfar = counts ? (int32_t*)carve(P * 4) : nullptr;
  int32_t* pnear = counts ? (int32_t*)carve(P * 4) : nullptr;

  int st = upload_i32(c, shape.target_off, toff);
  if (st) return st;
  const int grid = c.cu_count * 2;
  if (n_ex > 0 && ex_ptr) {
    hipLaunchKernelGGL(k_inverse, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, t.order,
                       inverse);
    hipLaunchKernelGGL(k_ex_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, (int)m,
                       ex_ptr, ex_seg, inverse, ex_pos);
  }
  const int32_t* ex_ptr_used = (n_ex > 0) ? ex_ptr : nullptr;
  hipLaunchKernelGGL(k_task_flags, dim3(grid), dim3(256), 0, c.stream, t, D, flag);
  scan_exclusive(c, t.cap + 1, flag, index, t.scan_part, nullptr);
  hipLaunchKernelGGL(k_task_fill, dim3(grid), dim3(256), 0, c.stream, t, flag, index, (int)task_cap, tasks);
  hipLaunchKernelGGL(k_work_plan, dim3(1), dim3(1024), 0, c.stream, t, index, toff, task_first, wave_off, part_off,
                     (int)wave_cap, (int)part_cap);
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  ForestTables ft{K, toff, task_first, wave_off, part_off, t.counters};
  if (wave_cap > 0) {
    dim3 egrid((unsigned)((wave_cap + 3) / 4));
    switch (model) {
      case 0: launch_eval<0>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
      case 1: launch_eval<1>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
      default: launch_eval<2>(precision, egrid, c.stream, ft, points, ex_ptr_used, ex_pos, tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
    }
  }
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  if (n > 0) {
    if (precision == OFW_FP32)
      hipLaunchKernelGGL(k_tree_reduce<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, ft, (int)n,
                         part, pflag, pfar, pnear, out, flags, far_counts, near_counts);
    else
      hipLaunchKernelGGL(k_tree_reduce<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, ft, (int)n,
                         part, pflag, pfar, pnear, out, flags, far_counts, near_counts);
  }
  OFW_HIP(hipGetLastError());
  c.stats_i[3] = D;
  return OFW_OK;
}

}  // namespace ofw
