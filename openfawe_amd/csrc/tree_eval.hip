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

template <int MODEL, int PREC>
__global__ __launch_bounds__(256) void k_tree_eval(int n, const double* __restrict__ pts,
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
  const int group = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int task = blockIdx.y;
  const int j = group * 64 + lane;
  if (group * 64 >= n) return;  // whole wave beyond the targets
  const bool valid = j < n;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  int cur = 0, cend = 0;
  if (valid) {
    x0 = pts[3 * (size_t)j];
    x1 = pts[3 * (size_t)j + 1];
    x2 = pts[3 * (size_t)j + 2];
    cur = ex_ptr[j];
    cend = ex_ptr[j + 1];
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
  for (int q = 0; q < nanc; ++q) {
    const int A = uniform(T.anc[q]);
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
        for (int k = lo; k < hi; ++k) {
          bool use = open;
          if (use && cur < cend && ex_pos[cur] == k) {
            ++cur;
            use = false;
          }
          if (!wave_any(use)) continue;
          if (use) {
            ++near;
            if (PREC == OFW_FP64) {
              double vx = 0.0, vy = 0.0, vz = 0.0;
              const int f = pair_fast64<MODEL>(x0, x1, x2, rec64 + (size_t)k * REC64, raise_singular, vx, vy, vz);
              if (f != OFW_FLAG_OK) {
                if (flag == OFW_FLAG_OK) flag = f;
              } else {
                ux += vx;
                uy += vy;
                uz += vz;
              }
            } else {
              const double* r = rec64 + (size_t)k * REC64;
              const float sx = (float)(x0 - r[R_SX]), sy = (float)(x1 - r[R_SY]), sz = (float)(x2 - r[R_SZ]);
              float vx = 0.0f, vy = 0.0f, vz = 0.0f;
              if (pair_f32<MODEL>(sx, sy, sz, rec32 + (size_t)k * REC32, vx, vy, vz)) {
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
    const size_t o = (size_t)task * n + j;
    part[3 * o] = (double)ux;
    part[3 * o + 1] = (double)uy;
    part[3 * o + 2] = (double)uz;
    pflag[o] = (int8_t)flag;
    pfar[o] = far;
    pnear[o] = near;
  }
}

template <bool WIDE>
__global__ __launch_bounds__(256) void k_tree_reduce(int n, int ntasks, const int32_t* __restrict__ perm,
                                                     const double* __restrict__ part,
                                                     const int8_t* __restrict__ pflag,
                                                     const int32_t* __restrict__ pfar,
                                                     const int32_t* __restrict__ pnear, double* __restrict__ out,
                                                     int8_t* __restrict__ flags, int64_t* __restrict__ far_counts,
                                                     int64_t* __restrict__ near_counts) {
  typedef typename std::conditional<WIDE, double, float>::type acc_t;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  acc_t ux = 0, uy = 0, uz = 0;
  int flag = 0;
  int64_t far = 0, near = 0;
  for (int q = 0; q < ntasks; ++q) {
    const size_t o = (size_t)q * n + j;
    ux += (acc_t)part[3 * o];
    uy += (acc_t)part[3 * o + 1];
    uz += (acc_t)part[3 * o + 2];
    if (flag == 0) flag = pflag[o];
    far += pfar[o];
    near += pnear[o];
  }
  if (!(isfinite((double)ux) && isfinite((double)uy) && isfinite((double)uz)) && flag == 0) flag = OFW_FLAG_NONFINITE;
  const int t = perm ? perm[j] : j;
  out[3 * (size_t)t] = (double)ux;
  out[3 * (size_t)t + 1] = (double)uy;
  out[3 * (size_t)t + 2] = (double)uz;
  flags[t] = (int8_t)flag;
  if (far_counts) far_counts[t] = far;
  if (near_counts) near_counts[t] = near;
}

// ------------------------------------------------------------------ exclusions in tree positions
__global__ void k_inverse(int m, const int32_t* __restrict__ order, int32_t* __restrict__ inverse) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) inverse[order[k]] = k;
}

// _positions (vortex_tree.py:571-580): per target, the sorted tree positions of its excluded segments
__global__ void k_ex_positions(int n, const int32_t* __restrict__ ex_ptr, const int32_t* __restrict__ ex_seg,
                               const int32_t* __restrict__ inverse, int32_t* __restrict__ pos) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int a = ex_ptr[p], b = ex_ptr[p + 1];
  for (int k = a; k < b; ++k) pos[k] = inverse[ex_seg[k]];
  for (int k = a + 1; k < b; ++k) {
    const int v = pos[k];
    int q = k - 1;
    while (q >= a && pos[q] > v) {
      pos[q + 1] = pos[q];
      --q;
    }
    pos[q + 1] = v;
  }
}

// ------------------------------------------------------------------ host driver
void frontier_tasks(const HostTree& t, int depth, std::vector<EvalTask>& tasks) {
  tasks.clear();
  const int count = (int)t.count;
  for (int i = 0; i < count; ++i) {
    const int d = t.pre_depth[i];
    const bool leaf = t.pre_skip[i] == i + 1;
    if (d == depth || (d < depth && leaf)) {
      EvalTask task{};
      task.node = i;
      task.nanc = d;
      int p = t.pre_parent[i];
      for (int q = d - 1; q >= 0; --q) {
        task.anc[q] = p;
        p = t.pre_parent[p];
      }
      tasks.push_back(task);
    }
  }
}

template <int MODEL>
static void launch_eval(int precision, dim3 grid, hipStream_t st, int n, const double* pts, const int32_t* ex_ptr,
                        const int32_t* ex_pos, const EvalTask* tasks, const TreeDevice& td, double theta, int order,
                        int raise, double* part, int8_t* pflag, int32_t* pfar, int32_t* pnear) {
  if (precision == OFW_FP64)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP64>), grid, dim3(256), 0, st, n, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
  else if (precision == OFW_MIXED)
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_MIXED>), grid, dim3(256), 0, st, n, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
  else
    hipLaunchKernelGGL((k_tree_eval<MODEL, OFW_FP32>), grid, dim3(256), 0, st, n, pts, ex_ptr, ex_pos, tasks, td.acc,
                       td.far64, td.far32, td.rec64, td.rec32, theta, order, raise, part, pflag, pfar, pnear);
}

int tree_evaluate(Ctx& c, const HostTree& t, const TreeDevice& td, const int32_t* order, int precision, int64_t n64,
                  const double* points, int model, int raise_singular, const int32_t* ex_ptr,
                  const int32_t* ex_seg, double theta, int expansion_order, int split_depth, double* out,
                  int8_t* flags, int64_t* far_counts, int64_t* near_counts, std::vector<EvalTask>& tasks) {
  const int n = (int)n64, m = (int)t.m;
  const int groups = (n + 63) / 64;
  int depth = split_depth;
  if (depth < 0) {
    depth = 0;
    const int want = c.cu_count * 4 * 8;
    while ((groups << depth) < want && depth < 8) ++depth;
  }
  depth = std::min(std::min(depth, t.depth), MAX_ANC);
  frontier_tasks(t, depth, tasks);
  const int ntasks = (int)tasks.size();
  // scratch
  int64_t e = 0;
  OFW_HIP(hipMemcpyAsync(&e, ex_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  e = (int32_t)e;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t need = al((size_t)m * 4) + al((size_t)(e > 0 ? e : 1) * 4) + al((size_t)ntasks * sizeof(EvalTask)) +
                al((size_t)ntasks * n * 3 * 8) + al((size_t)ntasks * n) + 2 * al((size_t)ntasks * n * 4);
  OFW_HIP(c.tree[2].reserve(need));
  char* base = (char*)c.tree[2].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (void*)p;
  };
  int32_t* inverse = (int32_t*)carve((size_t)m * 4);
  int32_t* ex_pos = (int32_t*)carve((size_t)(e > 0 ? e : 1) * 4);
  EvalTask* d_tasks = (EvalTask*)carve((size_t)ntasks * sizeof(EvalTask));
  double* part = (double*)carve((size_t)ntasks * n * 3 * 8);
  int8_t* pflag = (int8_t*)carve((size_t)ntasks * n);
  int32_t* pfar = (int32_t*)carve((size_t)ntasks * n * 4);
  int32_t* pnear = (int32_t*)carve((size_t)ntasks * n * 4);
  OFW_HIP(hipMemcpyAsync(d_tasks, tasks.data(), (size_t)ntasks * sizeof(EvalTask), hipMemcpyHostToDevice, c.stream));
  if (e > 0) {
    hipLaunchKernelGGL(k_inverse, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, order, inverse);
    hipLaunchKernelGGL(k_ex_positions, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, ex_ptr, ex_seg, inverse,
                       ex_pos);
  }
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  dim3 grid((unsigned)((groups + 3) / 4), (unsigned)ntasks);
  switch (model) {
    case 0: launch_eval<0>(precision, grid, c.stream, n, points, ex_ptr, ex_pos, d_tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
    case 1: launch_eval<1>(precision, grid, c.stream, n, points, ex_ptr, ex_pos, d_tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
    default: launch_eval<2>(precision, grid, c.stream, n, points, ex_ptr, ex_pos, d_tasks, td, theta, expansion_order, raise_singular, part, pflag, pfar, pnear); break;
  }
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  if (precision == OFW_FP32)
    hipLaunchKernelGGL(k_tree_reduce<false>, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, ntasks,
                       (const int32_t*)nullptr, part, pflag, pfar, pnear, out, flags, far_counts, near_counts);
  else
    hipLaunchKernelGGL(k_tree_reduce<true>, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, ntasks,
                       (const int32_t*)nullptr, part, pflag, pfar, pnear, out, flags, far_counts, near_counts);
  OFW_HIP(hipGetLastError());
  c.stats_i[2] = ntasks;
  c.stats_i[3] = depth;
  return OFW_OK;
}

}  // namespace ofw
