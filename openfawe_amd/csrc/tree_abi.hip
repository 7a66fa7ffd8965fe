// C ABI of the tree code: build, node quantities and the tree sum (include/ofw_hip.h).
#include <algorithm>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree.h"
#include "tree_eval.h"
#include "tree_records.h"

namespace ofw {

__global__ void k_aos_to_soa(int m, const double* __restrict__ aos, double* __restrict__ x, double* __restrict__ y,
                             double* __restrict__ z) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  x[s] = aos[3 * (size_t)s];
  y[s] = aos[3 * (size_t)s + 1];
  z[s] = aos[3 * (size_t)s + 2];
}

__global__ void k_i32_to_i64(int count, const int32_t* __restrict__ src, int64_t* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

// out_ref[pre_ref[i] * width + q] = in_pre[i * width + q]
__global__ void k_scatter_nodes(int count, int width, const int32_t* __restrict__ pre_ref,
                                const double* __restrict__ in, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const size_t r = pre_ref[i];
  for (int q = 0; q < width; ++q) out[r * width + q] = in[(size_t)i * width + q];
}

static size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// Shared front half of every tree entry point: midpoints, the build, and the tree order.
static int build_common(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths,
                        int64_t leaf, HostTree& t, int32_t** order_out) {
  OFW_HIP(c.tree[3].reserve(al((size_t)(m > 0 ? m : 1) * 8) * 3 + al((size_t)(m > 0 ? m : 1) * 4)));
  char* base = (char*)c.tree[3].ptr;
  double* mx = (double*)base;
  double* my = (double*)(base + al((size_t)m * 8));
  double* mz = (double*)(base + 2 * al((size_t)m * 8));
  int32_t* order = (int32_t*)(base + 3 * al((size_t)m * 8));
  int st = tree_prep(c, m, starts, directions, lengths, mx, my, mz);
  if (st) return st;
  st = build_tree(c, m, mx, my, mz, leaf, order, t);
  if (st) return st;
  *order_out = order;
  return OFW_OK;
}

}  // namespace ofw

using ofw::set_error;

extern "C" {

int ofw_build_tree(ofw_ctx* x, int64_t m, const double* mid, int64_t leaf_size, int64_t* order, int64_t* lo,
                   int64_t* hi, int64_t* left, int64_t* right, int64_t* count_host) {
  if (!x || m < 0 || leaf_size < 1 || m > ((int64_t)1 << 30) || !count_host) {
    set_error("ofw_build_tree: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  OFW_HIP(c.tree[3].reserve(ofw::al((size_t)(m > 0 ? m : 1) * 8) * 3 + ofw::al((size_t)(m > 0 ? m : 1) * 4)));
  char* base = (char*)c.tree[3].ptr;
  double* mx = (double*)base;
  double* my = (double*)(base + ofw::al((size_t)m * 8));
  double* mz = (double*)(base + 2 * ofw::al((size_t)m * 8));
  int32_t* ord32 = (int32_t*)(base + 3 * ofw::al((size_t)m * 8));
  if (m > 0)
    hipLaunchKernelGGL(ofw::k_aos_to_soa, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, mid, mx,
                       my, mz);
  ofw::HostTree t;
  int st = ofw::build_tree(c, m, mx, my, mz, leaf_size, ord32, t);
  if (st) return st;
  if (m > 0)
    hipLaunchKernelGGL(ofw::k_i32_to_i64, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, ord32,
                       order);
  const int64_t count = t.count;
  std::vector<int64_t> buf((size_t)count * 4);
  for (int64_t i = 0; i < count; ++i) {
    buf[i] = t.ref_lo[i];
    buf[count + i] = t.ref_hi[i];
    buf[2 * count + i] = t.ref_left[i];
    buf[3 * count + i] = t.ref_right[i];
  }
  OFW_HIP(hipMemcpyAsync(lo, buf.data(), count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(hi, buf.data() + count, count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(left, buf.data() + 2 * count, count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(right, buf.data() + 3 * count, count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  *count_host = count;
  return OFW_OK;
}

int ofw_tree_nodes(ofw_ctx* x, int64_t m, const double* starts, const double* directions, const double* lengths,
                   const double* gamma, const double* radius, int64_t leaf_size, double* centre, double* bound,
                   double* core, double* m0, double* m1, double* m2, double* w1, double* t2, double* w2,
                   int64_t* count_host) {
  if (!x || m < 1 || leaf_size < 1 || m > ((int64_t)1 << 30) || !count_host) {
    set_error("ofw_tree_nodes: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  ofw::HostTree t;
  int32_t* order = nullptr;
  int st = ofw::build_common(c, m, starts, directions, lengths, leaf_size, t, &order);
  if (st) return st;
  const int count = (int)t.count;
  // pre-order outputs in a scratch block, then scattered to the reference numbering
  const int widths[9] = {3, 1, 1, 3, 9, 27, 3, 3, 9};
  double* outs[9] = {centre, bound, core, m0, m1, m2, w1, t2, w2};
  size_t total = 0;
  for (int w : widths) total += ofw::al((size_t)count * w * 8);
  OFW_HIP(c.tree[4].reserve(total + ofw::al((size_t)count * 4)));
  char* base = (char*)c.tree[4].ptr;
  double* pre[9];
  for (int q = 0; q < 9; ++q) {
    pre[q] = (double*)base;
    base += ofw::al((size_t)count * widths[q] * 8);
  }
  int32_t* d_pre_ref = (int32_t*)base;
  ofw::TreeDevice td;
  td.out.bound = pre[1];
  td.out.m0 = pre[3];
  td.out.m1 = pre[4];
  td.out.m2 = pre[5];
  td.out.w1 = pre[6];
  td.out.t2 = pre[7];
  td.out.w2 = pre[8];
  st = ofw::tree_moments(c, t, order, starts, directions, lengths, gamma, radius, OFW_MODEL_NONE, 0.0, td, true);
  if (st) return st;
  OFW_HIP(hipMemcpyAsync(pre[0], td.centre, (size_t)count * 3 * 8, hipMemcpyDeviceToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(pre[2], td.core, (size_t)count * 8, hipMemcpyDeviceToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(d_pre_ref, t.pre_ref.data(), (size_t)count * 4, hipMemcpyHostToDevice, c.stream));
  for (int q = 0; q < 9; ++q)
    hipLaunchKernelGGL(ofw::k_scatter_nodes, dim3((count + 127) / 128), dim3(128), 0, c.stream, count, widths[q],
                       d_pre_ref, pre[q], outs[q]);
  OFW_HIP(hipGetLastError());
  OFW_HIP(hipStreamSynchronize(c.stream));
  *count_host = count;
  return OFW_OK;
}

int ofw_tree_sum(ofw_ctx* x, int precision, int64_t n, const double* points, int64_t m, const double* starts,
                 const double* directions, const double* lengths, const double* gamma, const double* radius,
                 int model, int raise_singular, const int64_t* ex_ptr, const int64_t* ex_seg, double theta,
                 int expansion_order, int64_t leaf_size, double clearance, int split_depth, double* out,
                 int8_t* flags, int64_t* far_counts, int64_t* near_counts) {
  if (!x || precision < 0 || precision > 2 || n < 0 || m < 0 || leaf_size < 1 || expansion_order < 0 ||
      expansion_order > 2 || !(theta > 0.0 && theta < 1.0) || !(clearance >= 0.0) || n > ((int64_t)1 << 30) ||
      m > ((int64_t)1 << 28)) {
    set_error("ofw_tree_sum: invalid arguments");
    return OFW_ERR_ARG;
  }
  if (model < 0 || model > 2) {
    set_error("tree backend supports core models NONE, RANKINE and LAMB_OSEEN only");
    return OFW_ERR_UNSUPPORTED;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0) return OFW_OK;
  if (m == 0) {
    OFW_HIP(hipMemsetAsync(out, 0, (size_t)n * 3 * 8, c.stream));
    OFW_HIP(hipMemsetAsync(flags, 0, (size_t)n, c.stream));
    if (far_counts) OFW_HIP(hipMemsetAsync(far_counts, 0, (size_t)n * 8, c.stream));
    if (near_counts) OFW_HIP(hipMemsetAsync(near_counts, 0, (size_t)n * 8, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    return OFW_OK;
  }
  OFW_HIP(c.mark());
  ofw::HostTree t;
  int32_t* order = nullptr;
  int st = ofw::build_common(c, m, starts, directions, lengths, leaf_size, t, &order);
  if (st) return st;
  OFW_HIP(c.mark());
  ofw::TreeDevice td;
  st = ofw::tree_moments(c, t, order, starts, directions, lengths, gamma, radius, model, clearance, td, false);
  if (st) return st;
  OFW_HIP(c.mark());
  int32_t *p32 = nullptr, *s32 = nullptr;
  st = ofw::csr_to_i32(c, n, ex_ptr, ex_seg, &p32, &s32);
  if (st) return st;
  std::vector<ofw::EvalTask> tasks;
  st = ofw::tree_evaluate(c, t, td, order, precision, n, points, model, raise_singular, p32, s32, theta,
                          expansion_order, split_depth, out, flags, far_counts, near_counts, tasks);
  if (st) return st;
  OFW_HIP(c.mark());
  OFW_HIP(hipStreamSynchronize(c.stream));
  c.stats_i[4] = t.count;
  c.stats_i[5] = t.depth;
  return OFW_OK;
}

}  // extern "C"
