// C ABI of the tree code: build, node quantities, the tree sum and the batched (graph-replayed) tree sum
// (include/ofw_hip.h).
#include <cstring>
#include <algorithm>
#include <climits>
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

template <typename T>
__global__ void k_fill(int64_t count, T* __restrict__ a, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = v;
}

static size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// Midpoints of the segments (scratch buffer 3 of the context).
static int midpoints(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths,
                     double** mx, double** my, double** mz) {
  const size_t M = (size_t)std::max<int64_t>(m, 1);
  OFW_HIP(c.tree[3].reserve(3 * al(M * 8)));
  char* base = (char*)c.tree[3].ptr;
  *mx = (double*)base;
  *my = (double*)(base + al(M * 8));
  *mz = (double*)(base + 2 * al(M * 8));
  return tree_prep(c, m, starts, directions, lengths, *mx, *my, *mz);
}

// The part of a tree sum that needs the segments only: midpoints, build and node quantities.
int tree_pipeline_build(Ctx& c, const TreeCall& tc, TreeArrays& t, TreeDevice& td) {
  OFW_HIP(c.mark());
  double *mx, *my, *mz;
  int st = midpoints(c, tc.m, tc.starts, tc.directions, tc.lengths, &mx, &my, &mz);
  if (st) return st;
  st = tree_layout(c, 0, tc.m, tc.K, tc.leaf, t);
  if (st) return st;
  st = upload_i32(c, tc.soff, t.roots_lo);
  if (st) return st;
  st = tree_build(c, t, mx, my, mz, tc.soff);
  if (st) return st;
  OFW_HIP(c.mark());
  // the traversal plan needs the tree only: outside a capture it runs on the second stream beside the node
  // quantities from their subtree kernel on (the plan's small kernels slow the permute more than the subtrees)
  if (!c.capturing) {
    if (!c.side) {
      if (hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&c.join_ev, hipEventDisableTiming) != hipSuccess) {
        c.release_side();
        (void)hipGetLastError();
      }
    }
    c.fork_pending = c.side != nullptr;
  }
  st = tree_moments(c, 1, t, tc.soff, tc.starts, tc.directions, tc.lengths, tc.gamma, tc.radius, tc.model,
                    tc.clearance, td, false, tc.precision == OFW_FP64);
  if (c.fork_pending) {  // not recorded by the node quantities (no segments, or an error)
    c.fork_pending = false;
    OFW_HIP(hipEventRecord(c.fork_ev, c.stream));
  }
  if (st) return st;
  OFW_HIP(c.mark());
  return OFW_OK;
}

// The part that needs the targets and their exclusions: the traversal.
static int traversal(Ctx& c, const TreeCall& tc, TreeArrays& t, const TreeDevice& td) {
  int32_t *p32 = nullptr, *s32 = nullptr;
  if (tc.ex_ptr && tc.n_ex > 0) {
    int st = tc.ex_capacity
                 ? csr_convert_checked(c, tc.n, tc.n_ex, tc.ex_ptr, tc.ex_seg, t.counters, &p32, &s32)
                 : csr_convert(c, tc.n, tc.n_ex, tc.ex_ptr, tc.ex_seg, &p32, &s32);
    if (st) return st;
  }
  EvalShape shape;
  shape.K = tc.K;
  shape.n = tc.n;
  shape.target_off = tc.toff;
  shape.seg_off = tc.soff;
  shape.split_depth = tc.split_depth;
  shape.chunk_slots = tc.chunk_slots;
  return tree_evaluate(c, 2, t, td, shape, tc.precision, tc.points, tc.model, tc.raise_singular, p32, s32,
                       p32 ? tc.n_ex : 0, tc.theta, tc.order, tc.out, tc.flags, tc.far_counts, tc.near_counts);
}

int tree_pipeline_eval(Ctx& c, const TreeCall& tc, TreeArrays& t, const TreeDevice& td) {
  // the plan's kernels go to the second stream, after the build; tree_evaluate joins it before the walk
  if (!c.capturing && c.side) {
    OFW_HIP(hipStreamWaitEvent(c.side, c.fork_ev, 0));
    c.plan_main = c.stream;
    c.plan_forked = true;
    c.stream = c.side;
  }
  int st = traversal(c, tc, t, td);
  if (c.plan_forked) {  // returned before tree_evaluate's join (an error): join here, the call's stream comes back
    const hipError_t e1 = hipEventRecord(c.join_ev, c.stream);
    c.stream = c.plan_main;
    c.plan_forked = false;
    const hipError_t e2 = hipStreamWaitEvent(c.stream, c.join_ev, 0);
    if (!st && (e1 != hipSuccess || e2 != hipSuccess)) {
      set_error("tree call: joining the plan stream failed");
      return OFW_ERR_HIP;
    }
  }
  if (st) return st;
  OFW_HIP(c.mark());
  return OFW_OK;
}

int tree_pipeline(Ctx& c, const TreeCall& tc, TreeArrays& t) {
  TreeDevice td;
  int st = tree_pipeline_build(c, tc, t, td);
  if (st) return st;
  return tree_pipeline_eval(c, tc, t, td);
}

// The counters of a finished call: an error when a capacity was exceeded, else the statistics.
static int check_counters(Ctx& c, const int32_t* counters) {
  if (counters[C_DUP]) {
    set_error("tree code: the recorded list build met equal midpoint coordinates (make the plan again)");
    return OFW_ERR_OVERFLOW;
  }
  if (counters[C_OVERFLOW]) {
    set_error("tree code exceeded a capacity of its build or traversal plan");
    return OFW_ERR_OVERFLOW;
  }
  c.stats_i[2] = counters[C_TASKS];
  c.stats_i[4] = counters[C_FINAL];
  c.stats_i[5] = counters[C_DEPTH];
  return OFW_OK;
}

// After the stream has finished: the build's counters, and an error when a capacity was exceeded.
static int read_counters(Ctx& c, const TreeArrays& t, int32_t* counters) {
  OFW_HIP(hipMemcpyAsync(counters, t.counters, C_COUNT * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  return check_counters(c, counters);
}

// A build for the bit-exactness entry points: the list build, made again with the per-level sorts when the
// midpoints have equal coordinates (one read-back of the counters).
static int build_once(Ctx& c, TreeArrays& t, int64_t m, int leaf, const double* mx, const double* my,
                      const double* mz, const std::vector<int64_t>& offsets) {
  const bool saved = c.sort_build;
  for (int attempt = 0; attempt < 2; ++attempt) {
    int st = tree_layout(c, 0, m, 1, leaf, t);
    if (st) return st;
    st = upload_i32(c, offsets, t.roots_lo);
    if (st) return st;
    st = tree_build(c, t, mx, my, mz, offsets);
    if (st) return st;
    int32_t dup = 0;
    OFW_HIP(hipMemcpyAsync(&dup, t.counters + C_DUP, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    if (!dup) break;
    c.sort_build = true;
  }
  c.sort_build = saved;
  return OFW_OK;
}

static int check_tree_args(const char* what, int precision, int64_t n, int64_t m, int64_t leaf_size, double theta,
                           int expansion_order, double clearance, int model, int split_depth) {
  if (precision < 0 || precision > 2 || n < 0 || m < 0 || leaf_size < 1 || leaf_size > (1 << 20) ||
      expansion_order < 0 || expansion_order > 2 || !(theta > 0.0 && theta < 1.0) || !(clearance >= 0.0) ||
      n > ((int64_t)1 << 30) || m > ((int64_t)1 << 28) || split_depth < -1) {
    set_error(std::string(what) + ": invalid arguments");
    return OFW_ERR_ARG;
  }
  if (model < 0 || model > 2) {
    set_error("tree backend supports core models NONE, RANKINE and LAMB_OSEEN only");
    return OFW_ERR_UNSUPPORTED;
  }
  return OFW_OK;
}

static int64_t target_groups(const std::vector<int64_t>& toff) {
  int64_t g = 0;
  for (size_t k = 0; k + 1 < toff.size(); ++k) g += (toff[k + 1] - toff[k] + 63) / 64;
  return g;
}

}  // namespace ofw

// The batched plan: its own scratch (so that the recorded graph's buffers stay valid), the call it replays,
// and the graph.
struct ofw_tree_batch {
  ofw::Ctx c;
  ofw::TreeCall call;
  ofw::TreeArrays t;
  hipStream_t user_stream = nullptr;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using ofw::set_error;

extern "C" {

int ofw_build_tree(ofw_ctx* x, int64_t m, const double* mid, int64_t leaf_size, int64_t* order, int64_t* lo,
                   int64_t* hi, int64_t* left, int64_t* right, int64_t* count_host) {
  if (!x || m < 0 || leaf_size < 1 || leaf_size > (1 << 20) || m > ((int64_t)1 << 28) || !count_host) {
    set_error("ofw_build_tree: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (m == 0) {
    // the reference's tree of no segments: one leaf root over the empty range
    const int64_t root[4] = {0, 0, -1, -1};
    OFW_HIP(hipMemcpyAsync(lo, &root[0], 8, hipMemcpyHostToDevice, c.stream));
    OFW_HIP(hipMemcpyAsync(hi, &root[1], 8, hipMemcpyHostToDevice, c.stream));
    OFW_HIP(hipMemcpyAsync(left, &root[2], 8, hipMemcpyHostToDevice, c.stream));
    OFW_HIP(hipMemcpyAsync(right, &root[3], 8, hipMemcpyHostToDevice, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    *count_host = 1;
    return OFW_OK;
  }
  OFW_HIP(c.tree[3].reserve(3 * ofw::al((size_t)m * 8)));
  char* base = (char*)c.tree[3].ptr;
  double* mx = (double*)base;
  double* my = (double*)(base + ofw::al((size_t)m * 8));
  double* mz = (double*)(base + 2 * ofw::al((size_t)m * 8));
  hipLaunchKernelGGL(ofw::k_aos_to_soa, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, mid, mx, my,
                     mz);
  ofw::TreeArrays t;
  const std::vector<int64_t> offsets = {0, m};
  int st = ofw::build_once(c, t, m, (int)leaf_size, mx, my, mz, offsets);
  if (st) return st;
  hipLaunchKernelGGL(ofw::k_i32_to_i64, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, t.order,
                     order);
  OFW_HIP(hipGetLastError());
  std::vector<int64_t> vlo, vhi, vleft, vright;
  std::vector<int32_t> ref_of_pre;
  int64_t count = 0;
  st = ofw::tree_reference_arrays(c, t, vlo, vhi, vleft, vright, ref_of_pre, count);
  if (st) return st;
  OFW_HIP(hipMemcpyAsync(lo, vlo.data(), count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(hi, vhi.data(), count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(left, vleft.data(), count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(right, vright.data(), count * 8, hipMemcpyHostToDevice, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  *count_host = count;
  return OFW_OK;
}

int ofw_tree_nodes(ofw_ctx* x, int64_t m, const double* starts, const double* directions, const double* lengths,
                   const double* gamma, const double* radius, int64_t leaf_size, double* centre, double* bound,
                   double* core, double* m0, double* m1, double* m2, double* w1, double* t2, double* w2,
                   int64_t* count_host) {
  if (!x || m < 1 || leaf_size < 1 || leaf_size > (1 << 20) || m > ((int64_t)1 << 28) || !count_host) {
    set_error("ofw_tree_nodes: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  double *mx, *my, *mz;
  int st = ofw::midpoints(c, m, starts, directions, lengths, &mx, &my, &mz);
  if (st) return st;
  ofw::TreeArrays t;
  const std::vector<int64_t> offsets = {0, m};
  st = ofw::build_once(c, t, m, (int)leaf_size, mx, my, mz, offsets);
  if (st) return st;
  std::vector<int64_t> vlo, vhi, vleft, vright;
  std::vector<int32_t> ref_of_pre;
  int64_t count64 = 0;
  st = ofw::tree_reference_arrays(c, t, vlo, vhi, vleft, vright, ref_of_pre, count64);
  if (st) return st;
  const int count = (int)count64;
  // pre-order outputs in a scratch block, then scattered to the reference numbering
  const int widths[9] = {3, 1, 1, 3, 9, 27, 3, 3, 9};
  double* outs[9] = {centre, bound, core, m0, m1, m2, w1, t2, w2};
  size_t total = 0;
  for (int w : widths) total += ofw::al((size_t)t.cap * w * 8);
  OFW_HIP(c.tree[4].reserve(total + ofw::al((size_t)count * 4)));
  char* base = (char*)c.tree[4].ptr;
  double* pre[9];
  for (int q = 0; q < 9; ++q) {
    pre[q] = (double*)base;
    base += ofw::al((size_t)t.cap * widths[q] * 8);
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
  st = ofw::tree_moments(c, 1, t, offsets, starts, directions, lengths, gamma, radius, OFW_MODEL_NONE, 0.0, td, true,
                         false);
  if (st) return st;
  OFW_HIP(hipMemcpyAsync(pre[0], td.centre, (size_t)count * 3 * 8, hipMemcpyDeviceToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(pre[2], td.core, (size_t)count * 8, hipMemcpyDeviceToDevice, c.stream));
  OFW_HIP(hipMemcpyAsync(d_pre_ref, ref_of_pre.data(), (size_t)count * 4, hipMemcpyHostToDevice, c.stream));
  for (int q = 0; q < 9; ++q)
    hipLaunchKernelGGL(ofw::k_scatter_nodes, dim3((count + 127) / 128), dim3(128), 0, c.stream, count, widths[q],
                       d_pre_ref, pre[q], outs[q]);
  OFW_HIP(hipGetLastError());
  int32_t counters[ofw::C_COUNT];
  st = ofw::read_counters(c, t, counters);
  if (st) return st;
  *count_host = count;
  return OFW_OK;
}

int ofw_tree_sum(ofw_ctx* x, int precision, int64_t n, const double* points, int64_t m, const double* starts,
                 const double* directions, const double* lengths, const double* gamma, const double* radius,
                 int model, int raise_singular, const int64_t* ex_ptr, const int64_t* ex_seg, double theta,
                 int expansion_order, int64_t leaf_size, double clearance, int split_depth, double* out,
                 int8_t* flags, int64_t* far_counts, int64_t* near_counts) {
  if (!x) {
    set_error("null context");
    return OFW_ERR_ARG;
  }
  int st = ofw::check_tree_args("ofw_tree_sum", precision, n, m, leaf_size, theta, expansion_order, clearance, model,
                                split_depth);
  if (st) return st;
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0) return OFW_OK;
  if (m == 0) {
    const unsigned g3 = (unsigned)((3 * n + 255) / 256), g1 = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(ofw::k_fill<double>, dim3(g3), dim3(256), 0, c.stream, 3 * n, out, 0.0);
    hipLaunchKernelGGL(ofw::k_fill<int8_t>, dim3(g1), dim3(256), 0, c.stream, n, flags, (int8_t)0);
    if (far_counts) hipLaunchKernelGGL(ofw::k_fill<int64_t>, dim3(g1), dim3(256), 0, c.stream, n, far_counts, (int64_t)0);
    if (near_counts)
      hipLaunchKernelGGL(ofw::k_fill<int64_t>, dim3(g1), dim3(256), 0, c.stream, n, near_counts, (int64_t)0);
    OFW_HIP(hipGetLastError());
    OFW_HIP(hipStreamSynchronize(c.stream));
    return OFW_OK;
  }
  ofw::TreeCall tc;
  tc.precision = precision;
  tc.K = 1;
  tc.n = n;
  tc.m = m;
  tc.toff = {0, n};
  tc.soff = {0, m};
  tc.points = points;
  tc.starts = starts;
  tc.directions = directions;
  tc.lengths = lengths;
  tc.gamma = gamma;
  tc.radius = radius;
  tc.model = model;
  tc.raise_singular = raise_singular;
  tc.ex_ptr = ex_ptr;
  tc.ex_seg = ex_seg;
  // the exclusion count is read back once per context; later calls assume the capacity kept from it (with margin)
  // and the device checks it, so that no call waits for a read-back before its first kernel
  auto keep_ex_cap = [&](int64_t e) { c.ex_cap = std::max(c.ex_cap, std::min<int64_t>(INT_MAX, e + e / 4 + 1024)); };
  if (ex_ptr) {
    if (c.ex_cap > 0) {
      tc.n_ex = c.ex_cap;
      tc.ex_capacity = true;
    } else {
      st = ofw::csr_count(c, n, ex_ptr, &tc.n_ex);
      if (st) return st;
      keep_ex_cap(tc.n_ex);
    }
  }
  tc.theta = theta;
  tc.order = expansion_order;
  tc.leaf = (int)leaf_size;
  tc.clearance = clearance;
  tc.split_depth = split_depth >= 0 ? split_depth : ofw::choose_split_depth(c, ofw::target_groups(tc.toff), 1);
  tc.out = out;
  tc.flags = flags;
  tc.far_counts = far_counts;
  tc.near_counts = near_counts;
  // the interaction-list slots of the previous call, with margin, spare this call's read-back of the count;
  // a call that needs more runs again with the count it measured
  // a call whose midpoints have equal coordinates runs again with the per-level sort build
  ofw::TreeArrays t;
  int32_t counters[ofw::C_COUNT];
  const bool saved_sort = c.sort_build;
  bool retried_slots = false;
  for (int attempt = 0;; ++attempt) {
    c.nev = 0;  // the phase timings describe the last attempt
    tc.chunk_slots = c.list_slots > 0 ? c.list_slots : -1;
    st = ofw::tree_pipeline(c, tc, t);
    if (st) {
      c.sort_build = saved_sort;
      return st;
    }
    // the counters through pinned memory: a direct copy, not a staged one
    if (!c.host_counters && hipHostMalloc((void**)&c.host_counters, sizeof(counters), hipHostMallocDefault) != hipSuccess) {
      c.host_counters = nullptr;
      (void)hipGetLastError();
    }
    int32_t* dst = c.host_counters ? c.host_counters : counters;
    OFW_HIP(hipMemcpyAsync(dst, t.counters, sizeof(counters), hipMemcpyDeviceToHost, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    if (dst != counters) std::memcpy(counters, dst, sizeof(counters));
    c.stats_i[7] = attempt + 1;  // runs of the pipeline this call took
    if (tc.ex_capacity && counters[ofw::C_EXNEED] != 0) {
      // more exclusions than the kept capacity (or a malformed CSR): count them and run again
      tc.ex_capacity = false;
      st = ofw::csr_count(c, n, ex_ptr, &tc.n_ex);
      if (st) {
        c.sort_build = saved_sort;
        return st;
      }
      keep_ex_cap(tc.n_ex);
      continue;
    }
    if (counters[ofw::C_DUP] && !c.sort_build) {
      c.sort_build = true;
      continue;
    }
    const int64_t used = counters[ofw::C_CHUNKS];
    const bool short_slots = counters[ofw::C_OVERFLOW] && tc.chunk_slots >= 0 && used > tc.chunk_slots;
    c.list_slots = used + used / 4 + 1024;
    if (!short_slots || retried_slots) break;
    retried_slots = true;
  }
  c.sort_build = saved_sort;
  c.stats_i[6] = counters[ofw::C_CHUNKS];
  return ofw::check_counters(c, counters);
}

int ofw_tree_batch_create(ofw_ctx* x, ofw_tree_batch** plan, int precision, int64_t K,
                          const int64_t* target_offsets_host, const int64_t* segment_offsets_host,
                          const double* points, const double* starts, const double* directions,
                          const double* lengths, const double* gamma, const double* radius, int model,
                          int raise_singular, const int64_t* ex_ptr, const int64_t* ex_seg, double theta,
                          int expansion_order, int64_t leaf_size, double clearance, int split_depth, int use_graph,
                          double* out, int8_t* flags, int64_t* far_counts, int64_t* near_counts) {
  if (!x || !plan || K < 1 || K > (1 << 20) || !target_offsets_host || !segment_offsets_host) {
    set_error("ofw_tree_batch_create: invalid arguments");
    return OFW_ERR_ARG;
  }
  *plan = nullptr;
  std::vector<int64_t> toff(target_offsets_host, target_offsets_host + K + 1);
  std::vector<int64_t> soff(segment_offsets_host, segment_offsets_host + K + 1);
  if (toff[0] != 0 || soff[0] != 0) {
    set_error("ofw_tree_batch_create: offsets must start at 0");
    return OFW_ERR_ARG;
  }
  for (int64_t k = 0; k < K; ++k)
    if (toff[k + 1] < toff[k] || soff[k + 1] < soff[k]) {
      set_error("ofw_tree_batch_create: offsets must not decrease");
      return OFW_ERR_ARG;
    }
  const int64_t n = toff[K], m = soff[K];
  int st = ofw::check_tree_args("ofw_tree_batch_create", precision, n, m, leaf_size, theta, expansion_order,
                                clearance, model, split_depth);
  if (st) return st;
  if (n == 0 || m == 0) {
    set_error("ofw_tree_batch_create: the batch needs targets and segments");
    return OFW_ERR_ARG;
  }
  ofw_tree_batch* b = new ofw_tree_batch();
  ofw::Ctx& c = b->c;
  c.device = x->c.device;
  c.cu_count = x->c.cu_count;
  c.stream = x->c.stream;
  b->user_stream = x->c.stream;
  auto fail = [&](int code) {
    ofw_tree_batch_destroy(b);
    return code;
  };
  for (auto& e : c.ev)
    if (hipEventCreate(&e) != hipSuccess) {
      set_error("ofw_tree_batch_create: hipEventCreate failed");
      return fail(OFW_ERR_HIP);
    }
  ofw::TreeCall& tc = b->call;
  tc.precision = precision;
  tc.K = (int)K;
  tc.n = n;
  tc.m = m;
  tc.toff = toff;
  tc.soff = soff;
  tc.points = points;
  tc.starts = starts;
  tc.directions = directions;
  tc.lengths = lengths;
  tc.gamma = gamma;
  tc.radius = radius;
  tc.model = model;
  tc.raise_singular = raise_singular;
  tc.ex_ptr = ex_ptr;
  tc.ex_seg = ex_seg;
  if (ex_ptr) {
    st = ofw::csr_count(c, n, ex_ptr, &tc.n_ex);
    if (st) return fail(st);
  }
  tc.theta = theta;
  tc.order = expansion_order;
  tc.leaf = (int)leaf_size;
  tc.clearance = clearance;
  tc.split_depth = split_depth >= 0 ? split_depth : ofw::choose_split_depth(c, ofw::target_groups(toff), (int)K);
  tc.out = out;
  tc.flags = flags;
  tc.far_counts = far_counts;
  tc.near_counts = near_counts;
  // one eager call measures the near-field work, a second one sizes every buffer for the fixed capacity
  // the recorded call keeps (half as much again as measured: the work varies with the contents of the inputs)
  c.nev = 0;
  st = ofw::tree_pipeline(c, tc, b->t);
  if (st) return fail(st);
  int32_t counters[ofw::C_COUNT];
  st = ofw::read_counters(c, b->t, counters);
  if (counters[ofw::C_DUP]) {  // equal midpoint coordinates: the plan builds with the per-level sorts
    c.sort_build = true;
    c.nev = 0;
    st = ofw::tree_pipeline(c, tc, b->t);
    if (st) return fail(st);
    st = ofw::read_counters(c, b->t, counters);
  }
  if (st) return fail(st);
  const int64_t measured = c.stats_i[6];
  tc.chunk_slots = measured + measured / 2 + 1024;
  c.nev = 0;
  st = ofw::tree_pipeline(c, tc, b->t);
  if (st) return fail(st);
  st = ofw::read_counters(c, b->t, counters);
  if (st) return fail(st);
  if (use_graph) {
    if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("ofw_tree_batch_create: hipStreamCreate failed");
      return fail(OFW_ERR_HIP);
    }
    c.stream = b->capture_stream;
    c.capturing = true;
    if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      set_error("ofw_tree_batch_create: hipStreamBeginCapture failed");
      return fail(OFW_ERR_HIP);
    }
    const int st_rec = ofw::tree_pipeline(c, tc, b->t);
    const hipError_t e_end = hipStreamEndCapture(c.stream, &b->graph);
    c.capturing = false;
    c.stream = b->user_stream;
    if (st_rec) return fail(st_rec);
    if (e_end != hipSuccess) {
      set_error(std::string("ofw_tree_batch_create: hipStreamEndCapture failed: ") + hipGetErrorString(e_end));
      return fail(OFW_ERR_HIP);
    }
    const hipError_t e_inst = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e_inst != hipSuccess) {
      set_error(std::string("ofw_tree_batch_create: hipGraphInstantiate failed: ") + hipGetErrorString(e_inst));
      return fail(OFW_ERR_HIP);
    }
  }
  *plan = b;
  return OFW_OK;
}

int ofw_tree_batch_run(ofw_tree_batch* b) {
  if (!b) {
    set_error("ofw_tree_batch_run: null plan");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = b->c;
  c.stream = b->user_stream;
  c.nev = 0;
  if (b->exec) {
    OFW_HIP(c.mark());
    OFW_HIP(hipGraphLaunch(b->exec, c.stream));
    OFW_HIP(c.mark());
  } else {
    int st = ofw::tree_pipeline(c, b->call, b->t);
    if (st) return st;
  }
  int32_t counters[ofw::C_COUNT];
  return ofw::read_counters(c, b->t, counters);
}

int ofw_tree_batch_timings(ofw_tree_batch* b, float* ms_host, int count) {
  if (!b || !ms_host) {
    set_error("ofw_tree_batch_timings: null argument");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = b->c;
  const int phases = c.nev > 1 ? c.nev - 1 : 0;
  for (int k = 0; k < count; ++k) {
    ms_host[k] = 0.0f;
    if (k < phases) OFW_HIP(hipEventElapsedTime(&ms_host[k], c.ev[k], c.ev[k + 1]));
  }
  return phases;
}

int ofw_tree_batch_stats(ofw_tree_batch* b, int64_t* stats_host, int count) {
  if (!b || !stats_host) {
    set_error("ofw_tree_batch_stats: null argument");
    return OFW_ERR_ARG;
  }
  int k = 0;
  for (; k < count && k < 8; ++k) stats_host[k] = b->c.stats_i[k];
  return k;
}

int ofw_tree_batch_destroy(ofw_tree_batch* b) {
  if (!b) return OFW_OK;
  ofw::Ctx& c = b->c;
  if (b->user_stream) (void)hipStreamSynchronize(b->user_stream);
  else (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture_stream) {
    (void)hipStreamSynchronize(b->capture_stream);
    (void)hipStreamDestroy(b->capture_stream);
  }
  c.rec64.release();
  c.rec32.release();
  c.part.release();
  c.pflag.release();
  c.expos.release();
  c.exptr32.release();
  c.exseg32.release();
  c.vjp.release();
  c.release_side();
  for (auto& buf : c.tree) buf.release();
  for (auto& e : c.ev)
    if (e) (void)hipEventDestroy(e);
  delete b;
  return OFW_OK;
}

}  // extern "C"
