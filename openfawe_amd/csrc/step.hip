// A free-wake geometry step as one recorded graph: the stage loop of convect_geometry (wake.py:586-616) over the
// target split of _network_convection (wake.py:404-491), for the tree code with one set of cores, a uniform
// inflow and frozen rows that keep it.  Every stage enqueues, with no host synchronisation:
//   the stage nodes (bound row, moved linearly by the displacement, then the free points), the segment geometry,
//   on a side stream the incident pairs (at most the plan's capacity) and the exclusion CSR of the targets (wing
//   rows [0, rows) then ring nodes [ring_lo, ring_hi), packed by a point map), meanwhile the tree build and node
//   quantities, then (joined) the traversal ('raise' for all targets: a singular pair adds nothing in either mode,
//   only its flag differs), the stage velocity (inflow plus the sums) and the RK arithmetic.
// The host reads one status block per step: capacity overflows (the plan is made again with the measured sizes)
// and the conditions the reference raises on or repeats a ring call for (row flags, ring flags, zero-length
// segments, non-finite velocities); those send the caller back to the call-by-call step for the reference's exact
// behaviour.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree.h"
#include "tree_eval.h"

namespace ofw {

namespace {
size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// status block: [0] status bits, [1] the largest incident-pair total, [2] the largest list work-item count,
// [3] zero (the pair total when a stage has no free point)
constexpr int SB_STATUS = 0, SB_PAIRS = 1, SB_CHUNKS = 2, SB_ZERO = 3, SB_WORDS = 4;
}  // namespace

__global__ void k_step_clear(unsigned long long* block) {
  if (threadIdx.x < SB_WORDS) block[threadIdx.x] = 0;
}

// nodes = (bound + f * displacement, or bound) then the free points
__global__ void k_stage_nodes(int64_t first, int64_t k, const double* __restrict__ bound,
                              const double* __restrict__ disp, double f, int moved, const double* __restrict__ pts,
                              double* __restrict__ nodes) {
#pragma clang fp contract(off)
  const int64_t nb = 3 * first, total = 3 * (first + k);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    nodes[i] = i < nb ? (moved ? bound[i] + f * disp[i] : bound[i]) : pts[i - nb];
}

// targets = rows [0, rows) then ring nodes [ring_lo, ring_lo + rings) of the free points
__global__ void k_gather_targets(int64_t rows, int64_t ring_lo, int64_t rings, const double* __restrict__ pts,
                                 double* __restrict__ targets) {
  const int64_t total = 3 * (rows + rings);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = i / 3, comp = i - 3 * q;
    const int64_t src = q < rows ? q : ring_lo + (q - rows);
    targets[i] = pts[3 * src + comp];
  }
}

// stage velocity of every free point: the inflow, plus the sum at the targets, plus the drift when there is one
// (wake.py:598); flags and non-finite sums into the status bits
__global__ void k_stage_velocity(int64_t k, const double* __restrict__ inflow, const double* __restrict__ drift,
                                 int64_t rows, int64_t ring_lo, int64_t rings, const double* __restrict__ out,
                                 const int8_t* __restrict__ flags, double* __restrict__ vel,
                                 unsigned long long* __restrict__ block) {
#pragma clang fp contract(off)
  unsigned long long bits = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < k; q += (int64_t)gridDim.x * blockDim.x) {
    double vx = inflow[0], vy = inflow[1], vz = inflow[2];
    const int64_t t = q < rows ? q : (q >= ring_lo && q < ring_lo + rings ? rows + (q - ring_lo) : -1);
    if (t >= 0) {
      const double ox = out[3 * t], oy = out[3 * t + 1], oz = out[3 * t + 2];
      if (flags[t]) bits |= t < rows ? OFW_STEP_ROW_FLAG : OFW_STEP_RING_FLAG;
      if (!(isfinite(ox) && isfinite(oy) && isfinite(oz))) bits |= OFW_STEP_NONFINITE;
      vx = vx + ox;
      vy = vy + oy;
      vz = vz + oz;
    }
    if (drift) {
      vx = vx + drift[3 * q];
      vy = vy + drift[3 * q + 1];
      vz = vz + drift[3 * q + 2];
    }
    vel[3 * q] = vx;
    vel[3 * q + 1] = vy;
    vel[3 * q + 2] = vz;
  }
  for (int o = 32; o > 0; o >>= 1) bits |= __shfl_xor(bits, o);
  if ((threadIdx.x & 63) == 0 && bits) atomicOr(&block[SB_STATUS], bits);
}

// the incident-pair status and the tree's counters of a stage into the step's block
__global__ void k_step_collect(const unsigned long long* __restrict__ pair_status, int64_t pair_cap,
                               const int32_t* __restrict__ counters, unsigned long long* __restrict__ block) {
  if (threadIdx.x != 0) return;
  unsigned long long bits = 0;
  const unsigned long long total = pair_status[0];
  if (pair_status[1]) bits |= OFW_STEP_EDGE_INDEX;
  if (total > (unsigned long long)pair_cap || total > (unsigned long long)INT_MAX) bits |= OFW_STEP_PAIR_OVERFLOW;
  atomicMax(&block[SB_PAIRS], total);
  if (counters) {  // the tree code's counters (none for the direct sum)
    if (counters[C_OVERFLOW]) bits |= OFW_STEP_TREE_OVERFLOW;
    atomicMax(&block[SB_CHUNKS], (unsigned long long)(counters[C_CHUNKS] > 0 ? counters[C_CHUNKS] : 0));
  }
  if (bits) atomicOr(&block[SB_STATUS], bits);
}

__global__ void k_copy_f64(int64_t count, const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_combine(int mode, int64_t count, double h, const double* x, const double* a, const double* b,
                          const double* cc, const double* d, double* out) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    double v;
    switch (mode) {
      case OFW_STAGE_STEP: v = x[i] + h * a[i]; break;
      case OFW_STAGE_HALF: v = x[i] + h * a[i] / 2.0; break;
      case OFW_STAGE_HEUN: v = x[i] + h * (a[i] + b[i]) / 2.0; break;
      default: {
        double s = a[i] + 2.0 * b[i];
        s = s + 2.0 * cc[i];
        s = s + d[i];
        v = x[i] + h * s / 6.0;
        break;
      }
    }
    out[i] = v;
  }
}

}  // namespace ofw

struct ofw_step {
  ofw::Ctx c;
  ofw::TreeCall tc;
  ofw::TreeArrays t;
  // caller's arrays
  const double *bound = nullptr, *disp = nullptr, *points_in = nullptr, *inflow = nullptr, *drift = nullptr;
  const int64_t* edges = nullptr;
  double* points_out = nullptr;
  int64_t first = 0, k = 0, m = 0, rows = 0, ring_lo = 0, rings = 0;
  double time = 0.0, dt_s = 0.0;
  int substeps = 1, integrator = 0;
  int engine = 0;                              // 0 the tree code, 1 the direct sum
  int precision = OFW_MIXED, threads = 256, model = 0;  // threads: the direct sum's workgroup size
  const double *gamma = nullptr, *radius = nullptr;
  // owned arrays
  ofw::DevBuf store;
  double *nodes = nullptr, *stage_pts = nullptr, *kv[4] = {}, *targets = nullptr, *out = nullptr;
  double *starts = nullptr, *dirs = nullptr, *lens = nullptr;
  int8_t* flags = nullptr;
  int64_t *pairs = nullptr, *ex_ptr = nullptr, *ex_seg = nullptr, *point_map = nullptr;
  unsigned long long* block = nullptr;
  int64_t pair_cap = 0;
  hipStream_t user_stream = nullptr, capture_stream = nullptr;
  hipStream_t side = nullptr;                  // the incident pairs and the exclusion CSR, beside the tree build
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace ofw {

static dim3 grid_for(const Ctx& c, int64_t items) {
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(c.cu_count * 8, (items + 255) / 256)));
}

// Owned arrays for the current pair capacity (made again when it grows).
static int step_arrays(ofw_step& p) {
  const int64_t n = p.first + p.k, T = p.rows + p.rings, K3 = 3 * p.k;
  const size_t need = al(3 * n * 8) + al(K3 * 8) + 4 * al(K3 * 8) + 2 * al(3 * T * 8) + al(3 * p.m * 8) * 2 +
                      al(p.m * 8) + al(std::max<int64_t>(T, 1)) + al(2 * p.pair_cap * 8) + al((T + 1) * 8) +
                      al(p.pair_cap * 8) + al((p.ring_lo + p.rings) * 8) + al(SB_WORDS * 8);
  OFW_HIP(p.store.reserve(need));
  char* base = (char*)p.store.ptr;
  auto carve = [&](size_t bytes) {
    char* q = base;
    base += al(bytes);
    return (void*)q;
  };
  p.nodes = (double*)carve(3 * n * 8);
  p.stage_pts = (double*)carve(K3 * 8);
  for (auto& v : p.kv) v = (double*)carve(K3 * 8);
  p.targets = (double*)carve(3 * T * 8);
  p.out = (double*)carve(3 * T * 8);
  p.starts = (double*)carve(3 * p.m * 8);
  p.dirs = (double*)carve(3 * p.m * 8);
  p.lens = (double*)carve(p.m * 8);
  p.flags = (int8_t*)carve(std::max<int64_t>(T, 1));
  p.pairs = (int64_t*)carve(2 * p.pair_cap * 8);
  p.ex_ptr = (int64_t*)carve((T + 1) * 8);
  p.ex_seg = (int64_t*)carve(p.pair_cap * 8);
  p.point_map = (int64_t*)carve((p.ring_lo + p.rings) * 8);
  p.block = (unsigned long long*)carve(SB_WORDS * 8);
  return OFW_OK;
}

__global__ void k_point_map(int64_t rows, int64_t ring_lo, int64_t rings, int64_t* __restrict__ map) {
  const int64_t total = ring_lo + rings;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x)
    map[q] = q < rows ? q : (q >= ring_lo ? rows + (q - ring_lo) : -1);
}

// The incident pairs and the exclusion CSR of a stage, on the side stream (they feed the traversal only).
static int stage_exclusions(ofw_step& p, const unsigned long long** pstat) {
  Ctx& c = p.c;
  const int64_t n = p.first + p.k, T = p.rows + p.rings;
  int32_t* status = (int32_t*)(p.block + SB_STATUS);
  int st = incident_pairs_enqueue(c, n, p.nodes, p.m, p.edges, p.first, p.pairs, p.pair_cap, pstat);
  if (st) return st;
  if (!*pstat) *pstat = p.block + SB_ZERO;
  return exclusion_subset_enqueue(c, p.pair_cap, p.pairs, *pstat, 0, p.ring_lo + p.rings, T, p.point_map, p.m,
                                  nullptr, 1, p.ex_ptr, p.ex_seg, status);
}

// One stage: the velocity kv[slot] of the free points `pts` at time t.  The incident pairs and the exclusion CSR
// run on a side stream while the tree is built on this one; they join before the traversal.
static int step_stage(ofw_step& p, const double* pts, double t, int slot) {
  Ctx& c = p.c;
  const int64_t n = p.first + p.k, T = p.rows + p.rings;
  const bool moved = p.disp != nullptr && t != p.time;
  const double f = (t - p.time) / p.dt_s;
  hipLaunchKernelGGL(k_stage_nodes, grid_for(c, 3 * n), dim3(256), 0, c.stream, p.first, p.k, p.bound, p.disp, f,
                     (int)moved, pts, p.nodes);
  int32_t* status = (int32_t*)(p.block + SB_STATUS);
  int st = network_segments_enqueue(c, n, p.nodes, p.m, p.edges, p.starts, p.dirs, p.lens, status);
  if (st) return st;
  const hipStream_t main = c.stream;
  OFW_HIP(hipEventRecord(p.fork_ev, main));
  OFW_HIP(hipStreamWaitEvent(p.side, p.fork_ev, 0));
  c.stream = p.side;
  const unsigned long long* pstat = nullptr;
  st = stage_exclusions(p, &pstat);
  const hipError_t e_join = hipEventRecord(p.join_ev, p.side);
  c.stream = main;
  if (st) return st;
  OFW_HIP(e_join);
  hipLaunchKernelGGL(k_gather_targets, grid_for(c, 3 * T), dim3(256), 0, c.stream, p.rows, p.ring_lo, p.rings, pts,
                     p.targets);
  p.tc.points = p.targets;
  p.tc.starts = p.starts;
  p.tc.directions = p.dirs;
  p.tc.lengths = p.lens;
  p.tc.ex_ptr = p.ex_ptr;
  p.tc.ex_seg = p.ex_seg;
  p.tc.n_ex = p.pair_cap;
  p.tc.out = p.out;
  p.tc.flags = p.flags;
  if (p.engine == 0) {
    TreeDevice td;
    st = tree_pipeline_build(c, p.tc, p.t, td);
    if (st) return st;
    OFW_HIP(hipStreamWaitEvent(main, p.join_ev, 0));
    st = tree_pipeline_eval(c, p.tc, p.t, td);
    if (st) return st;
  } else {
    // the direct sum: the segment records beside the exclusions, then the sum over every target
    st = pack_segments(c, p.m, p.starts, p.dirs, p.lens, p.gamma, p.radius, p.model, (double*)c.rec64.ptr,
                       (float*)c.rec32.ptr);
    if (st) return st;
    OFW_HIP(hipStreamWaitEvent(main, p.join_ev, 0));
    st = direct_sum(c, p.precision, p.threads, T, p.targets, p.m, (const double*)c.rec64.ptr,
                    (const float*)c.rec32.ptr, p.model, 1, p.ex_ptr, p.ex_seg, p.out, p.flags, nullptr);
    if (st) return st;
  }
  hipLaunchKernelGGL(k_step_collect, dim3(1), dim3(64), 0, c.stream, pstat, p.pair_cap,
                     p.engine == 0 ? (const int32_t*)p.t.counters : nullptr, p.block);
  hipLaunchKernelGGL(k_stage_velocity, grid_for(c, p.k), dim3(256), 0, c.stream, p.k, p.inflow, p.drift, p.rows,
                     p.ring_lo,
                     p.rings, p.out, p.flags, p.kv[slot], p.block);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

static void combine(ofw_step& p, int mode, double h, const double* x, const double* a, const double* b,
                    const double* cc, const double* d, double* out) {
  hipLaunchKernelGGL(k_combine, grid_for(p.c, 3 * p.k), dim3(256), 0, p.c.stream, mode, 3 * p.k, h, x, a, b, cc, d,
                     out);
}

// The whole step: points_out <- points_in advanced by `substeps` steps of the integrator.
static int step_enqueue(ofw_step& p) {
  Ctx& c = p.c;
  hipLaunchKernelGGL(k_step_clear, dim3(1), dim3(64), 0, c.stream, p.block);
  hipLaunchKernelGGL(k_point_map, grid_for(c, p.ring_lo + p.rings), dim3(256), 0, c.stream, p.rows, p.ring_lo,
                     p.rings, p.point_map);
  hipLaunchKernelGGL(k_copy_f64, grid_for(c, 3 * p.k), dim3(256), 0, c.stream, 3 * p.k, p.points_in, p.points_out);
  double* x = p.points_out;
  const double dt = p.dt_s / p.substeps;
  for (int sub = 0; sub < p.substeps; ++sub) {
    const double t = p.time + sub * dt;
    int st = step_stage(p, x, t, 0);
    if (st) return st;
    if (p.integrator == 0) {
      combine(p, OFW_STAGE_STEP, dt, x, p.kv[0], nullptr, nullptr, nullptr, x);
    } else if (p.integrator == 1) {
      combine(p, OFW_STAGE_STEP, dt, x, p.kv[0], nullptr, nullptr, nullptr, p.stage_pts);
      if ((st = step_stage(p, p.stage_pts, t + dt, 1))) return st;
      combine(p, OFW_STAGE_HEUN, dt, x, p.kv[0], p.kv[1], nullptr, nullptr, x);
    } else {
      combine(p, OFW_STAGE_HALF, dt, x, p.kv[0], nullptr, nullptr, nullptr, p.stage_pts);
      if ((st = step_stage(p, p.stage_pts, t + dt / 2, 1))) return st;
      combine(p, OFW_STAGE_HALF, dt, x, p.kv[1], nullptr, nullptr, nullptr, p.stage_pts);
      if ((st = step_stage(p, p.stage_pts, t + dt / 2, 2))) return st;
      combine(p, OFW_STAGE_STEP, dt, x, p.kv[2], nullptr, nullptr, nullptr, p.stage_pts);
      if ((st = step_stage(p, p.stage_pts, t + dt, 3))) return st;
      combine(p, OFW_STAGE_RK4, dt, x, p.kv[0], p.kv[1], p.kv[2], p.kv[3], x);
    }
  }
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

static int read_block(ofw_step& p, unsigned long long* host) {
  OFW_HIP(hipMemcpyAsync(host, p.block, SB_WORDS * 8, hipMemcpyDeviceToHost, p.c.stream));
  OFW_HIP(hipStreamSynchronize(p.c.stream));
  return OFW_OK;
}

}  // namespace ofw

using ofw::set_error;

extern "C" {

int ofw_step_destroy(ofw_step* p);

int ofw_step_create(ofw_ctx* x, ofw_step** plan, int precision, int64_t first, const double* bound,
                    const double* displacement, int64_t k, const double* points_in, double* points_out, int64_t m,
                    const int64_t* edges, const double* gamma, const double* radius, int model, int64_t rows,
                    int64_t ring_lo, int64_t ring_hi, double theta, int expansion_order, int64_t leaf_size,
                    double clearance, const double* inflow, const double* drift, double time, double dt_s,
                    int substeps, int integrator, int engine, int block, int use_graph, int64_t* stats_host) {
  const int max_model = engine == 0 ? 2 : 5;
  if (!x || !plan || first < 1 || k < 1 || m < 1 || !bound || !points_in || !points_out || !edges || !gamma ||
      engine < 0 || engine > 1 || !(block == 32 || block == 64 || block == 128 || block == 256) || model > max_model ||
      !radius || !inflow || rows < 0 || rows > k || ring_lo < rows || ring_hi < ring_lo || ring_hi > k ||
      rows + (ring_hi - ring_lo) < 1 || substeps < 1 || integrator < 0 || integrator > 2 || !(dt_s > 0.0) ||
      precision < 0 || precision > 2 || model < 0 || leaf_size < 1 || leaf_size > (1 << 20) ||
      expansion_order < 0 || expansion_order > 2 || !(theta > 0.0 && theta < 1.0) || !(clearance >= 0.0) ||
      first + k > (int64_t)1 << 28 || m > (int64_t)1 << 28) {
    set_error("ofw_step_create: invalid arguments");
    return OFW_ERR_ARG;
  }
  *plan = nullptr;
  ofw_step* p = new ofw_step();
  ofw::Ctx& c = p->c;
  c.device = x->c.device;
  c.cu_count = x->c.cu_count;
  c.stream = x->c.stream;
  p->user_stream = x->c.stream;
  auto fail = [&](int code) {
    ofw_step_destroy(p);
    return code;
  };
  for (auto& e : c.ev)
    if (hipEventCreate(&e) != hipSuccess) {
      set_error("ofw_step_create: hipEventCreate failed");
      return fail(OFW_ERR_HIP);
    }
  if (hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming) != hipSuccess) {
    set_error("ofw_step_create: side stream or events could not be created");
    return fail(OFW_ERR_HIP);
  }
  p->bound = bound;
  p->disp = displacement;
  p->points_in = points_in;
  p->points_out = points_out;
  p->inflow = inflow;
  p->drift = drift;
  p->edges = edges;
  p->first = first;
  p->k = k;
  p->m = m;
  p->rows = rows;
  p->ring_lo = ring_lo;
  p->rings = ring_hi - ring_lo;
  p->time = time;
  p->dt_s = dt_s;
  p->substeps = substeps;
  p->integrator = integrator;
  p->engine = engine;
  p->precision = precision;
  p->threads = block;
  p->model = model;
  p->gamma = gamma;
  p->radius = radius;
  if (engine == 1) {
    if (c.rec64.reserve((size_t)m * REC64 * sizeof(double)) != hipSuccess ||
        c.rec32.reserve((size_t)m * REC32 * sizeof(float)) != hipSuccess) {
      set_error("ofw_step_create: out of device memory");
      return fail(OFW_ERR_HIP);
    }
  }
  const int64_t T = rows + p->rings;
  ofw::TreeCall& tc = p->tc;
  tc.precision = precision;
  tc.K = 1;
  tc.n = T;
  tc.m = m;
  tc.toff = {0, T};
  tc.soff = {0, m};
  tc.gamma = gamma;
  tc.radius = radius;
  tc.model = model;
  tc.raise_singular = 1;
  tc.theta = theta;
  tc.order = expansion_order;
  tc.leaf = (int)leaf_size;
  tc.clearance = clearance;
  tc.split_depth = ofw::choose_split_depth(c, (T + 63) / 64, 1);
  // measure: a generous pair capacity and list slots read back per tree call, until no stage overflows
  p->pair_cap = 2 * m + (first + k);
  tc.chunk_slots = -1;
  unsigned long long host[ofw::SB_WORDS] = {};
  for (int attempt = 0;; ++attempt) {
    int st = ofw::step_arrays(*p);
    if (st) return fail(st);
    st = ofw::step_enqueue(*p);
    if (st) return fail(st);
    st = ofw::read_block(*p, host);
    if (st) return fail(st);
    if ((host[0] & OFW_STEP_TREE_OVERFLOW) && !c.sort_build && attempt < 3) {
      // the list build met equal midpoint coordinates (the only tree overflow of a measuring run): per-level sorts
      c.sort_build = true;
      continue;
    }
    if (!(host[0] & OFW_STEP_PAIR_OVERFLOW) || attempt >= 3) break;
    p->pair_cap = (int64_t)host[1] + (int64_t)host[1] / 4 + 1024;
  }
  if (host[0] & (OFW_STEP_PAIR_OVERFLOW | OFW_STEP_TREE_OVERFLOW)) {
    set_error("ofw_step_create: the step's capacities could not be measured");
    return fail(OFW_ERR_OVERFLOW);
  }
  // fixed capacities with margin; one eager run sizes every scratch buffer for them
  p->pair_cap = (int64_t)host[1] + (int64_t)host[1] / 4 + 1024;
  tc.chunk_slots = (int64_t)host[2] + (int64_t)host[2] / 2 + 1024;
  int st = ofw::step_arrays(*p);
  if (st) return fail(st);
  st = ofw::step_enqueue(*p);
  if (st) return fail(st);
  st = ofw::read_block(*p, host);
  if (st) return fail(st);
  if (use_graph) {
    if (hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("ofw_step_create: hipStreamCreate failed");
      return fail(OFW_ERR_HIP);
    }
    c.stream = p->capture_stream;
    c.capturing = true;
    if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      set_error("ofw_step_create: hipStreamBeginCapture failed");
      return fail(OFW_ERR_HIP);
    }
    const int st_rec = ofw::step_enqueue(*p);
    const hipError_t e_end = hipStreamEndCapture(c.stream, &p->graph);
    c.capturing = false;
    c.stream = p->user_stream;
    if (st_rec) return fail(st_rec);
    if (e_end != hipSuccess) {
      set_error(std::string("ofw_step_create: hipStreamEndCapture failed: ") + hipGetErrorString(e_end));
      return fail(OFW_ERR_HIP);
    }
    if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
      set_error("ofw_step_create: hipGraphInstantiate failed");
      return fail(OFW_ERR_HIP);
    }
  }
  if (stats_host) {
    stats_host[0] = (int64_t)host[0];
    stats_host[1] = p->pair_cap;
    stats_host[2] = tc.chunk_slots;
  }
  *plan = p;
  return OFW_OK;
}

int ofw_step_run(ofw_step* p, int64_t* status_host) {
  if (!p || !status_host) {
    set_error("ofw_step_run: null argument");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = p->c;
  c.stream = p->user_stream;
  c.nev = 0;
  OFW_HIP(c.mark());
  if (p->exec) {
    OFW_HIP(hipGraphLaunch(p->exec, c.stream));
  } else {
    const int st = ofw::step_enqueue(*p);
    if (st) return st;
  }
  OFW_HIP(c.mark());
  unsigned long long host[ofw::SB_WORDS] = {};
  const int st = ofw::read_block(*p, host);
  if (st) return st;
  status_host[0] = (int64_t)host[0];
  status_host[1] = (int64_t)host[1];
  status_host[2] = (int64_t)host[2];
  return OFW_OK;
}

int ofw_step_timings(ofw_step* p, float* ms_host) {
  if (!p || !ms_host) {
    set_error("ofw_step_timings: null argument");
    return OFW_ERR_ARG;
  }
  *ms_host = 0.0f;
  if (p->c.nev >= 2) OFW_HIP(hipEventElapsedTime(ms_host, p->c.ev[0], p->c.ev[1]));
  return OFW_OK;
}

int ofw_step_destroy(ofw_step* p) {
  if (!p) return OFW_OK;
  ofw::Ctx& c = p->c;
  if (p->user_stream)
    (void)hipStreamSynchronize(p->user_stream);
  else
    (void)hipDeviceSynchronize();
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->capture_stream) {
    (void)hipStreamSynchronize(p->capture_stream);
    (void)hipStreamDestroy(p->capture_stream);
  }
  if (p->side) {
    (void)hipStreamSynchronize(p->side);
    (void)hipStreamDestroy(p->side);
  }
  if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
  if (p->join_ev) (void)hipEventDestroy(p->join_ev);
  p->store.release();
  c.rec64.release();
  c.rec32.release();
  c.part.release();
  c.pflag.release();
  c.expos.release();
  c.exptr32.release();
  c.exseg32.release();
  c.vjp.release();
  c.release_side();
  for (auto& b : c.tree) b.release();
  for (auto& e : c.ev)
    if (e) (void)hipEventDestroy(e);
  delete p;
  return OFW_OK;
}

}  // extern "C"
