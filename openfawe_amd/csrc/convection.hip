// Device pieces of the free-wake geometry step (wake.py:404-491, _network_convection, and wake.py:559-616, the
// stage loop of convect_geometry), so that a whole step stays in HBM:
//   * ofw_network_segments -- the segment preparation of segments_velocity (vortex.py:458-465) from a node/edge
//     network: starts = nodes[edges[:,0]], delta = nodes[edges[:,1]] - starts, L = hypot(hypot(dx,dy),dz),
//     directions = delta / L, with the zero-length and index-range errors;
//   * ofw_exclusion_subset -- the exclusion CSR (vortex_numba.py:180-243: per point, the sorted unique excluded
//     segments) of the pairs whose point lies in [lo, hi), shifted by -lo or mapped through a point map (several
//     target ranges evaluated in one sum), optionally through a segment remap (the ring-edge subset of
//     wake.py:453-462), from an incident-pair list;
//   * ofw_stage_combine -- the RK stage arithmetic of wake.py:604-614 and the stage bound of wake.py:596-598, in
//     NumPy's evaluation order (left to right, every operation rounded on its own).
#include <climits>

#include "common.h"
#include "internal.h"
#include "tree.h"

namespace ofw {

namespace {
size_t al(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

// status: OFW_STEP_EDGE_INDEX when an edge names a node outside [0, n), OFW_STEP_ZERO_LENGTH for a zero-length
// segment.
__global__ void k_network_segments(int m, int64_t n, const double* __restrict__ nodes,
                                   const int64_t* __restrict__ edges, double* __restrict__ starts,
                                   double* __restrict__ dir, double* __restrict__ len, int32_t* __restrict__ status) {
#pragma clang fp contract(off)
  int bad = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < m; s += gridDim.x * blockDim.x) {
    const int64_t ia = edges[2 * (size_t)s], ib = edges[2 * (size_t)s + 1];
    if (ia < 0 || ia >= n || ib < 0 || ib >= n) {
      bad |= OFW_STEP_EDGE_INDEX;
      continue;
    }
    const double ax = nodes[3 * ia], ay = nodes[3 * ia + 1], az = nodes[3 * ia + 2];
    const double dx = nodes[3 * ib] - ax, dy = nodes[3 * ib + 1] - ay, dz = nodes[3 * ib + 2] - az;
    const double L = hypot(hypot(dx, dy), dz);
    if (L == 0.0) bad |= OFW_STEP_ZERO_LENGTH;
    starts[3 * (size_t)s] = ax;
    starts[3 * (size_t)s + 1] = ay;
    starts[3 * (size_t)s + 2] = az;
    len[s] = L;
    dir[3 * (size_t)s] = dx / L;
    dir[3 * (size_t)s + 1] = dy / L;
    dir[3 * (size_t)s + 2] = dz / L;
  }
  if (__builtin_amdgcn_ballot_w64(bad != 0) != 0) {
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o);
    if ((threadIdx.x & 63) == 0) atomicOr(status, bad);
  }
}

// pkey = point - lo for a kept pair (point in [lo, hi) and, with a remap, remap[segment] >= 0), else hi - lo;
// seg = the (remapped) segment, -1 when dropped; the first sort key is the segment.  status bit 0: a pair names a
// segment outside [0, m).
// With `live` (the pair count on the device, e.g. the incident-pair total), only pairs below it are read.
__global__ void k_exclusion_keys(int npairs, const int64_t* __restrict__ pairs, const unsigned long long* live,
                                 int64_t lo, int64_t hi, int count, const int64_t* __restrict__ point_map, int64_t m,
                                 const int64_t* __restrict__ remap, int32_t* __restrict__ key,
                                 int32_t* __restrict__ val, int32_t* __restrict__ pkey, int32_t* __restrict__ seg,
                                 int32_t* __restrict__ status) {
  bool bad = false;
  const int64_t valid = live ? (int64_t)(*live < (unsigned long long)npairs ? *live : (unsigned long long)npairs)
                             : (int64_t)npairs;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += gridDim.x * blockDim.x) {
    const bool real = i < valid;
    const int64_t p = real ? pairs[2 * (size_t)i] : -1;
    int64_t s = real ? pairs[2 * (size_t)i + 1] : -1;
    int64_t q = p - lo;
    bool inside = real && p >= lo && p < hi;
    if (inside && point_map) q = point_map[q];
    if (inside && q >= count) {
      bad = true;
      q = -1;
    }
    inside = inside && q >= 0;
    if (inside && (s < 0 || s >= m)) {
      bad = true;
      s = -1;
    } else if (inside && remap) {
      s = remap[s];
    }
    const bool keep = inside && s >= 0;
    pkey[i] = keep ? (int32_t)q : count;
    seg[i] = keep ? (int32_t)s : -1;
    key[i] = keep ? (int32_t)s : 0;
    val[i] = i;
  }
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(status, OFW_STEP_EXCLUSION);
}

__global__ void k_gather_keys(int npairs, const int32_t* __restrict__ val, const int32_t* __restrict__ pkey,
                              int32_t* __restrict__ key) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npairs; q += gridDim.x * blockDim.x) key[q] = pkey[val[q]];
}

// After the stable sorts by segment and then by point (each point's segments non-decreasing): keep the first of
// equal (point, segment) runs.
__global__ void k_exclusion_mark(int npairs, int count, const int32_t* __restrict__ key,
                                 const int32_t* __restrict__ val, const int32_t* __restrict__ seg,
                                 int32_t* __restrict__ keep, int32_t* __restrict__ per_point) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npairs; q += gridDim.x * blockDim.x) {
    const int k = key[q];
    int flag = k < count;
    if (flag && q > 0 && key[q - 1] == k && seg[val[q - 1]] == seg[val[q]]) flag = 0;
    keep[q] = flag;
    if (flag) atomicAdd(&per_point[k], 1);
  }
}

__global__ void k_exclusion_emit(int npairs, const int32_t* __restrict__ keep, const int32_t* __restrict__ pos,
                                 const int32_t* __restrict__ val, const int32_t* __restrict__ seg,
                                 int64_t* __restrict__ seg_out) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npairs; q += gridDim.x * blockDim.x)
    if (keep[q]) seg_out[pos[q]] = seg[val[q]];
}

__global__ void k_ptr_widen(int count, const int32_t* __restrict__ start, const int32_t* __restrict__ total,
                            int64_t* __restrict__ ptr) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= count; p += gridDim.x * blockDim.x)
    ptr[p] = p < count ? start[p] : *total;
}

__global__ void k_conv_fill(int count, int32_t* __restrict__ a, int32_t v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) a[i] = v;
}

// Stage arithmetic, element by element over `count` doubles (wake.py:596-614):
//   OFW_STAGE_STEP    out = x + h * a                      (Euler step, Heun predictor, RK4 stage 4)
//   OFW_STAGE_HALF    out = x + h * a / 2                  (RK4 stages 2 and 3)
//   OFW_STAGE_HEUN    out = x + h * (a + b) / 2
//   OFW_STAGE_RK4     out = x + h * (a + 2 b + 2 c + d) / 6
//   OFW_STAGE_ADD     out = x + a                          (ambient + induced velocity)
// x and out may be the same array (an in-place update): neither is declared restrict.
__global__ void k_stage_combine(int mode, int64_t count, double h, const double* x, const double* a,
                                const double* b, const double* cc, const double* d, double* out) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    double v;
    switch (mode) {
      case OFW_STAGE_STEP: v = x[i] + h * a[i]; break;
      case OFW_STAGE_HALF: v = x[i] + h * a[i] / 2.0; break;
      case OFW_STAGE_HEUN: v = x[i] + h * (a[i] + b[i]) / 2.0; break;
      case OFW_STAGE_RK4: {
        double s = a[i] + 2.0 * b[i];
        s = s + 2.0 * cc[i];
        s = s + d[i];
        v = x[i] + h * s / 6.0;
        break;
      }
      default: v = x[i] + a[i]; break;
    }
    out[i] = v;
  }
}

}  // namespace ofw

namespace ofw {

int network_segments_enqueue(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, double* starts,
                             double* directions, double* lengths, int32_t* status) {
  if (m == 0) return OFW_OK;
  const unsigned grid = (unsigned)std::min<int64_t>(c.cu_count * 8, (m + 255) / 256);
  hipLaunchKernelGGL(k_network_segments, dim3(grid), dim3(256), 0, c.stream, (int)m, n, nodes, edges, starts,
                     directions, lengths, status);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

int exclusion_subset_enqueue(Ctx& c, int64_t npairs, const int64_t* pairs, const unsigned long long* live, int64_t lo,
                             int64_t hi, int64_t count64, const int64_t* point_map, int64_t m, const int64_t* remap,
                             int segment_ordered, int64_t* ptr_out, int64_t* seg_out, int32_t* status) {
  const int P = (int)npairs, count = (int)count64;
  const int hist_n = std::max(radix_hist_size(P), 1);
  const int scan_max = std::max(std::max(P, count + 1), hist_n);
  const size_t spart = ((size_t)(scan_max + 4095) / 4096 + 8) * 4;
  const size_t need = 8 * al((size_t)std::max(P, 1) * 4) + 2 * al((size_t)(count + 1) * 4) +
                      2 * al((size_t)hist_n * 4) + al(spart) + al(4);
  OFW_HIP(c.tree[8].reserve(need));
  char* base = (char*)c.tree[8].ptr;
  auto carve = [&](size_t bytes) {
    char* p = base;
    base += al(bytes);
    return (int32_t*)p;
  };
  const size_t pb = (size_t)std::max(P, 1) * 4;
  int32_t* key[2] = {carve(pb), carve(pb)};
  int32_t* val[2] = {carve(pb), carve(pb)};
  int32_t* seg = carve(pb);
  int32_t* pkey = carve(pb);
  int32_t* keep = carve(pb);
  int32_t* pos = carve(pb);
  int32_t* per_point = carve((size_t)(count + 1) * 4);
  int32_t* start = carve((size_t)(count + 1) * 4);
  int32_t* hist = carve((size_t)hist_n * 4);
  int32_t* hist_off = carve((size_t)hist_n * 4);
  int32_t* part = carve(spart);
  int32_t* total = carve(4);
  const int grid = c.cu_count * 8;
  auto blocks = [&](int64_t items) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (items + 255) / 256)));
  };
  hipLaunchKernelGGL(k_conv_fill, blocks(count + 1), dim3(256), 0, c.stream, count + 1, per_point, 0);
  hipLaunchKernelGGL(k_conv_fill, dim3(1), dim3(64), 0, c.stream, 1, total, 0);
  if (P > 0) {
    hipLaunchKernelGGL(k_exclusion_keys, blocks(P), dim3(256), 0, c.stream, P, pairs, live, lo, hi, count, point_map,
                       m, remap, key[0], val[0], pkey, seg, status);
    // LSD: by segment (unless the pairs already come in segment order), then (stably) by point
    int cur = 0;
    if (!segment_ordered) cur = radix_sort(c, P, radix_passes(std::max<int64_t>(m - 1, 0)), key, val, hist, hist_off, part);
    hipLaunchKernelGGL(k_gather_keys, blocks(P), dim3(256), 0, c.stream, P, val[cur], pkey, key[cur]);
    int32_t* key2[2] = {key[cur], key[cur ^ 1]};
    int32_t* val2[2] = {val[cur], val[cur ^ 1]};
    cur = radix_sort(c, P, radix_passes(count), key2, val2, hist, hist_off, part);
    hipLaunchKernelGGL(k_exclusion_mark, blocks(P), dim3(256), 0, c.stream, P, count, key2[cur], val2[cur], seg, keep,
                       per_point);
    scan_exclusive(c, P, keep, pos, part, total);
    hipLaunchKernelGGL(k_exclusion_emit, blocks(P), dim3(256), 0, c.stream, P, keep, pos, val2[cur], seg, seg_out);
  }
  scan_exclusive(c, count + 1, per_point, start, part, nullptr);
  hipLaunchKernelGGL(k_ptr_widen, blocks(count + 1), dim3(256), 0, c.stream, count, start, total, ptr_out);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

// One status word in the context's flag buffer, cleared, for the synchronous entry points.
static int32_t* sync_status(Ctx& c) {
  if (c.pflag.reserve(sizeof(int32_t)) != hipSuccess) return nullptr;
  int32_t* status = (int32_t*)c.pflag.ptr;
  hipLaunchKernelGGL(k_conv_fill, dim3(1), dim3(64), 0, c.stream, 1, status, 0);
  return status;
}

static int read_status(Ctx& c, const int32_t* status, int32_t* host) {
  OFW_HIP(hipMemcpyAsync(host, status, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

}  // namespace ofw

extern "C" int ofw_network_segments(ofw_ctx* x, int64_t n, const double* nodes, int64_t m, const int64_t* edges,
                                    double* starts, double* directions, double* lengths) {
  if (!x || n < 0 || m < 0 || m > ((int64_t)1 << 28) || (m > 0 && (!nodes || !edges || !starts || !directions ||
                                                                  !lengths))) {
    ofw::set_error("ofw_network_segments: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (m == 0) return OFW_OK;
  int32_t* status = ofw::sync_status(c);
  if (!status) {
    ofw::set_error("ofw_network_segments: out of device memory");
    return OFW_ERR_HIP;
  }
  int st = ofw::network_segments_enqueue(c, n, nodes, m, edges, starts, directions, lengths, status);
  if (st) return st;
  int32_t host = 0;
  st = ofw::read_status(c, status, &host);
  if (st) return st;
  if (host & OFW_STEP_EDGE_INDEX) {
    ofw::set_error("connectivity index out of range");
    return OFW_ERR_ARG;
  }
  if (host & OFW_STEP_ZERO_LENGTH) {
    ofw::set_error("zero-length segment");
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

extern "C" int ofw_exclusion_subset(ofw_ctx* x, int64_t npairs, const int64_t* pairs, int64_t lo, int64_t hi,
                                    int64_t count64, const int64_t* point_map, int64_t m, const int64_t* remap,
                                    int segment_ordered, int64_t* ptr_out, int64_t* seg_out) {
  if (!x || npairs < 0 || npairs > ((int64_t)1 << 30) || lo < 0 || hi < lo || hi - lo > ((int64_t)1 << 30) ||
      count64 < 0 || count64 > ((int64_t)1 << 30) || (!point_map && count64 != hi - lo) || m < 0 ||
      m > ((int64_t)1 << 30) || !ptr_out || (npairs > 0 && (!pairs || !seg_out))) {
    ofw::set_error("ofw_exclusion_subset: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  int32_t* status = ofw::sync_status(c);
  if (!status) {
    ofw::set_error("ofw_exclusion_subset: out of device memory");
    return OFW_ERR_HIP;
  }
  int st = ofw::exclusion_subset_enqueue(c, npairs, pairs, nullptr, lo, hi, count64, point_map, m, remap,
                                         segment_ordered, ptr_out, seg_out, status);
  if (st) return st;
  int32_t host = 0;
  st = ofw::read_status(c, status, &host);
  if (st) return st;
  if (host) {
    ofw::set_error("exclusion index out of range");  // a segment outside [0, m) or a map beyond count
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

extern "C" int ofw_stage_combine(ofw_ctx* x, int mode, int64_t count, double h, const double* xv, const double* a,
                                 const double* b, const double* cc, const double* d, double* out) {
  const bool two = mode == OFW_STAGE_HEUN, four = mode == OFW_STAGE_RK4;
  if (!x || mode < OFW_STAGE_STEP || mode > OFW_STAGE_ADD || count < 0 ||
      (count > 0 && (!xv || !a || !out || ((two || four) && !b) || (four && (!cc || !d))))) {
    ofw::set_error("ofw_stage_combine: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (count == 0) return OFW_OK;
  const unsigned grid = (unsigned)std::min<int64_t>(c.cu_count * 8, (count + 255) / 256);
  hipLaunchKernelGGL(ofw::k_stage_combine, dim3(grid), dim3(256), 0, c.stream, mode, count, h, xv, a, b, cc, d, out);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}
