// Direct Biot-Savart sum on gfx950.
//
// Work decomposition: a 2-D grid of (target block) x (segment chunk).  A wavefront owns 64 targets (one
// per lane) and walks its chunk of segments in input order; each segment record is wave-uniform and is
// read with scalar loads, so the only per-lane operands are the target and the accumulators.  20,570
// targets are only 322 wavefronts, far too few for 256 CUs, so the segments are cut into chunks and the
// per-chunk partial sums are added in chunk order by a second kernel (fixed order: deterministic).
// Per target, the result is the reference's ordered sum over segments (vortex_numba.py:112-147),
// re-associated at chunk boundaries.
#include <climits>

#include "common.h"
#include "internal.h"
#include "tree.h"
#include "tree_records.h"

namespace ofw {

// ---------------------------------------------------------------- segment records
__global__ __launch_bounds__(256) void k_pack_segments(int m, const double* __restrict__ starts,
                                                       const double* __restrict__ directions,
                                                       const double* __restrict__ lengths,
                                                       const double* __restrict__ gamma,
                                                       const double* __restrict__ radius, int model,
                                                       double* __restrict__ rec64, float* __restrict__ rec32) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  pack_record(s, starts, directions, lengths, gamma, radius, model, rec64 + (size_t)s * REC64,
              rec32 + (size_t)s * REC32);
}

int pack_segments(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths,
                  const double* gamma, const double* radius, int model, double* rec64, float* rec32) {
  if (m == 0) return OFW_OK;
  hipLaunchKernelGGL(k_pack_segments, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m, starts,
                     directions, lengths, gamma, radius, model, rec64, rec32);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

// int64 CSR -> int32 CSR (the kernels index with 32-bit integers)
__global__ void k_i64_to_i32(int64_t count, const int64_t* __restrict__ src, int32_t* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (int32_t)src[i];
}

int csr_count(Ctx& c, int64_t n, const int64_t* ex_ptr, int64_t* e_out) {
  int64_t e = 0;
  OFW_HIP(hipMemcpyAsync(&e, ex_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (e < 0 || e > INT_MAX) {
    set_error("exclusion count out of range");
    return OFW_ERR_ARG;
  }
  *e_out = e;
  return OFW_OK;
}

int csr_convert(Ctx& c, int64_t n, int64_t e, const int64_t* ex_ptr, const int64_t* ex_seg, int32_t** ptr32,
                int32_t** seg32) {
  OFW_HIP(c.exptr32.reserve((size_t)(n + 1) * sizeof(int32_t)));
  OFW_HIP(c.exseg32.reserve((size_t)(e > 0 ? e : 1) * sizeof(int32_t)));
  *ptr32 = (int32_t*)c.exptr32.ptr;
  *seg32 = (int32_t*)c.exseg32.ptr;
  hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, c.stream, n + 1, ex_ptr,
                     *ptr32);
  if (e > 0)
    hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, c.stream, e, ex_seg, *seg32);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

__global__ void k_ex_seg_checked(int64_t n, int64_t cap, const int64_t* __restrict__ ex_ptr,
                                 const int64_t* __restrict__ src, int32_t* __restrict__ dst,
                                 int32_t* __restrict__ counters) {
  const int64_t e = ex_ptr[n];
  if (e < 0 || e > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      counters[C_EXNEED] = e < 0 ? -1 : (int32_t)(e < (int64_t)INT_MAX ? e : (int64_t)INT_MAX);
      counters[C_OVERFLOW] = 1;
    }
    return;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (int32_t)src[i];
}

int csr_convert_checked(Ctx& c, int64_t n, int64_t cap, const int64_t* ex_ptr, const int64_t* ex_seg,
                        int32_t* counters, int32_t** ptr32, int32_t** seg32) {
  OFW_HIP(c.exptr32.reserve((size_t)(n + 1) * sizeof(int32_t)));
  OFW_HIP(c.exseg32.reserve((size_t)(cap > 0 ? cap : 1) * sizeof(int32_t)));
  *ptr32 = (int32_t*)c.exptr32.ptr;
  *seg32 = (int32_t*)c.exseg32.ptr;
  hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, c.stream, n + 1, ex_ptr,
                     *ptr32);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((cap + 255) / 256, 8 * (int64_t)c.cu_count));
  hipLaunchKernelGGL(k_ex_seg_checked, dim3((unsigned)blocks), dim3(256), 0, c.stream, n, cap, ex_ptr, ex_seg, *seg32,
                     counters);
  OFW_HIP(hipGetLastError());
  return OFW_OK;
}

int csr_to_i32(Ctx& c, int64_t n, const int64_t* ex_ptr, const int64_t* ex_seg, int32_t** ptr32, int32_t** seg32) {
  int64_t e = 0;
  int st = csr_count(c, n, ex_ptr, &e);
  if (st) return st;
  return csr_convert(c, n, e, ex_ptr, ex_seg, ptr32, seg32);
}

// ---------------------------------------------------------------- direct kernels
// code of a chunk's first flag: -1 when none, else seg * 4 + flag
__device__ __forceinline__ int flag_code(const FlagState& st) { return st.flag ? st.seg * 4 + st.flag : -1; }

// FP64 direct sum: each lane owns two targets (blockDim apart, so loads stay coalesced) and walks the chunk's
// segments in order; the segment record is wave-uniform (scalar loads) and shared by the 128 targets of the
// wavefront, and the two independent pair evaluations of a lane interleave.  An excluded pair is evaluated
// and not added; the lane's next excluded segment advances on the wavefront's ballot.
template <int MODEL>
__global__ __launch_bounds__(256) void k_direct64(int n, const double* __restrict__ pts, int m,
                                                  const double* __restrict__ rec, int chunk,
                                                  const int64_t* __restrict__ ex_ptr,
                                                  const int64_t* __restrict__ ex_seg, int raise_singular,
                                                  double* __restrict__ part, int32_t* __restrict__ pflag) {
  const int pa = blockIdx.x * (2 * blockDim.x) + threadIdx.x;
  const int pb = pa + blockDim.x;
  const int c = blockIdx.y;
  const int s0 = c * chunk;
  const int s1 = min(m, s0 + chunk);
  const bool act_a = pa < n, act_b = pb < n;
  double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
  int64_t ka = 0, stop_a = 0, kb = 0, stop_b = 0;
  int next_a = INT_MAX, next_b = INT_MAX;
  if (act_a) {
    ax = pts[3 * (size_t)pa];
    ay = pts[3 * (size_t)pa + 1];
    az = pts[3 * (size_t)pa + 2];
    ka = ex_ptr[pa];
    stop_a = ex_ptr[pa + 1];
    while (ka < stop_a && ex_seg[ka] < s0) ka++;
    next_a = ka < stop_a ? (int)ex_seg[ka] : INT_MAX;
  }
  if (act_b) {
    bx = pts[3 * (size_t)pb];
    by = pts[3 * (size_t)pb + 1];
    bz = pts[3 * (size_t)pb + 2];
    kb = ex_ptr[pb];
    stop_b = ex_ptr[pb + 1];
    while (kb < stop_b && ex_seg[kb] < s0) kb++;
    next_b = kb < stop_b ? (int)ex_seg[kb] : INT_MAX;
  }
  double uax = 0.0, uay = 0.0, uaz = 0.0, ubx = 0.0, uby = 0.0, ubz = 0.0;
  FlagState sta{OFW_FLAG_OK, INT_MAX}, stb{OFW_FLAG_OK, INT_MAX};
  for (int s = s0; s < s1; ++s) {
    const double* r = rec + (size_t)s * REC64;
    const bool hit_a = s == next_a, hit_b = s == next_b;
    if (__builtin_amdgcn_ballot_w64(hit_a || hit_b) != 0) {
      if (hit_a) {
        ++ka;
        next_a = ka < stop_a ? (int)ex_seg[ka] : INT_MAX;
      }
      if (hit_b) {
        ++kb;
        next_b = kb < stop_b ? (int)ex_seg[kb] : INT_MAX;
      }
    }
    const int fa = pair_fast64<MODEL>(ax, ay, az, r, raise_singular, act_a && !hit_a, uax, uay, uaz);
    const int fb = pair_fast64<MODEL>(bx, by, bz, r, raise_singular, act_b && !hit_b, ubx, uby, ubz);
    if (__builtin_amdgcn_ballot_w64(fa != OFW_FLAG_OK || fb != OFW_FLAG_OK) != 0) {
      note_flag(sta, fa, s);
      note_flag(stb, fb, s);
    }
  }
  if (act_a) {
    const size_t o = (size_t)c * n + pa;
    part[3 * o] = uax;
    part[3 * o + 1] = uay;
    part[3 * o + 2] = uaz;
    pflag[o] = flag_code(sta);
  }
  if (act_b) {
    const size_t o = (size_t)c * n + pb;
    part[3 * o] = ubx;
    part[3 * o + 1] = uby;
    part[3 * o + 2] = ubz;
    pflag[o] = flag_code(stb);
  }
}

// Mixed / fp32: FP32 pair arithmetic on the target-minus-start vector formed in FP64; near pairs in FP64.
template <int MODEL, bool WIDE>
__global__ __launch_bounds__(256) void k_direct_mixed(int n, const double* __restrict__ pts, int m,
                                                      const double* __restrict__ rec,
                                                      const float* __restrict__ rec32, int chunk,
                                                      const int64_t* __restrict__ ex_ptr,
                                                      const int64_t* __restrict__ ex_seg, int raise_singular,
                                                      double* __restrict__ part, int32_t* __restrict__ pflag) {
  typedef typename std::conditional<WIDE, double, float>::type acc_t;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  const int s0 = c * chunk;
  const int s1 = min(m, s0 + chunk);
  const bool active = p < n;
  double px = 0.0, py = 0.0, pz = 0.0;
  int64_t k = 0, stop = 0;
  int next = INT_MAX;
  if (active) {
    px = pts[3 * (size_t)p];
    py = pts[3 * (size_t)p + 1];
    pz = pts[3 * (size_t)p + 2];
    k = ex_ptr[p];
    stop = ex_ptr[p + 1];
    while (k < stop && ex_seg[k] < s0) k++;
    next = k < stop ? (int)ex_seg[k] : INT_MAX;
  }
  acc_t ux = 0, uy = 0, uz = 0;
  FlagState st{OFW_FLAG_OK, INT_MAX};
  for (int s = s0; s < s1; ++s) {
    const double* r = rec + (size_t)s * REC64;
    const float* f = rec32 + (size_t)s * REC32;
    if (s == next) {
      ++k;
      next = k < stop ? (int)ex_seg[k] : INT_MAX;
      continue;
    }
    const float rx = (float)(px - r[R_SX]), ry = (float)(py - r[R_SY]), rz = (float)(pz - r[R_SZ]);
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (pair_f32<MODEL>(rx, ry, rz, load_rec32(f), vx, vy, vz)) {
      double wx = 0.0, wy = 0.0, wz = 0.0;
      const int fl = pair_fast64<MODEL>(px, py, pz, r, raise_singular, wx, wy, wz);
      note_flag(st, fl, s);
      ux += (acc_t)wx;
      uy += (acc_t)wy;
      uz += (acc_t)wz;
    } else {
      ux += (acc_t)vx;
      uy += (acc_t)vy;
      uz += (acc_t)vz;
    }
  }
  if (active) {
    const size_t o = (size_t)c * n + p;
    part[3 * o] = (double)ux;
    part[3 * o + 1] = (double)uy;
    part[3 * o + 2] = (double)uz;
    pflag[o] = flag_code(st);
  }
}

// Sum of the chunk partials of every target; first flag in segment order; the reference's final non-finite
// check (vortex_numba.py:142-143).  LANES consecutive lanes share a target: lane j adds the chunks j, j + LANES,
// j + 2 LANES, ... in order, then the lanes' sums are combined by a butterfly over the group (lane j adds lane
// j ^ o for o = LANES/2 ... 1).  Both partners of a butterfly step form the same sum, so every lane of the group
// ends with the same bits, and the association is fixed by (chunks, LANES): the result is deterministic.  The
// first flag is the smallest flag code (seg * 4 + flag), i.e. the flagged pair with the lowest segment index.
template <bool WIDE, int LANES>
__global__ __launch_bounds__(256) void k_direct_reduce(int n, int chunks, const double* __restrict__ part,
                                                       const int32_t* __restrict__ pflag, double* __restrict__ out,
                                                       int8_t* __restrict__ flags, int32_t* __restrict__ flag_seg) {
  static_assert(LANES >= 1 && LANES <= 64 && (64 % LANES) == 0, "a target's lanes share one wavefront");
  typedef typename std::conditional<WIDE, double, float>::type acc_t;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = g / LANES;
  const int j = g % LANES;
  acc_t ux = 0, uy = 0, uz = 0;
  int code = INT_MAX;
  if (p < n) {
    for (int c = j; c < chunks; c += LANES) {
      const size_t o = (size_t)c * n + p;
      ux += (acc_t)part[3 * o];
      uy += (acc_t)part[3 * o + 1];
      uz += (acc_t)part[3 * o + 2];
      const int pc = pflag[o];
      code = (pc >= 0 && pc < code) ? pc : code;
    }
  }
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) {
    ux += __shfl_xor(ux, o);
    uy += __shfl_xor(uy, o);
    uz += __shfl_xor(uz, o);
    code = min(code, __shfl_xor(code, o));
  }
  if (p >= n || j != 0) return;
  int flag = code == INT_MAX ? OFW_FLAG_OK : (code & 3);
  int seg = code == INT_MAX ? -1 : (code >> 2);
  if (!(isfinite((double)ux) && isfinite((double)uy) && isfinite((double)uz)) && flag == OFW_FLAG_OK)
    flag = OFW_FLAG_NONFINITE;
  out[3 * (size_t)p] = (double)ux;
  out[3 * (size_t)p + 1] = (double)uy;
  out[3 * (size_t)p + 2] = (double)uz;
  flags[p] = (int8_t)flag;
  if (flag_seg) flag_seg[p] = seg;
}

template <int MODEL>
static void launch_direct(int precision, dim3 grid, dim3 block, hipStream_t st, int n, const double* pts, int m,
                          const double* rec64, const float* rec32, int chunk, const int64_t* ex_ptr,
                          const int64_t* ex_seg, int raise, double* part, int32_t* pflag) {
  if (precision == OFW_FP64)
    hipLaunchKernelGGL(k_direct64<MODEL>, grid, block, 0, st, n, pts, m, rec64, chunk, ex_ptr, ex_seg, raise, part,
                       pflag);
  else if (precision == OFW_MIXED)
    hipLaunchKernelGGL((k_direct_mixed<MODEL, true>), grid, block, 0, st, n, pts, m, rec64, rec32, chunk, ex_ptr,
                       ex_seg, raise, part, pflag);
  else
    hipLaunchKernelGGL((k_direct_mixed<MODEL, false>), grid, block, 0, st, n, pts, m, rec64, rec32, chunk, ex_ptr,
                       ex_seg, raise, part, pflag);
}

template <bool WIDE>
static void launch_reduce(int lanes, hipStream_t st, int n, int chunks, const double* part, const int32_t* pflag,
                          double* out, int8_t* flags, int32_t* flag_seg) {
  const dim3 grid((unsigned)(((int64_t)n * lanes + 255) / 256));
  switch (lanes) {
    case 1: hipLaunchKernelGGL((k_direct_reduce<WIDE, 1>), grid, dim3(256), 0, st, n, chunks, part, pflag, out, flags, flag_seg); break;
    case 4: hipLaunchKernelGGL((k_direct_reduce<WIDE, 4>), grid, dim3(256), 0, st, n, chunks, part, pflag, out, flags, flag_seg); break;
    case 16: hipLaunchKernelGGL((k_direct_reduce<WIDE, 16>), grid, dim3(256), 0, st, n, chunks, part, pflag, out, flags, flag_seg); break;
    default: hipLaunchKernelGGL((k_direct_reduce<WIDE, 64>), grid, dim3(256), 0, st, n, chunks, part, pflag, out, flags, flag_seg); break;
  }
}

DirectPlan direct_plan(int cu_count, int precision, int block, int64_t n, int64_t m, int64_t max_chunks) {
  DirectPlan pl{};
  // A workgroup wider than the targets only adds idle lanes: halve it (down to one wavefront) while the
  // targets still fit, so a call with few targets spends its lanes on more segment chunks instead.
  const int per_lane = precision == OFW_FP64 ? 2 : 1;
  int eff = block;
  while (eff > 64 && (int64_t)per_lane * (eff / 2) >= n) eff /= 2;
  pl.block = eff;
  const int64_t per_block = (int64_t)per_lane * eff;
  pl.bx = (n + per_block - 1) / per_block;
  const int64_t waves_per_block = (eff + 63) / 64;
  // enough wavefronts for every SIMD of every CU to hold several (MI355X: 256 CUs x 4 SIMDs)
  const int64_t want_waves = (int64_t)cu_count * 4 * 8;
  int64_t chunks = (want_waves + pl.bx * waves_per_block - 1) / (pl.bx * waves_per_block);
  const int64_t min_chunk = 64;
  const int64_t most = (m + min_chunk - 1) / min_chunk;
  if (chunks > most) chunks = most;
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  if (chunks > 65535) chunks = 65535;
  int64_t chunk = (m + chunks - 1) / chunks;
  if (chunk < 1) chunk = 1;
  pl.chunk = chunk;
  pl.chunks = m > 0 ? (m + chunk - 1) / chunk : 0;
  pl.lanes = pl.chunks <= 8 ? 1 : pl.chunks <= 32 ? 4 : pl.chunks <= 128 ? 16 : 64;
  return pl;
}

int direct_run(Ctx& c, const DirectPlan& pl, int precision, int64_t n, const double* points, int64_t m,
               const double* rec64, const float* rec32, int model, int raise_singular, const int64_t* ex_ptr,
               const int64_t* ex_seg, double* part, int32_t* pflag, double* out, int8_t* flags, int32_t* flag_seg) {
  if (n == 0) return OFW_OK;
  if (pl.chunks > 0) {
    dim3 grid((unsigned)pl.bx, (unsigned)pl.chunks);
    dim3 blk((unsigned)pl.block);
    const int chunk = (int)pl.chunk;
    switch (model) {
      case 0: launch_direct<0>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
      case 1: launch_direct<1>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
      case 2: launch_direct<2>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
      case 3: launch_direct<3>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
      case 4: launch_direct<4>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
      default: launch_direct<5>(precision, grid, blk, c.stream, (int)n, points, (int)m, rec64, rec32, chunk, ex_ptr, ex_seg, raise_singular, part, pflag); break;
    }
    OFW_HIP(hipGetLastError());
  }
  OFW_HIP(c.mark());
  if (precision == OFW_FP32)
    launch_reduce<false>(pl.lanes, c.stream, (int)n, (int)pl.chunks, part, pflag, out, flags, flag_seg);
  else
    launch_reduce<true>(pl.lanes, c.stream, (int)n, (int)pl.chunks, part, pflag, out, flags, flag_seg);
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  c.stats_i[0] = pl.chunks;
  c.stats_i[1] = pl.chunk;
  return OFW_OK;
}

int direct_sum(Ctx& c, int precision, int block, int64_t n, const double* points, int64_t m, const double* rec64,
               const float* rec32, int model, int raise_singular, const int64_t* ex_ptr, const int64_t* ex_seg,
               double* out, int8_t* flags, int32_t* flag_seg) {
  if (n == 0) return OFW_OK;
  const DirectPlan pl = direct_plan(c.cu_count, precision, block, n, m, INT64_MAX);
  const int64_t slots = pl.chunks > 0 ? pl.chunks : 1;
  OFW_HIP(c.part.reserve((size_t)slots * n * 3 * sizeof(double)));
  OFW_HIP(c.pflag.reserve((size_t)slots * n * sizeof(int32_t)));
  return direct_run(c, pl, precision, n, points, m, rec64, rec32, model, raise_singular, ex_ptr, ex_seg,
                    (double*)c.part.ptr, (int32_t*)c.pflag.ptr, out, flags, flag_seg);
}

}  // namespace ofw
