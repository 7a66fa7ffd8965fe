// Workgroup-wide exclusive scan shared by the build, moment and traversal planning kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace ofw {

// Exclusive scan of one int per thread over the workgroup (blockDim a multiple of 64, at most 1024); every
// thread of the workgroup must call it.  *total receives the workgroup's sum.
__device__ inline int block_exclusive_scan(int v, int* total) {
  __shared__ int wave_sums[17];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wave_sums[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < nw; ++k) {
      const int s = wave_sums[k];
      wave_sums[k] = acc;
      acc += s;
    }
    wave_sums[16] = acc;
  }
  __syncthreads();
  const int result = x - v + wave_sums[w];
  *total = wave_sums[16];
  __syncthreads();  // the next call may overwrite wave_sums
  return result;
}

// Exclusive scans of N ints per thread over the workgroup together, for the barriers of one; every thread of
// the workgroup must call it.  total[q] receives the workgroup's sum of value q.
template <int N>
__device__ inline void block_exclusive_scan_n(const int (&v)[N], int (&r)[N], int (&total)[N]) {
  __shared__ int wave_sums[N][17];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  int x[N];
#pragma unroll
  for (int q = 0; q < N; ++q) x[q] = v[q];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const int y = __shfl_up(x[q], off, 64);
      if (lane >= off) x[q] += y;
    }
  if (lane == 63)
#pragma unroll
    for (int q = 0; q < N; ++q) wave_sums[q][w] = x[q];
  __syncthreads();
  if ((int)threadIdx.x < N) {  // thread q scans the wavefront sums of value q
    const int q = threadIdx.x;
    int acc = 0;
    for (int k = 0; k < nw; ++k) {
      const int s = wave_sums[q][k];
      wave_sums[q][k] = acc;
      acc += s;
    }
    wave_sums[q][16] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    r[q] = x[q] - v[q] + wave_sums[q][w];
    total[q] = wave_sums[q][16];
  }
  __syncthreads();  // the next call may overwrite wave_sums
}

// true when the predicate holds on any / every active lane of the wavefront
__device__ __forceinline__ bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }
__device__ __forceinline__ bool wave_all(bool v) {
  return __builtin_amdgcn_ballot_w64(v) == __builtin_amdgcn_read_exec();
}

// LDS written by some lanes of a wavefront and read by others: the wavefront's LDS operations execute in
// order, so only the compiler must keep them in program order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- single-pass prefix over ordered work units (decoupled look-back)
// Each unit publishes a 64-bit status word: its aggregate first, then its inclusive prefix.  Word layout: bits
// 63..33 the caller's tag (words of another tag read as not ready, so the status array is cleared once and
// reused with a new tag), bit 32 the inclusive flag, bits 31..0 the value (non-negative).  Units take their
// indices from an atomic ticket in launch order, so every unit a look-back waits for was claimed earlier by a
// running workgroup and finishes without waiting on a later one.
__device__ __forceinline__ void lb_publish(unsigned long long* word, unsigned tag, bool inclusive, int value) {
  const unsigned long long w = ((unsigned long long)tag << 33) | ((unsigned long long)(inclusive ? 1 : 0) << 32) |
                               (unsigned long long)(uint32_t)value;
  __hip_atomic_store(word, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static constexpr int LB_SPIN_LIMIT = 1 << 20;  // polls of one window before the look-back gives up

// Exclusive prefix of unit i over the units [first, i) whose status words are status[stride * u]; called by a
// whole wavefront, the result on every lane.  Windows of 64 predecessors are read together; a window is used
// once every unit back to its nearest inclusive word is ready.  *ok turns false (and the prefix is not the
// sum) only when a window was polled LB_SPIN_LIMIT times: every wait ends.
__device__ inline int lb_lookback(const unsigned long long* status, int stride, int first, int i, unsigned tag,
                                  bool* ok) {
  const int lane = threadIdx.x & 63;
  int prefix = 0;
  int j = i - 1;
  int spins = 0;
  while (j >= first) {
    const int u = j - lane;
    const bool in = u >= first;
    unsigned long long w = 0;
    if (in) w = __hip_atomic_load(status + (size_t)stride * u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool ready = in && (unsigned)(w >> 33) == tag;
    const unsigned long long incl = __builtin_amdgcn_ballot_w64(ready && ((w >> 32) & 1ull));
    const unsigned long long inmask = __builtin_amdgcn_ballot_w64(in);
    // the lanes needed: up to the nearest inclusive word, or the whole window
    const unsigned long long need =
        incl ? (__builtin_ctzll(incl) == 63 ? ~0ull : ((2ull << __builtin_ctzll(incl)) - 1ull)) & inmask : inmask;
    if ((__builtin_amdgcn_ballot_w64(ready) & need) != need) {
      if (++spins >= LB_SPIN_LIMIT) {
        *ok = false;
        return prefix;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    int v = ((need >> lane) & 1ull) ? (int)(uint32_t)w : 0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    prefix += v;
    if (incl) break;
    j -= 64;
  }
  return prefix;
}

// Workgroup index remapped so that each XCD (workgroups are dispatched to the 8 XCDs round-robin) receives a
// contiguous range of logical indices: neighbouring work shares one XCD's L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb >> 3;
  if (b >= (per << 3)) return b;
  return (b & 7) * per + (b >> 3);
}

}  // namespace ofw
