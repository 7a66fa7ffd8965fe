// Host-side internals shared by the translation units of libofw_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ofw_hip.h"

namespace ofw {

void set_error(const std::string& msg);

// A growable device buffer owned by a context.  Every call of the library that needs scratch takes it
// from its context; nothing is kept across calls except the capacity.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t need);
  void release();
};

struct Ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  int cu_count = 256;
  DevBuf rec64, rec32, part, pflag, expos, exptr32, exseg32, vjp;
  DevBuf tree[24];           // tree build / moments / evaluation scratch
  std::vector<char> host;    // host staging
  hipEvent_t ev[12] = {};
  int nev = 0;               // events recorded by the last call (phase k = ev[k] -> ev[k+1])
  int64_t stats_i[8] = {};
  int64_t list_slots = 0;    // interaction-list work-item slots of the next tree call (0: measure)
  int64_t ex_cap = 0;        // exclusion entries a tree call assumes without reading the count (0: read it)
  bool capturing = false;    // enqueuing into a graph capture: no events are recorded
  bool sort_build = false;   // build trees with the per-level sorts (set when the list build met equal coordinates)
  void* static_scratch = nullptr;  // this device's copy of the no-allocation scratch (resolved address)
  // A tree call's second stream: the traversal plan (which needs the tree, not its node quantities) runs on it
  // beside the node quantities; created on first use, default priority.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool fork_pending = false;          // the node quantities record fork_ev once their subtree kernel is enqueued
  bool plan_forked = false;           // the plan is being enqueued on `side`
  hipStream_t plan_main = nullptr;    // the call's stream meanwhile (the null stream is a valid one)
  int32_t* host_counters = nullptr;   // pinned host copy of a tree call's counters (allocated on first use)
  hipError_t mark() { return (!capturing && nev < 12) ? hipEventRecord(ev[nev++], stream) : hipSuccess; }
  void release_side() {
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    side = nullptr;
    fork_ev = join_ev = nullptr;
  }
};

#define OFW_HIP(expr)                                                                              \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess) {                                                                        \
      ::ofw::set_error(std::string(#expr " failed: ") + hipGetErrorString(_e));                    \
      return OFW_ERR_HIP;                                                                          \
    }                                                                                              \
  } while (0)

// direct.hip
// Launch shape of a direct sum: workgroup width, target blocks, segment chunks (0 when there are no segments),
// segments per chunk and lanes per target of the chunk reduction.  A function of (CU count, precision, block,
// n, m, chunk cap) only, so a call's association is fixed by its configuration.
struct DirectPlan {
  int block;
  int64_t bx, chunks, chunk;
  int lanes;
};
DirectPlan direct_plan(int cu_count, int precision, int block, int64_t n, int64_t m, int64_t max_chunks);
// the sum with caller-provided partial slots (max(pl.chunks, 1) * n of each); the exclusion CSR is the caller's
// int64 one (ptr (n+1), per-target sorted unique segment ids)
int direct_run(Ctx& c, const DirectPlan& pl, int precision, int64_t n, const double* points, int64_t m,
               const double* rec64, const float* rec32, int model, int raise_singular, const int64_t* ex_ptr,
               const int64_t* ex_seg, double* part, int32_t* pflag, double* out, int8_t* flags, int32_t* flag_seg);
int pack_segments(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths,
                  const double* gamma, const double* radius, int model, double* rec64, float* rec32);
int direct_sum(Ctx& c, int precision, int block, int64_t n, const double* points, int64_t m, const double* rec64,
               const float* rec32, int model, int raise_singular, const int64_t* ex_ptr, const int64_t* ex_seg,
               double* out, int8_t* flags, int32_t* flag_seg);
int csr_to_i32(Ctx& c, int64_t n, const int64_t* ex_ptr, const int64_t* ex_seg, int32_t** ptr32, int32_t** seg32);
// the two halves of csr_to_i32: the entry count (read back, synchronising) and the enqueued conversion
int csr_count(Ctx& c, int64_t n, const int64_t* ex_ptr, int64_t* e_out);
int csr_convert(Ctx& c, int64_t n, int64_t e, const int64_t* ex_ptr, const int64_t* ex_seg, int32_t** ptr32,
                int32_t** seg32);
// csr_convert for at most cap entries, the count read on the device: a larger count converts nothing and is left
// in counters[C_EXNEED] with counters[C_OVERFLOW] set
int csr_convert_checked(Ctx& c, int64_t n, int64_t cap, const int64_t* ex_ptr, const int64_t* ex_seg,
                        int32_t* counters, int32_t** ptr32, int32_t** seg32);

}  // namespace ofw

struct ofw_ctx {
  ofw::Ctx c;
};
