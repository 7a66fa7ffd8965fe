// C ABI of libofw_hip.so: contexts, errors and the entry points declared in include/ofw_hip.h.
#include <cstring>
#include <mutex>
#include <string>

#include "common.h"
#include "internal.h"

namespace ofw {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

hipError_t DevBuf::reserve(size_t need) {
  if (need == 0) need = 256;
  if (bytes >= need) return hipSuccess;
  if (ptr) {
    hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return e;
    ptr = nullptr;
    bytes = 0;
  }
  size_t grow = need + need / 4;  // headroom so that slightly larger calls do not reallocate
  hipError_t e = hipMalloc(&ptr, grow);
  if (e != hipSuccess) {
    ptr = nullptr;
    return e;
  }
  bytes = grow;
  return hipSuccess;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

// Scratch of ofw_direct_sum_static: a fixed device array, one copy per GPU, that lives until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: its device address is resolved once per
// device (with that device current) by hipGetSymbolAddress, and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[OFW_STATIC_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device.
static int static_scratch(int device, void** ptr_out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("static scratch: device index out of range");
    return OFW_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    OFW_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = ptr;
  }
  *ptr_out = g_scratch_ptr[device];
  return OFW_OK;
}

}  // namespace ofw

using ofw::set_error;

extern "C" {

int ofw_abi_version(void) { return OFW_ABI_VERSION; }

const char* ofw_last_error(void) { return ofw::g_error.c_str(); }

int ofw_ctx_create(ofw_ctx** out, void* stream) {
  if (!out) {
    set_error("ofw_ctx_create: null output pointer");
    return OFW_ERR_ARG;
  }
  ofw_ctx* x = new ofw_ctx();
  x->c.stream = (hipStream_t)stream;
  OFW_HIP(hipGetDevice(&x->c.device));
  int cus = 0;
  OFW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, x->c.device));
  x->c.cu_count = cus > 0 ? cus : 256;
  for (auto& e : x->c.ev) OFW_HIP(hipEventCreate(&e));
  // set-up of the no-allocation entry point: this device's copy of the static scratch
  int st = ofw::static_scratch(x->c.device, &x->c.static_scratch);
  if (st) return st;
  *out = x;
  return OFW_OK;
}

int ofw_ctx_destroy(ofw_ctx* x) {
  if (!x) return OFW_OK;
  ofw::Ctx& c = x->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  else (void)hipDeviceSynchronize();
  c.rec64.release();
  c.rec32.release();
  c.part.release();
  c.pflag.release();
  c.expos.release();
  c.exptr32.release();
  c.exseg32.release();
  c.vjp.release();
  c.release_side();
  if (c.host_counters) (void)hipHostFree(c.host_counters);
  c.host_counters = nullptr;
  for (auto& b : c.tree) b.release();
  for (auto& e : c.ev)
    if (e) (void)hipEventDestroy(e);
  delete x;
  return OFW_OK;
}

int ofw_ctx_set_stream(ofw_ctx* x, void* stream) {
  if (!x) {
    set_error("null context");
    return OFW_ERR_ARG;
  }
  x->c.stream = (hipStream_t)stream;
  return OFW_OK;
}

int ofw_ctx_timings(ofw_ctx* x, float* ms_host, int count) {
  if (!x || !ms_host) {
    set_error("ofw_ctx_timings: null argument");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  int phases = c.nev > 1 ? c.nev - 1 : 0;
  for (int k = 0; k < count; ++k) {
    ms_host[k] = 0.0f;
    if (k < phases) OFW_HIP(hipEventElapsedTime(&ms_host[k], c.ev[k], c.ev[k + 1]));
  }
  return phases;
}

int ofw_ctx_stats(ofw_ctx* x, int64_t* stats_host, int count) {
  if (!x || !stats_host) {
    set_error("ofw_ctx_stats: null argument");
    return OFW_ERR_ARG;
  }
  int k = 0;
  for (; k < count && k < 8; ++k) stats_host[k] = x->c.stats_i[k];
  return k;
}

int ofw_device_info(char* name_host, int name_len, int* cu_count_host) {
  int dev = 0;
  OFW_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  OFW_HIP(hipGetDeviceProperties(&prop, dev));
  if (name_host && name_len > 0) {
    std::strncpy(name_host, prop.gcnArchName, (size_t)name_len - 1);
    name_host[name_len - 1] = 0;
  }
  if (cu_count_host) *cu_count_host = prop.multiProcessorCount;
  return OFW_OK;
}

static int direct_args(ofw_ctx* x, const char* what, int precision, int block, int64_t n, int64_t m, int model) {
  if (!x) {
    set_error("null context");
    return OFW_ERR_ARG;
  }
  if (precision < 0 || precision > 2 || model < 0 || model > 5 || n < 0 || m < 0 ||
      !(block == 32 || block == 64 || block == 128 || block == 256)) {
    set_error(std::string(what) + ": invalid precision, model, block or size");
    return OFW_ERR_ARG;
  }
  if (n > (int64_t)1 << 30 || m > (int64_t)1 << 28) {
    set_error(std::string(what) + ": too many targets or segments for one call");
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

int ofw_direct_sum(ofw_ctx* x, int precision, int block, int64_t n, const double* points, int64_t m,
                   const double* starts, const double* directions, const double* lengths, const double* gamma,
                   const double* radius, int model, int raise_singular, const int64_t* ex_ptr,
                   const int64_t* ex_seg, double* out, int8_t* flags, int32_t* flag_seg) {
  int st = direct_args(x, "ofw_direct_sum", precision, block, n, m, model);
  if (st) return st;
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0) return OFW_OK;
  OFW_HIP(c.mark());
  OFW_HIP(c.rec64.reserve((size_t)(m > 0 ? m : 1) * REC64 * sizeof(double)));
  OFW_HIP(c.rec32.reserve((size_t)(m > 0 ? m : 1) * REC32 * sizeof(float)));
  st = ofw::pack_segments(c, m, starts, directions, lengths, gamma, radius, model, (double*)c.rec64.ptr,
                          (float*)c.rec32.ptr);
  if (st) return st;
  OFW_HIP(c.mark());
  st = ofw::direct_sum(c, precision, block, n, points, m, (const double*)c.rec64.ptr, (const float*)c.rec32.ptr,
                       model, raise_singular, ex_ptr, ex_seg, out, flags, flag_seg);
  if (st) return st;
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

// ofw_direct_sum on the static scratch: the same kernels and launch plan, with the records and the chunk
// partials carved from this device's copy of `scratch` (its resolved address) instead of the context's buffers.
int ofw_direct_sum_static(ofw_ctx* x, int precision, int block, int64_t n, const double* points, int64_t m,
                          const double* starts, const double* directions, const double* lengths,
                          const double* gamma, const double* radius, int model, int raise_singular,
                          const int64_t* ex_ptr, const int64_t* ex_seg, double* out, int8_t* flags,
                          int32_t* flag_seg) {
  int st = direct_args(x, "ofw_direct_sum_static", precision, block, n, m, model);
  if (st) return st;
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0) return OFW_OK;
  if (!c.static_scratch) {
    set_error("ofw_direct_sum_static: no static scratch on this device");
    return OFW_ERR_HIP;
  }
  auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
  const int64_t records = m > 0 ? m : 1;
  const int64_t rec64_bytes = up(records * REC64 * (int64_t)sizeof(double));
  const int64_t rec32_bytes = up(records * REC32 * (int64_t)sizeof(float));
  const int64_t slot_bytes = n * (3 * (int64_t)sizeof(double) + (int64_t)sizeof(int32_t));  // one chunk's partials
  const int64_t fixed = rec64_bytes + rec32_bytes + 256;
  const int64_t room = OFW_STATIC_SCRATCH_BYTES - fixed;
  if (room < slot_bytes) {
    set_error("ofw_direct_sum_static: " + std::to_string(n) + " targets x " + std::to_string(m) + " segments need " +
              std::to_string(fixed + slot_bytes) + " bytes of scratch, more than the " +
              std::to_string((int64_t)OFW_STATIC_SCRATCH_BYTES) + " of the static scratch");
    return OFW_ERR_ARG;
  }
  const ofw::DirectPlan pl = ofw::direct_plan(c.cu_count, precision, block, n, m, room / slot_bytes);
  char* base = (char*)c.static_scratch;
  double* rec64 = (double*)base;
  float* rec32 = (float*)(base + rec64_bytes);
  double* part = (double*)(base + rec64_bytes + rec32_bytes);
  const int64_t slots = pl.chunks > 0 ? pl.chunks : 1;
  int32_t* pflag = (int32_t*)((char*)part + up(slots * n * 3 * (int64_t)sizeof(double)));
  if ((char*)(pflag + slots * n) > base + OFW_STATIC_SCRATCH_BYTES) {
    set_error("ofw_direct_sum_static: scratch layout exceeds the static scratch");
    return OFW_ERR_ARG;
  }
  OFW_HIP(c.mark());
  st = ofw::pack_segments(c, m, starts, directions, lengths, gamma, radius, model, rec64, rec32);
  if (st) return st;
  OFW_HIP(c.mark());
  st = ofw::direct_run(c, pl, precision, n, points, m, rec64, rec32, model, raise_singular, ex_ptr, ex_seg, part,
                       pflag, out, flags, flag_seg);
  if (st) return st;
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

}  // extern "C"
