// C ABI of libofw_hip.so: contexts, errors and the entry points declared in include/ofw_hip.h.
#include <cstring>
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

}  // extern "C"
