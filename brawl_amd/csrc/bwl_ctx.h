// bwl_ctx.h -- the context behind the C-ABI's opaque bwl_ctx handle, and
// the error plumbing shared by the C-ABI source files.
#pragma once

#include <hip/hip_runtime.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/brawl_hip.h"
#include "bwl_host.h"
#include "bwl_kernels.h"

namespace bwl {
// message of the last failure on this thread (bwl_last_error)
extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace bwl

#define HIP_OK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bwl::fail(BWL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct bwl_ctx {
  bwl_lattice_desc desc;
  bwl::Geo g;
  std::vector<int32_t> off;
  std::vector<double> vex;
  hipStream_t stream = nullptr;
  int device = 0;
  int8_t *d_compact = nullptr;
  int8_t *d_ref = nullptr;
  int32_t *d_off = nullptr;
  double *d_vex = nullptr;
  void *d_scratch = nullptr;  // small per-call outputs
  size_t scratch_bytes = 0;
  bool has_config = false;
  bwl::MT19937 mt;
  unsigned long long phase = 0;  // parallel-sampler round counter
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace bwl {

inline int ensure_scratch(bwl_ctx *c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return BWL_OK;
  if (c->d_scratch) HIP_OK(hipFree(c->d_scratch));
  c->d_scratch = nullptr;
  c->scratch_bytes = 0;
  if (hipMalloc(&c->d_scratch, bytes) != hipSuccess) return fail(BWL_ERR_NOMEM, "device allocation failed");
  c->scratch_bytes = bytes;
  return BWL_OK;
}

inline int check_ctx(bwl_ctx *c, bool need_config) {
  if (!c) return fail(BWL_ERR_ARG, "null context");
  if (need_config && !c->has_config) return fail(BWL_ERR_STATE, "no configuration uploaded");
  HIP_OK(hipSetDevice(c->device));
  return BWL_OK;
}

inline int lds_limit() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kLdsBytes;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && v > 0)
    return v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
    return v;
  return kLdsBytes;
}

// Device buffer freed at scope exit.
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace bwl
