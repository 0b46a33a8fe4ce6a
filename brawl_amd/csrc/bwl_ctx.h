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
  bool borrowed = false;  // device memory and stream belong to someone else
  int32_t layout = 0;     // parallel-sampler layout: 0 by rule, else BWL_LAYOUT_*
  int32_t threads = 0;    // parallel-sampler threads per chain: 0 by rule
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

// A context whose device buffers live in caller-provided memory (`bytes` >=
// ctx_device_bytes) and whose uploads go on `stream`; bwl_destroy frees
// neither.
size_t ctx_device_bytes(const bwl_lattice_desc &d);
int ctx_create_borrowed(const bwl_lattice_desc &d, const double *v_ex, char *mem, size_t bytes,
                        hipStream_t stream, bwl_ctx **out);

// A batch of anneals prepared on the host (anneal_plan), its arguments and
// buffers placed in caller-provided device memory (anneal_upload), its
// launches enqueued one group per stream (anneal_launch), and its outputs
// copied out after completion (anneal_collect).
struct AnnealPlan {
  struct Item {
    int idx = 0;
    const void *fn = nullptr;
    int threads = 0;
    size_t lds = 0;
    ParArgs a;
    ParPlan plan;
    size_t T = 0, off_muls = 0, off_out = 0, off_cnt = 0, off_rho = 0, rho_n = 0;
    double work = 0.0;
  };
  struct Group {
    const void *fn;
    int threads;
    size_t lds;
    std::vector<int> members;
    size_t off_args;
  };
  std::vector<Item> items;
  std::vector<Group> groups;
  std::vector<std::vector<ParArgs>> host_args;  // kept alive until the uploads finish
  size_t scratch_bytes = 0;
  char *base = nullptr;
};
// The parallel-sampler kernel variant, block size and LDS for a lattice:
// layout 0 takes the first that fits of fixed pitch, own pitch, per
// sublattice; a BWL_LAYOUT_* value takes only that one.  threads_req > 0
// asks for that many threads per chain instead of the rule's count (layout
// 0 then takes the first layout that can run them).
int anneal_kernel(const Geo &g, ParPlan &plan, NatGeo &nat, const void *&fn, int &threads,
                  size_t &lds, bool *natural = nullptr, int layout = 0, int threads_req = 0);
int anneal_plan(int n, bwl_ctx *const *ctxs, const bwl_anneal_params *params, AnnealPlan &P);
// zero_by_kernel: clear the counters with a kernel instead of a memset
int anneal_upload(AnnealPlan &P, char *scratch_dev, hipStream_t s, bool zero_by_kernel = false);
int anneal_launch(const AnnealPlan &P, const hipStream_t *streams);
int anneal_collect(const AnnealPlan &P, bwl_ctx *const *ctxs, double *const *energies_of_T,
                   double *const *C_of_T, double *const *acceptance_of_T, bwl_run_stats *stats,
                   double kernel_ms, hipStream_t s, double *const *rho_of_T = nullptr);

// Device buffer freed at scope exit.
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace bwl
