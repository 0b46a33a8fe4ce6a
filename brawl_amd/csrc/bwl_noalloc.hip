// bwl_noalloc.hip -- the parallel anneal without device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving loops
// that forbid it).  Its scratch (round multipliers, per-T outputs, counters,
// launch arguments) lives in one file-scope device array per GPU instead of
// a per-call allocation.  Host code reaches the array only through the
// address hipGetSymbolAddress gives with the device current: on the host the
// array's name is not a device address.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "bwl_ctx.h"

using namespace bwl;

namespace {

// Every call on a GPU shares this array; calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current; looked
// up on the first call per device and kept.
int scratch_ptr(int device, char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(BWL_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BWL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (char *)p;
  }
  *out = g_scratch_ptr[device];
  return BWL_OK;
}

}  // namespace

extern "C" {

int bwl_metropolis_anneal_noalloc(bwl_ctx *c, const bwl_anneal_params *p, double *energies_of_T,
                                  double *C_of_T, double *acceptance_of_T, bwl_run_stats *stats) {
  if (!c || !p || !energies_of_T || !C_of_T || !acceptance_of_T)
    return fail(BWL_ERR_ARG, "null argument");
  bwl_anneal_params q = *p;
  q.asro_every = 0;
  bwl_ctx *cs[1] = {c};
  AnnealPlan P;
  int rc = anneal_plan(1, cs, &q, P);  // makes the context's device current
  if (rc) return rc;
  if (P.scratch_bytes > kNoAllocScratchBytes)
    return fail(BWL_ERR_ARG, "the anneal needs " + std::to_string(P.scratch_bytes) +
                                 " bytes of scratch, more than the " +
                                 std::to_string(kNoAllocScratchBytes) +
                                 " of the no-allocation entry point (fewer T steps fit)");
  char *base = nullptr;
  rc = scratch_ptr(c->device, &base);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev0, c->stream));
  rc = anneal_upload(P, base, c->stream, true);  // counters cleared by a kernel
  if (rc) return rc;
  rc = anneal_launch(P, &c->stream);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev1, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  double *e[1] = {energies_of_T}, *cv[1] = {C_of_T}, *ac[1] = {acceptance_of_T};
  return anneal_collect(P, cs, e, cv, ac, stats, ms, c->stream);
}

}  // extern "C"
