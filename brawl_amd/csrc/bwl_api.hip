// bwl_api.hip -- the C-ABI (include/brawl_hip.h) over the gfx950 kernels.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bwl_ctx.h"

using namespace bwl;

namespace bwl {
thread_local std::string g_err;
}

extern "C" {

const char *bwl_last_error(void) { return g_err.c_str(); }

int bwl_device_count(int32_t *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (count) *count = n;
  return e == hipSuccess ? BWL_OK : fail(BWL_ERR_DEVICE, hipGetErrorString(e));
}

int bwl_lds_bandwidth(int32_t iters, double *bytes_per_s, int32_t *cus) {
  if (iters < 1 || !bytes_per_s || !cus) return fail(BWL_ERR_ARG, "bad argument");
  int dev = 0, n_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      n_cu < 1)
    return fail(BWL_ERR_DEVICE, "no HIP device");
  // two 64 KiB, 1024-thread blocks per CU run at once; 8 blocks per CU in all
  const int blocks = 8 * n_cu;
  DevBuf sink;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ok = hipMalloc(&sink.p, sizeof(uint32_t) * blocks) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
            launch_lds_bandwidth(blocks, 16, (uint32_t *)sink.p, 0) == hipSuccess &&  // warm-up
            hipEventRecord(e0, 0) == hipSuccess &&
            launch_lds_bandwidth(blocks, iters, (uint32_t *)sink.p, 0) == hipSuccess &&
            hipEventRecord(e1, 0) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
  float ms = 0.0f;
  if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.0f;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!ok) return fail(BWL_ERR_DEVICE, "LDS bandwidth probe failed");
  *bytes_per_s = (double)blocks * 1024.0 * (double)iters * 16.0 / ((double)ms * 1e-3);
  *cus = n_cu;
  return BWL_OK;
}

int bwl_create(const bwl_lattice_desc *desc, const double *v_ex, bwl_ctx **out) {
  if (!desc || !v_ex || !out) return fail(BWL_ERR_ARG, "null argument");
  *out = nullptr;
  bwl_ctx *c = new bwl_ctx();
  c->desc = *desc;
  int rc = make_geo(*desc, c->g, c->off, g_err);
  if (rc) {
    delete c;
    return rc;
  }
  const Geo &g = c->g;
  c->vex.assign(v_ex, v_ex + (size_t)g.S * g.S * g.R);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    delete c;
    return fail(BWL_ERR_DEVICE, "no HIP device available");
  }
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  size_t ref_bytes = (size_t)g.m[0] * g.m[1] * g.m[2];
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&c->d_compact, g.compact_bytes) == hipSuccess &&
            hipMalloc(&c->d_ref, ref_bytes) == hipSuccess &&
            hipMalloc(&c->d_off, sizeof(int32_t) * c->off.size()) == hipSuccess &&
            hipMalloc(&c->d_vex, sizeof(double) * c->vex.size()) == hipSuccess &&
            hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
  if (!ok) {
    bwl_destroy(c);
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  }
  if (hipMemcpy(c->d_off, c->off.data(), sizeof(int32_t) * c->off.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->d_vex, c->vex.data(), sizeof(double) * c->vex.size(), hipMemcpyHostToDevice) !=
          hipSuccess) {
    bwl_destroy(c);
    return fail(BWL_ERR_DEVICE, "upload failed");
  }
  c->mt.seed(110179u);  // f90_init_genrand static seed, rank 0 (mt19937ar.c:138)
  *out = c;
  return BWL_OK;
}

int bwl_destroy(bwl_ctx *c) {
  if (!c) return BWL_OK;
  (void)hipSetDevice(c->device);
  if (c->stream && !c->borrowed) (void)hipStreamSynchronize(c->stream);
  if (!c->borrowed) {
    (void)hipFree(c->d_compact);
    (void)hipFree(c->d_ref);
    (void)hipFree(c->d_off);
    (void)hipFree(c->d_vex);
  }
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream && !c->borrowed) (void)hipStreamDestroy(c->stream);
  delete c;
  return BWL_OK;
}

int bwl_n_atoms(const bwl_ctx *c, int32_t *n_atoms) {
  if (!c || !n_atoms) return fail(BWL_ERR_ARG, "null argument");
  *n_atoms = c->g.n_atoms;
  return BWL_OK;
}

int bwl_set_config(bwl_ctx *c, const int8_t *config) {
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!config) return fail(BWL_ERR_ARG, "null config");
  rc = check_config(c->g, config, g_err);
  if (rc) return rc;
  size_t ref_bytes = (size_t)c->g.m[0] * c->g.m[1] * c->g.m[2];
  HIP_OK(hipMemcpyAsync(c->d_ref, config, ref_bytes, hipMemcpyHostToDevice, c->stream));
  HIP_OK(launch_pack(c->d_ref, c->d_compact, c->g, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->has_config = true;
  return BWL_OK;
}

int bwl_get_config(bwl_ctx *c, int8_t *config) {
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!config) return fail(BWL_ERR_ARG, "null config");
  size_t ref_bytes = (size_t)c->g.m[0] * c->g.m[1] * c->g.m[2];
  HIP_OK(launch_unpack(c->d_compact, c->d_ref, c->g, c->stream));
  HIP_OK(hipMemcpyAsync(config, c->d_ref, ref_bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return BWL_OK;
}

int bwl_pair_histogram(bwl_ctx *c, int64_t *hist) {
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!hist) return fail(BWL_ERR_ARG, "null hist");
  const Geo &g = c->g;
  const int threads = 512;
  if (hist_lds_bytes(g, threads) > (size_t)lds_limit())
    return fail(BWL_ERR_UNSUPPORTED, "lattice does not fit in LDS");
  const size_t hbytes = sizeof(long long) * g.S * g.S * g.R;
  rc = ensure_scratch(c, sizeof(HistArgs) + hbytes + 64);
  if (rc) return rc;
  HistArgs a;
  a.g = g;
  a.compact = c->d_compact;
  a.off = c->d_off;
  a.hist = (long long *)((char *)c->d_scratch + ((sizeof(HistArgs) + 63) & ~63));
  HIP_OK(hipMemcpyAsync(c->d_scratch, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemsetAsync(a.hist, 0, hbytes, c->stream));
  HIP_OK(launch_histogram((const HistArgs *)c->d_scratch, 1, g, threads, c->stream));
  HIP_OK(hipMemcpyAsync(hist, a.hist, hbytes, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return BWL_OK;
}

int bwl_total_energy(bwl_ctx *c, double *energy) {
  if (!c || !energy) return fail(BWL_ERR_ARG, "null argument");
  std::vector<int64_t> h((size_t)c->g.S * c->g.S * c->g.R);
  int rc = bwl_pair_histogram(c, h.data());
  if (rc) return rc;
  const int S = c->g.S;
  double e = 0.0;
  for (int n = 0; n < c->g.R; ++n)
    for (int b = 0; b < S; ++b)
      for (int a = 0; a < S; ++a) {
        size_t i = a + (size_t)S * (b + (size_t)S * n);
        e += (double)h[i] * c->vex[i];
      }
  *energy = 0.5 * e;
  return BWL_OK;
}

static int site_pair(bwl_ctx *c, const int32_t *s1, const int32_t *s2, double *out2) {
  int rc = check_ctx(c, true);
  if (rc) return rc;
  const Geo &g = c->g;
  if (site_lds_bytes(g) > (size_t)lds_limit())
    return fail(BWL_ERR_UNSUPPORTED, "lattice does not fit in LDS");
  const int32_t *sites[2] = {s1, s2};
  SiteArgs a;
  a.g = g;
  for (int w = 0; w < 2; ++w) {
    const int32_t *s = sites[w];
    if (s[0] != 1) return fail(BWL_ERR_ARG, "basis index must be 1");
    int x = s[1] - 1, y = s[2] - 1, z = s[3] - 1;
    if (x < 0 || y < 0 || z < 0 || x >= g.m[0] || y >= g.m[1] || z >= g.m[2])
      return fail(BWL_ERR_ARG, "site index out of range");
    bool occ = g.lattice == BWL_FCC ? (((x + y + z) & 1) == 0)
                                    : ((x & 1) == (y & 1) && (y & 1) == (z & 1));
    if (!occ) return fail(BWL_ERR_ARG, "site is not a lattice site");
    a.x[w] = x; a.y[w] = y; a.z[w] = z;
  }
  rc = ensure_scratch(c, 64);
  if (rc) return rc;
  a.compact = c->d_compact;
  a.off = c->d_off;
  a.vex = c->d_vex;
  a.out = (double *)c->d_scratch;
  HIP_OK(launch_site_energy(a, c->stream));
  HIP_OK(hipMemcpyAsync(out2, a.out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return BWL_OK;
}

int bwl_site_energy(bwl_ctx *c, int32_t site_b, int32_t site_i, int32_t site_j, int32_t site_k,
                    double *energy) {
  if (!energy) return fail(BWL_ERR_ARG, "null energy");
  int32_t s[4] = {site_b, site_i, site_j, site_k};
  double o[2];
  int rc = site_pair(c, s, s, o);
  if (rc) return rc;
  *energy = o[0];
  return BWL_OK;
}

int bwl_pair_energy(bwl_ctx *c, const int32_t *site1, const int32_t *site2, double *energy) {
  if (!energy || !site1 || !site2) return fail(BWL_ERR_ARG, "null argument");
  double o[2];
  int rc = site_pair(c, site1, site2, o);
  if (rc) return rc;
  *energy = o[1];
  return BWL_OK;
}

int bwl_genrand_seed(bwl_ctx *c, uint32_t seed) {
  if (!c) return fail(BWL_ERR_ARG, "null context");
  c->mt.seed(seed);
  return BWL_OK;
}

int bwl_genrand_get_state(bwl_ctx *c, uint32_t *mt, int32_t *mti) {
  if (!c || !mt || !mti) return fail(BWL_ERR_ARG, "null argument");
  memcpy(mt, c->mt.mt, sizeof(c->mt.mt));
  *mti = c->mt.mti;
  return BWL_OK;
}

int bwl_genrand_set_state(bwl_ctx *c, const uint32_t *mt, int32_t mti) {
  if (!c || !mt) return fail(BWL_ERR_ARG, "null argument");
  if (mti < 0 || mti > 625) return fail(BWL_ERR_ARG, "mti out of range");
  memcpy(c->mt.mt, mt, sizeof(c->mt.mt));
  c->mt.mti = mti;
  return BWL_OK;
}

// A lattice site of the context given as (b, i, j, k), 1-based.
static int check_site(const Geo &g, const int32_t *s) {
  if (s[0] != 1) return fail(BWL_ERR_ARG, "basis index must be 1");
  const int x = s[1] - 1, y = s[2] - 1, z = s[3] - 1;
  if (x < 0 || y < 0 || z < 0 || x >= g.m[0] || y >= g.m[1] || z >= g.m[2])
    return fail(BWL_ERR_ARG, "site index out of range");
  const bool occ = g.lattice == BWL_FCC ? (((x + y + z) & 1) == 0)
                                        : ((x & 1) == (y & 1) && (y & 1) == (z & 1));
  if (!occ) return fail(BWL_ERR_ARG, "site is not a lattice site");
  return BWL_OK;
}

// Proposal draws on the device from the context's MT19937 stream: the
// state goes up, the draws are made by one wavefront, the state comes back.
static int draw_sites(bwl_ctx *c, int mode, int32_t n, const int32_t *site, int32_t *out) {
  int rc = check_ctx(c, false);
  if (rc) return rc;
  if (!out || n < 1 || n > (1 << 20)) return fail(BWL_ERR_ARG, "bad count or null output");
  if (mode == 1) {
    if (!site) return fail(BWL_ERR_ARG, "null site");
    rc = check_site(c->g, site);
    if (rc) return rc;
  }
  if (c->mt.mti == 625) c->mt.seed(5489u);
  const size_t vals = (size_t)n * (mode >= 2 ? 8 : 4);
  rc = ensure_scratch(c, 624 * 4 + 64 + sizeof(int32_t) * vals);
  if (rc) return rc;
  char *base = (char *)c->d_scratch;
  DrawArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.mt = (uint32_t *)base;
  a.mt_index = (int32_t *)(base + 624 * 4);
  a.out = (int32_t *)(base + 624 * 4 + 64);
  a.mode = mode;
  a.n = mode == 1 ? 1 : n;
  if (site) memcpy(a.site, site, sizeof(a.site));
  int32_t mti = c->mt.mti;
  HIP_OK(hipMemcpyAsync(a.mt, c->mt.mt, 624 * 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(a.mt_index, &mti, 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(launch_draw_sites(a, c->stream));
  HIP_OK(hipMemcpyAsync(c->mt.mt, a.mt, 624 * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&mti, a.mt_index, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(out, a.out, sizeof(int32_t) * vals, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->mt.mti = mti;
  return BWL_OK;
}

int bwl_random_site(bwl_ctx *c, int32_t n, int32_t *sites) {
  return draw_sites(c, 0, n, nullptr, sites);
}

int bwl_random_nbr(bwl_ctx *c, const int32_t *site, int32_t *nbr) {
  return draw_sites(c, 1, 1, site, nbr);
}

int bwl_select_sites(bwl_ctx *c, int32_t nbr_swap, int32_t n, int32_t *sites) {
  return draw_sites(c, nbr_swap ? 3 : 2, n, nullptr, sites);
}

int bwl_initial_setup(bwl_ctx *c, const int64_t *species_numbers, const double *concentrations) {
  int rc = check_ctx(c, false);
  if (rc) return rc;
  std::vector<int8_t> cfg((size_t)c->g.m[0] * c->g.m[1] * c->g.m[2]);
  rc = initial_setup_host(c->g, species_numbers, concentrations, c->mt, cfg.data(), g_err);
  if (rc) return rc;
  return bwl_set_config(c, cfg.data());
}

int bwl_mc_steps(bwl_ctx *c, double beta, int32_t nbr_swap, int64_t n_trials, int64_t sample_every,
                 double *energies, int64_t *accepted) {
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (n_trials < 0) return fail(BWL_ERR_ARG, "negative trial count");
  if (energies && sample_every <= 0) return fail(BWL_ERR_ARG, "sample_every must be positive");
  const Geo &g = c->g;
  if (seq_lds_bytes(g) > (size_t)lds_limit())
    return fail(BWL_ERR_UNSUPPORTED, "lattice does not fit in LDS");
  if (2 * g.R > 64) return fail(BWL_ERR_UNSUPPORTED, "too many shells");
  // the MT state must be initialised (mti == 625 means "seed with 5489")
  if (c->mt.mti == 625) c->mt.seed(5489u);
  const long long n_samples = energies ? n_trials / sample_every : 0;
  const size_t need = 624 * 4 + 64 + sizeof(long long) + sizeof(double) * (size_t)n_samples + 256;
  rc = ensure_scratch(c, need);
  if (rc) return rc;
  char *base = (char *)c->d_scratch;
  SeqArgs a;
  a.g = g;
  a.compact = c->d_compact;
  a.off = c->d_off;
  a.vex = c->d_vex;
  a.mt = (uint32_t *)base;
  a.mt_index = (int32_t *)(base + 624 * 4);
  a.accepted = (long long *)(base + 624 * 4 + 16);
  a.energies = n_samples ? (double *)(base + 624 * 4 + 64) : nullptr;
  a.beta = beta;
  a.nbr_swap = nbr_swap ? 1 : 0;
  a.n_trials = n_trials;
  a.sample_every = energies ? sample_every : 0;
  int32_t mti = c->mt.mti;
  HIP_OK(hipMemcpyAsync(a.mt, c->mt.mt, 624 * 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemcpyAsync(a.mt_index, &mti, 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(launch_metro_seq(a, c->stream));
  long long acc = 0;
  HIP_OK(hipMemcpyAsync(c->mt.mt, a.mt, 624 * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&mti, a.mt_index, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(&acc, a.accepted, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
  if (n_samples)
    HIP_OK(hipMemcpyAsync(energies, a.energies, sizeof(double) * n_samples, hipMemcpyDeviceToHost,
                          c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->mt.mti = mti;
  if (accepted) *accepted = acc;
  return BWL_OK;
}

int bwl_mc_steps_batch(int32_t n, bwl_ctx *const *ctxs, const double *beta, int32_t nbr_swap,
                       int64_t n_trials, int64_t sample_every, double *const *energies,
                       int64_t *accepted, double *kernel_ms) {
  if (n < 1 || !ctxs || !beta) return fail(BWL_ERR_ARG, "null argument or empty batch");
  if (n_trials < 0) return fail(BWL_ERR_ARG, "negative trial count");
  if (energies && sample_every <= 0) return fail(BWL_ERR_ARG, "sample_every must be positive");
  size_t lds = 0;
  int key = 0;
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (c->device != ctxs[0]->device) return fail(BWL_ERR_ARG, "all chains of a batch must be on one device");
    if (2 * c->g.R > 64) return fail(BWL_ERR_UNSUPPORTED, "too many shells");
    lds = std::max(lds, seq_lds_bytes(c->g));
    const int k = pair_sum_key(c->g.lattice, c->g.R);
    key = i == 0 ? k : (k == key ? key : -1);
  }
  if (lds > (size_t)lds_limit()) return fail(BWL_ERR_UNSUPPORTED, "lattice does not fit in LDS");
  HIP_OK(hipSetDevice(ctxs[0]->device));
  // per chain: MT state, index, accepted count, sampled energies
  const long long n_samples = energies ? n_trials / sample_every : 0;
  const size_t stride = ((624 * 4 + 64 + sizeof(double) * (size_t)n_samples) + 255) & ~(size_t)255;
  const size_t args_bytes = (sizeof(SeqArgs) * (size_t)n + 255) & ~(size_t)255;
  DevBuf buf;
  if (hipMalloc(&buf.p, args_bytes + stride * (size_t)n) != hipSuccess)
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  char *base = (char *)buf.p;
  std::vector<SeqArgs> args((size_t)n);
  std::vector<char> host(stride * (size_t)n, 0);
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    if (c->mt.mti == 625) c->mt.seed(5489u);
    HIP_OK(hipStreamSynchronize(c->stream));  // the chain's own earlier work
    char *h = host.data() + stride * (size_t)i;
    memcpy(h, c->mt.mt, 624 * 4);
    memcpy(h + 624 * 4, &c->mt.mti, 4);
    char *d = base + args_bytes + stride * (size_t)i;
    SeqArgs &a = args[(size_t)i];
    a.g = c->g;
    a.compact = c->d_compact;
    a.off = c->d_off;
    a.vex = c->d_vex;
    a.mt = (uint32_t *)d;
    a.mt_index = (int32_t *)(d + 624 * 4);
    a.accepted = (long long *)(d + 624 * 4 + 16);
    a.energies = n_samples ? (double *)(d + 624 * 4 + 64) : nullptr;
    a.beta = beta[i];
    a.nbr_swap = nbr_swap ? 1 : 0;
    a.n_trials = n_trials;
    a.sample_every = energies ? sample_every : 0;
  }
  hipStream_t s = ctxs[0]->stream;
  HIP_OK(hipMemcpyAsync(base, args.data(), sizeof(SeqArgs) * (size_t)n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(base + args_bytes, host.data(), host.size(), hipMemcpyHostToDevice, s));
  HIP_OK(hipEventRecord(ctxs[0]->ev0, s));
  HIP_OK(launch_metro_seq_batch((const SeqArgs *)base, n, lds, s, key));
  HIP_OK(hipEventRecord(ctxs[0]->ev1, s));
  HIP_OK(hipMemcpyAsync(host.data(), base + args_bytes, host.size(), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctxs[0]->ev0, ctxs[0]->ev1);
  if (kernel_ms) *kernel_ms = ms;
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    const char *h = host.data() + stride * (size_t)i;
    memcpy(c->mt.mt, h, 624 * 4);
    memcpy(&c->mt.mti, h + 624 * 4, 4);
    long long acc = 0;
    memcpy(&acc, h + 624 * 4 + 16, sizeof(acc));
    if (accepted) accepted[i] = acc;
    if (n_samples && energies[i]) memcpy(energies[i], h + 624 * 4 + 64, sizeof(double) * n_samples);
  }
  return BWL_OK;
}

int bwl_set_layout(bwl_ctx *c, int32_t layout) {
  if (!c) return fail(BWL_ERR_ARG, "null context");
  if (layout < BWL_LAYOUT_AUTO || layout > BWL_LAYOUT_FIXED) return fail(BWL_ERR_ARG, "unknown layout");
  c->layout = layout;
  return BWL_OK;
}

int bwl_set_threads(bwl_ctx *c, int32_t threads) {
  if (!c) return fail(BWL_ERR_ARG, "null context");
  if (threads < 0 || threads > kMaxNatThreads || threads % 64 != 0)
    return fail(BWL_ERR_ARG, "threads per chain must be 0 (by rule) or a multiple of 64 up to 1024");
  c->threads = threads;
  return BWL_OK;
}

int bwl_layout_info(bwl_ctx *c, int32_t layout, int32_t *threads, int64_t *lds_bytes) {
  if (!c) return fail(BWL_ERR_ARG, "null context");
  if (layout < BWL_LAYOUT_AUTO || layout > BWL_LAYOUT_FIXED) return fail(BWL_ERR_ARG, "unknown layout");
  ParPlan plan;
  NatGeo nat;
  const void *fn = nullptr;
  int t = 0;
  size_t lds = 0;
  int rc = anneal_kernel(c->g, plan, nat, fn, t, lds, nullptr, layout);
  if (rc) return rc;
  if (threads) *threads = t;
  if (lds_bytes) *lds_bytes = (int64_t)lds;
  return BWL_OK;
}

int bwl_layouts(bwl_ctx *c, int32_t *mask) {
  if (!c || !mask) return fail(BWL_ERR_ARG, "null argument");
  *mask = 0;
  for (int layout = BWL_LAYOUT_COMPACT; layout <= BWL_LAYOUT_FIXED; ++layout)
    if (bwl_layout_info(c, layout, nullptr, nullptr) == BWL_OK) *mask |= 1 << (layout - 1);
  bwl::g_err.clear();
  return BWL_OK;
}

int bwl_metropolis_anneal(bwl_ctx *c, const bwl_anneal_params *p, double *energies_of_T,
                          double *C_of_T, double *acceptance_of_T, bwl_run_stats *stats) {
  if (!p) return fail(BWL_ERR_ARG, "null argument");
  bwl_anneal_params q = *p;
  q.asro_every = 0;
  bwl_ctx *cs[1] = {c};
  double *e[1] = {energies_of_T}, *cv[1] = {C_of_T}, *ac[1] = {acceptance_of_T};
  return bwl_metropolis_anneal_batch(1, cs, &q, e, cv, ac, stats, nullptr);
}

int bwl_metropolis_anneal_rho(bwl_ctx *c, const bwl_anneal_params *p, double *energies_of_T,
                              double *C_of_T, double *acceptance_of_T, double *rho_of_T,
                              bwl_run_stats *stats) {
  if (!c || !p || !energies_of_T || !C_of_T || !acceptance_of_T || !rho_of_T)
    return fail(BWL_ERR_ARG, "null argument");
  if (p->asro_every < 1) return fail(BWL_ERR_ARG, "asro_every must be positive");
  bwl_ctx *cs[1] = {c};
  AnnealPlan P;
  int rc = anneal_plan(1, cs, p, P);
  if (rc) return rc;
  DevBuf scratch;
  if (hipMalloc(&scratch.p, P.scratch_bytes) != hipSuccess)
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  HIP_OK(hipEventRecord(c->ev0, c->stream));
  rc = anneal_upload(P, (char *)scratch.p, c->stream);
  if (rc) return rc;
  rc = anneal_launch(P, &c->stream);
  if (rc) return rc;
  HIP_OK(hipEventRecord(c->ev1, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  double *e[1] = {energies_of_T}, *cv[1] = {C_of_T}, *ac[1] = {acceptance_of_T}, *rh[1] = {rho_of_T};
  return anneal_collect(P, cs, e, cv, ac, stats, ms, c->stream, rh);
}

}  // extern "C"

namespace bwl {

size_t ctx_device_bytes(const bwl_lattice_desc &d) {
  Geo g;
  std::vector<int32_t> off;
  std::string err;
  if (make_geo(d, g, off, err)) return 0;
  const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
  auto a = [](size_t v) { return (v + 255) & ~(size_t)255; };
  return a((size_t)g.compact_bytes) + a(ref) + a(sizeof(int32_t) * off.size()) +
         a(sizeof(double) * (size_t)g.S * g.S * g.R);
}

int ctx_create_borrowed(const bwl_lattice_desc &d, const double *v_ex, char *mem, size_t bytes,
                        hipStream_t stream, bwl_ctx **out) {
  *out = nullptr;
  bwl_ctx *c = new bwl_ctx();
  c->desc = d;
  int rc = make_geo(d, c->g, c->off, g_err);
  if (rc) {
    delete c;
    return rc;
  }
  if (bytes < ctx_device_bytes(d)) {
    delete c;
    return fail(BWL_ERR_ARG, "borrowed block too small for the lattice");
  }
  const Geo &g = c->g;
  c->vex.assign(v_ex, v_ex + (size_t)g.S * g.S * g.R);
  c->borrowed = true;
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  auto a = [](size_t v) { return (v + 255) & ~(size_t)255; };
  char *q = mem;
  c->d_compact = (int8_t *)q;
  q += a((size_t)g.compact_bytes);
  c->d_ref = (int8_t *)q;
  q += a((size_t)g.m[0] * g.m[1] * g.m[2]);
  c->d_off = (int32_t *)q;
  q += a(sizeof(int32_t) * c->off.size());
  c->d_vex = (double *)q;
  c->stream = stream;
  bool ok = hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
            hipMemcpyAsync(c->d_off, c->off.data(), sizeof(int32_t) * c->off.size(),
                           hipMemcpyHostToDevice, stream) == hipSuccess &&
            hipMemcpyAsync(c->d_vex, c->vex.data(), sizeof(double) * c->vex.size(),
                           hipMemcpyHostToDevice, stream) == hipSuccess;
  if (!ok) {
    bwl_destroy(c);
    return fail(BWL_ERR_DEVICE, "borrowed context set-up failed");
  }
  c->mt.seed(110179u);
  *out = c;
  return BWL_OK;
}

}  // namespace bwl

// ----------------------------------------------------------------------
// Batched anneal in three steps (used by bwl_metropolis_anneal_batch and by
// the request server): plan on the host, upload + launch on streams, collect
// after completion.
// ----------------------------------------------------------------------
namespace bwl {

int anneal_kernel(const Geo &g, ParPlan &plan, NatGeo &nat, const void *&fn, int &threads,
                  size_t &lds, bool *natural, int layout, int threads_req) {
  const int lim = lds_limit();
  int rc = plan_parallel(g, plan, g_err);
  if (rc) return rc;
  int max_pairs = 0;
  for (const RoundVar &v : plan.vars) max_pairs = v.n2 > max_pairs ? v.n2 : max_pairs;
  // fewest passes over a round's pairs, then the fewest threads for that;
  // with two or more passes on the 1024-thread kernels, all their threads:
  // the extra wavefronts hide the gathers' latency (bcc 32^3 anneal 2.68 ->
  // 2.43 s, 256-chain batch 165 -> 157 ms)
  auto block_size = [&](int cap) {
    const int passes = (max_pairs + cap - 1) / cap;
    int t = (((max_pairs + passes - 1) / passes + 63) / 64) * 64;
    if (cap == kMaxNatThreads && passes > 1) t = cap;
    if (t > cap) t = cap;
    return t < 64 ? 64 : t;
  };
  // Layouts in order of preference: the fixed-pitch natural layout (one
  // kernel with compile-time offsets for all sizes of a lattice and shell
  // count), unless the shape has a kernel of its own; the natural layout
  // with this size's own pitch; the per-sublattice layout.  Packed counts
  // (par_count_words): 32-bit words with 8-bit fields up to 4 species,
  // 6-bit fields for 5, else 64-bit words.
  NatGeo per{}, fix{};
  const bool nat_ok = make_nat(g, per);
  const bool fix_ok = nat_ok && !par_specialised(g) && make_nat_fixed(g, fix) &&
                      par_fixed_kernel(g.lattice, g.R, par_count_words(g, true)) != nullptr;
  const int modes[3] = {fix_ok ? 2 : -1, nat_ok ? 1 : -1, 0};
  // BWL_LAYOUT_COMPACT / _NATURAL / _FIXED are modes 0 / 1 / 2
  for (int mode : modes) {
    if (mode < 0 || (layout != 0 && mode != layout - 1)) continue;
    const bool use_nat = mode != 0, fixed = mode == 2;
    const NatGeo *ng = fixed ? &fix : (use_nat ? &per : nullptr);
    const int cap = par_max_threads(g, use_nat, fixed);
    if (threads_req > cap) continue;
    threads = threads_req > 0 ? threads_req : block_size(cap);
    lds = par_lds_bytes(g, ng, threads);
    if (lds > (size_t)lim) continue;
    if (ng) nat = *ng;
    else memset(&nat, 0, sizeof(nat));
    fn = par_kernel_for(g, use_nat, fixed);
    if (!fn) return fail(BWL_ERR_UNSUPPORTED, "no kernel for this lattice");
    if (natural) *natural = use_nat;
    return BWL_OK;
  }
  if (threads_req > 0)
    return fail(BWL_ERR_UNSUPPORTED, "no layout of this lattice runs the requested threads per chain");
  if (layout != 0) return fail(BWL_ERR_UNSUPPORTED, "the requested layout is not available for this lattice");
  return fail(BWL_ERR_UNSUPPORTED, "lattice does not fit in LDS");
}

int anneal_plan(int n, bwl_ctx *const *ctxs, const bwl_anneal_params *params, AnnealPlan &P) {
  P.items.assign((size_t)n, AnnealPlan::Item());
  P.groups.clear();
  size_t total = 0;
  const int lim = lds_limit();
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (i > 0 && c->device != ctxs[0]->device)
      return fail(BWL_ERR_ARG, "all chains of a batch must be on one device");
    const bwl_anneal_params *p = &params[i];
    if (p->T_steps < 1 || p->n_sweeps < 1 || p->burn_in_sweeps < 0)
      return fail(BWL_ERR_ARG, "T_steps and n_sweeps must be positive");
    AnnealPlan::Item &it = P.items[(size_t)i];
    it.idx = i;
    const Geo &g = c->g;
    NatGeo nat;
    int threads = 0;
    size_t lds = 0;
    bool natural = false;
    rc = anneal_kernel(g, it.plan, nat, it.fn, threads, lds, &natural, c->layout, c->threads);
    if (rc) return rc;
    if (p->asro_every < 0 || p->asro_every > p->n_sweeps)
      return fail(BWL_ERR_ARG, "asro_every must be in 0..n_sweeps");
    if (p->asro_every > 0) {
      if (!natural)
        return fail(BWL_ERR_UNSUPPORTED,
                    "radial densities in the anneal need a cubic supercell held in lattice coordinates");
      const int w = std::min(std::min(g.n[0], g.n[1]), std::min(g.n[2], 5));
      if (g.rcut > w)
        return fail(BWL_ERR_UNSUPPORTED,
                    "an interaction shell reaches outside the radial-density window (supercell too small)");
    }
    it.threads = threads;
    it.lds = lds;
    ParArgs &a = it.a;
    memset(&a, 0, sizeof(a));
    a.g = g;
    a.nat = nat;
    a.compact = c->d_compact;
    a.off = c->d_off;
    a.vex = c->d_vex;
    a.stride = it.plan.stride;
    a.L = it.plan.L;
    for (int k = 0; k < 3; ++k) {
      a.K[k] = it.plan.K[k];
      a.n_units[k] = it.plan.n_units[k];
      for (int u = 0; u < 16; ++u) a.units[k][u] = it.plan.units[k][u];
    }
    a.n_var = (int32_t)it.plan.vars.size();
    for (size_t k = 0; k < it.plan.vars.size(); ++k) a.var[k] = it.plan.vars[k];
    a.fast_halo = it.plan.fast_halo ? 1 : 0;
    a.min_trials_per_sweep = g.n_atoms;
    a.threads = threads;
    a.T0 = p->T;
    a.dT = p->delta_T;
    a.T_steps = p->T_steps;
    a.n_sweeps = p->n_sweeps;
    a.burn_sweeps = p->burn_in_sweeps;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.chain_id = 0;  // the stream depends on the seed and position only, not on the batch
    a.phase0 = c->phase;
    a.e_trace = nullptr;
    a.asro_every = p->asro_every;
    a.rho_of_T = nullptr;
    it.T = (size_t)p->T_steps;
    it.off_muls = total;
    total += 256 * ((it.plan.muls.size() * 4 + 255) / 256);
    it.off_out = total;
    total += 256 * ((5 * it.T * sizeof(double) + 255) / 256);
    it.off_cnt = total;
    total += 256;
    it.rho_n = p->asro_every > 0 ? (size_t)g.S * g.S * (g.R + 1) * it.T : 0;
    it.off_rho = total;
    total += 256 * ((it.rho_n * sizeof(double) + 255) / 256);
    it.work = (double)g.n_atoms * (double)p->T_steps * (double)(p->n_sweeps + p->burn_in_sweeps);
  }
  // group chains that run the same kernel (one launch each, the block size
  // of its largest member; each chain loops over its own pairs with the
  // threads it has), largest work first inside a group.  Few groups keep
  // every launch on its own hardware queue, so all groups run at once.
  std::vector<int> order((size_t)n);
  for (int i = 0; i < n; ++i) order[(size_t)i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    const AnnealPlan::Item &A = P.items[(size_t)x], &B = P.items[(size_t)y];
    if (A.fn != B.fn) return A.fn < B.fn;
    return A.work > B.work;
  });
  for (int i : order) {
    const AnnealPlan::Item &it = P.items[(size_t)i];
    if (P.groups.empty() || P.groups.back().fn != it.fn)
      P.groups.push_back(AnnealPlan::Group{it.fn, it.threads, 0, {}, 0});
    AnnealPlan::Group &gr = P.groups.back();
    gr.members.push_back(i);
    if (it.threads > gr.threads) gr.threads = it.threads;
  }
  // LDS per launch: the largest member's need at the group's block size
  for (AnnealPlan::Group &gr : P.groups)
    for (int i : gr.members) {
      const AnnealPlan::Item &it = P.items[(size_t)i];
      const size_t need = it.lds + sizeof(double) * (size_t)(gr.threads - it.threads) / 64;
      if (need > gr.lds) gr.lds = need;
    }
  // With no more chains than compute units, give every chain a CU of its
  // own: ask each block for more than half the LDS so no two share a CU
  // (the largest chains set the makespan; sharing a CU would halve them).
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctxs[0]->device) !=
      hipSuccess)
    n_cu = 0;
  if (n_cu > 0 && n <= n_cu) {
    const size_t exclusive = (size_t)lim / 2 + 1024;
    for (AnnealPlan::Group &gr : P.groups)
      if (gr.lds < exclusive && exclusive <= (size_t)lim) gr.lds = exclusive;
  }
  for (AnnealPlan::Group &gr : P.groups) {
    gr.off_args = total;
    total += 256 * ((sizeof(ParArgs) * gr.members.size() + 255) / 256);
  }
  P.scratch_bytes = total;
  return BWL_OK;
}

int anneal_upload(AnnealPlan &P, char *base, hipStream_t s, bool zero_by_kernel) {
  P.base = base;
  P.host_args.clear();
  for (AnnealPlan::Item &it : P.items) {
    it.a.muls = (const uint32_t *)(base + it.off_muls);
    it.a.energies_of_T = (double *)(base + it.off_out);
    it.a.C_of_T = it.a.energies_of_T + it.T;
    it.a.acc_of_T = it.a.C_of_T + it.T;
    it.a.e_drift = it.a.acc_of_T + it.T;
    it.a.counters = (long long *)(base + it.off_cnt);
    it.a.rho_of_T = it.rho_n ? (double *)(base + it.off_rho) : nullptr;
    HIP_OK(hipMemcpyAsync(base + it.off_muls, it.plan.muls.data(), it.plan.muls.size() * 4,
                          hipMemcpyHostToDevice, s));
    if (zero_by_kernel)
      HIP_OK(launch_zero_i64(it.a.counters, 3, s));
    else
      HIP_OK(hipMemsetAsync(it.a.counters, 0, 3 * sizeof(long long), s));
  }
  for (AnnealPlan::Group &gr : P.groups) {
    std::vector<ParArgs> host;
    for (int i : gr.members) host.push_back(P.items[(size_t)i].a);
    P.host_args.push_back(std::move(host));
    HIP_OK(hipMemcpyAsync(base + gr.off_args, P.host_args.back().data(),
                          sizeof(ParArgs) * P.host_args.back().size(), hipMemcpyHostToDevice, s));
  }
  return BWL_OK;
}

int anneal_launch(const AnnealPlan &P, const hipStream_t *streams) {
  for (size_t k = 0; k < P.groups.size(); ++k) {
    const AnnealPlan::Group &gr = P.groups[k];
    HIP_OK(launch_metro_par_fn(gr.fn, (const ParArgs *)(P.base + gr.off_args),
                               (int)gr.members.size(), gr.lds, gr.threads, streams[k]));
  }
  return BWL_OK;
}

int anneal_collect(const AnnealPlan &P, bwl_ctx *const *ctxs, double *const *energies_of_T,
                   double *const *C_of_T, double *const *acceptance_of_T, bwl_run_stats *stats,
                   double kernel_ms, hipStream_t s, double *const *rho_of_T) {
  const size_t n = P.items.size();
  std::vector<std::vector<double>> out(n);
  std::vector<long long> cnt(3 * n);
  for (size_t i = 0; i < n; ++i) {
    const AnnealPlan::Item &it = P.items[i];
    out[i].resize(5 * it.T);
    HIP_OK(hipMemcpyAsync(out[i].data(), it.a.energies_of_T, 5 * it.T * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&cnt[3 * i], it.a.counters, 3 * sizeof(long long), hipMemcpyDeviceToHost, s));
    if (rho_of_T && rho_of_T[i] && it.rho_n)
      HIP_OK(hipMemcpyAsync(rho_of_T[i], it.a.rho_of_T, it.rho_n * sizeof(double),
                            hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; ++i) {
    const size_t T = P.items[i].T;
    if (energies_of_T && energies_of_T[i]) memcpy(energies_of_T[i], out[i].data(), T * sizeof(double));
    if (C_of_T && C_of_T[i]) memcpy(C_of_T[i], out[i].data() + T, T * sizeof(double));
    if (acceptance_of_T && acceptance_of_T[i])
      memcpy(acceptance_of_T[i], out[i].data() + 2 * T, T * sizeof(double));
    if (stats) {
      stats[i].trials = cnt[3 * i + 2];
      stats[i].accepted = cnt[3 * i + 1];
      stats[i].kernel_ms = kernel_ms;
      // largest difference between the carried energy and its recount over
      // the T steps, absolute (Ry) and relative to the recounted energy
      double da = 0.0, dr = 0.0;
      for (size_t t = 0; t < T; ++t) {
        const double d = fabs(out[i][3 * T + t]), e = fabs(out[i][4 * T + t]);
        da = d > da ? d : da;
        if (e > 0.0 && d / e > dr) dr = d / e;
      }
      stats[i].e_drift_abs = da;
      stats[i].e_drift_rel = dr;
    }
    ctxs[i]->phase = (unsigned long long)cnt[3 * i];
  }
  return BWL_OK;
}

}  // namespace bwl

extern "C" {

int bwl_metropolis_anneal_batch(int32_t n, bwl_ctx *const *ctxs, const bwl_anneal_params *params,
                                double *const *energies_of_T, double *const *C_of_T,
                                double *const *acceptance_of_T, bwl_run_stats *stats,
                                double *makespan_ms) {
  if (n < 1 || !ctxs || !params || !energies_of_T || !C_of_T || !acceptance_of_T)
    return fail(BWL_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i)
    if (!energies_of_T[i] || !C_of_T[i] || !acceptance_of_T[i]) return fail(BWL_ERR_ARG, "null output");
  AnnealPlan P;
  int rc = anneal_plan(n, ctxs, params, P);
  if (rc) return rc;
  bwl_ctx *c0 = ctxs[0];
  DevBuf scratch;
  if (hipMalloc(&scratch.p, P.scratch_bytes) != hipSuccess)
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  // every group on its own stream so groups overlap on the device
  std::vector<hipStream_t> streams;
  struct StreamFreer {
    std::vector<hipStream_t> *v;
    ~StreamFreer() {
      for (size_t k = 1; k < v->size(); ++k) (void)hipStreamDestroy((*v)[k]);
    }
  } sfree{&streams};
  streams.push_back(c0->stream);
  for (size_t k = 1; k < P.groups.size(); ++k) {
    hipStream_t s = nullptr;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    streams.push_back(s);
  }
  hipEvent_t ready = nullptr;
  HIP_OK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  struct EventFreer {
    hipEvent_t e;
    ~EventFreer() { (void)hipEventDestroy(e); }
  } efree{ready};
  HIP_OK(hipEventRecord(c0->ev0, c0->stream));
  rc = anneal_upload(P, (char *)scratch.p, c0->stream);
  if (rc) return rc;
  HIP_OK(hipEventRecord(ready, c0->stream));
  for (size_t k = 1; k < streams.size(); ++k) HIP_OK(hipStreamWaitEvent(streams[k], ready, 0));
  rc = anneal_launch(P, streams.data());
  if (rc) return rc;
  for (size_t k = 1; k < streams.size(); ++k) {
    hipEvent_t done = nullptr;
    HIP_OK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    HIP_OK(hipEventRecord(done, streams[k]));
    HIP_OK(hipStreamWaitEvent(c0->stream, done, 0));
    (void)hipEventDestroy(done);
  }
  HIP_OK(hipEventRecord(c0->ev1, c0->stream));
  HIP_OK(hipStreamSynchronize(c0->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c0->ev0, c0->ev1);
  if (makespan_ms) *makespan_ms = ms;
  return anneal_collect(P, ctxs, energies_of_T, C_of_T, acceptance_of_T, stats, ms, c0->stream);
}

}  // extern "C"
