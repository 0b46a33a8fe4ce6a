// bwl_api_walkers.hip -- C-ABI of the radial densities, nested sampling and
// Wang-Landau walkers (include/brawl_hip.h) over the gfx950 kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bwl_ctx.h"

using namespace bwl;

namespace {

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Upload a reference-layout config and pack it into `compact`.
int upload_packed(const Geo &g, const int8_t *cfg, int8_t *d_ref, int8_t *compact, hipStream_t s) {
  const size_t ref_bytes = (size_t)g.m[0] * g.m[1] * g.m[2];
  HIP_OK(hipMemcpyAsync(d_ref, cfg, ref_bytes, hipMemcpyHostToDevice, s));
  HIP_OK(launch_pack(d_ref, compact, g, s));
  return BWL_OK;
}

bool has_two_species(const Geo &g, const int64_t *sn, const double *cc) {
  int nz = 0;
  if (sn) {
    long long t = 0;
    for (int i = 0; i < g.S; ++i) t += sn[i];
    if (t == g.n_atoms) {
      for (int i = 0; i < g.S; ++i) nz += sn[i] > 0;
      return nz >= 2;
    }
  }
  if (!cc) return false;
  for (int i = 0; i < g.S; ++i) nz += cc[i] > 0.0;
  return nz >= 2;
}

}  // namespace

// ----------------------------------------------------------------------
// Wang-Landau walker set
// ----------------------------------------------------------------------
struct bwl_wl {
  bwl_ctx *ctx = nullptr;  // geometry, V_ex, stream; its own lattice is unused
  bwl_wl_params p;
  int W = 0;               // walkers
  int hstride = 0;         // largest window width
  std::vector<int32_t> windows;
  std::vector<double> edges;
  std::vector<MT19937> mt;  // host copies (kept in step with the device)
  void *d = nullptr;        // one allocation for everything below
  int8_t *d_walkers = nullptr;
  uint32_t *d_mt = nullptr;
  int32_t *d_mti = nullptr;
  double *d_edges = nullptr, *d_lng = nullptr, *d_hist = nullptr, *d_e = nullptr;
  long long *d_isteps = nullptr, *d_trials = nullptr, *d_acc = nullptr;
  int32_t *d_resume = nullptr, *d_status = nullptr;
  // replica exchange: windows, walker energies, decisions
  int32_t *d_windows = nullptr, *d_partner = nullptr, *d_xacc = nullptr, *d_swaps = nullptr;
  double *d_xe = nullptr;
  WlArgs *d_args = nullptr;
  WlEnterArgs *d_eargs = nullptr;
  bool has_config = false;
  // composition for the re-set-ups of enter_energy_window
  std::vector<int64_t> species_numbers;
  std::vector<double> concentrations;
};

namespace {

int wl_pull_mt(bwl_wl *w) {
  std::vector<uint32_t> st((size_t)w->W * 624);
  std::vector<int32_t> mti((size_t)w->W);
  hipStream_t s = w->ctx->stream;
  HIP_OK(hipMemcpyAsync(st.data(), w->d_mt, st.size() * 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(mti.data(), w->d_mti, mti.size() * 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (int k = 0; k < w->W; ++k) {
    memcpy(w->mt[(size_t)k].mt, &st[(size_t)k * 624], 624 * 4);
    w->mt[(size_t)k].mti = mti[(size_t)k];
  }
  return BWL_OK;
}

int wl_push_mt(bwl_wl *w) {
  std::vector<uint32_t> st((size_t)w->W * 624);
  std::vector<int32_t> mti((size_t)w->W);
  for (int k = 0; k < w->W; ++k) {
    memcpy(&st[(size_t)k * 624], w->mt[(size_t)k].mt, 624 * 4);
    mti[(size_t)k] = w->mt[(size_t)k].mti;
  }
  hipStream_t s = w->ctx->stream;
  HIP_OK(hipMemcpyAsync(w->d_mt, st.data(), st.size() * 4, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(w->d_mti, mti.data(), mti.size() * 4, hipMemcpyHostToDevice, s));
  HIP_OK(hipStreamSynchronize(s));
  return BWL_OK;
}

// The per-walker arguments of k_wl_sweeps (also read by k_wl_energies).
void wl_fill_args(bwl_wl *w, double wl_f, std::vector<WlArgs> &args) {
  const Geo &g = w->ctx->g;
  const int W = w->W;
  args.resize((size_t)W);
  for (int k = 0; k < W; ++k) {
    const int win = k / w->p.walkers_per_window;
    WlArgs &a = args[(size_t)k];
    memset(&a, 0, sizeof(a));
    a.g = g;
    a.off = w->ctx->d_off;
    a.vex = w->ctx->d_vex;
    a.compact = w->d_walkers + (size_t)g.compact_bytes * k;
    a.mt = w->d_mt + (size_t)624 * k;
    a.mt_index = w->d_mti + k;
    a.nbr_swap = w->p.nbr_swap ? 1 : 0;
    a.mc_sweeps = w->p.mc_sweeps;
    a.bins = w->p.bins;
    a.start_idx = w->windows[(size_t)2 * win];
    a.end_idx = w->windows[(size_t)2 * win + 1];
    a.wl_f = wl_f;
    a.bin_edges = w->d_edges;
    a.lng = w->d_lng + (size_t)w->p.bins * k;
    a.hist = w->d_hist + (size_t)w->hstride * k;
    a.accepted = w->d_acc + k;
  }
}

int wl_check(bwl_wl *w, bool need_config) {
  if (!w) return fail(BWL_ERR_ARG, "null Wang-Landau handle");
  if (need_config && !w->has_config) return fail(BWL_ERR_STATE, "walkers have no configuration");
  HIP_OK(hipSetDevice(w->ctx->device));
  return BWL_OK;
}

}  // namespace

extern "C" {

// ----------------------------------------------------------------------
// radial densities
// ----------------------------------------------------------------------
int bwl_lattice_shells(bwl_ctx *c, int32_t wc_range, double *shells) {
  if (!c || !shells) return fail(BWL_ERR_ARG, "null argument");
  if (wc_range < 1 || wc_range > kMaxAsroShells) return fail(BWL_ERR_ARG, "wc_range must be in 1..16");
  lattice_shells_host(c->g, wc_range, shells);
  return BWL_OK;
}

int bwl_radial_densities(bwl_ctx *c, int32_t wc_range, double *rho) {
  int rc = check_ctx(c, true);
  if (rc) return rc;
  if (!rho) return fail(BWL_ERR_ARG, "null rho");
  if (wc_range < 1 || wc_range > kMaxAsroShells) return fail(BWL_ERR_ARG, "wc_range must be in 1..16");
  const Geo &g = c->g;
  std::vector<double> shells((size_t)wc_range);
  lattice_shells_host(g, wc_range, shells.data());
  AsroArgs a;
  memset(&a, 0, sizeof(a));
  std::vector<int32_t> off3;
  asro_offsets_host(g, wc_range, shells.data(), off3, a.shell_start);
  const int S = g.S;
  const size_t hn = (size_t)S * S * wc_range;
  rc = ensure_scratch(c, align256(sizeof(unsigned long long) * hn) + sizeof(int32_t) * (off3.size() + 3));
  if (rc) return rc;
  char *base = (char *)c->d_scratch;
  a.g = g;
  a.compact = c->d_compact;
  a.wc = wc_range;
  a.hist = (unsigned long long *)base;
  a.off = (const int32_t *)(base + align256(sizeof(unsigned long long) * hn));
  HIP_OK(hipMemsetAsync(a.hist, 0, sizeof(unsigned long long) * hn, c->stream));
  if (!off3.empty())
    HIP_OK(hipMemcpyAsync((void *)a.off, off3.data(), sizeof(int32_t) * off3.size(),
                          hipMemcpyHostToDevice, c->stream));
  HIP_OK(launch_asro(a, c->stream));
  std::vector<unsigned long long> h(hn);
  HIP_OK(hipMemcpyAsync(h.data(), a.hist, sizeof(unsigned long long) * hn, hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  // the on-site shell (distance 0, offset 0 only) counts each species
  for (int l = 0; l < wc_range; ++l)
    for (int b = 0; b < S; ++b)
      for (int x = 0; x < S; ++x) {
        const size_t i = x + (size_t)S * (b + (size_t)S * l);
        const double count = (double)h[x + (size_t)S * x];
        rho[i] = (double)h[i] / count;
      }
  return BWL_OK;
}

// ----------------------------------------------------------------------
// nested sampling
// ----------------------------------------------------------------------
int bwl_nested_sampling(bwl_ctx *c, const bwl_ns_params *p, const int64_t *species_numbers,
                        const double *concentrations, double *e_limits, bwl_run_stats *stats) {
  bwl_ctx *cs[1] = {c};
  const int64_t *sn[1] = {species_numbers};
  const double *cc[1] = {concentrations};
  double *el[1] = {e_limits};
  return bwl_nested_sampling_batch(1, cs, p, sn, cc, el, stats, nullptr);
}

int bwl_nested_sampling_batch(int32_t n, bwl_ctx *const *ctxs, const bwl_ns_params *params,
                              const int64_t *const *species_numbers,
                              const double *const *concentrations, double *const *e_limits,
                              bwl_run_stats *stats, double *makespan_ms) {
  if (n < 1 || !ctxs || !params || !e_limits || !species_numbers || !concentrations)
    return fail(BWL_ERR_ARG, "null argument");
  const int lim = lds_limit();
  size_t lds = 0, total = align256(sizeof(NsArgs) * (size_t)n);
  std::vector<size_t> offs((size_t)n);
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    int rc = check_ctx(c, false);
    if (rc) return rc;
    if (c->device != ctxs[0]->device) return fail(BWL_ERR_ARG, "all runs of a batch must be on one device");
    const bwl_ns_params *p = &params[i];
    if (!e_limits[i]) return fail(BWL_ERR_ARG, "null e_limits");
    if (p->n_walkers < 2 || p->n_steps < 1 || p->n_iter < 1)
      return fail(BWL_ERR_ARG, "n_walkers >= 2, n_steps >= 1 and n_iter >= 1 are required");
    if (!has_two_species(c->g, species_numbers[i], concentrations[i]))
      return fail(BWL_ERR_ARG, "nested sampling needs at least two species present");
    if (2 * c->g.R > 64) return fail(BWL_ERR_UNSUPPORTED, "too many shells");
    lds = std::max(lds, ns_lds_bytes(c->g, p->n_walkers));
    offs[(size_t)i] = total;
    const size_t ref = (size_t)c->g.m[0] * c->g.m[1] * c->g.m[2];
    total += align256((size_t)c->g.compact_bytes * p->n_walkers) +
             align256(sizeof(double) * 2 * p->n_walkers) + align256(624 * 4 + 64) +
             align256(sizeof(long long) * 3) + align256(sizeof(double) * p->n_iter) +
             align256(ref * p->n_walkers);
  }
  if (lds > (size_t)lim) return fail(BWL_ERR_UNSUPPORTED, "lattice and walker energies do not fit in LDS");
  bwl_ctx *c0 = ctxs[0];
  DevBuf buf;
  if (hipMalloc(&buf.p, total) != hipSuccess) return fail(BWL_ERR_NOMEM, "device allocation failed");
  char *base = (char *)buf.p;
  std::vector<NsArgs> args((size_t)n);
  std::vector<std::vector<double>> rnde((size_t)n);
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    const Geo &g = c->g;
    const bwl_ns_params *p = &params[i];
    NsArgs &a = args[(size_t)i];
    memset(&a, 0, sizeof(a));
    char *q = base + offs[(size_t)i];
    a.g = g;
    a.off = c->d_off;
    a.vex = c->d_vex;
    a.walkers = (int8_t *)q;
    q += align256((size_t)g.compact_bytes * p->n_walkers);
    a.rnde = (const double *)q;
    a.E = (double *)q + p->n_walkers;
    q += align256(sizeof(double) * 2 * p->n_walkers);
    a.mt = (uint32_t *)q;
    a.mt_index = (int32_t *)(q + 624 * 4);
    q += align256(624 * 4 + 64);
    a.counters = (long long *)q;
    q += align256(sizeof(long long) * 3);
    a.e_limits = (double *)q;
    q += align256(sizeof(double) * p->n_iter);
    int8_t *staging = (int8_t *)q;  // the walkers in the reference layout
    a.n_walkers = p->n_walkers;
    a.n_steps = p->n_steps;
    a.n_iter = p->n_iter;
    // the reference's atom count for the walk-length test (nested_sampling.f90:82)
    a.n_at = g.n[0] * g.n[1] * g.n[2] * 1 * g.S;
    a.max_redraws = 1LL << 24;
    // walkers from the run's MT19937 stream: initial_setup, then one draw
    // for the energy jitter (nested_sampling.f90:88-96)
    if (c->mt.mti == 625) c->mt.seed(5489u);
    const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
    std::vector<int8_t> cfgs(ref * (size_t)p->n_walkers);
    rnde[(size_t)i].resize((size_t)p->n_walkers);
    for (int w = 0; w < p->n_walkers; ++w) {
      int rc = initial_setup_host(g, species_numbers[i], concentrations[i], c->mt,
                                  cfgs.data() + ref * (size_t)w, g_err);
      if (rc) return rc;
      rnde[(size_t)i][(size_t)w] = c->mt.genrand();
    }
    HIP_OK(hipMemcpyAsync(staging, cfgs.data(), cfgs.size(), hipMemcpyHostToDevice, c0->stream));
    HIP_OK(launch_pack_many(staging, a.walkers, g, p->n_walkers, c0->stream));
    HIP_OK(hipMemcpyAsync((void *)a.rnde, rnde[(size_t)i].data(), sizeof(double) * p->n_walkers,
                          hipMemcpyHostToDevice, c0->stream));
    HIP_OK(hipMemcpyAsync(a.mt, c->mt.mt, 624 * 4, hipMemcpyHostToDevice, c0->stream));
    int32_t mti = c->mt.mti;
    HIP_OK(hipMemcpyAsync(a.mt_index, &mti, 4, hipMemcpyHostToDevice, c0->stream));
    HIP_OK(hipStreamSynchronize(c0->stream));
  }
  HIP_OK(hipMemcpyAsync(base, args.data(), sizeof(NsArgs) * (size_t)n, hipMemcpyHostToDevice, c0->stream));
  HIP_OK(hipEventRecord(c0->ev0, c0->stream));
  int max_w = 0;
  size_t lds_e = 0;
  for (int i = 0; i < n; ++i) {
    max_w = std::max(max_w, params[i].n_walkers);
    lds_e = std::max(lds_e, ns_energies_lds_bytes(ctxs[i]->g));
  }
  // one pair-sum variant for the launch when every run shares it
  int key = pair_sum_key(ctxs[0]->g.lattice, ctxs[0]->g.R);
  for (int i = 1; i < n; ++i)
    if (pair_sum_key(ctxs[i]->g.lattice, ctxs[i]->g.R) != key) key = -1;
  HIP_OK(launch_ns((const NsArgs *)base, n, lds, max_w, lds_e, c0->stream, key));
  HIP_OK(hipEventRecord(c0->ev1, c0->stream));
  HIP_OK(hipStreamSynchronize(c0->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c0->ev0, c0->ev1);
  if (makespan_ms) *makespan_ms = ms;
  int status = BWL_OK;
  for (int i = 0; i < n; ++i) {
    bwl_ctx *c = ctxs[i];
    const NsArgs &a = args[(size_t)i];
    long long cnt[3];
    int32_t mti = 0;
    HIP_OK(hipMemcpy(e_limits[i], a.e_limits, sizeof(double) * a.n_iter, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(cnt, a.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(c->mt.mt, a.mt, 624 * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&mti, a.mt_index, 4, hipMemcpyDeviceToHost));
    c->mt.mti = mti;
    if (stats) {
      stats[i].trials = cnt[0];
      stats[i].accepted = cnt[1];
      stats[i].kernel_ms = ms;
      stats[i].e_drift_abs = stats[i].e_drift_rel = 0.0;
    }
    if (cnt[2] != 0) status = fail(BWL_ERR_STATE, "nested sampling: no second species found for a trial");
  }
  return status;
}

// ----------------------------------------------------------------------
// Wang-Landau
// ----------------------------------------------------------------------
int bwl_wl_create(const bwl_lattice_desc *desc, const double *v_ex, const bwl_wl_params *p,
                  bwl_wl **out) {
  if (!desc || !v_ex || !p || !out) return fail(BWL_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->bins < 1 || p->num_windows < 1 || p->walkers_per_window < 1 || p->mc_sweeps < 1 ||
      p->num_windows > p->bins)
    return fail(BWL_ERR_ARG, "bins, num_windows (<= bins), walkers_per_window, mc_sweeps must be positive");
  if (!(p->energy_max > p->energy_min)) return fail(BWL_ERR_ARG, "energy_max must exceed energy_min");
  bwl_ctx *c = nullptr;
  int rc = bwl_create(desc, v_ex, &c);
  if (rc) return rc;
  bwl_wl *w = new bwl_wl();
  w->ctx = c;
  w->p = *p;
  w->W = p->num_windows * p->walkers_per_window;
  const Geo &g = c->g;
  w->windows.resize((size_t)2 * p->num_windows);
  w->edges.resize((size_t)p->bins + 1);
  wl_setup_host(g, p->bins, p->num_windows, p->bin_overlap, p->energy_min, p->energy_max,
                w->windows.data(), w->edges.data());
  for (int k = 0; k < p->num_windows; ++k) {
    const int s = w->windows[(size_t)2 * k], e = w->windows[(size_t)2 * k + 1];
    if (s < 1 || e > p->bins || e < s) {
      bwl_wl_destroy(w);
      return fail(BWL_ERR_ARG, "window layout out of range (too many windows for the bins)");
    }
    w->hstride = std::max(w->hstride, e - s + 1);
  }
  if (wl_lds_bytes(g, p->bins, w->hstride) > (size_t)lds_limit() || 2 * g.R > 64) {
    bwl_wl_destroy(w);
    return fail(BWL_ERR_UNSUPPORTED, "lattice and ln g do not fit in LDS");
  }
  // the reference seeds each rank's stream with 110179 + 11*rank
  // (mt19937ar.c:136-139); walker k is rank rank_offset + k
  w->mt.resize((size_t)w->W);
  for (int k = 0; k < w->W; ++k) w->mt[(size_t)k].seed(110179u + 11u * (p->rank_offset + (uint32_t)k));
  const size_t W = (size_t)w->W;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o += align256(bytes);
    return at;
  };
  const size_t o_wk = take((size_t)g.compact_bytes * W), o_mt = take(W * 624 * 4),
               o_mti = take(W * 4), o_edges = take(sizeof(double) * (p->bins + 1)),
               o_lng = take(sizeof(double) * W * p->bins),
               o_hist = take(sizeof(double) * W * w->hstride), o_e = take(sizeof(double) * W),
               o_is = take(8 * W), o_tr = take(8 * W), o_acc = take(8 * W), o_res = take(4 * W),
               o_st = take(4 * W), o_win = take(8 * (size_t)p->num_windows),
               o_part = take(4 * W), o_xacc = take(4 * W), o_swaps = take(4 * (W + 3)),
               o_xe = take(sizeof(double) * W), o_args = take(sizeof(WlArgs) * W),
               o_eargs = take(sizeof(WlEnterArgs) * W);
  if (hipMalloc(&w->d, o) != hipSuccess) {
    bwl_wl_destroy(w);
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  }
  char *b = (char *)w->d;
  w->d_walkers = (int8_t *)(b + o_wk);
  w->d_mt = (uint32_t *)(b + o_mt);
  w->d_mti = (int32_t *)(b + o_mti);
  w->d_edges = (double *)(b + o_edges);
  w->d_lng = (double *)(b + o_lng);
  w->d_hist = (double *)(b + o_hist);
  w->d_e = (double *)(b + o_e);
  w->d_isteps = (long long *)(b + o_is);
  w->d_trials = (long long *)(b + o_tr);
  w->d_acc = (long long *)(b + o_acc);
  w->d_resume = (int32_t *)(b + o_res);
  w->d_status = (int32_t *)(b + o_st);
  w->d_windows = (int32_t *)(b + o_win);
  w->d_partner = (int32_t *)(b + o_part);
  w->d_xacc = (int32_t *)(b + o_xacc);
  w->d_swaps = (int32_t *)(b + o_swaps);
  w->d_xe = (double *)(b + o_xe);
  w->d_args = (WlArgs *)(b + o_args);
  w->d_eargs = (WlEnterArgs *)(b + o_eargs);
  bool ok = hipMemsetAsync(b, 0, o, c->stream) == hipSuccess &&
            hipMemcpyAsync(w->d_edges, w->edges.data(), sizeof(double) * (p->bins + 1),
                           hipMemcpyHostToDevice, c->stream) == hipSuccess &&
            hipMemcpyAsync(w->d_windows, w->windows.data(), 8 * (size_t)p->num_windows,
                           hipMemcpyHostToDevice, c->stream) == hipSuccess;
  if (!ok || wl_push_mt(w) != BWL_OK) {
    bwl_wl_destroy(w);
    return fail(BWL_ERR_DEVICE, "Wang-Landau set-up upload failed");
  }
  *out = w;
  return BWL_OK;
}

int bwl_wl_destroy(bwl_wl *w) {
  if (!w) return BWL_OK;
  if (w->ctx) {
    (void)hipSetDevice(w->ctx->device);
    (void)hipStreamSynchronize(w->ctx->stream);
    if (w->d) (void)hipFree(w->d);
    bwl_destroy(w->ctx);
  }
  delete w;
  return BWL_OK;
}

int bwl_wl_n_walkers(const bwl_wl *w, int32_t *n_walkers) {
  if (!w || !n_walkers) return fail(BWL_ERR_ARG, "null argument");
  *n_walkers = w->W;
  return BWL_OK;
}

int bwl_wl_windows(const bwl_wl *w, int32_t *windows, double *bin_edges) {
  if (!w) return fail(BWL_ERR_ARG, "null Wang-Landau handle");
  if (windows) memcpy(windows, w->windows.data(), sizeof(int32_t) * w->windows.size());
  if (bin_edges) memcpy(bin_edges, w->edges.data(), sizeof(double) * w->edges.size());
  return BWL_OK;
}

int bwl_wl_initial_setup(bwl_wl *w, const int64_t *species_numbers, const double *concentrations) {
  int rc = wl_check(w, false);
  if (rc) return rc;
  const Geo &g = w->ctx->g;
  w->species_numbers.assign(species_numbers ? species_numbers : nullptr,
                            species_numbers ? species_numbers + g.S : nullptr);
  w->concentrations.assign(concentrations ? concentrations : nullptr,
                           concentrations ? concentrations + g.S : nullptr);
  std::vector<int8_t> cfg((size_t)g.m[0] * g.m[1] * g.m[2]);
  for (int k = 0; k < w->W; ++k) {
    rc = initial_setup_host(g, species_numbers, concentrations, w->mt[(size_t)k], cfg.data(), g_err);
    if (rc) return rc;
    rc = upload_packed(g, cfg.data(), w->ctx->d_ref, w->d_walkers + (size_t)g.compact_bytes * k,
                       w->ctx->stream);
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(w->ctx->stream));
  }
  rc = wl_push_mt(w);
  if (rc) return rc;
  w->has_config = true;
  return BWL_OK;
}

int bwl_wl_set_walker(bwl_wl *w, int32_t k, const int8_t *config) {
  int rc = wl_check(w, false);
  if (rc) return rc;
  if (k < 0 || k >= w->W || !config) return fail(BWL_ERR_ARG, "walker index out of range or null config");
  const Geo &g = w->ctx->g;
  rc = check_config(g, config, g_err);
  if (rc) return rc;
  rc = upload_packed(g, config, w->ctx->d_ref, w->d_walkers + (size_t)g.compact_bytes * k, w->ctx->stream);
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(w->ctx->stream));
  // re-set-ups keep the composition of the uploaded configuration
  w->species_numbers.assign((size_t)g.S, 0);
  const size_t cells = (size_t)g.m[0] * g.m[1] * g.m[2];
  for (size_t i = 0; i < cells; ++i)
    if (config[i] > 0) w->species_numbers[(size_t)config[i] - 1]++;
  w->concentrations.clear();
  w->has_config = true;
  return BWL_OK;
}

int bwl_wl_get_walker(bwl_wl *w, int32_t k, int8_t *config) {
  int rc = wl_check(w, true);
  if (rc) return rc;
  if (k < 0 || k >= w->W || !config) return fail(BWL_ERR_ARG, "walker index out of range or null config");
  const Geo &g = w->ctx->g;
  const size_t ref_bytes = (size_t)g.m[0] * g.m[1] * g.m[2];
  HIP_OK(launch_unpack(w->d_walkers + (size_t)g.compact_bytes * k, w->ctx->d_ref, g, w->ctx->stream));
  HIP_OK(hipMemcpyAsync(config, w->ctx->d_ref, ref_bytes, hipMemcpyDeviceToHost, w->ctx->stream));
  HIP_OK(hipStreamSynchronize(w->ctx->stream));
  return BWL_OK;
}

int bwl_wl_genrand_get_state(bwl_wl *w, int32_t k, uint32_t *mt, int32_t *mti) {
  if (!w || !mt || !mti || k < 0 || k >= w->W) return fail(BWL_ERR_ARG, "bad argument");
  memcpy(mt, w->mt[(size_t)k].mt, 624 * 4);
  *mti = w->mt[(size_t)k].mti;
  return BWL_OK;
}

int bwl_wl_genrand_set_state(bwl_wl *w, int32_t k, const uint32_t *mt, int32_t mti) {
  int rc = wl_check(w, false);
  if (rc) return rc;
  if (!mt || k < 0 || k >= w->W || mti < 0 || mti > 625) return fail(BWL_ERR_ARG, "bad argument");
  memcpy(w->mt[(size_t)k].mt, mt, 624 * 4);
  w->mt[(size_t)k].mti = mti;
  return wl_push_mt(w);
}

int bwl_wl_get_lng(bwl_wl *w, double *lng, double *hist) {
  int rc = wl_check(w, false);
  if (rc) return rc;
  const size_t W = (size_t)w->W;
  if (lng)
    HIP_OK(hipMemcpyAsync(lng, w->d_lng, sizeof(double) * W * w->p.bins, hipMemcpyDeviceToHost,
                          w->ctx->stream));
  if (hist)
    HIP_OK(hipMemcpyAsync(hist, w->d_hist, sizeof(double) * W * w->hstride, hipMemcpyDeviceToHost,
                          w->ctx->stream));
  HIP_OK(hipStreamSynchronize(w->ctx->stream));
  return BWL_OK;
}

int bwl_wl_set_lng(bwl_wl *w, const double *lng, const double *hist) {
  int rc = wl_check(w, false);
  if (rc) return rc;
  const size_t W = (size_t)w->W;
  if (lng)
    HIP_OK(hipMemcpyAsync(w->d_lng, lng, sizeof(double) * W * w->p.bins, hipMemcpyHostToDevice,
                          w->ctx->stream));
  if (hist)
    HIP_OK(hipMemcpyAsync(w->d_hist, hist, sizeof(double) * W * w->hstride, hipMemcpyHostToDevice,
                          w->ctx->stream));
  HIP_OK(hipStreamSynchronize(w->ctx->stream));
  return BWL_OK;
}

int bwl_wl_hist_stride(const bwl_wl *w, int32_t *stride) {
  if (!w || !stride) return fail(BWL_ERR_ARG, "null argument");
  *stride = w->hstride;
  return BWL_OK;
}

int bwl_wl_enter_windows(bwl_wl *w, int64_t max_trials, int64_t *trials, int32_t *entered) {
  int rc = wl_check(w, true);
  if (rc) return rc;
  if (max_trials < 0) return fail(BWL_ERR_ARG, "negative trial bound");
  const Geo &g = w->ctx->g;
  hipStream_t s = w->ctx->stream;
  const int W = w->W;
  const long long per_launch = 1LL << 18;  // bounds one launch to well under a second
  // bias width: REAL(4) product, FP64 division (wang-landau.F90:724-725)
  float sf = 0.0025f * fabsf(w->p.energy_max - w->p.energy_min);
  sf = sf * (float)g.n_atoms;
  const double sigma = (double)sf / (13.605693122 * 1000.0);
  const double denom = 2.0 * (sigma * sigma);
  std::vector<WlEnterArgs> args((size_t)W);
  std::vector<int32_t> status((size_t)W, 2), resume((size_t)W, 0);
  std::vector<long long> tr((size_t)W, 0), zeros((size_t)W, 0);
  HIP_OK(hipMemcpyAsync(w->d_isteps, zeros.data(), 8 * W, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(w->d_trials, zeros.data(), 8 * W, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(w->d_resume, resume.data(), 4 * W, hipMemcpyHostToDevice, s));
  const size_t lds = wl_enter_lds_bytes(g);
  bool first = true;
  std::vector<int8_t> cfg((size_t)g.m[0] * g.m[1] * g.m[2]);
  for (;;) {
    // walkers still walking (status 1 or 2) and under the trial bound
    std::vector<int> act;
    for (int k = 0; k < W; ++k)
      if (status[(size_t)k] != 0 && tr[(size_t)k] < max_trials) act.push_back(k);
    if (act.empty()) break;
    for (size_t q = 0; q < act.size(); ++q) {
      const int k = act[q];
      const int win = k / w->p.walkers_per_window;
      WlEnterArgs &a = args[q];
      memset(&a, 0, sizeof(a));
      a.g = g;
      a.off = w->ctx->d_off;
      a.vex = w->ctx->d_vex;
      a.compact = w->d_walkers + (size_t)g.compact_bytes * k;
      a.mt = w->d_mt + (size_t)624 * k;
      a.mt_index = w->d_mti + k;
      // window energy range: its first and last bin edges
      a.min_e = w->edges[(size_t)w->windows[(size_t)2 * win] - 1];
      a.max_e = w->edges[(size_t)w->windows[(size_t)2 * win + 1]];
      a.denom = denom;
      a.period = (long long)g.n_atoms * 50 * 5;
      a.budget = std::min(per_launch, (long long)max_trials - tr[(size_t)k]);
      a.init_e = first ? 1 : 0;
      a.resume = w->d_resume + k;
      a.e_unswapped = w->d_e + k;
      a.i_steps = w->d_isteps + k;
      a.trials = w->d_trials + k;
      a.status = w->d_status + k;
    }
    HIP_OK(hipMemcpyAsync(w->d_eargs, args.data(), sizeof(WlEnterArgs) * act.size(),
                          hipMemcpyHostToDevice, s));
    HIP_OK(launch_wl_enter(w->d_eargs, (int)act.size(), lds, s, pair_sum_key(g.lattice, g.R)));
    std::vector<int32_t> st((size_t)W);
    HIP_OK(hipMemcpyAsync(st.data(), w->d_status, 4 * W, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(tr.data(), w->d_trials, 8 * W, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    first = false;
    bool reinit = false;
    for (int k : act) {
      status[(size_t)k] = st[(size_t)k];
      if (st[(size_t)k] == 1) reinit = true;
    }
    if (reinit) {
      // initial_setup from the walker's own stream, as the reference does
      // every n_atoms*250 iterations outside the window (wang-landau.F90:672-675)
      rc = wl_pull_mt(w);
      if (rc) return rc;
      for (int k : act) {
        if (st[(size_t)k] != 1) continue;
        rc = initial_setup_host(g, w->species_numbers.empty() ? nullptr : w->species_numbers.data(),
                                w->concentrations.empty() ? nullptr : w->concentrations.data(),
                                w->mt[(size_t)k], cfg.data(), g_err);
        if (rc) return rc;
        rc = upload_packed(g, cfg.data(), w->ctx->d_ref, w->d_walkers + (size_t)g.compact_bytes * k, s);
        if (rc) return rc;
        HIP_OK(hipStreamSynchronize(s));
      }
      rc = wl_push_mt(w);
      if (rc) return rc;
    }
  }
  rc = wl_pull_mt(w);
  if (rc) return rc;
  for (int k = 0; k < W; ++k) {
    if (trials) trials[k] = tr[(size_t)k];
    if (entered) entered[k] = status[(size_t)k] == 0 ? 1 : 0;
  }
  return BWL_OK;
}

int bwl_wl_sweeps(bwl_wl *w, double wl_f, int32_t average, int64_t *accepted, double *kernel_ms) {
  int rc = wl_check(w, true);
  if (rc) return rc;
  const Geo &g = w->ctx->g;
  hipStream_t s = w->ctx->stream;
  const int W = w->W;
  std::vector<WlArgs> args;
  wl_fill_args(w, wl_f, args);
  HIP_OK(hipMemcpyAsync(w->d_args, args.data(), sizeof(WlArgs) * W, hipMemcpyHostToDevice, s));
  HIP_OK(hipEventRecord(w->ctx->ev0, s));
  HIP_OK(launch_wl_sweeps(w->d_args, W, wl_lds_bytes(g, w->p.bins, w->hstride), s,
                          pair_sum_key(g.lattice, g.R)));
  if (average)
    HIP_OK(launch_wl_average(w->d_lng, w->d_hist, w->p.num_windows, w->p.walkers_per_window,
                             w->p.bins, w->hstride, s));
  HIP_OK(hipEventRecord(w->ctx->ev1, s));
  std::vector<long long> acc((size_t)W);
  HIP_OK(hipMemcpyAsync(acc.data(), w->d_acc, 8 * W, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, w->ctx->ev0, w->ctx->ev1);
  if (kernel_ms) *kernel_ms = ms;
  if (accepted)
    for (int k = 0; k < W; ++k) accepted[k] = acc[(size_t)k];
  return wl_pull_mt(w);
}

int bwl_wl_replica_exchange(bwl_wl *w, int32_t *partner, int32_t *exchanged, double *kernel_ms) {
  int rc = wl_check(w, true);
  if (rc) return rc;
  const Geo &g = w->ctx->g;
  hipStream_t s = w->ctx->stream;
  const int W = w->W, wpw = w->p.walkers_per_window;
  const size_t xlds = wl_exchange_lds_bytes(W, wpw);
  if (xlds > (size_t)lds_limit())
    return fail(BWL_ERR_UNSUPPORTED, "too many walkers for the replica-exchange kernel");
  std::vector<WlArgs> args;
  wl_fill_args(w, 0.0, args);
  WlExchangeArgs x;
  x.E = w->d_xe;
  x.bin_edges = w->d_edges;
  x.windows = w->d_windows;
  x.lng = w->d_lng;
  x.mt = w->d_mt;
  x.mt_index = w->d_mti;
  x.partner = w->d_partner;
  x.accepted = w->d_xacc;
  x.swaps = w->d_swaps;
  x.n_swaps = w->d_swaps + W + 1;
  x.W = W;
  x.wpw = wpw;
  x.num_windows = w->p.num_windows;
  x.bins = w->p.bins;
  HIP_OK(hipMemcpyAsync(w->d_args, args.data(), sizeof(WlArgs) * W, hipMemcpyHostToDevice, s));
  HIP_OK(hipEventRecord(w->ctx->ev0, s));
  HIP_OK(launch_wl_energies(w->d_args, W, w->d_xe, ns_energies_lds_bytes(g), s));
  HIP_OK(launch_wl_exchange(x, xlds, s));
  HIP_OK(launch_wl_swap_configs(w->d_walkers, (long long)g.compact_bytes, x.swaps, x.n_swaps,
                                W / 2, s));
  HIP_OK(hipEventRecord(w->ctx->ev1, s));
  std::vector<int32_t> part((size_t)W), acc((size_t)W);
  HIP_OK(hipMemcpyAsync(part.data(), w->d_partner, 4 * (size_t)W, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(acc.data(), w->d_xacc, 4 * (size_t)W, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, w->ctx->ev0, w->ctx->ev1);
  if (kernel_ms) *kernel_ms = ms;
  for (int k = 0; k < W; ++k) {
    if (partner) partner[k] = part[(size_t)k];
    if (exchanged) exchanged[k] = acc[(size_t)k];
  }
  return wl_pull_mt(w);
}

}  // extern "C"
