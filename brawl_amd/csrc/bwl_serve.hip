// bwl_serve.hip -- a request server on one device (continuous batching).
//
// Sampling runs (Metropolis anneals, nested-sampling runs, short Wang-Landau
// runs) are submitted with their host inputs and queue on the host.  Each bwl_server_progress call
// first retires finished launches (results copied to the callers' buffers,
// device memory returned), then admits queued requests in arrival order,
// skipping ahead past any that do not fit, while device memory (a fixed
// arena: the admission budget) and launch slots (streams) are free.  The
// requests admitted together are grouped by kernel variant; each group is
// one launch on its own stream, so groups admitted at different times run
// side by side on the device.  A request's lattice, RNG state and outputs
// stay in device memory from admission to completion.  Every request keeps
// its own random stream (Philox keyed by its seed; MT19937 for nested
// sampling and Wang-Landau walkers), so its result does not depend on what it
// was batched with.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "bwl_ctx.h"

using namespace bwl;

namespace {

size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

// First-fit allocator over one device allocation.
struct Arena {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  size_t used = 0;

  void init(char *b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    free_[0] = n;
    used = 0;
  }
  char *alloc(size_t n) {
    n = al(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      used += n;
      return base + off;
    }
    return nullptr;
  }
  void release(char *p, size_t n) {
    if (!p) return;
    n = al(n);
    size_t off = (size_t)(p - base);
    used -= n;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
};

enum { KIND_ANNEAL = 1, KIND_NS = 2, KIND_WL = 3 };
enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2, ST_FAILED = 3 };

struct Request {
  int64_t id = 0;
  int kind = 0;
  int state = ST_QUEUED;
  int error = BWL_OK;
  bwl_lattice_desc desc;
  Geo g;
  std::vector<double> vex;
  // anneal
  std::vector<int8_t> config;
  bwl_anneal_params ap;
  double *E = nullptr, *Cv = nullptr, *acc = nullptr;
  int8_t *config_out = nullptr;
  const void *fn = nullptr;  // kernel variant (grouping key)
  // nested sampling: walkers prepared on the host at submission
  bwl_ns_params np;
  std::vector<int8_t> walkers_ref;  // n_walkers reference-layout configs
  std::vector<double> rnde;
  uint32_t mt[624];
  int32_t mti = 0;
  double *e_limits = nullptr;
  // Wang-Landau window: walkers prepared on the host at submission
  bwl_wl_params wp;
  double wl_f = 0.0;
  int32_t hstride = 0;
  std::vector<int32_t> windows;
  std::vector<double> edges;
  std::vector<uint32_t> wl_mt;   // 624 words per walker
  std::vector<int32_t> wl_mti;
  double *lng_out = nullptr, *hist_out = nullptr;
  // device residency
  size_t block_bytes = 0;
  char *block = nullptr;
  bwl_ctx *ctx = nullptr;
  // accounting
  double t_submit = 0, t_admit = 0, t_done = 0, kernel_ms = 0;
  int64_t trials = 0, accepted = 0;
};

struct Launch {
  int kind = 0;
  int slot = -1;
  std::vector<Request *> reqs;
  std::unique_ptr<AnnealPlan> plan;
  std::vector<NsArgs> ns_args;
  std::vector<WlArgs> wl_args;
  std::vector<size_t> wl_first;  // first walker of each request in wl_args
  char *scratch = nullptr;
  size_t scratch_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bwl_server {
  int device = 0;
  Arena arena;
  void *arena_mem = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::deque<Request *> queue;
  std::vector<std::unique_ptr<Launch>> running;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  std::vector<int64_t> done_unpolled;
  int64_t next_id = 1;
  double batch_wait_ms = 0.0;
  std::chrono::steady_clock::time_point t0;
};

namespace {

size_t ns_block_bytes(const Request &r) {
  const Geo &g = r.g;
  const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
  const size_t W = (size_t)r.np.n_walkers;
  return ctx_device_bytes(r.desc) + al((size_t)g.compact_bytes * W) + al(ref * W) +
         al(sizeof(double) * 2 * W) + al(624 * 4 + 64) + al(sizeof(long long) * 3) +
         al(sizeof(double) * (size_t)r.np.n_iter);
}

size_t wl_block_bytes(const Request &r) {
  const Geo &g = r.g;
  const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
  const size_t W = (size_t)r.wp.num_windows * r.wp.walkers_per_window;
  return ctx_device_bytes(r.desc) + al((size_t)g.compact_bytes * W) + al(ref * W) + al(W * 624 * 4) +
         al(W * 4) + al(sizeof(double) * (r.wp.bins + 1)) + al(sizeof(double) * W * r.wp.bins) +
         al(sizeof(double) * W * r.hstride) + al(8 * W);
}

int free_slot(bwl_server *s) {
  for (size_t k = 0; k < s->slot_busy.size(); ++k)
    if (!s->slot_busy[k]) return (int)k;
  return -1;
}

void release_request(bwl_server *s, Request *r) {
  if (r->ctx) {
    bwl_destroy(r->ctx);
    r->ctx = nullptr;
  }
  s->arena.release(r->block, r->block_bytes);
  r->block = nullptr;
}

// Retire one finished launch: copy results out, free its device memory.
int retire(bwl_server *s, Launch &L) {
  hipStream_t st = s->streams[(size_t)L.slot];
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, L.ev0, L.ev1);
  const double t = now_ms(s->t0);
  int status = BWL_OK;
  if (L.kind == KIND_ANNEAL) {
    const size_t n = L.reqs.size();
    std::vector<bwl_ctx *> ctxs(n);
    std::vector<double *> E(n), C(n), A(n);
    std::vector<bwl_run_stats> stats(n);
    for (size_t i = 0; i < n; ++i) {
      ctxs[i] = L.reqs[i]->ctx;
      E[i] = L.reqs[i]->E;
      C[i] = L.reqs[i]->Cv;
      A[i] = L.reqs[i]->acc;
    }
    status = anneal_collect(*L.plan, ctxs.data(), E.data(), C.data(), A.data(), stats.data(), ms, st);
    for (size_t i = 0; i < n; ++i) {
      Request *r = L.reqs[i];
      r->trials = stats[i].trials;
      r->accepted = stats[i].accepted;
      if (status == BWL_OK && r->config_out) {
        const Geo &g = r->g;
        const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
        if (launch_unpack(r->ctx->d_compact, r->ctx->d_ref, g, st) != hipSuccess ||
            hipMemcpyAsync(r->config_out, r->ctx->d_ref, ref, hipMemcpyDeviceToHost, st) != hipSuccess)
          status = fail(BWL_ERR_DEVICE, "result download failed");
      }
    }
  } else if (L.kind == KIND_WL) {
    for (size_t i = 0; i < L.reqs.size(); ++i) {
      Request *r = L.reqs[i];
      const size_t f = L.wl_first[i];
      const int nwin = r->wp.num_windows, wpw = r->wp.walkers_per_window;
      const size_t W = (size_t)nwin * wpw;
      std::vector<long long> acc(W);
      for (int w = 0; w < nwin; ++w) {
        const WlArgs &a = L.wl_args[f + (size_t)w * wpw];
        if (hipMemcpyAsync(r->lng_out + (size_t)w * r->wp.bins, a.lng, sizeof(double) * r->wp.bins,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(r->hist_out + (size_t)w * r->wp.bins, a.hist, sizeof(double) * r->hstride,
                           hipMemcpyDeviceToHost, st) != hipSuccess)
          status = fail(BWL_ERR_DEVICE, "result download failed");
      }
      if (hipMemcpyAsync(acc.data(), L.wl_args[f].accepted, 8 * W, hipMemcpyDeviceToHost, st) !=
              hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        status = fail(BWL_ERR_DEVICE, "result download failed");
        continue;
      }
      r->accepted = 0;
      for (long long a : acc) r->accepted += a;
      r->trials = (int64_t)W * r->wp.mc_sweeps * r->g.n_atoms;
    }
  } else {
    for (size_t i = 0; i < L.reqs.size(); ++i) {
      Request *r = L.reqs[i];
      const NsArgs &a = L.ns_args[i];
      long long cnt[3] = {0, 0, 0};
      if (hipMemcpyAsync(r->e_limits, a.e_limits, sizeof(double) * a.n_iter, hipMemcpyDeviceToHost,
                         st) != hipSuccess ||
          hipMemcpyAsync(cnt, a.counters, sizeof(cnt), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        status = fail(BWL_ERR_DEVICE, "result download failed");
        continue;
      }
      r->trials = cnt[0];
      r->accepted = cnt[1];
      if (cnt[2] != 0) r->error = BWL_ERR_STATE;
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess && status == BWL_OK)
    status = fail(BWL_ERR_DEVICE, "stream synchronisation failed");
  for (Request *r : L.reqs) {
    r->kernel_ms = ms;
    r->t_done = t;
    if (status != BWL_OK && r->error == BWL_OK) r->error = status;
    r->state = r->error == BWL_OK ? ST_DONE : ST_FAILED;
    release_request(s, r);
    s->done_unpolled.push_back(r->id);
  }
  s->arena.release(L.scratch, L.scratch_bytes);
  (void)hipEventDestroy(L.ev0);
  (void)hipEventDestroy(L.ev1);
  s->slot_busy[(size_t)L.slot] = false;
  return status;
}

// Device set-up of an admitted request on stream `st` (its block is
// allocated): borrowed context, configuration or walkers.
int stage_request(Request *r, hipStream_t st) {
  const size_t cb = ctx_device_bytes(r->desc);
  int rc = ctx_create_borrowed(r->desc, r->vex.data(), r->block, cb, st, &r->ctx);
  if (rc) return rc;
  if (r->kind == KIND_ANNEAL) {
    rc = check_config(r->g, r->config.data(), g_err);
    if (rc) return rc;
    const size_t ref = (size_t)r->g.m[0] * r->g.m[1] * r->g.m[2];
    HIP_OK(hipMemcpyAsync(r->ctx->d_ref, r->config.data(), ref, hipMemcpyHostToDevice, st));
    HIP_OK(launch_pack(r->ctx->d_ref, r->ctx->d_compact, r->g, st));
    r->ctx->has_config = true;
  }
  return BWL_OK;
}

int launch_anneal_group(bwl_server *s, std::vector<Request *> &grp, int slot) {
  hipStream_t st = s->streams[(size_t)slot];
  auto L = std::make_unique<Launch>();
  L->kind = KIND_ANNEAL;
  L->slot = slot;
  L->reqs = grp;
  L->plan = std::make_unique<AnnealPlan>();
  std::vector<bwl_ctx *> ctxs;
  std::vector<bwl_anneal_params> params;
  for (Request *r : grp) {
    ctxs.push_back(r->ctx);
    params.push_back(r->ap);
  }
  int rc = anneal_plan((int)grp.size(), ctxs.data(), params.data(), *L->plan);
  if (rc) return rc;
  L->scratch_bytes = L->plan->scratch_bytes;
  L->scratch = s->arena.alloc(L->scratch_bytes);
  if (!L->scratch) return BWL_ERR_NOMEM;  // caller requeues
  HIP_OK(hipEventCreate(&L->ev0));
  HIP_OK(hipEventCreate(&L->ev1));
  HIP_OK(hipEventRecord(L->ev0, st));
  rc = anneal_upload(*L->plan, L->scratch, st);
  if (rc) return rc;
  rc = anneal_launch(*L->plan, &st);
  if (rc) return rc;
  HIP_OK(hipEventRecord(L->ev1, st));
  s->slot_busy[(size_t)slot] = true;
  s->running.push_back(std::move(L));
  return BWL_OK;
}

int launch_ns_group(bwl_server *s, std::vector<Request *> &grp, int slot) {
  hipStream_t st = s->streams[(size_t)slot];
  auto L = std::make_unique<Launch>();
  L->kind = KIND_NS;
  L->slot = slot;
  L->reqs = grp;
  size_t lds = 0;
  for (Request *r : grp) lds = std::max(lds, ns_lds_bytes(r->g, r->np.n_walkers));
  L->scratch_bytes = al(sizeof(NsArgs) * grp.size());
  L->scratch = s->arena.alloc(L->scratch_bytes);
  if (!L->scratch) return BWL_ERR_NOMEM;
  for (Request *r : grp) {
    const Geo &g = r->g;
    const size_t W = (size_t)r->np.n_walkers;
    const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
    char *q = r->block + ctx_device_bytes(r->desc);
    NsArgs a;
    memset(&a, 0, sizeof(a));
    a.g = g;
    a.off = r->ctx->d_off;
    a.vex = r->ctx->d_vex;
    a.walkers = (int8_t *)q;
    q += al((size_t)g.compact_bytes * W);
    int8_t *staging = (int8_t *)q;
    q += al(ref * W);
    a.rnde = (const double *)q;
    a.E = (double *)q + W;
    q += al(sizeof(double) * 2 * W);
    a.mt = (uint32_t *)q;
    a.mt_index = (int32_t *)(q + 624 * 4);
    q += al(624 * 4 + 64);
    a.counters = (long long *)q;
    q += al(sizeof(long long) * 3);
    a.e_limits = (double *)q;
    a.n_walkers = r->np.n_walkers;
    a.n_steps = r->np.n_steps;
    a.n_iter = r->np.n_iter;
    a.n_at = g.n[0] * g.n[1] * g.n[2] * 1 * g.S;  // nested_sampling.f90:82
    a.max_redraws = 1LL << 24;
    HIP_OK(hipMemcpyAsync(staging, r->walkers_ref.data(), ref * W, hipMemcpyHostToDevice, st));
    HIP_OK(launch_pack_many(staging, a.walkers, g, (int)W, st));
    HIP_OK(hipMemcpyAsync((void *)a.rnde, r->rnde.data(), sizeof(double) * W, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(a.mt, r->mt, 624 * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(a.mt_index, &r->mti, 4, hipMemcpyHostToDevice, st));
    L->ns_args.push_back(a);
  }
  HIP_OK(hipMemcpyAsync(L->scratch, L->ns_args.data(), sizeof(NsArgs) * grp.size(),
                        hipMemcpyHostToDevice, st));
  HIP_OK(hipEventCreate(&L->ev0));
  HIP_OK(hipEventCreate(&L->ev1));
  HIP_OK(hipEventRecord(L->ev0, st));
  int max_w = 0;
  size_t lds_e = 0;
  for (Request *r : grp) {
    max_w = std::max(max_w, r->np.n_walkers);
    lds_e = std::max(lds_e, ns_energies_lds_bytes(r->g));
  }
  int key = pair_sum_key(grp[0]->g.lattice, grp[0]->g.R);
  for (Request *r : grp)
    if (pair_sum_key(r->g.lattice, r->g.R) != key) key = -1;
  HIP_OK(launch_ns((const NsArgs *)L->scratch, (int)grp.size(), lds, max_w, lds_e, st, key));
  HIP_OK(hipEventRecord(L->ev1, st));
  s->slot_busy[(size_t)slot] = true;
  s->running.push_back(std::move(L));
  return BWL_OK;
}

int launch_wl_group(bwl_server *s, std::vector<Request *> &grp, int slot) {
  hipStream_t st = s->streams[(size_t)slot];
  auto L = std::make_unique<Launch>();
  L->kind = KIND_WL;
  L->slot = slot;
  L->reqs = grp;
  size_t lds = 0, total_w = 0;
  for (Request *r : grp) {
    lds = std::max(lds, wl_lds_bytes(r->g, r->wp.bins, r->hstride));
    total_w += (size_t)r->wp.num_windows * r->wp.walkers_per_window;
  }
  L->scratch_bytes = al(sizeof(WlArgs) * total_w);
  L->scratch = s->arena.alloc(L->scratch_bytes);
  if (!L->scratch) return BWL_ERR_NOMEM;
  for (Request *r : grp) {
    const Geo &g = r->g;
    const int wpw = r->wp.walkers_per_window;
    const size_t W = (size_t)r->wp.num_windows * wpw;
    const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
    char *q = r->block + ctx_device_bytes(r->desc);
    int8_t *walkers = (int8_t *)q;
    q += al((size_t)g.compact_bytes * W);
    int8_t *staging = (int8_t *)q;
    q += al(ref * W);
    uint32_t *mt = (uint32_t *)q;
    q += al(W * 624 * 4);
    int32_t *mti = (int32_t *)q;
    q += al(W * 4);
    double *edges = (double *)q;
    q += al(sizeof(double) * (r->wp.bins + 1));
    double *lng = (double *)q;
    q += al(sizeof(double) * W * r->wp.bins);
    double *hist = (double *)q;
    q += al(sizeof(double) * W * r->hstride);
    long long *acc = (long long *)q;
    HIP_OK(hipMemcpyAsync(staging, r->walkers_ref.data(), ref * W, hipMemcpyHostToDevice, st));
    HIP_OK(launch_pack_many(staging, walkers, g, (int)W, st));
    HIP_OK(hipMemcpyAsync(mt, r->wl_mt.data(), W * 624 * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(mti, r->wl_mti.data(), W * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(edges, r->edges.data(), sizeof(double) * (r->wp.bins + 1),
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemsetAsync(lng, 0, sizeof(double) * W * r->wp.bins, st));
    HIP_OK(hipMemsetAsync(hist, 0, sizeof(double) * W * r->hstride, st));
    L->wl_first.push_back(L->wl_args.size());
    for (size_t k = 0; k < W; ++k) {
      const int win = (int)(k / (size_t)wpw);
      WlArgs a;
      memset(&a, 0, sizeof(a));
      a.g = g;
      a.off = r->ctx->d_off;
      a.vex = r->ctx->d_vex;
      a.compact = walkers + (size_t)g.compact_bytes * k;
      a.mt = mt + 624 * k;
      a.mt_index = mti + k;
      a.nbr_swap = r->wp.nbr_swap ? 1 : 0;
      a.mc_sweeps = r->wp.mc_sweeps;
      a.bins = r->wp.bins;
      a.start_idx = r->windows[(size_t)2 * win];
      a.end_idx = r->windows[(size_t)2 * win + 1];
      a.wl_f = r->wl_f;
      a.bin_edges = edges;
      a.lng = lng + (size_t)r->wp.bins * k;
      a.hist = hist + (size_t)r->hstride * k;
      a.accepted = acc + k;
      L->wl_args.push_back(a);
    }
  }
  HIP_OK(hipMemcpyAsync(L->scratch, L->wl_args.data(), sizeof(WlArgs) * total_w,
                        hipMemcpyHostToDevice, st));
  HIP_OK(hipEventCreate(&L->ev0));
  HIP_OK(hipEventCreate(&L->ev1));
  HIP_OK(hipEventRecord(L->ev0, st));
  // one pair-sum variant for the launch when every walker shares it
  int key = pair_sum_key(L->wl_args[0].g.lattice, L->wl_args[0].g.R);
  for (const WlArgs &a : L->wl_args)
    if (pair_sum_key(a.g.lattice, a.g.R) != key) key = -1;
  HIP_OK(launch_wl_sweeps((const WlArgs *)L->scratch, (int)total_w, lds, st, key));
  for (size_t i = 0; i < grp.size(); ++i) {
    Request *r = grp[i];
    const WlArgs &a0 = L->wl_args[L->wl_first[i]];
    HIP_OK(launch_wl_average(a0.lng, a0.hist, r->wp.num_windows, r->wp.walkers_per_window,
                             r->wp.bins, r->hstride, st));
  }
  HIP_OK(hipEventRecord(L->ev1, st));
  s->slot_busy[(size_t)slot] = true;
  s->running.push_back(std::move(L));
  return BWL_OK;
}

int fail_request(bwl_server *s, Request *r, int rc) {
  r->error = rc;
  r->state = ST_FAILED;
  r->t_done = now_ms(s->t0);
  release_request(s, r);
  s->done_unpolled.push_back(r->id);
  return BWL_OK;
}

// Admit queued requests while memory and slots last, grouped by variant.
// With a batching window, nothing is admitted until the oldest queued
// request has waited that long (so requests arriving close together share
// launches) unless no launch is running at all.
int admit(bwl_server *s) {
  if (s->queue.empty() || free_slot(s) < 0) return BWL_OK;
  if (s->batch_wait_ms > 0.0 && !s->running.empty() &&
      now_ms(s->t0) - s->queue.front()->t_submit < s->batch_wait_ms)
    return BWL_OK;
  // candidates in arrival order; a request that does not fit is skipped
  std::vector<Request *> taken;
  std::deque<Request *> keep;
  for (Request *r : s->queue) {
    r->block = s->arena.alloc(r->block_bytes);
    if (r->block)
      taken.push_back(r);
    else
      keep.push_back(r);
  }
  // groups: anneals by kernel variant, all nested-sampling runs together;
  // the earliest-arriving groups get the free slots
  std::vector<std::vector<Request *>> groups;
  for (Request *r : taken) {
    bool placed = false;
    for (auto &gr : groups)
      if (gr[0]->kind == r->kind && gr[0]->fn == r->fn) {
        gr.push_back(r);
        placed = true;
        break;
      }
    if (!placed) groups.push_back({r});
  }
  std::vector<Request *> back;  // not launched this time: return to the queue
  for (auto &gr : groups) {
    const int slot = free_slot(s);
    if (slot < 0) {
      back.insert(back.end(), gr.begin(), gr.end());
      continue;
    }
    hipStream_t st = s->streams[(size_t)slot];
    std::vector<Request *> ok;
    for (Request *r : gr) {
      int rc = stage_request(r, st);
      if (rc) {
        fail_request(s, r, rc);
        continue;
      }
      r->state = ST_RUNNING;
      r->t_admit = now_ms(s->t0);
      ok.push_back(r);
    }
    if (ok.empty()) continue;
    const int rc = ok[0]->kind == KIND_ANNEAL ? launch_anneal_group(s, ok, slot)
                   : ok[0]->kind == KIND_NS   ? launch_ns_group(s, ok, slot)
                                              : launch_wl_group(s, ok, slot);
    if (rc == BWL_ERR_NOMEM) {
      for (Request *r : ok) {
        release_request(s, r);
        r->state = ST_QUEUED;
        back.push_back(r);
      }
    } else if (rc != BWL_OK) {
      for (Request *r : ok) fail_request(s, r, rc);
    }
  }
  for (Request *r : back) {
    release_request(s, r);
    r->state = ST_QUEUED;
  }
  // queue order: arrival order among those still waiting
  std::vector<Request *> waiting(keep.begin(), keep.end());
  waiting.insert(waiting.end(), back.begin(), back.end());
  std::sort(waiting.begin(), waiting.end(),
            [](const Request *a, const Request *b) { return a->id < b->id; });
  s->queue.assign(waiting.begin(), waiting.end());
  return BWL_OK;
}

Request *new_request(bwl_server *s, int kind, const bwl_lattice_desc *desc, const double *v_ex,
                     int &rc) {
  auto r = std::make_unique<Request>();
  r->kind = kind;
  r->desc = *desc;
  std::vector<int32_t> off;
  rc = make_geo(*desc, r->g, off, g_err);
  if (rc) return nullptr;
  r->vex.assign(v_ex, v_ex + (size_t)r->g.S * r->g.S * r->g.R);
  r->id = s->next_id++;
  r->t_submit = now_ms(s->t0);
  Request *p = r.get();
  s->reqs[p->id] = std::move(r);
  return p;
}

}  // namespace

extern "C" {

int bwl_server_create(const bwl_server_params *p, bwl_server **out) {
  if (!p || !out) return fail(BWL_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->device_bytes < (1 << 20) || p->max_launches < 1 || p->max_launches > 64 ||
      p->batch_wait_us < 0)
    return fail(BWL_ERR_ARG,
                "device_bytes >= 1 MiB, 1 <= max_launches <= 64 and batch_wait_us >= 0 are required");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(BWL_ERR_DEVICE, "no HIP device available");
  auto s = std::make_unique<bwl_server>();
  if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
  if (hipMalloc(&s->arena_mem, (size_t)p->device_bytes) != hipSuccess)
    return fail(BWL_ERR_NOMEM, "device allocation failed");
  s->arena.init((char *)s->arena_mem, (size_t)p->device_bytes);
  for (int k = 0; k < p->max_launches; ++k) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      bwl_server_destroy(s.release());
      return fail(BWL_ERR_DEVICE, "stream creation failed");
    }
    s->streams.push_back(st);
    s->slot_busy.push_back(false);
  }
  s->batch_wait_ms = p->batch_wait_us / 1000.0;
  s->t0 = std::chrono::steady_clock::now();
  *out = s.release();
  return BWL_OK;
}

int bwl_server_destroy(bwl_server *s) {
  if (!s) return BWL_OK;
  (void)hipSetDevice(s->device);
  for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
  for (auto &L : s->running) {
    (void)hipEventDestroy(L->ev0);
    (void)hipEventDestroy(L->ev1);
  }
  for (auto &kv : s->reqs)
    if (kv.second->ctx) bwl_destroy(kv.second->ctx);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  delete s;
  return BWL_OK;
}

int bwl_server_submit_anneal(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                             const int8_t *config, const bwl_anneal_params *p, double *energies_of_T,
                             double *C_of_T, double *acceptance_of_T, int8_t *config_out,
                             int64_t *id) {
  if (!s || !desc || !v_ex || !config || !p || !energies_of_T || !C_of_T || !acceptance_of_T || !id)
    return fail(BWL_ERR_ARG, "null argument");
  if (p->T_steps < 1 || p->n_sweeps < 1 || p->burn_in_sweeps < 0)
    return fail(BWL_ERR_ARG, "T_steps and n_sweeps must be positive");
  int rc = BWL_OK;
  Request *r = new_request(s, KIND_ANNEAL, desc, v_ex, rc);
  if (!r) return rc;
  const Geo &g = r->g;
  rc = check_config(g, config, g_err);
  if (rc) {
    s->reqs.erase(r->id);
    return rc;
  }
  ParPlan plan;
  NatGeo nat;
  int threads = 0;
  size_t lds = 0;
  rc = anneal_kernel(g, plan, nat, r->fn, threads, lds);
  if (rc) {
    s->reqs.erase(r->id);
    return rc;
  }
  r->config.assign(config, config + (size_t)g.m[0] * g.m[1] * g.m[2]);
  r->ap = *p;
  r->E = energies_of_T;
  r->Cv = C_of_T;
  r->acc = acceptance_of_T;
  r->config_out = config_out;
  r->block_bytes = ctx_device_bytes(*desc);
  if (r->block_bytes > s->arena.size) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_NOMEM, "request larger than the server's device memory");
  }
  s->queue.push_back(r);
  *id = r->id;
  return BWL_OK;
}

int bwl_server_submit_ns(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                         uint32_t seed, const bwl_ns_params *p, const int64_t *species_numbers,
                         const double *concentrations, double *e_limits, int64_t *id) {
  if (!s || !desc || !v_ex || !p || !e_limits || !id) return fail(BWL_ERR_ARG, "null argument");
  if (p->n_walkers < 2 || p->n_steps < 1 || p->n_iter < 1)
    return fail(BWL_ERR_ARG, "n_walkers >= 2, n_steps >= 1 and n_iter >= 1 are required");
  int rc = BWL_OK;
  Request *r = new_request(s, KIND_NS, desc, v_ex, rc);
  if (!r) return rc;
  const Geo &g = r->g;
  if (ns_lds_bytes(g, p->n_walkers) > (size_t)lds_limit() || 2 * g.R > 64) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_UNSUPPORTED, "lattice and walker energies do not fit in LDS");
  }
  // walkers from the run's own MT19937 stream (nested_sampling.f90:88-96)
  MT19937 mt;
  mt.seed(seed);
  const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
  r->walkers_ref.assign(ref * (size_t)p->n_walkers, 0);
  r->rnde.resize((size_t)p->n_walkers);
  bool two = false;
  for (int w = 0; w < p->n_walkers; ++w) {
    rc = initial_setup_host(g, species_numbers, concentrations, mt, r->walkers_ref.data() + ref * w,
                            g_err);
    if (rc) {
      s->reqs.erase(r->id);
      return rc;
    }
    r->rnde[(size_t)w] = mt.genrand();
    if (w == 0) {
      int first = 0;
      for (size_t i = 0; i < ref && !two; ++i) {
        const int v = r->walkers_ref[i];
        if (!v) continue;
        if (!first) first = v;
        else if (v != first) two = true;
      }
    }
  }
  if (!two) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_ARG, "nested sampling needs at least two species present");
  }
  memcpy(r->mt, mt.mt, sizeof(r->mt));
  r->mti = mt.mti;
  r->np = *p;
  r->e_limits = e_limits;
  r->block_bytes = ns_block_bytes(*r);
  if (r->block_bytes > s->arena.size) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_NOMEM, "request larger than the server's device memory");
  }
  s->queue.push_back(r);
  *id = r->id;
  return BWL_OK;
}

int bwl_server_submit_wl(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                         const bwl_wl_params *p, const int64_t *species_numbers,
                         const double *concentrations, double wl_f, double *lng, double *hist,
                         int64_t *id) {
  if (!s || !desc || !v_ex || !p || !lng || !hist || !id) return fail(BWL_ERR_ARG, "null argument");
  if (p->bins < 1 || p->num_windows < 1 || p->walkers_per_window < 1 || p->mc_sweeps < 1 ||
      p->num_windows > p->bins || !(p->energy_max > p->energy_min))
    return fail(BWL_ERR_ARG, "invalid Wang-Landau parameters");
  int rc = BWL_OK;
  Request *r = new_request(s, KIND_WL, desc, v_ex, rc);
  if (!r) return rc;
  const Geo &g = r->g;
  r->wp = *p;
  r->wl_f = wl_f;
  r->windows.resize((size_t)2 * p->num_windows);
  r->edges.resize((size_t)p->bins + 1);
  wl_setup_host(g, p->bins, p->num_windows, p->bin_overlap, p->energy_min, p->energy_max,
                r->windows.data(), r->edges.data());
  for (int k = 0; k < p->num_windows; ++k) {
    const int a = r->windows[(size_t)2 * k], b = r->windows[(size_t)2 * k + 1];
    if (a < 1 || b > p->bins || b < a) {
      s->reqs.erase(r->id);
      return fail(BWL_ERR_ARG, "window layout out of range (too many windows for the bins)");
    }
    r->hstride = std::max(r->hstride, b - a + 1);
  }
  if (wl_lds_bytes(g, p->bins, r->hstride) > (size_t)lds_limit() || 2 * g.R > 64) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_UNSUPPORTED, "lattice and ln g do not fit in LDS");
  }
  // each walker from its own stream, seeded as rank rank_offset + k
  const size_t W = (size_t)p->num_windows * p->walkers_per_window;
  const size_t ref = (size_t)g.m[0] * g.m[1] * g.m[2];
  r->walkers_ref.assign(ref * W, 0);
  r->wl_mt.resize(W * 624);
  r->wl_mti.resize(W);
  for (size_t k = 0; k < W; ++k) {
    MT19937 mt;
    mt.seed(110179u + 11u * (p->rank_offset + (uint32_t)k));
    rc = initial_setup_host(g, species_numbers, concentrations, mt, r->walkers_ref.data() + ref * k,
                            g_err);
    if (rc) {
      s->reqs.erase(r->id);
      return rc;
    }
    memcpy(&r->wl_mt[624 * k], mt.mt, 624 * 4);
    r->wl_mti[k] = mt.mti;
  }
  r->lng_out = lng;
  r->hist_out = hist;
  r->block_bytes = wl_block_bytes(*r);
  if (r->block_bytes > s->arena.size) {
    s->reqs.erase(r->id);
    return fail(BWL_ERR_NOMEM, "request larger than the server's device memory");
  }
  s->queue.push_back(r);
  *id = r->id;
  return BWL_OK;
}

int bwl_server_progress(bwl_server *s, int32_t *n_running, int32_t *n_queued) {
  if (!s) return fail(BWL_ERR_ARG, "null server");
  HIP_OK(hipSetDevice(s->device));
  int status = BWL_OK;
  for (size_t k = 0; k < s->running.size();) {
    const hipError_t q = hipEventQuery(s->running[k]->ev1);
    if (q == hipErrorNotReady) {
      ++k;
      continue;
    }
    if (q != hipSuccess) return fail(BWL_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(q));
    const int rc = retire(s, *s->running[k]);
    if (rc != BWL_OK) status = rc;
    s->running.erase(s->running.begin() + (long)k);
  }
  const int rc = admit(s);
  if (rc != BWL_OK) status = rc;
  if (n_running) {
    int nr = 0;
    for (auto &L : s->running) nr += (int)L->reqs.size();
    *n_running = nr;
  }
  if (n_queued) *n_queued = (int32_t)s->queue.size();
  return status;
}

int bwl_server_poll(bwl_server *s, int64_t *ids, int32_t max_ids, int32_t *n) {
  if (!s || !n || (max_ids > 0 && !ids)) return fail(BWL_ERR_ARG, "null argument");
  const int32_t k = (int32_t)std::min<size_t>((size_t)std::max(max_ids, 0), s->done_unpolled.size());
  for (int32_t i = 0; i < k; ++i) ids[i] = s->done_unpolled[(size_t)i];
  s->done_unpolled.erase(s->done_unpolled.begin(), s->done_unpolled.begin() + k);
  *n = k;
  return BWL_OK;
}

int bwl_server_info(bwl_server *s, int64_t id, bwl_request_info *info) {
  if (!s || !info) return fail(BWL_ERR_ARG, "null argument");
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(BWL_ERR_ARG, "unknown request id");
  const Request &r = *it->second;
  info->kind = r.kind;
  info->state = r.state;
  info->error = r.error;
  info->reserved = 0;
  info->trials = r.trials;
  info->accepted = r.accepted;
  info->t_submit_ms = r.t_submit;
  info->t_admit_ms = r.t_admit;
  info->t_done_ms = r.t_done;
  info->kernel_ms = r.kernel_ms;
  return BWL_OK;
}

int bwl_server_drain(bwl_server *s) {
  if (!s) return fail(BWL_ERR_ARG, "null server");
  int status = BWL_OK;
  int32_t nr = 0, nq = 0;
  do {
    const int rc = bwl_server_progress(s, &nr, &nq);
    if (rc != BWL_OK) {
      status = rc;
      if (rc == BWL_ERR_DEVICE) return rc;
    }
    if (nr > 0 || nq > 0) {
      // wait for the oldest launch instead of spinning
      if (!s->running.empty()) (void)hipEventSynchronize(s->running.front()->ev1);
      else if (nq > 0 && s->running.empty()) break;  // nothing runs and nothing fits
    }
  } while (nr > 0 || nq > 0);
  if (nq > 0) return fail(BWL_ERR_NOMEM, "queued requests can never be admitted");
  return status;
}

int bwl_server_memory(const bwl_server *s, int64_t *used_bytes, int64_t *total_bytes) {
  if (!s) return fail(BWL_ERR_ARG, "null server");
  if (used_bytes) *used_bytes = (int64_t)s->arena.used;
  if (total_bytes) *total_bytes = (int64_t)s->arena.size;
  return BWL_OK;
}

}  // extern "C"
