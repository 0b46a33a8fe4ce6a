// bwl_kernels.h -- kernel argument blocks and launchers (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "bwl_core.h"

namespace bwl {

struct HistArgs {
  Geo g;
  const int8_t *compact;  // lattice, compact layout
  const int32_t *off;     // [nsub][z] padded-layout neighbour offsets
  long long *hist;        // H(S,S,R), Fortran layout
};

struct SiteArgs {
  Geo g;
  const int8_t *compact;
  const int32_t *off;
  const double *vex;  // V_ex(S,S,R)
  int x[2], y[2], z[2];  // 0-based doubled coordinates of two sites
  double *out;           // out[0] = e(site 0), out[1] = e(site 0) + e(site 1)
};

struct SeqArgs {
  Geo g;
  int8_t *compact;
  const int32_t *off;
  const double *vex;
  uint32_t *mt;        // 624 words, in/out
  int32_t *mt_index;   // in/out
  double beta;
  int32_t nbr_swap;
  long long n_trials;
  long long sample_every;
  double *energies;    // optional, n_trials / sample_every values
  long long *accepted;
};

// Proposal draws from a context's MT19937 stream (random_site.f90:127-225,
// wang-landau.F90:1527-1539), one wavefront:
//   mode 0: n sites (rdm_site), mode 1: one neighbour of `site` (rdm_nbr),
//   mode 2: n pairs (rdm_site, rdm_site), mode 3: n pairs (rdm_site, rdm_nbr).
// Sites are written as 4 int32 (b, i, j, k), 1-based.
struct DrawArgs {
  Geo g;
  uint32_t *mt;        // 624 words, in/out
  int32_t *mt_index;   // in/out
  int32_t mode, n;
  int32_t site[4];     // mode 1 input
  int32_t *out;        // 4 * n (modes 0, 1) or 8 * n (modes 2, 3) values
};

struct ParArgs {
  Geo g;
  NatGeo nat;            // used by the natural-layout kernel
  int8_t *compact;       // in/out lattice
  const int32_t *off;
  const double *vex;
  // rounds: two interleaved active grids (see RoundVar)
  int32_t stride;        // even stride of both grids on the two free axes
  int32_t L;             // layer period on the round's axis
  int32_t K[3];          // stride-`stride` grid positions per axis
  int32_t n_var;         // usable round kinds
  RoundVar var[kMaxRoundVars];
  const uint32_t *muls;  // per round kind: multipliers coprime with its nL1
  uint32_t units[3][16]; // per axis: multipliers coprime with K[axis]
  int32_t n_units[3];
  int32_t fast_halo;     // n >= 2h on every axis
  long long min_trials_per_sweep;  // trials per sweep (>= n_atoms)
  int32_t threads;       // threads that do the chain's work (<= block size)
  // schedule
  double T0, dT;
  int32_t T_steps, n_sweeps, burn_sweeps;
  uint32_t seed_lo, seed_hi, chain_id;
  unsigned long long phase0;  // first round counter (resumable streams)
  // outputs
  double *energies_of_T, *C_of_T, *acc_of_T;
  double *e_drift;       // 2 x T_steps: per T step, the carried energy minus its recount at
                         // the step's end, then the recounts
  double *e_trace;       // optional [T_steps][n_sweeps]
  int32_t asro_every;    // sampled sweeps between radial-density samples (0: none)
  double *rho_of_T;      // (S, S, R+1, T_steps) when asro_every > 0
  long long *counters;   // [0] = next round counter, [1] += accepted, [2] += trials
};

// Nested sampling (nested_sampling.f90:45-206), one wavefront per run.
struct NsArgs {
  Geo g;
  const int32_t *off;
  const double *vex;
  int8_t *walkers;        // n_walkers compact lattices, g.compact_bytes apart
  const double *rnde;     // per walker: the genrand() drawn after its set-up
  double *E;              // walker energies: initial (k_ns_energies), then final
  uint32_t *mt;           // 624 words, in/out
  int32_t *mt_index;      // in/out
  int32_t n_walkers, n_steps, n_iter, n_at;
  long long max_redraws;  // bound on second-site redraws of one trial
  double *e_limits;       // n_iter energy limits
  long long *counters;    // [0] trials, [1] accepted, [2] status (0 ok, 1 redraw bound hit)
};

// One `sweeps` call of one Wang-Landau walker (wang-landau.F90:539-632).
struct WlArgs {
  Geo g;
  const int32_t *off;
  const double *vex;
  int8_t *compact;
  uint32_t *mt;
  int32_t *mt_index;
  int32_t nbr_swap, mc_sweeps, bins, start_idx, end_idx;  // 1-based window bins
  double wl_f;
  const double *bin_edges;  // bins + 1
  double *lng;              // bins, in/out
  double *hist;             // end_idx - start_idx + 1, in/out
  long long *accepted;      // out
};

// replica_exchange (wang-landau.F90:1392-1501) over a whole Wang-Landau
// walker set: walker k stands for MPI rank k.
struct WlExchangeArgs {
  const double *E;          // walker energies in the reference's order, W
  const double *bin_edges;  // bins + 1
  const int32_t *windows;   // num_windows (start, end) pairs, 1-based bins
  const double *lng;        // W rows of bins
  uint32_t *mt;             // W x 624, in/out
  int32_t *mt_index;        // W, in/out
  int32_t *partner;         // W out: walker matched with, -1 if none
  int32_t *accepted;        // W out: 1 if the matched pair exchanged
  int32_t *swaps;           // out: exchanged pairs (a, b), at most W/2
  int32_t *n_swaps;         // out
  int32_t W, wpw, num_windows, bins;
};

// enter_energy_window (wang-landau.F90:643-741) of one walker, resumable:
// the kernel stops when the walker is inside its window (status 0), when
// the reference would re-run initial_setup (status 1; the host does that
// and relaunches with resume = 1), or after `budget` trials (status 2).
struct WlEnterArgs {
  Geo g;
  const int32_t *off;
  const double *vex;
  int8_t *compact;
  uint32_t *mt;
  int32_t *mt_index;
  double min_e, max_e, denom;
  long long period;        // n_atoms * 250 iterations between re-set-ups
  long long budget;        // trials this launch may make
  int32_t init_e;          // recount the energy first
  int32_t *resume;         // in/out: skip the first iteration's count
  double *e_unswapped;     // in/out
  long long *i_steps;      // in/out
  long long *trials;       // in/out, cumulative
  int32_t *status;         // out
};

// Windowed pair-count histogram of radial_densities (analytics.f90:293-404).
struct AsroArgs {
  Geo g;
  const int8_t *compact;
  int32_t wc;                 // shells including the on-site one
  int32_t shell_start[17];    // offsets of shell l: [shell_start[l], shell_start[l+1])
  const int32_t *off;         // 3 ints (dx, dy, dz) per entry, doubled-grid units
  unsigned long long *hist;   // H(S, S, wc), zeroed by the caller
};
constexpr int kMaxAsroShells = 16;

hipError_t launch_pack(const int8_t *ref, int8_t *compact, const Geo &g, hipStream_t s);
// `count` reference-layout lattices back to back -> compact lattices back to back
hipError_t launch_pack_many(const int8_t *ref, int8_t *compact, const Geo &g, int count,
                            hipStream_t s);
hipError_t launch_unpack(const int8_t *compact, int8_t *ref, const Geo &g, hipStream_t s);
// LDS read-bandwidth probe: blocks x 1024 threads, iters 16-byte reads each
hipError_t launch_lds_bandwidth(int blocks, int iters, uint32_t *sink, hipStream_t s);
size_t hist_lds_bytes(const Geo &g, int threads);
hipError_t launch_histogram(const HistArgs *args_dev, int n, const Geo &g_max, int threads,
                            hipStream_t s);
size_t site_lds_bytes(const Geo &g);
hipError_t launch_site_energy(const SiteArgs &a, hipStream_t s);
size_t seq_lds_bytes(const Geo &g);
hipError_t launch_metro_seq(const SeqArgs &a, hipStream_t s);
hipError_t launch_draw_sites(const DrawArgs &a, hipStream_t s);
// Sets n 64-bit words to zero with a kernel (no memset node in a captured
// graph).
hipError_t launch_zero_i64(long long *p, int n, hipStream_t s);
// n chains (args_dev[0..n)), one wavefront each; key as for launch_wl_sweeps
hipError_t launch_metro_seq_batch(const SeqArgs *args_dev, int n, size_t lds, hipStream_t s,
                                  int key);
size_t par_lds_bytes(const Geo &g, const NatGeo *nat, int threads);
// Block size limit of the parallel-sampler kernels: the natural-layout
// kernels with compile-time neighbour offsets and 32-bit count words run 16
// wavefronts per CU (128 registers each; what the compiler then spills is
// in the radial-density reduction and the per-round set-up, not the pair
// loop), the others are limited to 8.
constexpr int kMaxNatThreads = 1024;
// fixed: the fixed-pitch natural layout (nat_dims_fixed)
int par_max_threads(const Geo &g, bool nat, bool fixed);
// packed count words of the sampler for a lattice (see par_count_words in
// bwl_metro_par.hip): 0 = 32-bit / 8-bit fields, 1 = 32-bit / 6-bit, 2 = 64-bit
int par_count_words(const Geo &g, bool nat);
const void *par_kernel_for(const Geo &g, bool nat, bool fixed);
// whether a kernel with compile-time offsets exists for this exact shape
bool par_specialised(const Geo &g);
// the fixed-pitch kernel for (lattice, R, count words), nullptr if none
const void *par_fixed_kernel(int lattice, int R, int words);
hipError_t launch_metro_par_fn(const void *fn, const ParArgs *args_dev, int n_chains, size_t lds,
                               int threads, hipStream_t s);

size_t ns_lds_bytes(const Geo &g, int n_walkers);
// Initial walker energies (all walkers at once), then the runs (one
// wavefront each).
// key: the pair-sum variant shared by every run (pair_sum_key), -1 if mixed
hipError_t launch_ns(const NsArgs *args_dev, int n, size_t lds, int max_walkers, size_t lds_energies,
                     hipStream_t s, int key);
size_t ns_energies_lds_bytes(const Geo &g);
size_t wl_lds_bytes(const Geo &g, int bins, int window_bins);
// key: the pair-sum variant shared by every walker (pair_sum_key), -1 if mixed
hipError_t launch_wl_sweeps(const WlArgs *args_dev, int n, size_t lds, hipStream_t s, int key);
hipError_t launch_wl_average(double *lng, double *hist, int n_windows, int walkers_per_window,
                             int bins, int hist_stride, hipStream_t s);
// Walker energies (ns_energies_lds_bytes of LDS), the exchange decisions
// (one wavefront) and the exchange of configurations.
hipError_t launch_wl_energies(const WlArgs *args_dev, int n, double *E, size_t lds, hipStream_t s);
size_t wl_exchange_lds_bytes(int W, int wpw);
hipError_t launch_wl_exchange(const WlExchangeArgs &a, size_t lds, hipStream_t s);
hipError_t launch_wl_swap_configs(int8_t *walkers, long long bytes, const int32_t *swaps,
                                  const int32_t *n_swaps, int max_pairs, hipStream_t s);
size_t wl_enter_lds_bytes(const Geo &g);
hipError_t launch_wl_enter(const WlEnterArgs *args_dev, int n, size_t lds, hipStream_t s, int key);
hipError_t launch_asro(const AsroArgs &a, hipStream_t s);

}  // namespace bwl
