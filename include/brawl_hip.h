/*
 * brawl_hip.h -- C-ABI of the MI355X (gfx950) lattice-sampling library.
 *
 * Plain C types only (no torch, no HIP types).  Every entry point returns an
 * int status (BWL_OK == 0); bwl_last_error() gives the message of the last
 * failure on the calling thread.  A context (bwl_ctx) is one lattice chain
 * with its device-resident state and its own HIP stream; contexts are not
 * re-entrant (one host thread per context at a time).
 *
 * Layouts at the boundary are the reference's own:
 *   config  int8 config(1, 2*n1, 2*n2, 2*n3), Fortran column-major
 *           (shared_data.f90:30, allocated at initialise.F90:295);
 *           element (1,i,j,k), 1-based, at (i-1) + 2*n1*((j-1) + 2*n2*(k-1)).
 *   v_ex    FP64 V_ex(S, S, R), Fortran column-major, Rydberg
 *           (shared_data.f90:37, read at io.f90:399-402).
 *   hist    int64 H(S, S, R), same layout: ordered (site, listed neighbour)
 *           pairs with species (a, b) on shell n.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef BRAWL_HIP_H
#define BRAWL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BWL_OK 0
#define BWL_ERR_ARG (-1)         /* invalid argument or descriptor */
#define BWL_ERR_UNSUPPORTED (-2) /* lattice/shell/species combination not built */
#define BWL_ERR_DEVICE (-3)      /* HIP runtime error */
#define BWL_ERR_NOMEM (-4)       /* device allocation failed */
#define BWL_ERR_STATE (-5)       /* call order (e.g. no configuration uploaded) */

#define BWL_FCC 1
#define BWL_BCC 2

/* Lattice descriptor: the fields of run_params the path reads
 * (derived_types.f90:56-100: lattice, n_1..n_3, n_species,
 * interaction_range). */
typedef struct {
  int32_t lattice;   /* BWL_FCC or BWL_BCC */
  int32_t n1, n2, n3; /* conventional cells per axis */
  int32_t n_species;  /* S, 1..8 */
  int32_t n_shells;   /* R: fcc 1..6, bcc 1..10 */
} bwl_lattice_desc;

typedef struct bwl_ctx bwl_ctx;

/* ---------------------------------------------------------------- */
/* set-up                                                            */
/* ---------------------------------------------------------------- */
const char *bwl_last_error(void);
int bwl_device_count(int32_t *count);

/* Replaces initialise_interaction + initialise_function_pointers +
 * initialise_local_arrays (initialise.F90:153-298): allocates the device
 * lattice for `desc` and uploads V_ex. */
int bwl_create(const bwl_lattice_desc *desc, const double *v_ex, bwl_ctx **out);
int bwl_destroy(bwl_ctx *ctx);
/* Number of occupied sites (run_params%n_atoms, initialise.F90:174-212) */
int bwl_n_atoms(const bwl_ctx *ctx, int32_t *n_atoms);
/* Upload / download the reference-layout config (config array,
 * shared_data.f90:30). */
int bwl_set_config(bwl_ctx *ctx, const int8_t *config);
int bwl_get_config(bwl_ctx *ctx, int8_t *config);

/* ---------------------------------------------------------------- */
/* Hamiltonian (bw_hamiltonian.f90)                                  */
/* ---------------------------------------------------------------- */
/* Integer pair-count histogram over the reference neighbour lists,
 * counted on the device. */
int bwl_pair_histogram(bwl_ctx *ctx, int64_t *hist);
/* `hamiltonian` interface, bound as setup%full_energy (derived_types.f90:
 * 373-384; total_energy, bw_hamiltonian.f90:58-81): 0.5*sum H*V_ex
 * contracted in FP64 from the device histogram. */
int bwl_total_energy(bwl_ctx *ctx, double *energy);
/* `neighbour` interface, bound as setup%nbr_energy (derived_types.f90:
 * 401-413; {fcc,bcc}_energy_Rshells, bw_hamiltonian.f90:898-1910).
 * Evaluated on the device in the reference's FP64 summation order, so the
 * result is bit-identical.  site_b/i/j/k are 1-based. */
int bwl_site_energy(bwl_ctx *ctx, int32_t site_b, int32_t site_i,
                    int32_t site_j, int32_t site_k, double *energy);
/* pair_energy (bw_hamiltonian.f90:99-114); sites are int(4) (b,i,j,k). */
int bwl_pair_energy(bwl_ctx *ctx, const int32_t *site1, const int32_t *site2,
                    double *energy);

/* ---------------------------------------------------------------- */
/* Reference random stream (mt19937ar.c:58-142) -- sequential mode   */
/* ---------------------------------------------------------------- */
/* init_genrand(seed) on the context's MT19937 state. */
int bwl_genrand_seed(bwl_ctx *ctx, uint32_t seed);
/* Copy the 624-word state and index (the globals mt[], mti of
 * mt19937ar.c:58-59) in or out, so a host that also draws numbers stays in
 * step with the device. */
int bwl_genrand_get_state(bwl_ctx *ctx, uint32_t *mt, int32_t *mti);
int bwl_genrand_set_state(bwl_ctx *ctx, const uint32_t *mt, int32_t mti);
/* `rand_site` interface, bound as setup%rdm_site (derived_types.f90:
 * 424-434; bcc_random_site / fcc_random_site, random_site.f90:127-139,
 * 182-193): n consecutive draws from the context's MT19937 stream, each
 * made as the reference makes it (genrand for k, then i, then j).  sites
 * receives int32 (4, n): (b, i, j, k) per site, 1-based. */
int bwl_random_site(bwl_ctx *ctx, int32_t n, int32_t *sites);
/* `rand_neighbour` interface, bound as setup%rdm_nbr (derived_types.f90:
 * 447-458; bcc_random_nbr / fcc_random_nbr, random_site.f90:152-171,
 * 206-225): one draw from the context's stream picks a shell-1 neighbour
 * of `site` (b, i, j, k), wrapped into the box; nbr receives 4 int32. */
int bwl_random_nbr(bwl_ctx *ctx, const int32_t *site, int32_t *nbr);
/* wl_params%mc_select (derived_types.f90:349, bound at wang-landau.F90:
 * 110-114; select_sites / select_sites_nbr, :1527-1539): n consecutive
 * site pairs (rdm1 = rdm_site, rdm2 = rdm_site, or rdm_nbr(rdm1) when
 * nbr_swap != 0) from the context's stream; sites receives int32
 * (4, 2, n). */
int bwl_select_sites(bwl_ctx *ctx, int32_t nbr_swap, int32_t n, int32_t *sites);
/* initial_setup (initialise.F90:434-617) from the context's stream: uses
 * species_numbers (length S) when they sum to n_atoms, else
 * concentrations (length S).  Result is uploaded as the configuration. */
int bwl_initial_setup(bwl_ctx *ctx, const int64_t *species_numbers,
                      const double *concentrations);

/* `monte_carlo` interface, bound as setup%mc_step (derived_types.f90:
 * 474-486; monte_carlo_step_lattice metropolis.F90:751-813, or
 * monte_carlo_step_nbr :829-891 when nbr_swap != 0): n_trials sequential
 * trials on the device consuming the context's MT19937 stream exactly as
 * the reference does.  accepted counts trials returning 1.  If energies is
 * not NULL, the total energy is evaluated after every `sample_every`
 * trials (metropolis.F90:357-368) and written there
 * (n_trials / sample_every values). */
int bwl_mc_steps(bwl_ctx *ctx, double beta, int32_t nbr_swap, int64_t n_trials,
                 int64_t sample_every, double *energies, int64_t *accepted);
/* Batched form of bwl_mc_steps: n contexts on one device, each making
 * n_trials sequential trials at beta[i] on its own MT19937 stream, in one
 * launch (one wavefront per context) -- the reference's MPI replica
 * ensemble (comms.F90:122-160, one chain per rank), every chain
 * bit-identical to the same context's bwl_mc_steps.  accepted (n values),
 * energies (optional: n pointers of n_trials/sample_every values, NULL
 * entries skipped) and *kernel_ms (optional, device time) are outputs. */
int bwl_mc_steps_batch(int32_t n, bwl_ctx *const *ctxs, const double *beta, int32_t nbr_swap,
                       int64_t n_trials, int64_t sample_every, double *const *energies,
                       int64_t *accepted, double *kernel_ms);

/* ---------------------------------------------------------------- */
/* Parallel Metropolis sampler (statistical parity)                  */
/* ---------------------------------------------------------------- */
/* Replaces the temperature loop of metropolis_simulated_annealing
 * (metropolis.F90:193-437) for lattice-wide swaps.  Each sweep makes at
 * least n_atoms swap trials as rounds of simultaneous, non-interacting
 * pair trials (sites of a round are further apart than the interaction
 * range); proposals and acceptances use Philox4x32-10 keyed by `seed`,
 * continuing the context's stream position across calls.  The energy is
 * recounted from the lattice before the first T step and at the end of
 * each, and carried through the sweeps as the sum of the accepted dE,
 * sampled after every sweep (stats: the carried value's largest difference
 * from the recounts); outputs per T step follow metropolis.F90:414-437 (<E>/N, C,
 * acceptance rate with equal-species trials counted as accepted). */
typedef struct {
  double T;            /* start temperature (K) */
  double delta_T;      /* step (K) */
  int32_t T_steps;
  int32_t n_sweeps;    /* sweeps per T step (samples per T step) */
  int32_t burn_in_sweeps; /* unsampled sweeps before each T step's samples */
  int32_t asro_every;  /* radial-density sample every asro_every sampled sweeps
                          (0: none; bwl_metropolis_anneal_rho only) */
  uint64_t seed;
} bwl_anneal_params;

typedef struct {
  int64_t trials;        /* swap trials made (all T steps) */
  int64_t accepted;      /* accepted, equal species counted as accepted */
  double kernel_ms;      /* device time of the sampler kernel */
  /* Parallel sampler only (0 elsewhere): the carried energy (recount at the
   * start, plus the accepted dE) against an FP64 recount of the lattice at
   * the end of every T step, the largest |difference| in Ry and the largest
   * |difference| / |recount|. */
  double e_drift_abs, e_drift_rel;
} bwl_run_stats;

int bwl_metropolis_anneal(bwl_ctx *ctx, const bwl_anneal_params *p,
                          double *energies_of_T, double *C_of_T,
                          double *acceptance_of_T, bwl_run_stats *stats);

/* As bwl_metropolis_anneal, also averaging the radial densities
 * (analytics.f90:293-404; metropolis.F90:377-384, 439-441) over samples
 * taken every p->asro_every sampled sweeps (1 <= asro_every <= n_sweeps):
 * rho_of_T is FP64 (S, S, R+1, T_steps), Fortran layout, the reference's
 * rho_of_T with wc_range = R+1 (shell 0 on-site, shells 1..R the
 * interaction shells).  Counted from the device's integer shell pair
 * counts; equal to radial_densities when every shell offset lies in the
 * reference's +-min(n,5) window (BWL_ERR_UNSUPPORTED otherwise, and for
 * supercells that are not cubic). */
int bwl_metropolis_anneal_rho(bwl_ctx *ctx, const bwl_anneal_params *p, double *energies_of_T,
                              double *C_of_T, double *acceptance_of_T, double *rho_of_T,
                              bwl_run_stats *stats);

/* No counterpart in the reference: bwl_metropolis_anneal for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  Same arguments, kernels and results as
 * bwl_metropolis_anneal on the same context, seed and stream position; the
 * call's scratch (round multipliers, per-T outputs, counters, launch
 * arguments) is a 1 MiB array in device memory that the library holds per
 * GPU until the process ends, so it makes no hipMalloc, hipFree or other
 * allocating call.  Inputs whose scratch exceeds 1 MiB (about 25,000 T
 * steps) fail with BWL_ERR_ARG before anything runs.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several host
 * threads or contexts at once) race; calls made one after another do not. */
int bwl_metropolis_anneal_noalloc(bwl_ctx *ctx, const bwl_anneal_params *p,
                                  double *energies_of_T, double *C_of_T,
                                  double *acceptance_of_T, bwl_run_stats *stats);

/* Layout of the parallel sampler's lattice in LDS, the tunable part of the
 * kernel choice (no counterpart in the reference).  BWL_LAYOUT_AUTO (the
 * default) takes the first of FIXED, NATURAL, COMPACT that exists for the
 * lattice and fits in LDS.  NATURAL and FIXED run the same sampler on the
 * lattice held in lattice coordinates (NATURAL with this size's own row
 * pitch, or the kernel compiled for this exact shape where there is one;
 * FIXED with a pitch shared by all sizes of the lattice type and shell
 * count and compile-time neighbour offsets) and give identical results;
 * COMPACT holds the lattice per sublattice and draws its random numbers per
 * pair, so its chain differs.  bwl_set_layout applies to the context's
 * later anneals (an unavailable layout fails there with
 * BWL_ERR_UNSUPPORTED); bwl_layouts sets bit (layout - 1) of *mask for
 * every layout available to the context's lattice; bwl_layout_info gives
 * the threads per chain and the LDS bytes of one layout (AUTO: the rule's
 * pick), or BWL_ERR_UNSUPPORTED.  A chain's random numbers are keyed by
 * its thread and step: layouts run with the same thread count give the
 * same chain. */
#define BWL_LAYOUT_AUTO 0
#define BWL_LAYOUT_COMPACT 1
#define BWL_LAYOUT_NATURAL 2
#define BWL_LAYOUT_FIXED 3
int bwl_set_layout(bwl_ctx *ctx, int32_t layout);
int bwl_layouts(bwl_ctx *ctx, int32_t *mask);
int bwl_layout_info(bwl_ctx *ctx, int32_t layout, int32_t *threads, int64_t *lds_bytes);

/* Threads per chain of the parallel sampler for the context's later
 * anneals (no counterpart in the reference: a tuning knob of this
 * sampler): 0 (the default) takes the rule's count (the fewest threads
 * that need the fewest passes over a round's pairs); otherwise a multiple
 * of 64 up to 1024, else BWL_ERR_ARG.  With BWL_LAYOUT_AUTO the first
 * layout that can run that many threads is taken; an anneal whose layout
 * cannot (the fixed-pitch kernels take up to 1024 threads, the others up
 * to 512) fails with BWL_ERR_UNSUPPORTED.  The thread count is part of a
 * chain's definition (its random numbers are keyed by thread and step):
 * equal thread counts give the same chain in either lattice-coordinate
 * layout, different counts give different, equally exact chains. */
int bwl_set_threads(bwl_ctx *ctx, int32_t threads);

/* Batched form: n independent chains (contexts on one device, uneven
 * lattice sizes, compositions and schedules) annealed in as few launches
 * as possible -- one workgroup per chain, chains with the same kernel
 * variant share a launch, variants run concurrently.  Per-chain outputs
 * and stats are as for bwl_metropolis_anneal; *makespan_ms (optional) is
 * the device time of the whole batch.  Each chain's random stream depends
 * only on its seed and its own position in that stream, so a chain gives
 * the same result alone and in any batch.  This is the MPI replica
 * ensemble of the reference (comms.F90:122-160, one chain per rank) held
 * on one GPU. */
int bwl_metropolis_anneal_batch(int32_t n, bwl_ctx *const *ctxs,
                                const bwl_anneal_params *params,
                                double *const *energies_of_T,
                                double *const *C_of_T,
                                double *const *acceptance_of_T,
                                bwl_run_stats *stats, double *makespan_ms);

/* ---------------------------------------------------------------- */
/* Radial densities (analytics.f90:205-404)                          */
/* ---------------------------------------------------------------- */
/* lattice_shells (analytics.f90:205-275): the first wc_range distinct
 * REAL(4) distances (doubled-grid units) from site (1,1,1) to the occupied
 * sites, 0 first.  wc_range 1..16. */
int bwl_lattice_shells(bwl_ctx *ctx, int32_t wc_range, double *shells);
/* radial_densities (analytics.f90:293-404) of the context's configuration:
 * rho(a, b, l) = number of species-b sites at shell-l distance (minimum
 * image, inside the reference's +-min(n,5) doubled-cell window, counted
 * with the window's multiplicity) from species-a sites, divided by the
 * number of species-a sites.  rho is FP64 (S, S, wc_range), Fortran
 * layout.  Pair counts are integer on the device, so the result is
 * bit-identical to the reference's. */
int bwl_radial_densities(bwl_ctx *ctx, int32_t wc_range, double *rho);

/* ---------------------------------------------------------------- */
/* Nested sampling (nested_sampling.f90:45-206)                      */
/* ---------------------------------------------------------------- */
typedef struct {
  int32_t n_walkers;  /* >= 2 */
  int32_t n_steps;    /* walk length (trials per iteration before adaptation) */
  int32_t n_iter;     /* iterations = energy limits written */
  int32_t reserved;
} bwl_ns_params;

/* The serial nested-sampling run of the reference on the device: walkers
 * are made by initial_setup from the context's MT19937 stream (plus one
 * draw each for the energy jitter), then each iteration discards the
 * highest-energy walker, clones a random one over it and walks it with
 * lattice-wide swaps of different species accepted iff E + dE < E_limit,
 * with the reference's walk-length adaptation.  Consumes the context's
 * stream exactly as the reference does, so e_limits (n_iter values) are
 * bit-identical to the reference's for the same stream.  The context's
 * own configuration is not used or changed.  stats: trials and accepted
 * walk moves. */
int bwl_nested_sampling(bwl_ctx *ctx, const bwl_ns_params *p, const int64_t *species_numbers,
                        const double *concentrations, double *e_limits, bwl_run_stats *stats);
/* n independent runs (one wavefront each, each on its own context and
 * stream of random numbers) in one launch; each run's result is the same
 * as alone.  species_numbers[i] / concentrations[i] may be NULL as for
 * bwl_initial_setup. */
int bwl_nested_sampling_batch(int32_t n, bwl_ctx *const *ctxs, const bwl_ns_params *params,
                              const int64_t *const *species_numbers,
                              const double *const *concentrations, double *const *e_limits,
                              bwl_run_stats *stats, double *makespan_ms);

/* ---------------------------------------------------------------- */
/* Wang-Landau walkers (wang-landau.F90)                             */
/* ---------------------------------------------------------------- */
/* The reference's wl_params fields the walker moves read
 * (derived_types.f90 wl_params; read in io.f90). */
typedef struct {
  int32_t bins;
  int32_t num_windows;
  int32_t walkers_per_window; /* the reference's MPI ranks per window */
  float bin_overlap;          /* REAL(4), as read */
  float energy_min;           /* meV/atom, REAL(4) */
  float energy_max;
  int32_t nbr_swap;           /* mc_select: 0 lattice-wide, 1 nearest neighbour */
  int32_t mc_sweeps;          /* sweeps per `sweeps` call */
  uint32_t rank_offset;       /* walker k draws from MT seed 110179 + 11*(rank_offset+k) */
} bwl_wl_params;

typedef struct bwl_wl bwl_wl;

/* A set of num_windows * walkers_per_window walkers (walker k in window
 * k / walkers_per_window) with windows and bin edges from divide_range,
 * create_overlap and create_energy_bins (wang-landau.F90:855-992). */
int bwl_wl_create(const bwl_lattice_desc *desc, const double *v_ex, const bwl_wl_params *p,
                  bwl_wl **out);
int bwl_wl_destroy(bwl_wl *wl);
int bwl_wl_n_walkers(const bwl_wl *wl, int32_t *n_walkers);
/* windows: (start, end) 1-based bins per window (2*num_windows); bin_edges:
 * bins+1 energies in Ry for the whole lattice.  Either may be NULL. */
int bwl_wl_windows(const bwl_wl *wl, int32_t *windows, double *bin_edges);
/* Width of the per-walker histogram rows (the widest window). */
int bwl_wl_hist_stride(const bwl_wl *wl, int32_t *stride);
/* initial_setup for every walker from its own stream. */
int bwl_wl_initial_setup(bwl_wl *wl, const int64_t *species_numbers, const double *concentrations);
int bwl_wl_set_walker(bwl_wl *wl, int32_t walker, const int8_t *config);
int bwl_wl_get_walker(bwl_wl *wl, int32_t walker, int8_t *config);
int bwl_wl_genrand_get_state(bwl_wl *wl, int32_t walker, uint32_t *mt, int32_t *mti);
int bwl_wl_genrand_set_state(bwl_wl *wl, int32_t walker, const uint32_t *mt, int32_t mti);
/* ln g (walkers x bins) and visit histograms (walkers x stride), FP64.
 * Either pointer may be NULL. */
int bwl_wl_get_lng(bwl_wl *wl, double *lng, double *hist);
int bwl_wl_set_lng(bwl_wl *wl, const double *lng, const double *hist);
/* enter_energy_window (wang-landau.F90:643-741) for every walker, each on
 * its own: biased swaps towards the middle of its window until its
 * recounted energy is inside the window's inner 80%, re-running
 * initial_setup every n_atoms*250 iterations as the reference does.  The
 * reference's hand-off of the first walker's configuration to the rest of
 * its window (MPI timing) is not made.  Stops a walker after max_trials
 * trials; entered[k] = 1 if walker k is inside. */
int bwl_wl_enter_windows(bwl_wl *wl, int64_t max_trials, int64_t *trials, int32_t *entered);
/* One `sweeps` call (wang-landau.F90:539-632) for every walker at once
 * (mc_sweeps*n_atoms trials each), then, if average != 0, each window's ln g
 * and histograms replaced by their mean over the window's walkers, summed
 * in walker order (the reference's MPI_Allreduce, :628-631).  accepted
 * (optional) per walker; kernel_ms the device time. */
int bwl_wl_sweeps(bwl_wl *wl, double wl_f, int32_t average, int64_t *accepted, double *kernel_ms);
/* replica_exchange (wang-landau.F90:1392-1501) over the whole walker set,
 * walker k standing for MPI rank k (walker 0's stream, as rank 0's, shuffles
 * the candidates).  Every walker's energy is recounted (full_energy); a
 * walker whose bin lies in the overlap with the window above (below) is a
 * candidate for that pair of windows; candidates of windows i and i+1 are
 * shuffled, matched in order, and each matched pair exchanges
 * configurations iff genrand() of the lower walker's stream is below
 * exp(lng(bin_lower) - lng(bin_upper)) with the lower walker's ln g.
 * partner[k] (optional) = walker matched with k or -1; exchanged[k]
 * (optional) = 1 if that pair exchanged; kernel_ms the device time. */
int bwl_wl_replica_exchange(bwl_wl *wl, int32_t *partner, int32_t *exchanged, double *kernel_ms);

/* ---------------------------------------------------------------- */
/* Request server (continuous batching on one device)                */
/* ---------------------------------------------------------------- */
/* No counterpart in the reference (there, one MPI rank runs one chain).
 * Requests are submitted with host inputs and queue on the host;
 * bwl_server_progress retires finished launches (results written to the
 * callers' buffers, device memory returned) and admits queued requests in
 * arrival order -- skipping ahead past any that do not fit -- while the
 * device-memory budget and launch slots allow.  Requests admitted together
 * are grouped by kernel variant, one launch per group on its own stream, so
 * work admitted at different times overlaps on the device.  Each request
 * keeps its own random stream, so its result is the same as when run alone
 * (bwl_metropolis_anneal on a fresh context with the same seed, or
 * bwl_nested_sampling after bwl_genrand_seed(seed)).  Output buffers must
 * stay valid until the request is done.  One host thread per server. */
typedef struct {
  int64_t device_bytes;  /* admission budget: one device arena of this size */
  int32_t max_launches;  /* launch slots (streams), 1..64 */
  int32_t batch_wait_us; /* batching window: while launches run, admit only
                            once the oldest queued request has waited this
                            long (0: admit at once) */
} bwl_server_params;

typedef struct {
  int32_t kind;          /* 1 anneal, 2 nested sampling, 3 Wang-Landau */
  int32_t state;         /* 0 queued, 1 running, 2 done, 3 failed */
  int32_t error;         /* status of a failed request */
  int32_t reserved;
  int64_t trials, accepted;
  double t_submit_ms, t_admit_ms, t_done_ms;  /* server clock */
  double kernel_ms;      /* device time of the launch the request ran in */
} bwl_request_info;

typedef struct bwl_server bwl_server;

int bwl_server_create(const bwl_server_params *p, bwl_server **out);
int bwl_server_destroy(bwl_server *s);
/* A Metropolis anneal (as bwl_metropolis_anneal) of `config` (reference
 * layout); config_out (optional) receives the final configuration. */
int bwl_server_submit_anneal(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                             const int8_t *config, const bwl_anneal_params *p,
                             double *energies_of_T, double *C_of_T, double *acceptance_of_T,
                             int8_t *config_out, int64_t *id);
/* A nested-sampling run (as bwl_nested_sampling) from MT19937 seed `seed`;
 * its walkers are set up on the host at submission. */
int bwl_server_submit_ns(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                         uint32_t seed, const bwl_ns_params *p, const int64_t *species_numbers,
                         const double *concentrations, double *e_limits, int64_t *id);
/* A short Wang-Landau run: num_windows x walkers_per_window walkers (walker
 * k seeded as rank rank_offset + k, set up by initial_setup), ln g and
 * histograms starting at 0, one `sweeps` call at modification factor wl_f,
 * then the window average (as bwl_wl_initial_setup + bwl_wl_sweeps(wl_f,
 * 1)).  lng receives num_windows x bins values (each window's averaged
 * ln g), hist num_windows x bins values (each window's averaged visit
 * histogram in its first end-start+1 entries; the rest is left as is). */
int bwl_server_submit_wl(bwl_server *s, const bwl_lattice_desc *desc, const double *v_ex,
                         const bwl_wl_params *p, const int64_t *species_numbers,
                         const double *concentrations, double wl_f, double *lng, double *hist,
                         int64_t *id);
/* Non-blocking: retire finished launches, admit what fits. */
int bwl_server_progress(bwl_server *s, int32_t *n_running, int32_t *n_queued);
/* Ids of requests finished since the last poll (up to max_ids). */
int bwl_server_poll(bwl_server *s, int64_t *ids, int32_t max_ids, int32_t *n);
int bwl_server_info(bwl_server *s, int64_t id, bwl_request_info *info);
/* Blocking: progress until every request is done. */
int bwl_server_drain(bwl_server *s);
int bwl_server_memory(const bwl_server *s, int64_t *used_bytes, int64_t *total_bytes);

/* ---------------------------------------------------------------- */
/* measurement                                                        */
/* ---------------------------------------------------------------- */
/* No counterpart in the reference: the LDS read bandwidth of the current
 * device, for the LDS roofline of the samplers (whose lattices live in
 * LDS).  Every CU streams conflict-free 16-byte LDS reads (8 blocks of 1024
 * threads per CU, `iters` reads per thread); bytes_per_s is the aggregate
 * over all CUs, cus the number of CUs. */
int bwl_lds_bandwidth(int32_t iters, double *bytes_per_s, int32_t *cus);

#ifdef __cplusplus
}
#endif
#endif
