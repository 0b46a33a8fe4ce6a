/*
 * brawl_oracle.h -- CPU restatement of BraWl's lattice-sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (brawl_amd) never links or
 * calls it.
 *
 * Every routine restates the reference's Fortran/C in the reference's own
 * order of operations (random-number consumption, FP64 summation order,
 * REAL(4) intermediates), so that its outputs are bit-identical to the
 * reference's.  Each function cites the reference lines it follows.
 *
 * Pinning (what checks each function independently of this code):
 *   - MT19937: C++11 std::mt19937 known answers (10000th output of seed
 *     5489 is 4123659995) and the reference's own golden trajectories.
 *   - initial_setup: golden initial configurations (tests/99_ref 01, 02).
 *   - site/total/pair energy, random site, MC step: golden per-trial energy
 *     trajectories and final configurations (tests/99_ref 01, 02), plus a
 *     brute-force O(N^2) pair enumeration on tiny lattices.
 *   - radial densities / lattice shells: golden rho_of_T.nc and ASRO
 *     trajectories.
 *   - nested sampling: golden .energies file (tests/99_ref 03,
 *     statistically: that golden is stale), and bit for bit against the
 *     reference's loop (nested_sampling.f90:78-200) run over the
 *     reference's own compiled routines in oracle/_ref.
 *   - Wang-Landau sweeps and window entry: the reference test is
 *     time-seeded; bit for bit against the reference's loops
 *     (wang-landau.F90:539-741, MPI left out) run over its own compiled
 *     routines in oracle/_ref, and closed-form invariants.
 *   - Wang-Landau replica exchange: no golden and no reference build (its
 *     file needs an MPI module); pinned by what the algorithm fixes: the
 *     exchange probability exp(d ln g), a uniformly random matching of the
 *     shuffled candidates, the draws each stream gives, overlap-only
 *     pairing (tests/test_oracle_invariants.py).
 *
 * Conventions
 *   config   int8, the reference's config(1, 2*n1, 2*n2, 2*n3) array in
 *            Fortran column-major order: element (1,i,j,k), 1-based, is at
 *            (i-1) + 2*n1*((j-1) + 2*n2*(k-1)).  0 = empty site.
 *   v_ex     FP64 V_ex(S,S,R) column-major: V_ex(a,b,n) at
 *            (a-1) + S*((b-1) + S*(n-1)), Rydberg.
 *   hist     int64 H(S,S,R), same layout: number of ordered (site, listed
 *            neighbour) pairs with species (a,b) on shell n.
 */
#ifndef BRAWL_ORACLE_H
#define BRAWL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { BWO_FCC = 1, BWO_BCC = 2 };

typedef struct {
  int32_t lattice;    /* BWO_FCC or BWO_BCC */
  int32_t n1, n2, n3; /* conventional cells per axis */
  int32_t n_species;  /* S */
  int32_t n_shells;   /* R (interaction_range) */
  const double *v_ex; /* V_ex(S,S,R) */
} bwo_lattice;

typedef struct {
  uint32_t mt[624];
  int32_t mti;
} bwo_mt;

/* ---- MT19937 (mt19937ar.c:63-125) ---- */
void bwo_mt_init(bwo_mt *g, uint32_t seed);
uint32_t bwo_mt_u32(bwo_mt *g);
double bwo_genrand(bwo_mt *g);
/* Static seed of f90_init_genrand (mt19937ar.c:136-139): 110179 + 11*rank */
uint32_t bwo_static_seed(int32_t rank);

/* ---- geometry ---- */
int32_t bwo_n_atoms(const bwo_lattice *L);
int32_t bwo_max_shells(int32_t lattice);
int32_t bwo_shell_size(int32_t lattice, int32_t shell);
/* neighbour offsets (doubled-grid units) of shell (1-based) in list order */
const int8_t *bwo_shell_offsets(int32_t lattice, int32_t shell);

/* ---- Hamiltonian (bw_hamiltonian.f90) ---- */
double bwo_site_energy(const bwo_lattice *L, const int8_t *cfg, int32_t i,
                       int32_t j, int32_t k);
double bwo_total_energy(const bwo_lattice *L, const int8_t *cfg);
double bwo_pair_energy(const bwo_lattice *L, const int8_t *cfg,
                       const int32_t *s1, const int32_t *s2);
void bwo_pair_histogram(const bwo_lattice *L, const int8_t *cfg, int64_t *hist);
double bwo_energy_from_histogram(const bwo_lattice *L, const int64_t *hist);

/* ---- proposals (random_site.f90) ---- */
void bwo_random_site(const bwo_lattice *L, bwo_mt *g, int32_t *site);
void bwo_random_nbr(const bwo_lattice *L, bwo_mt *g, const int32_t *site,
                    int32_t *nbr);
void bwo_pair_swap(const bwo_lattice *L, int8_t *cfg, const int32_t *s1,
                   const int32_t *s2);

/* ---- initial configuration (initialise.F90:434-617) ----
 * species_numbers (length S) is used when it sums to the atom count,
 * otherwise concentrations (length S) are used.  Returns 0 on success. */
int32_t bwo_initial_setup(const bwo_lattice *L, const int64_t *species_numbers,
                          const double *concentrations, bwo_mt *g,
                          int8_t *cfg);

/* ---- Metropolis (metropolis.F90) ---- */
int32_t bwo_mc_step(const bwo_lattice *L, int8_t *cfg, double beta,
                    int32_t nbr_swap, bwo_mt *g);

typedef struct {
  double T, delta_T;
  int32_t T_steps;
  int32_t n_mc_steps;
  int32_t n_sample_steps;
  int32_t burn_in_start, burn_in;
  int32_t n_burn_in_steps;
  int32_t nbr_swap;
  int32_t calculate_asro;
  int32_t n_sample_steps_asro;
  int32_t wc_range;
} bwo_metropolis_params;

/* Simulated annealing loop (metropolis.F90:193-437).  Optional outputs
 * (NULL to skip):
 *   e_traj     energy at step 0 and after every sweep of every T step,
 *              (n_sweeps+1)*T_steps values;
 *   rho_of_T   S*S*wc_range*T_steps averaged radial densities.
 * Always written: energies_of_T, C_of_T, acceptance_of_T (T_steps each). */
int32_t bwo_metropolis_anneal(const bwo_lattice *L, int8_t *cfg, bwo_mt *g,
                              const bwo_metropolis_params *p, double *e_traj,
                              double *energies_of_T, double *C_of_T,
                              double *acceptance_of_T, double *rho_of_T);

/* ---- analytics (analytics.f90:205-404) ---- */
void bwo_lattice_shells(const bwo_lattice *L, const int8_t *cfg,
                        int32_t wc_range, double *shells);
void bwo_radial_densities(const bwo_lattice *L, const int8_t *cfg,
                          int32_t wc_range, const double *shells,
                          double *rho);

/* ---- nested sampling (nested_sampling.f90:45-206) ----
 * Runs the whole serial NS loop from the given MT state.  Writes n_iter
 * energy limits.  Returns 0 on success. */
int32_t bwo_nested_sampling(const bwo_lattice *L, const int64_t *species_numbers,
                            const double *concentrations, bwo_mt *g,
                            int32_t n_walkers, int32_t n_steps, int32_t n_iter,
                            double *e_limits);

/* ---- Wang-Landau (wang-landau.F90) ---- */
int32_t bwo_wl_bin_index(double energy, const double *bin_edges, int32_t bins);
/* One `sweeps` call for one walker (wang-landau.F90:539-632) without the
 * window-sub-communicator average.  lng has `bins` entries (global bins),
 * hist has (end_idx-start_idx+1).  Histogram cadence and window test follow
 * the reference.  Returns accepted count. */
int64_t bwo_wl_sweeps(const bwo_lattice *L, int8_t *cfg, bwo_mt *g,
                      int32_t nbr_swap, int32_t mc_sweeps,
                      const double *bin_edges, int32_t bins, int32_t start_idx,
                      int32_t end_idx, double wl_f, double *lng, double *hist);
/* Biased walk into [min_e, max_e] (wang-landau.F90:643-741) for one walker
 * without MPI hand-off; returns the number of trials made.  energy_min and
 * energy_max are the REAL(4) wl_params bounds in meV/atom. */
int64_t bwo_wl_enter_window(const bwo_lattice *L, int8_t *cfg, bwo_mt *g,
                            double min_e, double max_e, float energy_min,
                            float energy_max, const int64_t *species_numbers,
                            const double *concentrations, int64_t max_trials);
/* divide_range + create_overlap + create_energy_bins (wang-landau.F90:855-992)
 * windows: 2*num_windows ints (start,end per window, 1-based bins, with
 * overlap); bin_edges: bins+1 doubles. */
void bwo_wl_setup(const bwo_lattice *L, int32_t bins, int32_t num_windows,
                  float bin_overlap, float energy_min, float energy_max,
                  int32_t *windows, double *bin_edges);
/* replica_exchange (wang-landau.F90:1392-1501) over W = num_windows*wpw
 * walkers (walker k = rank k; walker 0's stream shuffles the candidates).
 * lng: W rows of `bins` values.  Exchanges the configurations of accepted
 * pairs; partner[k] = walker matched with k or -1, accepted[k] = 1 if that
 * pair exchanged.  Returns the number of exchanged pairs. */
int32_t bwo_wl_replica_exchange(const bwo_lattice *L, int8_t *const *cfgs,
                                bwo_mt *const *gens, const double *lng,
                                int32_t bins, const double *bin_edges,
                                const int32_t *windows, int32_t num_windows,
                                int32_t wpw, int32_t *partner, int32_t *accepted);

#ifdef __cplusplus
}
#endif
#endif
