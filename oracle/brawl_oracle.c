/*
 * brawl_oracle.c -- CPU restatement of BraWl's lattice-sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Never
 * linked into or called by the product.
 *
 * Written from the reference's behaviour (file:line citations into the
 * BraWl sources).  Order of floating-point operations and of random-number
 * draws follows the reference exactly, so outputs are bit-identical.
 * Build with -ffp-contract=off (no FMA contraction): the reference's sums
 * are plain left-to-right FP64 additions.
 */
#include "brawl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* constants.f90:41-49                                                 */
/* ------------------------------------------------------------------ */
/* The reference's Boltzmann constant carries a typo (8.167e-5 instead of
 * 8.617e-5 eV/K); it is kept because every reference beta uses it. */
static const double K_B_IN_RY = 8.167333262e-5 / 13.605693122990;
static const double RY_TO_EV = 13.605693122;

/* ------------------------------------------------------------------ */
/* MT19937, from Matsumoto & Nishimura's published algorithm           */
/* (mt19937ar.c:63-125 is the reference's copy)                        */
/* ------------------------------------------------------------------ */
void bwo_mt_init(bwo_mt *g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i) {
    uint32_t prev = g->mt[i - 1];
    g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
  }
  g->mti = 624;
}

static void mt_refill(bwo_mt *g) {
  uint32_t *s = g->mt;
  for (int k = 0; k < 624; ++k) {
    uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % 624] & 0x7fffffffu);
    uint32_t v = s[(k + 397) % 624] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    s[k] = v;
  }
  g->mti = 0;
}

uint32_t bwo_mt_u32(bwo_mt *g) {
  if (g->mti >= 624) {
    if (g->mti == 625) bwo_mt_init(g, 5489u);
    mt_refill(g);
  }
  uint32_t y = g->mt[g->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* genrand(): 32-bit output scaled by 2^-32 (mt19937ar.c:121-125) */
double bwo_genrand(bwo_mt *g) { return (double)bwo_mt_u32(g) * (1.0 / 4294967296.0); }

uint32_t bwo_static_seed(int32_t rank) { return 110179u + 11u * (uint32_t)rank; }

/* ------------------------------------------------------------------ */
/* Neighbour-shell offset lists, doubled-grid units (dx,dy,dz), in the */
/* reference's list order -- the list order is the FP64 summation     */
/* order of each shell energy.                                          */
/* ------------------------------------------------------------------ */
static const int8_t FCC_OFF[][3] = {
  /* shell 1: bw_hamiltonian.f90:1310-1321 */
  { 0,-1, 1}, { 0,-1,-1}, { 0, 1, 1}, { 0, 1,-1}, { 1,-1, 0}, { 1, 1, 0},
  { 1, 0, 1}, { 1, 0,-1}, {-1,-1, 0}, {-1, 1, 0}, {-1, 0, 1}, {-1, 0,-1},
  /* shell 2: bw_hamiltonian.f90:1377-1382 */
  { 2, 0, 0}, {-2, 0, 0}, { 0, 2, 0}, { 0,-2, 0}, { 0, 0, 2}, { 0, 0,-2},
  /* shell 3: bw_hamiltonian.f90:1445-1468 */
  {-2, 1, 1}, {-2, 1,-1}, {-2,-1, 1}, {-2,-1,-1}, { 2, 1, 1}, { 2, 1,-1},
  { 2,-1, 1}, { 2,-1,-1}, { 1, 2, 1}, {-1, 2, 1}, { 1, 2,-1}, {-1, 2,-1},
  { 1,-2, 1}, {-1,-2, 1}, { 1,-2,-1}, {-1,-2,-1}, { 1, 1, 2}, {-1, 1, 2},
  { 1,-1, 2}, {-1,-1, 2}, { 1, 1,-2}, {-1, 1,-2}, { 1,-1,-2}, {-1,-1,-2},
  /* shell 4: bw_hamiltonian.f90:1524-1535 */
  { 2, 0,-2}, {-2, 0,-2}, { 0, 2,-2}, { 0,-2,-2}, { 2, 2, 0}, {-2, 2, 0},
  { 2,-2, 0}, {-2,-2, 0}, { 2, 0, 2}, {-2, 0, 2}, { 0, 2, 2}, { 0,-2, 2},
  /* shell 5: bw_hamiltonian.f90:1598-1621 */
  { 1, 0,-3}, {-1, 0,-3}, { 0,-1,-3}, { 0, 1,-3}, { 0, 3,-1}, { 3, 0,-1},
  {-3, 0,-1}, { 0,-3,-1}, { 1, 3, 0}, {-1, 3, 0}, { 3, 1, 0}, {-3, 1, 0},
  { 3,-1, 0}, {-3,-1, 0}, { 1,-3, 0}, {-1,-3, 0}, { 0, 3, 1}, { 3, 0, 1},
  {-3, 0, 1}, { 0,-3, 1}, { 1, 0, 3}, {-1, 0, 3}, { 0,-1, 3}, { 0, 1, 3},
  /* shell 6: bw_hamiltonian.f90:1676-1683 */
  { 2, 2,-2}, {-2, 2,-2}, { 2,-2,-2}, {-2,-2,-2}, { 2, 2, 2}, {-2, 2, 2},
  { 2,-2, 2}, {-2,-2, 2},
};
static const int FCC_SHELL_SIZE[6] = {12, 6, 24, 12, 24, 8};

static const int8_t BCC_OFF[][3] = {
  /* shell 1: bw_hamiltonian.f90:162-169 */
  { 1, 1, 1}, { 1, 1,-1}, { 1,-1, 1}, { 1,-1,-1}, {-1, 1, 1}, {-1, 1,-1},
  {-1,-1, 1}, {-1,-1,-1},
  /* shell 2: bw_hamiltonian.f90:225-230 */
  { 2, 0, 0}, {-2, 0, 0}, { 0,-2, 0}, { 0, 2, 0}, { 0, 0, 2}, { 0, 0,-2},
  /* shell 3: bw_hamiltonian.f90:286-297 */
  { 0,-2,-2}, { 2, 0,-2}, {-2, 0,-2}, { 0, 2,-2}, { 2,-2, 0}, {-2,-2, 0},
  { 2, 2, 0}, {-2, 2, 0}, { 0,-2, 2}, { 2, 0, 2}, {-2, 0, 2}, { 0, 2, 2},
  /* shell 4: bw_hamiltonian.f90:361-384 */
  { 1, 1, 3}, {-1, 1, 3}, { 1,-1, 3}, {-1,-1, 3}, { 1, 1,-3}, {-1, 1,-3},
  { 1,-1,-3}, {-1,-1,-3}, { 1, 3, 1}, {-1, 3, 1}, { 1, 3,-1}, {-1, 3,-1},
  { 1,-3, 1}, {-1,-3, 1}, { 1,-3,-1}, {-1,-3,-1}, { 3, 1, 1}, { 3,-1, 1},
  { 3, 1,-1}, { 3,-1,-1}, {-3, 1, 1}, {-3,-1, 1}, {-3, 1,-1}, {-3,-1,-1},
  /* shell 5: bw_hamiltonian.f90:440-447 */
  { 2, 2, 2}, {-2, 2, 2}, { 2,-2, 2}, {-2,-2, 2}, { 2, 2,-2}, {-2, 2,-2},
  { 2,-2,-2}, {-2,-2,-2},
  /* shell 6: bw_hamiltonian.f90:504-509 */
  { 4, 0, 0}, {-4, 0, 0}, { 0, 4, 0}, { 0,-4, 0}, { 0, 0, 4}, { 0, 0,-4},
  /* shell 7: bw_hamiltonian.f90:572-595 */
  { 1, 3, 3}, {-1, 3, 3}, { 3, 1, 3}, {-3, 1, 3}, { 3,-1, 3}, {-3,-1, 3},
  { 1,-3, 3}, {-1,-3, 3}, { 3, 3, 1}, {-3, 3, 1}, { 3,-3, 1}, {-3,-3, 1},
  { 3, 3,-1}, {-3, 3,-1}, { 3,-3,-1}, {-3,-3,-1}, { 1, 3,-3}, {-1, 3,-3},
  { 3, 1,-3}, {-3, 1,-3}, { 3,-1,-3}, {-3,-1,-3}, { 1,-3,-3}, {-1,-3,-3},
  /* shell 8: bw_hamiltonian.f90:658-681 (entries 15-16 are shell-6 vectors
   * in the reference; kept, they define its Hamiltonian) */
  { 0, 2, 4}, { 2, 0, 4}, {-2, 0, 4}, { 0,-2, 4}, { 0, 4, 2}, {-4, 0, 2},
  { 4, 0, 2}, { 0,-4, 2}, { 2, 4, 0}, {-2, 4, 0}, { 4, 2, 0}, {-4, 2, 0},
  { 4,-2, 0}, {-4,-2, 0}, { 0, 4, 0}, {-4, 0, 0}, { 0, 4,-2}, {-4, 0,-2},
  { 4, 0,-2}, { 0,-4,-2}, { 0, 2,-4}, { 2, 0,-4}, {-2, 0,-4}, { 0,-2,-4},
  /* shell 9: bw_hamiltonian.f90:744-767 */
  { 2, 2, 4}, {-2, 2, 4}, { 2,-2, 4}, {-2,-2, 4}, { 2, 4, 2}, {-2, 4, 2},
  { 4, 2, 2}, {-4, 2, 2}, { 4,-2, 2}, {-4,-2, 2}, { 2,-4, 2}, {-2,-4, 2},
  { 2, 4,-2}, {-2, 4,-2}, { 4, 2,-2}, {-4, 2,-2}, { 4,-2,-2}, {-4,-2,-2},
  { 2,-4,-2}, {-2,-4,-2}, { 2, 2,-4}, {-2, 2,-4}, { 2,-2,-4}, {-2,-2,-4},
  /* shell 10: bw_hamiltonian.f90:837-868 (duplicated entries kept) */
  { 1, 1, 5}, {-1, 1, 5}, { 1,-1, 5}, {-1,-1, 5}, { 3, 3, 3}, {-3, 3, 3},
  { 3,-3, 3}, {-3,-3, 3}, { 1,-5, 1}, {-1,-5, 1}, { 5,-1, 1}, {-5,-1, 1},
  { 5, 1, 1}, {-5, 1, 1}, { 1,-5, 1}, {-1,-5, 1}, { 1,-5,-1}, {-1,-5,-1},
  { 5,-1,-1}, {-5,-1,-1}, { 5, 1,-1}, {-5, 1,-1}, { 1,-5,-1}, {-1,-5,-1},
  { 3, 3,-3}, {-3, 3,-3}, { 3,-3,-3}, {-3,-3,-3}, { 1, 1,-5}, {-1, 1,-5},
  { 1,-1,-5}, {-1,-1,-5},
};
static const int BCC_SHELL_SIZE[10] = {8, 6, 12, 24, 8, 6, 24, 24, 24, 32};

int32_t bwo_max_shells(int32_t lattice) {
  return lattice == BWO_FCC ? 6 : lattice == BWO_BCC ? 10 : 0;
}

int32_t bwo_shell_size(int32_t lattice, int32_t shell) {
  if (shell < 1 || shell > bwo_max_shells(lattice)) return 0;
  return lattice == BWO_FCC ? FCC_SHELL_SIZE[shell - 1] : BCC_SHELL_SIZE[shell - 1];
}

const int8_t *bwo_shell_offsets(int32_t lattice, int32_t shell) {
  if (shell < 1 || shell > bwo_max_shells(lattice)) return NULL;
  const int *sz = lattice == BWO_FCC ? FCC_SHELL_SIZE : BCC_SHELL_SIZE;
  const int8_t(*tab)[3] = lattice == BWO_FCC ? FCC_OFF : BCC_OFF;
  int start = 0;
  for (int s = 1; s < shell; ++s) start += sz[s - 1];
  return tab[start];
}

/* initialise_function_pointers (initialise.F90:174-212): atoms per cell */
int32_t bwo_n_atoms(const bwo_lattice *L) {
  int32_t cells = L->n1 * L->n2 * L->n3;
  return L->lattice == BWO_FCC ? 4 * cells : 2 * cells;
}

/* Fortran modulo for a positive divisor */
static inline int fmod_pos(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

static inline int64_t cfg_index(const bwo_lattice *L, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)2 * L->n1 * ((j - 1) + (int64_t)2 * L->n2 * (k - 1));
}

static inline int8_t cfg_at(const bwo_lattice *L, const int8_t *cfg, int i, int j, int k) {
  return cfg[cfg_index(L, i, j, k)];
}

static inline double vex(const bwo_lattice *L, int a, int b, int n) {
  const int S = L->n_species;
  return L->v_ex[(a - 1) + S * ((b - 1) + S * (n - 1))];
}

/* {fcc,bcc}_shellN_energy (bw_hamiltonian.f90:134-877, 1282-1691): sum of
 * V_ex(species, nbr, n) from 0.0 in list order, neighbours wrapped with
 * modulo(site +- d - 1, 2n) + 1. */
static double shell_energy(const bwo_lattice *L, const int8_t *cfg, int i, int j, int k,
                           int species, int shell) {
  const int8_t *off = bwo_shell_offsets(L->lattice, shell);
  const int cnt = bwo_shell_size(L->lattice, shell);
  const int m1 = 2 * L->n1, m2 = 2 * L->n2, m3 = 2 * L->n3;
  double e = 0.0;
  for (int t = 0; t < cnt; ++t) {
    int ni = fmod_pos(i - 1 + off[3 * t + 0], m1) + 1;
    int nj = fmod_pos(j - 1 + off[3 * t + 1], m2) + 1;
    int nk = fmod_pos(k - 1 + off[3 * t + 2], m3) + 1;
    int nb = cfg_at(L, cfg, ni, nj, nk);
    e = e + vex(L, species, nb, shell);
  }
  return e;
}

/* {fcc,bcc}_energy_Rshells (bw_hamiltonian.f90:898-1262, 1712-1910):
 * shell1 + shell2 + ... evaluated left to right. */
double bwo_site_energy(const bwo_lattice *L, const int8_t *cfg, int32_t i, int32_t j,
                       int32_t k) {
  int species = cfg_at(L, cfg, i, j, k);
  double e = shell_energy(L, cfg, i, j, k, species, 1);
  for (int n = 2; n <= L->n_shells; ++n) e = e + shell_energy(L, cfg, i, j, k, species, n);
  return e;
}

/* total_energy (bw_hamiltonian.f90:58-81): z outer, then y, then x; empty
 * sites skipped; halved at the end. */
double bwo_total_energy(const bwo_lattice *L, const int8_t *cfg) {
  double e = 0.0;
  for (int l = 1; l <= 2 * L->n3; ++l)
    for (int k = 1; k <= 2 * L->n2; ++k)
      for (int j = 1; j <= 2 * L->n1; ++j) {
        if (cfg_at(L, cfg, j, k, l) == 0) continue;
        e = e + bwo_site_energy(L, cfg, j, k, l);
      }
  return 0.5 * e;
}

/* pair_energy (bw_hamiltonian.f90:99-114) */
double bwo_pair_energy(const bwo_lattice *L, const int8_t *cfg, const int32_t *s1,
                       const int32_t *s2) {
  return bwo_site_energy(L, cfg, s1[0], s1[1], s1[2]) +
         bwo_site_energy(L, cfg, s2[0], s2[1], s2[2]);
}

/* Integer pair-count histogram over the same neighbour lists; the total
 * energy is 0.5 * sum H(a,b,n) V_ex(a,b,n). */
void bwo_pair_histogram(const bwo_lattice *L, const int8_t *cfg, int64_t *hist) {
  const int S = L->n_species;
  memset(hist, 0, sizeof(int64_t) * (size_t)S * S * L->n_shells);
  const int m1 = 2 * L->n1, m2 = 2 * L->n2, m3 = 2 * L->n3;
  for (int k = 1; k <= m3; ++k)
    for (int j = 1; j <= m2; ++j)
      for (int i = 1; i <= m1; ++i) {
        int a = cfg_at(L, cfg, i, j, k);
        if (a == 0) continue;
        for (int n = 1; n <= L->n_shells; ++n) {
          const int8_t *off = bwo_shell_offsets(L->lattice, n);
          const int cnt = bwo_shell_size(L->lattice, n);
          for (int t = 0; t < cnt; ++t) {
            int b = cfg_at(L, cfg, fmod_pos(i - 1 + off[3 * t], m1) + 1,
                           fmod_pos(j - 1 + off[3 * t + 1], m2) + 1,
                           fmod_pos(k - 1 + off[3 * t + 2], m3) + 1);
            hist[(a - 1) + S * ((b - 1) + S * (n - 1))] += 1;
          }
        }
      }
}

/* Fixed-order FP64 contraction: shells outer, then b, then a. */
double bwo_energy_from_histogram(const bwo_lattice *L, const int64_t *hist) {
  const int S = L->n_species;
  double e = 0.0;
  for (int n = 1; n <= L->n_shells; ++n)
    for (int b = 1; b <= S; ++b)
      for (int a = 1; a <= S; ++a) {
        int64_t h = hist[(a - 1) + S * ((b - 1) + S * (n - 1))];
        e += (double)h * vex(L, a, b, n);
      }
  return 0.5 * e;
}

/* ------------------------------------------------------------------ */
/* proposals: random_site.f90                                          */
/* ------------------------------------------------------------------ */
/* bcc_random_site (:127-139) and fcc_random_site (:182-193): three draws,
 * for k, then i, then j. */
void bwo_random_site(const bwo_lattice *L, bwo_mt *g, int32_t *site) {
  int k = (int)floor(2.0 * bwo_genrand(g) * (double)L->n3) + 1;
  if (L->lattice == BWO_BCC) {
    int i = 2 * (int)floor(bwo_genrand(g) * (double)L->n1) + 2 - fmod_pos(k, 2);
    int j = 2 * (int)floor(bwo_genrand(g) * (double)L->n2) + 2 - fmod_pos(k, 2);
    site[0] = i; site[1] = j; site[2] = k;
  } else {
    int i = (int)floor(2.0 * bwo_genrand(g) * (double)L->n1) + 1;
    int j = 2 * (int)floor(bwo_genrand(g) * (double)L->n2) + 1 + fmod_pos(i - fmod_pos(k, 2), 2);
    site[0] = i; site[1] = j; site[2] = k;
  }
}

/* nearest-neighbour tables of random_site.f90:38-61 */
static const int8_t RS_BCC_NBRS[8][3] = {
  { 1, 1, 1}, { 1, 1,-1}, { 1,-1, 1}, { 1,-1,-1},
  {-1, 1, 1}, {-1, 1,-1}, {-1,-1, 1}, {-1,-1,-1}};
static const int8_t RS_FCC_NBRS[12][3] = {
  { 0, 1, 1}, { 0, 1,-1}, { 0,-1, 1}, { 0,-1,-1}, { 1, 1, 0}, { 1,-1, 0},
  { 1, 0, 1}, { 1, 0,-1}, {-1, 1, 0}, {-1,-1, 0}, {-1, 0, 1}, {-1, 0,-1}};

/* bcc_random_nbr (:152-171), fcc_random_nbr (:206-225): one draw */
void bwo_random_nbr(const bwo_lattice *L, bwo_mt *g, const int32_t *site, int32_t *nbr) {
  const int8_t *d;
  if (L->lattice == BWO_BCC) {
    int n = (int)floor(8.0 * bwo_genrand(g));
    d = RS_BCC_NBRS[n];
  } else {
    int n = (int)floor(12.0 * bwo_genrand(g));
    d = RS_FCC_NBRS[n];
  }
  nbr[0] = fmod_pos(site[0] + d[0] - 1, 2 * L->n1) + 1;
  nbr[1] = fmod_pos(site[1] + d[1] - 1, 2 * L->n2) + 1;
  nbr[2] = fmod_pos(site[2] + d[2] - 1, 2 * L->n3) + 1;
}

/* pair_swap (random_site.f90:238-250) */
void bwo_pair_swap(const bwo_lattice *L, int8_t *cfg, const int32_t *s1, const int32_t *s2) {
  int64_t a = cfg_index(L, s1[0], s1[1], s1[2]);
  int64_t b = cfg_index(L, s2[0], s2[1], s2[2]);
  int8_t t = cfg[a];
  cfg[a] = cfg[b];
  cfg[b] = t;
}

/* ------------------------------------------------------------------ */
/* initial_setup (initialise.F90:434-617)                              */
/* ------------------------------------------------------------------ */
int32_t bwo_initial_setup(const bwo_lattice *L, const int64_t *species_numbers,
                          const double *concentrations, bwo_mt *g, int8_t *cfg) {
  const int S = L->n_species;
  const int n_atoms = bwo_n_atoms(L);
  const int n_sites = n_atoms;
  int64_t count[16], check[16];
  double conc[17]; /* species_concentrations(0:S), element 0 = 0 */
  if (S < 1 || S > 15) return -1;
  conc[0] = 0.0;
  int64_t sum_numbers = 0;
  if (species_numbers)
    for (int i = 0; i < S; ++i) sum_numbers += species_numbers[i];
  if (species_numbers && sum_numbers == n_atoms) {
    /* initialise.F90:468-472: concentrations in REAL(4) */
    for (int i = 0; i < S; ++i) {
      count[i] = species_numbers[i];
      conc[i + 1] = (double)((float)count[i] / (float)n_atoms);
    }
  } else {
    if (!concentrations) return -2;
    for (int i = 0; i < S; ++i) conc[i + 1] = concentrations[i];
    /* initialise.F90:480-504 */
    int64_t tot = 0;
    for (int i = 0; i < S; ++i) {
      double x = (double)(float)n_sites * conc[i + 1];
      count[i] = (int64_t)lround(x); /* nint: half away from zero */
      tot += count[i];
    }
    int64_t round_err = tot - n_sites;
    int64_t inc = (-round_err) >= 0 ? 1 : -1; /* sign(1, -round_err) */
    int64_t round_check = round_err;
    int guard = 0;
    while (round_check != 0) {
      for (int j = 0; j < S; ++j) {
        if (count[j] == 0) continue;
        if (round_check == 0) continue;
        count[j] += inc;
        round_check += inc;
      }
      if (++guard > 1000000) return -3;
    }
  }
  int64_t tot = 0;
  for (int i = 0; i < S; ++i) { tot += count[i]; check[i] = 0; }
  if (tot != n_sites) return -4;

  /* cumulative bounds sum(conc(0:l)), summed left to right from element 0 */
  double cum[17];
  for (int l = 0; l <= S; ++l) {
    double s = 0.0;
    for (int m = 0; m <= l; ++m) s = s + conc[m];
    cum[l] = s;
  }

  memset(cfg, 0, (size_t)8 * L->n1 * L->n2 * L->n3);
  const int d2 = 2 * L->n1, d3 = 2 * L->n2, d4 = 2 * L->n3;
  if (L->lattice == BWO_BCC) {
    /* initialise.F90:540-570 */
    for (int k = 1; k <= d4; ++k)
      for (int j = 1; j <= d3 / 2; ++j)
        for (int i = 1; i <= d2 / 2; ++i) {
          int mk = fmod_pos(k, 2);
          int64_t idx = cfg_index(L, 2 * i - mk, 2 * j - mk, k);
          while (cfg[idx] == 0) {
            double r = bwo_genrand(g);
            for (int l = 1; l <= S; ++l) {
              if (r >= cum[l - 1] && r <= cum[l]) {
                if (check[l - 1] < count[l - 1]) {
                  cfg[idx] = (int8_t)l;
                  check[l - 1] += 1;
                }
              }
            }
          }
        }
  } else {
    /* initialise.F90:572-611 */
    for (int k = 1; k <= d4; ++k)
      for (int j = 1; j <= d3; ++j)
        for (int i = 1; i <= d2 / 2; ++i) {
          int idx_i = 2 * i - fmod_pos(k, 2) * fmod_pos(j, 2) - fmod_pos(k + 1, 2) * fmod_pos(j + 1, 2);
          int64_t idx = cfg_index(L, idx_i, j, k);
          while (cfg[idx] == 0) {
            double r = bwo_genrand(g);
            for (int l = 1; l <= S; ++l) {
              if (r >= cum[l - 1] && r <= cum[l]) {
                if (check[l - 1] < count[l - 1]) {
                  cfg[idx] = (int8_t)l;
                  check[l - 1] += 1;
                }
              }
            }
          }
        }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Metropolis-Kawasaki step (metropolis.F90:751-813, 829-891)          */
/* ------------------------------------------------------------------ */
int32_t bwo_mc_step(const bwo_lattice *L, int8_t *cfg, double beta, int32_t nbr_swap,
                    bwo_mt *g) {
  int32_t r1[3], r2[3];
  bwo_random_site(L, g, r1);
  if (nbr_swap)
    bwo_random_nbr(L, g, r1, r2);
  else
    bwo_random_site(L, g, r2);
  int s1 = cfg[cfg_index(L, r1[0], r1[1], r1[2])];
  int s2 = cfg[cfg_index(L, r2[0], r2[1], r2[2])];
  if (s1 == s2) return 1; /* counted as accepted, no further draws */
  double e_unswapped = bwo_pair_energy(L, cfg, r1, r2);
  bwo_pair_swap(L, cfg, r1, r2);
  double e_swapped = bwo_pair_energy(L, cfg, r1, r2);
  double delta_e = e_swapped - e_unswapped;
  if (delta_e < 0.0) return 1;
  if (bwo_genrand(g) < exp(-beta * delta_e)) return 1;
  bwo_pair_swap(L, cfg, r1, r2);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Simulated annealing statistics loop (metropolis.F90:193-437)        */
/* ------------------------------------------------------------------ */
int32_t bwo_metropolis_anneal(const bwo_lattice *L, int8_t *cfg, bwo_mt *g,
                              const bwo_metropolis_params *p, double *e_traj,
                              double *energies_of_T, double *C_of_T,
                              double *acceptance_of_T, double *rho_of_T) {
  const int S = L->n_species;
  const int n_atoms = bwo_n_atoms(L);
  if (p->n_sample_steps <= 0 || p->n_mc_steps < p->n_sample_steps) return -1;
  /* metropolis.F90:158-159: REAL(4) division, floored */
  const int n_save_energy = (int)floorf((float)p->n_mc_steps / (float)p->n_sample_steps);
  const int n_save_asro =
      p->calculate_asro ? (int)floorf((float)p->n_mc_steps / (float)p->n_sample_steps_asro) : 1;
  const int n_sweeps = p->n_mc_steps / p->n_sample_steps;
  const int n_sweep_steps = p->n_mc_steps / n_sweeps;
  const int wc = p->wc_range;
  double *shells = NULL, *rho = NULL, *rsum = NULL;
  if (p->calculate_asro) {
    shells = (double *)calloc((size_t)wc, sizeof(double));
    rho = (double *)calloc((size_t)S * S * wc, sizeof(double));
    rsum = (double *)calloc((size_t)S * S * wc, sizeof(double));
    bwo_lattice_shells(L, cfg, wc, shells);
  }
  int64_t traj_pos = 0;
  for (int jt = 1; jt <= p->T_steps; ++jt) {
    double step_E = 0.0, step_Esq = 0.0, acceptance = 0.0;
    double temp = p->T + (double)(jt - 1) * p->delta_T;
    double sim_temp = temp * K_B_IN_RY;
    double beta = 1.0 / sim_temp;
    /* burn-in (metropolis.F90:214-238) */
    int do_burn = (jt == 1) ? p->burn_in_start : p->burn_in;
    if (do_burn)
      for (int i = 0; i < p->n_burn_in_steps; ++i)
        acceptance += bwo_mc_step(L, cfg, beta, p->nbr_swap, g);
    double current = bwo_total_energy(L, cfg);
    if (e_traj) e_traj[traj_pos++] = current;
    if (p->calculate_asro) {
      for (size_t q = 0; q < (size_t)S * S * wc; ++q) rsum[q] = 0.0;
      bwo_radial_densities(L, cfg, wc, shells, rho);
    }
    acceptance = 0.0;
    for (int i = 1; i <= n_sweeps; ++i) {
      for (int kk = 0; kk < n_sweep_steps; ++kk)
        acceptance += bwo_mc_step(L, cfg, beta, p->nbr_swap, g);
      int step_n = i * n_sweep_steps;
      current = bwo_total_energy(L, cfg);
      step_E = step_E + current;
      step_Esq = step_Esq + current * current;
      if (e_traj) e_traj[traj_pos++] = current;
      if (p->calculate_asro && step_n % p->n_sample_steps_asro == 0) {
        bwo_radial_densities(L, cfg, wc, shells, rho);
        for (size_t q = 0; q < (size_t)S * S * wc; ++q) rsum[q] = rsum[q] + rho[q];
      }
    }
    acceptance_of_T[jt - 1] = acceptance / (double)(float)p->n_mc_steps;
    energies_of_T[jt - 1] = step_E / n_save_energy / n_atoms;
    double C = (step_Esq / n_save_energy - (step_E / n_save_energy) * (step_E / n_save_energy)) /
               (sim_temp * temp) / n_atoms;
    if (C < 0.0) C = 0.0;
    if (temp <= 0.0) C = 0.0;
    C_of_T[jt - 1] = C;
    if (p->calculate_asro && rho_of_T)
      for (size_t q = 0; q < (size_t)S * S * wc; ++q)
        rho_of_T[(size_t)(jt - 1) * S * S * wc + q] = rsum[q] / n_save_asro;
  }
  free(shells);
  free(rho);
  free(rsum);
  return 0;
}

/* ------------------------------------------------------------------ */
/* analytics.f90                                                       */
/* ------------------------------------------------------------------ */
static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* lattice_shells (analytics.f90:205-275): REAL(4) distances from site
 * (1,1,1) to every occupied site, sorted, first wc_range distinct values
 * (tolerance 1e-3), the zero shell first. */
void bwo_lattice_shells(const bwo_lattice *L, const int8_t *cfg, int32_t wc_range,
                        double *shells) {
  const int m1 = 2 * L->n1, m2 = 2 * L->n2, m3 = 2 * L->n3;
  size_t total = (size_t)m1 * m2 * m3 + 1;
  double *all = (double *)calloc(total, sizeof(double));
  size_t l = 0;
  for (int k = 1; k <= m3; ++k)
    for (int j = 1; j <= m2; ++j)
      for (int i = 1; i <= m1; ++i) {
        if (cfg_at(L, cfg, i, j, k) == 0) continue;
        float d2 = (float)((k - 1) * (k - 1)) + (float)((j - 1) * (j - 1));
        d2 = d2 + (float)((i - 1) * (i - 1));
        all[l++] = (double)sqrtf(d2);
      }
  qsort(all, total, sizeof(double), cmp_double);
  for (int s = 0; s < wc_range; ++s) shells[s] = 0.0;
  int cnt = 1;
  for (size_t i = 0; i + 1 < total; ++i) {
    if (fabs(all[i] - all[i + 1]) < 1e-3) continue;
    shells[cnt - 1] = all[i];
    cnt++;
    if (cnt > wc_range) break;
  }
  free(all);
}

/* radial_densities (analytics.f90:293-404): ordered pair counts inside a
 * +-min(n,5) doubled-cell window with minimum-image distances, binned by
 * the lattice_shells distances, divided by the count of the first species. */
void bwo_radial_densities(const bwo_lattice *L, const int8_t *cfg, int32_t wc_range,
                          const double *shells, double *rho) {
  const int S = L->n_species;
  const int m1 = 2 * L->n1, m2 = 2 * L->n2, m3 = 2 * L->n3;
  int counts[16] = {0};
  for (size_t q = 0; q < (size_t)S * S * wc_range; ++q) rho[q] = 0.0;
  for (int k = 1; k <= m3; ++k)
    for (int j = 1; j <= m2; ++j)
      for (int i = 1; i <= m1; ++i) {
        int s = cfg_at(L, cfg, i, j, k);
        if (s >= 1 && s <= S) counts[s - 1]++;
      }
  const int w1 = L->n1 < 5 ? L->n1 : 5, w2 = L->n2 < 5 ? L->n2 : 5, w3 = L->n3 < 5 ? L->n3 : 5;
  for (int i3 = 1; i3 <= m3; ++i3)
    for (int i2 = 1; i2 <= m2; ++i2)
      for (int i1 = 1; i1 <= m1; ++i1) {
        int si = cfg_at(L, cfg, i1, i2, i3);
        if (si == 0) continue;
        for (int jj3 = i3 - w3; jj3 <= i3 + w3; ++jj3) {
          int j3 = fmod_pos(jj3 - 1, m3) + 1;
          for (int jj2 = i2 - w2; jj2 <= i2 + w2; ++jj2) {
            int j2 = fmod_pos(jj2 - 1, m2) + 1;
            for (int jj1 = i1 - w1; jj1 <= i1 + w1; ++jj1) {
              int j1 = fmod_pos(jj1 - 1, m1) + 1;
              int sj = cfg_at(L, cfg, j1, j2, j3);
              if (sj == 0) continue;
              double dx = (double)(float)(i1 - j1);
              double dy = (double)(float)(i2 - j2);
              double dz = (double)(float)(i3 - j3);
              dx = dx - (double)(float)m1 * (double)lround(dx / (double)(float)m1);
              dy = dy - (double)(float)m2 * (double)lround(dy / (double)(float)m2);
              dz = dz - (double)(float)m3 * (double)lround(dz / (double)(float)m3);
              double dist = sqrt(dx * dx + dy * dy + dz * dz);
              for (int l = 0; l < wc_range; ++l)
                if (fabs(dist - shells[l]) < 1e-3)
                  rho[(si - 1) + S * ((sj - 1) + S * l)] += 1.0;
            }
          }
        }
      }
  for (int l = 0; l < wc_range; ++l)
    for (int a = 0; a < S; ++a)
      for (int b = 0; b < S; ++b) rho[a + S * (b + S * l)] = rho[a + S * (b + S * l)] / counts[a];
}

/* ------------------------------------------------------------------ */
/* Nested sampling (nested_sampling.f90:45-206)                        */
/* ------------------------------------------------------------------ */
int32_t bwo_nested_sampling(const bwo_lattice *L, const int64_t *species_numbers,
                            const double *concentrations, bwo_mt *g, int32_t n_walkers,
                            int32_t n_steps, int32_t n_iter, double *e_limits) {
  const size_t cells = (size_t)8 * L->n1 * L->n2 * L->n3;
  int8_t *walkers = (int8_t *)malloc(cells * (size_t)n_walkers);
  double *E = (double *)malloc(sizeof(double) * (size_t)n_walkers);
  if (!walkers || !E) { free(walkers); free(E); return -1; }
  const int n_at = L->n1 * L->n2 * L->n3 * 1 * L->n_species; /* :82 */
  for (int w = 0; w < n_walkers; ++w) {
    int8_t *c = walkers + cells * (size_t)w;
    int rc = bwo_initial_setup(L, species_numbers, concentrations, g, c);
    if (rc) { free(walkers); free(E); return rc; }
    double rnde = bwo_genrand(g);
    /* 1e-8 is a default-real (REAL(4)) literal in the reference (:95) */
    E[w] = bwo_total_energy(L, c) + rnde * (double)1e-8f;
  }
  int extra_steps = 0;
  int n_acc = 0;
  for (int it = 1; it <= n_iter; ++it) {
    int imax = 0; /* maxloc: first maximum */
    for (int w = 1; w < n_walkers; ++w)
      if (E[w] > E[imax]) imax = w;
    double ener_limit = E[imax];
    e_limits[it - 1] = ener_limit;
    /* walk-length adaptation (:129-144) */
    if (it % (int)((float)n_walkers / 2.0f) == 0) {
      if (((float)n_acc < (float)n_at * 0.05f) && (extra_steps < n_steps * 100)) extra_steps += n_steps;
    }
    double rnd = bwo_genrand(g);
    int irnd = (int)ceil(rnd * n_walkers);
    if (irnd < 1) irnd = 1; /* genrand() == 0 exactly; the reference indexes 0 */
    int8_t *cm = walkers + cells * (size_t)imax;
    if (irnd - 1 != imax) memcpy(cm, walkers + cells * (size_t)(irnd - 1), cells);
    E[imax] = E[irnd - 1];
    n_acc = 0;
    for (int st = 1; st <= n_steps + extra_steps; ++st) {
      int32_t r1[3], r2[3];
      bwo_random_site(L, g, r1);
      int s1 = cm[cfg_index(L, r1[0], r1[1], r1[2])];
      int s2;
      do {
        bwo_random_site(L, g, r2);
        s2 = cm[cfg_index(L, r2[0], r2[1], r2[2])];
      } while (s1 == s2);
      double eu = bwo_pair_energy(L, cm, r1, r2);
      bwo_pair_swap(L, cm, r1, r2);
      double es = bwo_pair_energy(L, cm, r1, r2);
      double de = es - eu;
      if (E[imax] + de < ener_limit) {
        E[imax] = E[imax] + de;
        n_acc++;
      } else {
        bwo_pair_swap(L, cm, r1, r2);
      }
    }
  }
  free(walkers);
  free(E);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Wang-Landau (wang-landau.F90)                                       */
/* ------------------------------------------------------------------ */
/* bin_index (:515-523): truncation toward zero, then +1 */
int32_t bwo_wl_bin_index(double energy, const double *bin_edges, int32_t bins) {
  double range = bin_edges[bins] - bin_edges[0];
  return (int32_t)(((energy - bin_edges[0]) / range) * (double)bins) + 1;
}

/* sweeps (:539-632) for one walker; the radial-density sampling and the
 * config_per_bin store consume no random numbers and are omitted. */
int64_t bwo_wl_sweeps(const bwo_lattice *L, int8_t *cfg, bwo_mt *g, int32_t nbr_swap,
                      int32_t mc_sweeps, const double *bin_edges, int32_t bins,
                      int32_t start_idx, int32_t end_idx, double wl_f, double *lng,
                      double *hist) {
  const int n_atoms = bwo_n_atoms(L);
  const int cadence = (int)(0.02 * (double)(float)n_atoms);
  double e_unswapped = bwo_total_energy(L, cfg);
  int64_t accepted = 0;
  const int64_t total = (int64_t)mc_sweeps * n_atoms;
  for (int64_t i = 1; i <= total; ++i) {
    int32_t r1[3], r2[3];
    bwo_random_site(L, g, r1);
    if (nbr_swap)
      bwo_random_nbr(L, g, r1, r2);
    else
      bwo_random_site(L, g, r2);
    int s1 = cfg[cfg_index(L, r1[0], r1[1], r1[2])];
    int s2 = cfg[cfg_index(L, r2[0], r2[1], r2[2])];
    double e_swapped = e_unswapped;
    double pu = bwo_pair_energy(L, cfg, r1, r2);
    bwo_pair_swap(L, cfg, r1, r2);
    if (s1 != s2) {
      double ps = bwo_pair_energy(L, cfg, r1, r2);
      e_swapped = e_unswapped - pu + ps;
    }
    int ibin = bwo_wl_bin_index(e_unswapped, bin_edges, bins);
    int jbin = bwo_wl_bin_index(e_swapped, bin_edges, bins);
    int hit = cadence > 0 && (i % cadence) == 0;
    if (jbin > start_idx - 1 && jbin < end_idx + 1) {
      double li = (ibin >= 1 && ibin <= bins) ? lng[ibin - 1] : 0.0;
      if (log(bwo_genrand(g)) < (li - lng[jbin - 1])) {
        accepted++;
        e_unswapped = e_swapped;
      } else {
        bwo_pair_swap(L, cfg, r1, r2);
        jbin = ibin;
      }
    } else {
      jbin = ibin;
      bwo_pair_swap(L, cfg, r1, r2);
    }
    /* a walker outside its window would index out of bounds in the
     * reference; such updates are dropped here */
    if (hit && jbin >= start_idx && jbin <= end_idx) hist[jbin - start_idx] += 1.0;
    if (jbin >= 1 && jbin <= bins) lng[jbin - 1] = lng[jbin - 1] + wl_f;
  }
  return accepted;
}

/* enter_energy_window (:643-741) for one walker, without the MPI hand-off
 * between walkers of a window. */
int64_t bwo_wl_enter_window(const bwo_lattice *L, int8_t *cfg, bwo_mt *g, double min_e,
                            double max_e, float energy_min, float energy_max,
                            const int64_t *species_numbers, const double *concentrations,
                            int64_t max_trials) {
  const int n_atoms = bwo_n_atoms(L);
  const double target = (min_e + max_e) / 2.0;
  const double condition = fabs(max_e - min_e) * 0.1;
  /* sigma of the bias: REAL(4) product, then FP64 division (:724-725) */
  float sf = 0.0025f * fabsf(energy_max - energy_min);
  sf = sf * (float)n_atoms;
  const double sigma = (double)sf / (RY_TO_EV * 1000.0);
  const double denom = 2.0 * (sigma * sigma);
  double e_unswapped = bwo_total_energy(L, cfg);
  int64_t i_steps = 0, trials = 0;
  for (;;) {
    i_steps++;
    if (i_steps % ((int64_t)n_atoms * 50 * 5) == 0) {
      i_steps = 0;
      bwo_initial_setup(L, species_numbers, concentrations, g, cfg);
    }
    if (e_unswapped < max_e - condition && e_unswapped > min_e + condition) {
      e_unswapped = bwo_total_energy(L, cfg);
      if (e_unswapped < max_e - condition && e_unswapped > min_e + condition) break;
      continue;
    }
    if (trials >= max_trials) break;
    trials++;
    int32_t r1[3], r2[3];
    bwo_random_site(L, g, r1);
    bwo_random_site(L, g, r2);
    int s1 = cfg[cfg_index(L, r1[0], r1[1], r1[2])];
    int s2 = cfg[cfg_index(L, r2[0], r2[1], r2[2])];
    if (s1 != s2) {
      double pu = bwo_pair_energy(L, cfg, r1, r2);
      bwo_pair_swap(L, cfg, r1, r2);
      double ps = bwo_pair_energy(L, cfg, r1, r2);
      double e_swapped = e_unswapped - pu + ps;
      double de = ((e_swapped - target) * (e_swapped - target) -
                   (e_unswapped - target) * (e_unswapped - target)) / denom;
      if (log(bwo_genrand(g)) < -de)
        e_unswapped = e_swapped;
      else
        bwo_pair_swap(L, cfg, r1, r2);
    }
  }
  return trials;
}

/* divide_range (:855-891), create_overlap (:934-955), create_energy_bins
 * (:969-992) */
void bwo_wl_setup(const bwo_lattice *L, int32_t bins, int32_t num_windows, float bin_overlap,
                  float energy_min, float energy_max, int32_t *windows, double *bin_edges) {
  int32_t iv[2 * 64];
  const double b = bins, n = num_windows, power = 1.0;
  const double factor = (b - 1.0) / (pow(n + 1.0, power) - 1.0);
  iv[0] = 1;
  iv[2 * (num_windows - 1) + 1] = bins;
  for (int i = 2; i <= num_windows; ++i) {
    iv[2 * (i - 2) + 1] = (int32_t)floor(factor * pow((double)(i - 1), power) + 1.0);
    iv[2 * (i - 1)] = iv[2 * (i - 2) + 1] + 1;
  }
  for (int q = 0; q < 2 * num_windows; ++q) windows[q] = iv[q];
  if (num_windows > 1) {
    for (int i = 2; i <= num_windows - 1; ++i) {
      int nb = windows[2 * (i - 2) + 1] - windows[2 * (i - 2)] + 1;
      int ov = (int)ceilf(bin_overlap * (float)nb);
      if (ov < 2) ov = 2;
      windows[2 * (i - 1)] = iv[2 * (i - 1)] - ov;
      windows[2 * (i - 1) + 1] = iv[2 * (i - 1) + 1];
    }
    int nw = num_windows;
    int nb = windows[2 * (nw - 2) + 1] - windows[2 * (nw - 2)];
    int ov = (int)ceilf(bin_overlap * (float)nb);
    if (ov < 2) ov = 2;
    windows[2 * (nw - 1)] = iv[2 * (nw - 1)] - ov;
    windows[2 * (nw - 1) + 1] = iv[2 * (nw - 1) + 1];
  }
  const double to_ry = (double)bwo_n_atoms(L) / (RY_TO_EV * 1000.0);
  const double width = (double)((energy_max - energy_min) / (float)bins) * to_ry;
  for (int i = 1; i <= bins + 1; ++i)
    bin_edges[i - 1] = (double)energy_min * to_ry + (double)(i - 1) * width;
}

/* replica_exchange (:1392-1501) over num_windows*wpw walkers, walker k in
 * window k/wpw + 1 standing for MPI rank k (walker 0 is rank 0, whose stream
 * shuffles the candidate rows).  lng is W rows of `bins` values; windows is
 * num_windows (start, end) pairs.  The configurations of accepted pairs are
 * exchanged in cfgs; partner/accepted (W each) record every matched pair.
 * Returns the number of exchanged pairs. */
static void shuffle_rows(int32_t *rows, int n, bwo_mt *g) {
  /* shuffle_rows (:1513-1524): Fisher-Yates from the last row down */
  for (int i = n; i >= 2; --i) {
    int r = 1 + (int)(bwo_genrand(g) * (double)i);
    if (r > n) r = n;
    int32_t t0 = rows[2 * (i - 1)], t1 = rows[2 * (i - 1) + 1];
    rows[2 * (i - 1)] = rows[2 * (r - 1)];
    rows[2 * (i - 1) + 1] = rows[2 * (r - 1) + 1];
    rows[2 * (r - 1)] = t0;
    rows[2 * (r - 1) + 1] = t1;
  }
}

int32_t bwo_wl_replica_exchange(const bwo_lattice *L, int8_t *const *cfgs, bwo_mt *const *gens,
                                const double *lng, int32_t bins, const double *bin_edges,
                                const int32_t *windows, int32_t num_windows, int32_t wpw,
                                int32_t *partner, int32_t *accepted) {
  const int W = num_windows * wpw;
  const size_t bytes = (size_t)8 * L->n1 * L->n2 * L->n3;
  int32_t *ibin = malloc(sizeof(int32_t) * (size_t)W);
  int32_t *loc = malloc(sizeof(int32_t) * (size_t)W);
  int32_t *lo = malloc(sizeof(int32_t) * 2 * (size_t)wpw);
  int32_t *up = malloc(sizeof(int32_t) * 2 * (size_t)wpw);
  int32_t *pr = malloc(sizeof(int32_t) * 2 * (size_t)wpw);
  int8_t *tmp = malloc(bytes);
  for (int k = 0; k < W; ++k) {
    /* e_unswapped and the overlap location of rank k (:1406-1431) */
    const int b = bwo_wl_bin_index(bwo_total_energy(L, cfgs[k]), bin_edges, bins);
    const int m = k / wpw + 1;
    int lower = 0, upper = 0;
    if (m > 1) lower = b < windows[2 * (m - 2) + 1] + 1 && b > windows[2 * (m - 1)] - 1;
    if (m < num_windows) upper = b > windows[2 * m] - 1 && b < windows[2 * (m - 1) + 1] + 1;
    ibin[k] = b;
    loc[k] = upper ? m : (lower ? m - 1 : 0);
    partner[k] = -1;
    accepted[k] = 0;
  }
  int32_t exchanged = 0;
  for (int i = 1; i < num_windows; ++i) {
    /* rank 0 matches the rows of windows i and i+1 (:1440-1470) */
    for (int j = 0; j < wpw; ++j) {
      lo[2 * j] = (i - 1) * wpw + j;
      lo[2 * j + 1] = loc[(i - 1) * wpw + j];
      up[2 * j] = i * wpw + j;
      up[2 * j + 1] = loc[i * wpw + j];
    }
    shuffle_rows(lo, wpw, gens[0]);
    shuffle_rows(up, wpw, gens[0]);
    int np = 0;
    for (int j = 0; j < wpw; ++j) {
      if (lo[2 * j + 1] == 0) continue;
      for (int k = 0; k < wpw; ++k) {
        if (up[2 * k + 1] == 0) continue;
        if (lo[2 * j + 1] == up[2 * k + 1]) {
          pr[2 * np] = lo[2 * j];
          pr[2 * np + 1] = up[2 * k];
          ++np;
          lo[2 * j] = lo[2 * j + 1] = 0;
          up[2 * k] = up[2 * k + 1] = 0;
        }
      }
    }
    /* each pair in order (:1476-1497): the lower walker decides */
    for (int j = 0; j < np; ++j) {
      const int a = pr[2 * j], b = pr[2 * j + 1];
      const double u = bwo_genrand(gens[a]);
      const double x = lng[(size_t)bins * a + ibin[a] - 1] - lng[(size_t)bins * a + ibin[b] - 1];
      const int acc = u < exp(x);
      partner[a] = b;
      partner[b] = a;
      accepted[a] = accepted[b] = acc;
      if (acc) {
        memcpy(tmp, cfgs[a], bytes);
        memcpy(cfgs[a], cfgs[b], bytes);
        memcpy(cfgs[b], tmp, bytes);
        ++exchanged;
      }
    }
  }
  free(ibin);
  free(loc);
  free(lo);
  free(up);
  free(pr);
  free(tmp);
  return exchanged;
}
