// bwl_host.h -- host-side logic of the library (no HIP calls): lattice
// geometry and neighbour offsets, the reference's MT19937 stream, the
// initial configuration, and the round plan of the parallel sampler.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "bwl_core.h"
#include "../../include/brawl_hip.h"

namespace bwl {

// Builds the geometry and the [nsub][z] padded-layout offset table.
// Returns BWL_OK or an error code with `err` set.
int make_geo(const bwl_lattice_desc &d, Geo &g, std::vector<int32_t> &off, std::string &err);

// Reference MT19937 (mt19937ar.c:63-125), 32-bit words.
struct MT19937 {
  uint32_t mt[624];
  int32_t mti = 625;
  void seed(uint32_t s);
  uint32_t next();
  double genrand() { return (double)next() * (1.0 / 4294967296.0); }
};

// initial_setup (initialise.F90:434-617) into a reference-layout config.
int initial_setup_host(const Geo &g, const int64_t *species_numbers, const double *concentrations,
                       MT19937 &mt, int8_t *config, std::string &err);

// Round plan of the parallel sampler (see RoundVar in bwl_core.h).
struct ParPlan {
  int32_t stride, L;
  int32_t K[3];
  std::vector<RoundVar> vars;
  std::vector<uint32_t> muls;      // per round kind, coprime with its nL1
  uint32_t units[3][16];           // per axis, coprime with K[axis]
  int32_t n_units[3];
  bool fast_halo;
};
int plan_parallel(const Geo &g, ParPlan &p, std::string &err);

// Natural layout of a cubic supercell; false if the lattice is not cubic or
// an axis is shorter than twice its halo.
bool make_nat(const Geo &g, NatGeo &nat);
// The fixed-pitch natural layout (nat_dims_fixed); false if the lattice is
// not cubic, too small for the halo or larger than the pitch covers.
bool make_nat_fixed(const Geo &g, NatGeo &nat);

// lattice_shells (analytics.f90:205-275): the first `wc` distinct REAL(4)
// distances from site (1,1,1) to the occupied sites (0 first), in doubled-
// grid units.  Shells the lattice does not have stay 0.
void lattice_shells_host(const Geo &g, int wc, double *shells);

// Window offsets of radial_densities (analytics.f90:293-404), grouped by
// shell: offsets (dx,dy,dz) in the +-min(n,5) doubled-cell window (listed
// with multiplicity when the window wraps a short axis) that join two
// occupied sites whose minimum-image distance is within 1e-3 of shell l.
// off3 gets 3 ints per entry; shell_start gets wc+1 entries.
void asro_offsets_host(const Geo &g, int wc, const double *shells, std::vector<int32_t> &off3,
                       int32_t *shell_start);

// divide_range + create_overlap + create_energy_bins (wang-landau.F90:
// 855-992): windows gets (start, end) 1-based bins per window, bin_edges
// bins+1 energies (Ry, whole lattice).
void wl_setup_host(const Geo &g, int bins, int num_windows, float bin_overlap, float energy_min,
                   float energy_max, int32_t *windows, double *bin_edges);

// Validate a reference-layout config: species 1..S on sites, 0 elsewhere.
int check_config(const Geo &g, const int8_t *config, std::string &err);

}  // namespace bwl
