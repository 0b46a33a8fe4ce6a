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

// Validate a reference-layout config: species 1..S on sites, 0 elsewhere.
int check_config(const Geo &g, const int8_t *config, std::string &err);

}  // namespace bwl
