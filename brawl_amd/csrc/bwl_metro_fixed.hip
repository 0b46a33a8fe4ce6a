// bwl_metro_fixed.hip -- the fixed-pitch natural-layout variants of the
// parallel Metropolis kernel (one per lattice, shell count <= 4 and count
// word): every cubic supercell up to nat_fixed_nmax cells per axis shares
// one row length and plane count, so the neighbour offsets are compile-time
// immediates as in the single-shape kernels.  Own translation unit so that
// it compiles in parallel with bwl_metro_par.hip.
#include "bwl_metro_kernels.h"

namespace bwl {

namespace {

template <int LAT, int R>
const void *fixed_of(int words) {
  if (words == 0) return (const void *)k_metro_nat<LAT, R, uint32_t, kFixedPitch, 8>;
  if (words == 1) return (const void *)k_metro_nat<LAT, R, uint32_t, kFixedPitch, 6>;
  return (const void *)k_metro_nat<LAT, R, uint64_t, kFixedPitch, 8>;
}

}  // namespace

const void *par_fixed_kernel(int lattice, int R, int words) {
  if (lattice == 1) {
    switch (R) {
      case 1: return fixed_of<1, 1>(words);
      case 2: return fixed_of<1, 2>(words);
      case 3: return fixed_of<1, 3>(words);
      case 4: return fixed_of<1, 4>(words);
    }
  } else {
    switch (R) {
      case 1: return fixed_of<2, 1>(words);
      case 2: return fixed_of<2, 2>(words);
      case 3: return fixed_of<2, 3>(words);
      case 4: return fixed_of<2, 4>(words);
    }
  }
  return nullptr;
}

}  // namespace bwl
