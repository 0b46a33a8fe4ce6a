// bwl_metro_par.hip -- kernel choice, LDS sizing and launch of the parallel
// Metropolis sampler (kernels in bwl_metro_kernels.h); the fixed-pitch
// natural-layout variants are instantiated in bwl_metro_fixed.hip.
#include "bwl_metro_kernels.h"

namespace bwl {

namespace {

// count words: 0 = 32-bit with 8-bit fields (4 species), 1 = 32-bit with
// 6-bit fields (5 species; shells of at most 31 sites), 2 = 64-bit with
// 8-bit fields (8 species)
template <int LAT, int R>
const void *pick(bool nat, int words) {
  if (nat) {
    if (words == 0) return (const void *)k_metro_nat<LAT, R, uint32_t, 0, 8>;
    if constexpr (R <= 6)
      if (words == 1) return (const void *)k_metro_nat<LAT, R, uint32_t, 0, 6>;
    return words == 2 ? (const void *)k_metro_nat<LAT, R, uint64_t, 0, 8> : nullptr;
  }
  return words == 2 ? (const void *)k_metro_cmp<LAT, R, uint64_t, 8>
                    : (const void *)k_metro_cmp<LAT, R, uint32_t, 6>;
}

// Shapes with compile-time neighbour offsets: the quaternary bcc anneal
// (bcc, 4 shells, 32^3 cells) and the binary fcc run (fcc, 2 shells, 8^3).
const void *specialised(int lattice, int R, int n, int words) {
  if (words != 0) return nullptr;
  if (lattice == 2 && R == 4 && n == 32) return (const void *)k_metro_nat<2, 4, uint32_t, 32, 8>;
  if (lattice == 1 && R == 2 && n == 8) return (const void *)k_metro_nat<1, 2, uint32_t, 8, 8>;
  return nullptr;
}

const void *par_kernel(int lattice, int R, int n, bool nat, int words) {
  if (nat) {
    const void *f = specialised(lattice, R, n, words);
    if (f) return f;
  }
  if (lattice == 1) {
    switch (R) {
      case 1: return pick<1, 1>(nat, words);
      case 2: return pick<1, 2>(nat, words);
      case 3: return pick<1, 3>(nat, words);
      case 4: return pick<1, 4>(nat, words);
      case 5: return pick<1, 5>(nat, words);
      case 6: return pick<1, 6>(nat, words);
    }
  } else {
    switch (R) {
      case 1: return pick<2, 1>(nat, words);
      case 2: return pick<2, 2>(nat, words);
      case 3: return pick<2, 3>(nat, words);
      case 4: return pick<2, 4>(nat, words);
      case 5: return pick<2, 5>(nat, words);
      case 6: return pick<2, 6>(nat, words);
      case 7: return pick<2, 7>(nat, words);
      case 8: return pick<2, 8>(nat, words);
      case 9: return pick<2, 9>(nat, words);
      case 10: return pick<2, 10>(nat, words);
    }
  }
  return nullptr;
}

}  // namespace

int par_count_words(const Geo &g, bool nat) {
  if (!nat) return g.S > 5 ? 2 : 1;
  if (g.S <= 4) return 0;
  return g.S == 5 && g.R <= 6 ? 1 : 2;  // the shells of R <= 6 hold <= 24 sites
}

size_t par_lds_bytes(const Geo &g, const NatGeo *nat, int threads) {
  const size_t red = sizeof(double) * (threads / 64 + 1);
  // dtab rows as in k_metro_nat: R * S1P doubles made odd, S1P the species
  // a count word holds, less one
  const int words = par_count_words(g, nat != nullptr);
  const size_t S1P = words == 0 ? 3 : (words == 1 ? 4 : 7);
  const size_t rstr = (g.R * S1P) | 1;
  if (nat)
    return (size_t)((nat->lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * rstr +
           sizeof(double) * g.S * g.R * g.S + sizeof(RoundP) * kRoundBatch +
           sizeof(int32_t) * ((g.z + 3) & ~3) + red + sizeof(unsigned long long) * g.S * g.S * g.R;
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * g.R + red +
         sizeof(int32_t) * g.nsub * g.z;
}

bool par_specialised(const Geo &g) {
  return specialised(g.lattice, g.R, g.n[0], par_count_words(g, true)) != nullptr;
}

int par_max_threads(const Geo &g, bool nat, bool fixed) {
  // the compile-time-offset kernels with 32-bit count words (launch bounds
  // of k_metro_nat)
  const bool narrow = par_count_words(g, nat) != 2;
  return narrow && (fixed || (nat && par_specialised(g))) ? kMaxNatThreads : 512;
}

const void *par_kernel_for(const Geo &g, bool nat, bool fixed) {
  if (fixed) return par_fixed_kernel(g.lattice, g.R, par_count_words(g, true));
  return par_kernel(g.lattice, g.R, g.n[0], nat, par_count_words(g, nat));
}

hipError_t launch_metro_par_fn(const void *fn, const ParArgs *args_dev, int n_chains, size_t lds,
                               int threads, hipStream_t s) {
  if (!fn) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void *args[] = {(void *)&args_dev};
  return hipLaunchKernel(fn, dim3(n_chains), dim3(threads), args, lds, s);
}


}  // namespace bwl
