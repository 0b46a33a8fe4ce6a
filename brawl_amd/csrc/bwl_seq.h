// bwl_seq.h -- device pieces of the sequential (one wavefront per chain)
// samplers: the reference's random-site draws from the MT19937 stream, site
// and pair energies in the reference's FP64 order, and whole-lattice
// energies (histogram order and reference order).
#pragma once

#include "bwl_device.h"

namespace bwl {

// ------------------------------------------------------------------
// One site's energy in the reference's order: lane n sums shell n+1 from
// 0.0 in list order (bw_hamiltonian.f90 *_shellN_energy), then the shells
// are added left to right (*_energy_Rshells).
// ------------------------------------------------------------------
__device__ inline double shell_sum_lane(const int8_t *lat, const Geo &g, const int32_t *off,
                                        const double *vex, int addr, int sub, int species, int n) {
  const int32_t *o = off + sub * g.z;
  double e = 0.0;
  for (int t = g.shell_start[n]; t < g.shell_start[n + 1]; ++t) {
    int nb = lat[addr + o[t]];
    e = e + vex[(species - 1) + g.S * ((nb - 1) + g.S * n)];
  }
  return e;
}

// Value of `v` in lane `l` (l uniform across the wavefront).
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Copy the [nsub][z] neighbour-offset table into LDS (`dst`).
__device__ inline void stage_offsets(int32_t *dst, const int32_t *__restrict__ src, const Geo &g) {
  for (int i = threadIdx.x; i < g.nsub * g.z; i += blockDim.x) dst[i] = src[i];
}

// Per-lane constants of pair_energy_lanes, computed once per kernel: the
// shell of each lane's entries and, per 64-entry pass, the mask of entries
// that end a shell.  Entries 0..z-1 are site A's neighbours in list order,
// z..2z-1 site B's.
struct SiteXYZ {
  int x, y, z;  // 0-based doubled-grid coordinates
};

// A doubled-coordinate difference (dx, dy, dz) reduced mod the periods m
// (each below 1024), 10 bits per axis: two sites' differences are equal mod
// the periodic box exactly when their packed words are equal.
__device__ inline int pack_dmod(int dx, int dy, int dz, const int32_t *m) {
  return pmod(dx, m[0]) | (pmod(dy, m[1]) << 10) | (pmod(dz, m[2]) << 20);
}

// The same for the difference of two sites' coordinates (each in [0, m)).
__device__ inline int pack_site_diff(SiteXYZ a, SiteXYZ b, const int32_t *m) {
  const int dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return (dx < 0 ? dx + m[0] : dx) | ((dy < 0 ? dy + m[1] : dy) << 10) |
         ((dz < 0 ? dz + m[2] : dz) << 20);
}

struct PairLanes {
  unsigned long long end_mask[6];
  int shell[6];
  int dmod[6];   // the lane's neighbour offset mod (m0, m1, m2), 10 bits each
  int chain[6];  // the lane's slot in the shell-chain layout (fixed_pair_sum2_chains)
  int passes;
  int fixed;  // compile-time summation variant (0: generic loop)
};

// First list entry of shell n (0-based) of a lattice, at compile time.
template <int LAT>
__host__ __device__ constexpr int shell_first(int n) {
  int s = 0;
  for (int k = 0; k < n; ++k) s += LAT == 1 ? kFccShellSize[k] : kBccShellSize[k];
  return s;
}

// Largest shell of the first r shells of a lattice (lattice 1: fcc).
__host__ __device__ constexpr int shell_max(int lattice, int r) {
  int m = 0;
  for (int k = 0; k < r; ++k) {
    const int v = lattice == 1 ? kFccShellSize[k] : kBccShellSize[k];
    m = v > m ? v : m;
  }
  return m;
}

// The ordered sum of pair_energy_lanes with the shell structure known at
// compile time: straight-line readlane/add code, no loop or branches.
template <int LAT, int R>
__device__ __forceinline__ double fixed_pair_sum(const double (&vals)[6], double *eA_out) {
  constexpr int z = shell_first<LAT>(R);
  double eA = 0.0, eB = 0.0;
#pragma unroll
  for (int n = 0; n < R; ++n) {
    double es = 0.0;
#pragma unroll
    for (int t = shell_first<LAT>(n); t < shell_first<LAT>(n + 1); ++t)
      es = es + readlane_d(vals[t >> 6], t & 63);
    eA = n == 0 ? es : eA + es;
  }
#pragma unroll
  for (int n = 0; n < R; ++n) {
    double es = 0.0;
#pragma unroll
    for (int t = z + shell_first<LAT>(n); t < z + shell_first<LAT>(n + 1); ++t)
      es = es + readlane_d(vals[t >> 6], t & 63);
    eB = n == 0 ? es : eB + es;
  }
  if (eA_out) *eA_out = eA;
  return eA + eB;
}

// Bytes of LDS for the sums' entries (fixed_pair_sum2_chains: 4R chains of
// a stride of shell_max | 1 doubles, at most 24 x 25).
constexpr int kSumBufBytes = 4800;

__device__ inline void init_pair_lanes(const Geo &g, PairLanes &pl) {
  const int lane = threadIdx.x & 63;
  const int z = g.z;
  int bounds[kMaxShells + 1];
#pragma unroll
  for (int i = 0; i <= kMaxShells; ++i) bounds[i] = g.shell_start[i];
  pl.passes = (2 * z + 63) / 64;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int t = 64 * k + lane;
    const int tt = t < z ? t : t - z;
    int n = 0;
    bool last = false;
#pragma unroll
    for (int i = 1; i <= kMaxShells; ++i) {
      if (i <= g.R && tt >= bounds[i]) n = i;
      if (i <= g.R && tt + 1 == bounds[i]) last = true;
    }
    pl.shell[k] = n;
    pl.end_mask[k] = __ballot(t < 2 * z && last);
    int d = 0;
    if (t < 2 * z) {
      const int8_t *v = g.lattice == 1 ? kFccOffsets[tt] : kBccOffsets[tt];
      d = pack_dmod(v[0], v[1], v[2], g.m);
    }
    pl.dmod[k] = d;
    // chain slot: (site * R + shell) * (shell_max | 1) + index within the shell
    const int st = shell_max(g.lattice, g.R) | 1;
    pl.chain[k] = t < 2 * z ? ((t < z ? 0 : g.R) + n) * st + (tt - bounds[n]) : 0;
  }
  // straight-line sums for every fcc shell count and bcc up to 6 shells
  pl.fixed = pair_sum_key(g.lattice, g.R);
}

// The ordered sum of pair_energy_lanes from the lanes' values (entries
// 0..z-1 site A's, z..2z-1 site B's): e_A + e_B, each site's shells left to
// right, each shell from 0.0 in list order.
template <class GEO>
__device__ inline double pair_sum_loop(const GEO &g, const PairLanes &pl,
                                       const double (&vals)[6], double *eA_out);

template <class GEO>
__device__ inline double pair_sum(const GEO &g, const PairLanes &pl, const double (&vals)[6],
                                  double *eA_out) {
  switch (pl.fixed) {
    case 16 + 1: return fixed_pair_sum<1, 1>(vals, eA_out);
    case 16 + 2: return fixed_pair_sum<1, 2>(vals, eA_out);
    case 16 + 3: return fixed_pair_sum<1, 3>(vals, eA_out);
    case 16 + 4: return fixed_pair_sum<1, 4>(vals, eA_out);
    case 16 + 5: return fixed_pair_sum<1, 5>(vals, eA_out);
    case 16 + 6: return fixed_pair_sum<1, 6>(vals, eA_out);
    case 32 + 1: return fixed_pair_sum<2, 1>(vals, eA_out);
    case 32 + 2: return fixed_pair_sum<2, 2>(vals, eA_out);
    case 32 + 3: return fixed_pair_sum<2, 3>(vals, eA_out);
    case 32 + 4: return fixed_pair_sum<2, 4>(vals, eA_out);
    case 32 + 5: return fixed_pair_sum<2, 5>(vals, eA_out);
    case 32 + 6: return fixed_pair_sum<2, 6>(vals, eA_out);
    default: break;
  }
  return pair_sum_loop(g, pl, vals, eA_out);
}

// The same sum for any shell structure, driven by the lanes' shell-end mask.
template <class GEO>
__device__ inline double pair_sum_loop(const GEO &g, const PairLanes &pl,
                                       const double (&vals)[6], double *eA_out) {
  const int z = g.z;
  double eA = 0.0, e = 0.0, es = 0.0;
  bool first = true;
  int t = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (k >= pl.passes) break;
    const int lim = 2 * z - 64 * k < 64 ? 2 * z - 64 * k : 64;
    const unsigned long long em = pl.end_mask[k];
    for (int l = 0; l < lim; ++l, ++t) {
      es = es + readlane_d(vals[k], l);
      if ((em >> l) & 1ull) {
        e = first ? es : e + es;
        es = 0.0;
        first = false;
        if (t == z - 1) {
          eA = e;
          first = true;
        }
      }
    }
  }
  if (eA_out) *eA_out = eA;
  return eA + e;
}

// fixed_pair_sum2 with the shells' sums formed side by side in lanes: the
// 4R chains (before / after the swap, site A / B, shell n) each run in a
// lane of their own over the shell's entries in list order, from 0.0, then
// every lane adds each site's shell sums left to right and e_A + e_B.  The
// entries are passed through `buf` (kSumBufBytes, its chain slots past each
// shell's length zero: adding +0.0 to a sum that started at +0.0 leaves it
// unchanged, since such a sum is never -0.0), so every addition is the one
// of the reference's order, on the same values, and the results are
// bit-identical, with 4R-fold fewer additions per wavefront.  One wavefront
// per workgroup (its LDS accesses complete in program order).
template <int LAT, int R>
__device__ __forceinline__ void fixed_pair_sum2_chains(const PairLanes &pl, const double (&v1)[6],
                                                       const double (&v2)[6], double *buf, double &e1,
                                                       double &e2) {
  constexpr int z = shell_first<LAT>(R);
  constexpr int P = (2 * z + 63) / 64;
  constexpr int ML = shell_max(LAT, R), ST = ML | 1, NC = 4 * R;
  static_assert(NC * ST * 8 <= kSumBufBytes && NC <= 64, "sum buffer too small");
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (64 * k + lane < 2 * z) {
      buf[pl.chain[k]] = v1[k];
      buf[2 * R * ST + pl.chain[k]] = v2[k];
    }
  }
  const double *cb = buf + (lane < NC ? lane : NC - 1) * ST;
  double c = 0.0;
#pragma unroll
  for (int j = 0; j < ML; ++j) c = c + cb[j];
  double a[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {  // (before, A), (before, B), (after, A), (after, B)
    a[w] = readlane_d(c, w * R);
#pragma unroll
    for (int n = 1; n < R; ++n) a[w] = a[w] + readlane_d(c, w * R + n);
  }
  e1 = a[0] + a[1];
  e2 = a[2] + a[3];
}

// Both ordered sums of pair_energy_swap_lanes, the shell structure chosen
// at run time (the chains for every fcc shell count and bcc up to 6
// shells, the generic loop beyond).
template <class GEO>
__device__ inline void pair_sum2(const GEO &g, const PairLanes &pl, const double (&v1)[6],
                                 const double (&v2)[6], double *buf, double &e1, double &e2) {
  switch (pl.fixed) {
    case 16 + 1: fixed_pair_sum2_chains<1, 1>(pl, v1, v2, buf, e1, e2); return;
    case 16 + 2: fixed_pair_sum2_chains<1, 2>(pl, v1, v2, buf, e1, e2); return;
    case 16 + 3: fixed_pair_sum2_chains<1, 3>(pl, v1, v2, buf, e1, e2); return;
    case 16 + 4: fixed_pair_sum2_chains<1, 4>(pl, v1, v2, buf, e1, e2); return;
    case 16 + 5: fixed_pair_sum2_chains<1, 5>(pl, v1, v2, buf, e1, e2); return;
    case 16 + 6: fixed_pair_sum2_chains<1, 6>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 1: fixed_pair_sum2_chains<2, 1>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 2: fixed_pair_sum2_chains<2, 2>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 3: fixed_pair_sum2_chains<2, 3>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 4: fixed_pair_sum2_chains<2, 4>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 5: fixed_pair_sum2_chains<2, 5>(pl, v1, v2, buf, e1, e2); return;
    case 32 + 6: fixed_pair_sum2_chains<2, 6>(pl, v1, v2, buf, e1, e2); return;
    default: break;
  }
  e1 = pair_sum(g, pl, v1, nullptr);
  e2 = pair_sum(g, pl, v2, nullptr);
}

// Clears the chain slots of the sum buffer (at kernel start: the chains'
// padding slots are never written afterwards).
__device__ inline void clear_sum_buf(double *buf) {
  for (int i = threadIdx.x; i < kSumBufBytes / 8; i += blockDim.x) buf[i] = 0.0;
}

// pair_sum2 for a variant fixed at compile time (KEY = pair_sum_key, which
// must match pl.fixed; the straight-line sums read the entries back from
// `sumbuf` in LDS), or chosen at run time (KEY < 0): a kernel compiled for
// one variant holds only that sum's code.
template <int KEY, class GEO>
__device__ inline void pair_sum2_t(const GEO &g, const PairLanes &pl, const double (&v1)[6],
                                   const double (&v2)[6], double *sumbuf, double &e1, double &e2) {
  if constexpr (KEY > 0) {
    fixed_pair_sum2_chains<KEY / 16, KEY % 16>(pl, v1, v2, sumbuf, e1, e2);
  } else if constexpr (KEY == 0) {
    e1 = pair_sum_loop(g, pl, v1, nullptr);
    e2 = pair_sum_loop(g, pl, v2, nullptr);
  } else {
    pair_sum2(g, pl, v1, v2, sumbuf, e1, e2);
  }
}

// The kernel of template family `kern` for pair-sum variant `key`
// (pair_sum_key), the run-time-choice instance (-1) for any other key.
#define BWL_KERNEL_FOR_PAIR_SUM_KEY(kern, key)            \
  switch (key) {                                          \
    case 0: return (const void *)kern<0>;                 \
    case 16 + 1: return (const void *)kern<16 + 1>;       \
    case 16 + 2: return (const void *)kern<16 + 2>;       \
    case 16 + 3: return (const void *)kern<16 + 3>;       \
    case 16 + 4: return (const void *)kern<16 + 4>;       \
    case 16 + 5: return (const void *)kern<16 + 5>;       \
    case 16 + 6: return (const void *)kern<16 + 6>;       \
    case 32 + 1: return (const void *)kern<32 + 1>;       \
    case 32 + 2: return (const void *)kern<32 + 2>;       \
    case 32 + 3: return (const void *)kern<32 + 3>;       \
    case 32 + 4: return (const void *)kern<32 + 4>;       \
    case 32 + 5: return (const void *)kern<32 + 5>;       \
    case 32 + 6: return (const void *)kern<32 + 6>;       \
    default: return (const void *)kern<-1>;               \
  }

// Energies of sites A and B in the reference's FP64 order (pair_energy,
// bw_hamiltonian.f90:99-114): every lane gathers neighbour V entries (one
// per 64-entry pass), then the entries are added in list order from 0.0
// per shell, shells left to right per site, and e_A + e_B -- on every lane,
// from the lanes' values.  `off` is the offset table and `vex` V_ex, both
// in LDS.
__device__ inline double pair_energy_lanes(const int8_t *lat, const Geo &g, const int32_t *off,
                                           const double *vex, const PairLanes &pl, int addrA,
                                           int subA, int addrB, int subB, double *eA_out) {
  const int lane = threadIdx.x & 63;
  const int z = g.z, S = g.S;
  const int spA = lat[addrA], spB = lat[addrB];
  double vals[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    vals[k] = 0.0;
    const int t = 64 * k + lane;
    if (k < pl.passes && t < 2 * z) {
      const bool isA = t < z;
      const int tt = isA ? t : t - z;
      const int addr = isA ? addrA : addrB, sub = isA ? subA : subB, sp = isA ? spA : spB;
      const int nb = lat[addr + off[sub * z + tt]];
      vals[k] = vex[(sp - 1) + S * ((nb - 1) + S * pl.shell[k])];
    }
  }
  return pair_sum(g, pl, vals, eA_out);
}

// pair_energy of sites A and B (species spA, spB, doubled coordinates sA,
// sB) before and after the exchange of their species, in the reference's
// order, from one gather of the lattice as it is: before the swap every
// entry takes its neighbour's species; after it, an entry whose neighbour
// is A itself takes spB and one whose neighbour is B takes spA.  A lane of
// A's list sees B when its offset equals B - A mod the box, and A itself
// when its offset is 0 mod the box (possible on boxes shorter than the
// interaction range); likewise for B's list.  Both sums are bit-identical
// to gathers before and after a real swap.  For kernels of one wavefront
// per workgroup whose `sumbuf` (kSumBufBytes of LDS) was cleared with
// clear_sum_buf at kernel start.
template <int KEY = -1, class GEO>
__device__ inline void pair_energy_swap_lanes(const int8_t *lat, const GEO &g, const int32_t *off,
                                              const double *vex, const PairLanes &pl, int addrA,
                                              int subA, SiteXYZ sA, int addrB, int subB, SiteXYZ sB,
                                              int spA, int spB, double *sumbuf, double &e_before,
                                              double &e_after) {
  const int lane = threadIdx.x & 63;
  const int z = g.z, S = g.S;
  const int ab = pack_site_diff(sB, sA, g.m), ba = pack_site_diff(sA, sB, g.m);
  double vb[6], va[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    vb[k] = 0.0;
    va[k] = 0.0;
    const int t = 64 * k + lane;
    if (k < pl.passes && t < 2 * z) {
      const bool isA = t < z;
      const int tt = isA ? t : t - z;
      const int addr = isA ? addrA : addrB, sub = isA ? subA : subB;
      const int nb = lat[addr + off[sub * z + tt]];
      const bool self = pl.dmod[k] == 0, other = pl.dmod[k] == (isA ? ab : ba);
      const bool nA = isA ? self : other, nB = isA ? other : self;
      const int nba = nA ? spB : (nB ? spA : nb);
      const int sb = isA ? spA : spB, sa = isA ? spB : spA;
      vb[k] = vex[(sb - 1) + S * ((nb - 1) + S * pl.shell[k])];
      va[k] = vex[(sa - 1) + S * ((nba - 1) + S * pl.shell[k])];
    }
  }
  pair_sum2_t<KEY>(g, pl, vb, va, sumbuf, e_before, e_after);
}

// The trial loops' geometry as words in LDS (a region of kSeqGeoWords
// words at the end of each sequential kernel's carve, plus room for a few
// kernel-specific loop constants): written by lane 0 at set-up, read back
// into registers after the block barrier.  Values read from an argument
// block may be re-read from it at every use when scalar registers run
// short, each re-read a scalar memory round trip in the middle of a trial;
// values loaded from LDS are kept in registers instead.
constexpr int kSeqGeoWords = 16;   // 15 geometry words, rounded
constexpr int kSeqLoopWords = 16;  // kernel-specific loop constants after them

__host__ __device__ inline size_t seq_geo_region(size_t o) {
  return ((o + 15) & ~(size_t)15) + sizeof(int32_t) * (kSeqGeoWords + kSeqLoopWords);
}

__device__ inline void put_seq_geo(int32_t *gw, const Geo &g) {
  if ((threadIdx.x & 63) == 0) {
    gw[0] = g.lattice;
    for (int a = 0; a < 3; ++a) {
      gw[1 + a] = g.n[a];
      gw[4 + a] = g.m[a];
      gw[7 + a] = g.P[a];
    }
    gw[10] = g.h;
    gw[11] = g.sub_stride;
    gw[12] = g.S;
    gw[13] = g.R;
    gw[14] = g.z;
  }
}

// A 64-bit loop constant in two of the loop words (lane 0 writes).
__device__ inline void put_loop64(int32_t *gw, int k, unsigned long long v) {
  if ((threadIdx.x & 63) == 0) {
    gw[kSeqGeoWords + k] = (int32_t)(uint32_t)v;
    gw[kSeqGeoWords + k + 1] = (int32_t)(uint32_t)(v >> 32);
  }
}

__device__ inline unsigned long long get_loop64(const int32_t *gw, int k) {
  return (unsigned long long)(uint32_t)gw[kSeqGeoWords + k] |
         ((unsigned long long)(uint32_t)gw[kSeqGeoWords + k + 1] << 32);
}

__device__ inline SeqGeo get_seq_geo(const int32_t *gw) {
  SeqGeo sg;
  sg.lattice = gw[0];
  for (int a = 0; a < 3; ++a) {
    sg.n[a] = gw[1 + a];
    sg.m[a] = gw[4 + a];
    sg.P[a] = gw[7 + a];
  }
  sg.h = gw[10];
  sg.sub_stride = gw[11];
  sg.S = gw[12];
  sg.R = gw[13];
  sg.z = gw[14];
  return sg;
}


// A cell's periodic images in the padded box, one per lane: with a halo no
// wider than the box there are at most 3 images per axis, and lane
// 9 iz + 3 iy + ix (< 27) takes image (ix, iy, iz), which lies ox = ix n0,
// oy = iy n1, oz = iz n2 past the first image (p0 = (c + h) mod n) and
// exists when p0 + o stays inside the padded extent on every axis.  The
// lane's address step dq is fixed for a geometry, so a write is a few
// compares and one add per lane.
struct ImageLanes {
  int ox, oy, oz, dq;
  bool fast;  // halo no wider than the box
};

template <class GEO>
__device__ inline ImageLanes image_lanes(const GEO &g) {
  const int lane = threadIdx.x & 63;
  const int ix = lane % 3, iy = (lane / 3) % 3, iz = lane / 9;  // iz >= 3: no image
  ImageLanes il;
  il.fast = g.h <= g.n[0] && g.h <= g.n[1] && g.h <= g.n[2];
  il.ox = ix * g.n[0];
  il.oy = iy * g.n[1];
  il.oz = lane < 27 ? iz * g.n[2] : 1 << 29;  // past every extent
  il.dq = lane < 27 ? (il.oz * g.P[1] + il.oy) * g.P[0] + il.ox : 0;
  return il;
}

// Store v at interior cell (s, cx, cy, cz) and all its images (the same
// cells as write_site), one image per lane.
template <class GEO>
__device__ __forceinline__ void write_site_lanes(int8_t *lat, const GEO &g, const ImageLanes &il, int s,
                                                 int cx, int cy, int cz, int8_t v) {
  if (!il.fast) {
    write_site(lat, g, s, cx, cy, cz, v, threadIdx.x & 63, 64);
    return;
  }
  const int p0x = cx + g.h >= g.n[0] ? cx + g.h - g.n[0] : cx + g.h;
  const int p0y = cy + g.h >= g.n[1] ? cy + g.h - g.n[1] : cy + g.h;
  const int p0z = cz + g.h >= g.n[2] ? cz + g.h - g.n[2] : cz + g.h;
  if (p0x + il.ox < g.P[0] && p0y + il.oy < g.P[1] && p0z + il.oz < g.P[2])
    lat[s * g.sub_stride + (p0z * g.P[1] + p0y) * g.P[0] + p0x + il.dq] = v;
}

// Write species va at site a and vb at site b (every image of each; the
// whole wavefront helps).  For kernels of one wavefront per workgroup: the
// writes are ordered after every earlier LDS access and before every later
// one (wave_lds_order), with no barrier.  With ImageLanes each lane writes
// one fixed image (the walker kernels); without, the lanes split the images
// of write_site (the sequential Metropolis kernel, where the image lanes
// measured slower).
template <class GEO>
__device__ inline void put_pair(int8_t *lat, const GEO &g, const ImageLanes &il, SiteXYZ a, int suba,
                                int8_t va, SiteXYZ b, int subb, int8_t vb) {
  wave_lds_order();
  write_site_lanes(lat, g, il, suba, a.x >> 1, a.y >> 1, a.z >> 1, va);
  write_site_lanes(lat, g, il, subb, b.x >> 1, b.y >> 1, b.z >> 1, vb);
  wave_lds_order();
}

template <class GEO>
__device__ inline void put_pair(int8_t *lat, const GEO &g, SiteXYZ a, int suba, int8_t va, SiteXYZ b,
                                int subb, int8_t vb) {
  const int lane = threadIdx.x & 63;
  wave_lds_order();
  write_site(lat, g, suba, a.x >> 1, a.y >> 1, a.z >> 1, va, lane, 64);
  write_site(lat, g, subb, b.x >> 1, b.y >> 1, b.z >> 1, vb, lane, 64);
  wave_lds_order();
}

// bcc_random_site / fcc_random_site (random_site.f90:127-139, 182-193):
// floor(c*genrand()*n) equals (r*c*n) >> 32 exactly since r*c*n < 2^53.
template <class GEO>
__device__ inline SiteXYZ site_of_draws(const GEO &g, uint32_t r1, uint32_t r2, uint32_t r3) {
  int k = (int)(((uint64_t)r1 * (uint64_t)(2 * g.n[2])) >> 32) + 1;
  SiteXYZ s;
  if (g.lattice == 2) {
    int i = 2 * (int)(((uint64_t)r2 * (uint64_t)g.n[0]) >> 32) + 2 - (k & 1);
    int j = 2 * (int)(((uint64_t)r3 * (uint64_t)g.n[1]) >> 32) + 2 - (k & 1);
    s.x = i - 1; s.y = j - 1; s.z = k - 1;
  } else {
    int i = (int)(((uint64_t)r2 * (uint64_t)(2 * g.n[0])) >> 32) + 1;
    int j = 2 * (int)(((uint64_t)r3 * (uint64_t)g.n[1]) >> 32) + 1 + pmod(i - (k & 1), 2);
    s.x = i - 1; s.y = j - 1; s.z = k - 1;
  }
  return s;
}

template <class GEO>
__device__ inline SiteXYZ random_site_dev(const GEO &g, uint32_t *st, int &mti) {
  uint32_t r1, r2, r3;
  mt_next3(st, mti, r1, r2, r3);
  return site_of_draws(g, r1, r2, r3);
}

template <class GEO>
__device__ inline SiteXYZ random_site_dev(const GEO &g, uint32_t *st, int &mti, MtWin &win) {
  uint32_t r1, r2, r3;
  mt_next3(st, mti, win, r1, r2, r3);
  return site_of_draws(g, r1, r2, r3);
}

// random_site.f90:38-61 nearest-neighbour tables (one draw)
static __constant__ int8_t kRsBcc[8][3] = {{1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
                                    {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
static __constant__ int8_t kRsFcc[12][3] = {{0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1},
                                     {1, 1, 0},  {1, -1, 0},  {1, 0, 1},  {1, 0, -1},
                                     {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}};

template <class GEO>
__device__ inline SiteXYZ nbr_of_draw(const GEO &g, SiteXYZ s, uint32_t r) {
  const int8_t *d;
  if (g.lattice == 2)
    d = kRsBcc[((uint64_t)r * 8u) >> 32];
  else
    d = kRsFcc[((uint64_t)r * 12u) >> 32];
  SiteXYZ t;
  t.x = pmod(s.x + d[0], g.m[0]);
  t.y = pmod(s.y + d[1], g.m[1]);
  t.z = pmod(s.z + d[2], g.m[2]);
  return t;
}

template <class GEO>
__device__ inline SiteXYZ random_nbr_dev(const GEO &g, SiteXYZ s, uint32_t *st, int &mti) {
  return nbr_of_draw(g, s, mt_next(st, mti));
}

template <class GEO>
__device__ inline SiteXYZ random_nbr_dev(const GEO &g, SiteXYZ s, uint32_t *st, int &mti, MtWin &win) {
  return nbr_of_draw(g, s, mt_next(st, mti, win));
}

// Whole-lattice energy by the wavefront: 0.5 * sum of site energies, each
// site's packed neighbour counts contracted with V_ex in FP64.
__device__ inline double wave_total_energy(const int8_t *lat, const Geo &g, const int32_t *off,
                                           const double *vex, double *red) {
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int nsites = g.nsub * cells;
  double part = 0.0;
  for (int i = threadIdx.x; i < nsites; i += blockDim.x) {
    int s = i / cells;
    int r = i - s * cells;
    int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    int cy = r / g.n[0];
    int cx = r - cy * g.n[0];
    int addr = padded_addr(g, s, cx, cy, cz);
    int a = lat[addr];
    const int32_t *o = off + s * g.z;
    for (int n = 0; n < g.R; ++n) {
      uint64_t c = 0;
      for (int t = g.shell_start[n]; t < g.shell_start[n + 1]; ++t)
        c += (uint64_t)1 << (8 * (lat[addr + o[t]] - 1));
      for (int b = 0; b < g.S; ++b) {
        uint32_t k = (uint32_t)(c >> (8 * b)) & 0xffu;
        part += (double)k * vex[(a - 1) + g.S * (b + g.S * n)];
      }
    }
  }
  return 0.5 * block_sum(part, red);
}


// total_energy (bw_hamiltonian.f90:58-81) in the reference's order: sites in
// doubled-grid order (x fastest, then y, then z), empty cells skipped, site
// energies added left to right from 0.0, halved at the end.  The lanes take
// 64 consecutive occupied sites (a doubled-grid row holds every other x:
// fcc in every row, from x = (y+z) mod 2; bcc in the rows with y = z mod 2,
// from x = z mod 2), and lane 0 adds their energies in order.  Every lane
// returns the result.
__device__ inline double ref_total_energy(const int8_t *lat, const Geo &g, const int32_t *off,
                                          const double *vex) {
  const int lane = threadIdx.x & 63;
  const int hx = g.m[0] >> 1;  // occupied sites per occupied row
  const int hy = g.m[1] >> 1;
  const int total = g.n_atoms;
  double e = 0.0;
  for (int base = 0; base < total; base += 64) {
    const int idx = base + lane;
    double v = 0.0;
    const bool occ = idx < total;
    if (occ) {
      const int row = idx / hx;
      const int xh = idx - row * hx;
      int x, y, z;
      if (g.lattice == 1) {
        z = row / g.m[1];
        y = row - z * g.m[1];
        x = 2 * xh + ((y + z) & 1);
      } else {
        z = row / hy;
        y = 2 * (row - z * hy) + (z & 1);
        x = 2 * xh + (z & 1);
      }
      const int addr = padded_addr_xyz(g, x, y, z);
      const int sub = sub_of(g.lattice, x & 1, y & 1);
      const int sp = lat[addr];
      v = shell_sum_lane(lat, g, off, vex, addr, sub, sp, 0);
      for (int n = 1; n < g.R; ++n) v = v + shell_sum_lane(lat, g, off, vex, addr, sub, sp, n);
    }
    const unsigned long long mask = __ballot(occ);
    for (int l = 0; l < 64; ++l)
      if ((mask >> l) & 1ull) e = e + readlane_d(v, l);
  }
  return 0.5 * e;
}

}  // namespace bwl
