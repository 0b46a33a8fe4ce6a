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

// Energy of sites A and B (lanes [0,R) and [R,2R)); returns e_A + e_B on
// every lane, each in reference order.
__device__ inline double pair_energy_lanes(const int8_t *lat, const Geo &g, const int32_t *off,
                                           const double *vex, int addrA, int subA,
                                           int addrB, int subB, double *eA_out) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
  if (lane < 2 * g.R) {
    bool isA = lane < g.R;
    int n = isA ? lane : lane - g.R;
    int addr = isA ? addrA : addrB;
    v = shell_sum_lane(lat, g, off, vex, addr, isA ? subA : subB, lat[addr], n);
  }
  double eA = __shfl(v, 0), eB = __shfl(v, g.R);
  for (int n = 1; n < g.R; ++n) {
    eA = eA + __shfl(v, n);
    eB = eB + __shfl(v, g.R + n);
  }
  if (eA_out) *eA_out = eA;
  return eA + eB;
}

struct SiteXYZ {
  int x, y, z;  // 0-based doubled-grid coordinates
};

// bcc_random_site / fcc_random_site (random_site.f90:127-139, 182-193):
// floor(c*genrand()*n) equals (r*c*n) >> 32 exactly since r*c*n < 2^53.
__device__ inline SiteXYZ random_site_dev(const Geo &g, uint32_t *st, int &mti) {
  uint32_t r1 = mt_next(st, mti);
  int k = (int)(((uint64_t)r1 * (uint64_t)(2 * g.n[2])) >> 32) + 1;
  SiteXYZ s;
  if (g.lattice == 2) {
    uint32_t r2 = mt_next(st, mti);
    uint32_t r3 = mt_next(st, mti);
    int i = 2 * (int)(((uint64_t)r2 * (uint64_t)g.n[0]) >> 32) + 2 - (k & 1);
    int j = 2 * (int)(((uint64_t)r3 * (uint64_t)g.n[1]) >> 32) + 2 - (k & 1);
    s.x = i - 1; s.y = j - 1; s.z = k - 1;
  } else {
    uint32_t r2 = mt_next(st, mti);
    uint32_t r3 = mt_next(st, mti);
    int i = (int)(((uint64_t)r2 * (uint64_t)(2 * g.n[0])) >> 32) + 1;
    int j = 2 * (int)(((uint64_t)r3 * (uint64_t)g.n[1]) >> 32) + 1 + pmod(i - (k & 1), 2);
    s.x = i - 1; s.y = j - 1; s.z = k - 1;
  }
  return s;
}

// random_site.f90:38-61 nearest-neighbour tables (one draw)
static __constant__ int8_t kRsBcc[8][3] = {{1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
                                    {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
static __constant__ int8_t kRsFcc[12][3] = {{0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1},
                                     {1, 1, 0},  {1, -1, 0},  {1, 0, 1},  {1, 0, -1},
                                     {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}};

__device__ inline SiteXYZ random_nbr_dev(const Geo &g, SiteXYZ s, uint32_t *st, int &mti) {
  uint32_t r = mt_next(st, mti);
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


// Value of `v` in lane `l` (l uniform across the wavefront).
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// total_energy (bw_hamiltonian.f90:58-81) in the reference's order: sites in
// doubled-grid order (x fastest, then y, then z), empty cells skipped, site
// energies added left to right from 0.0, halved at the end.  Lanes compute
// 64 consecutive site energies, lane 0 adds them in order.  Every lane
// returns the result.
__device__ inline double ref_total_energy(const int8_t *lat, const Geo &g, const int32_t *off,
                                          const double *vex) {
  const int lane = threadIdx.x & 63;
  const int total = g.m[0] * g.m[1] * g.m[2];
  double e = 0.0;
  for (int base = 0; base < total; base += 64) {
    const int idx = base + lane;
    double v = 0.0;
    bool occ = false;
    if (idx < total) {
      const int x = idx % g.m[0];
      const int y = (idx / g.m[0]) % g.m[1];
      const int z = idx / (g.m[0] * g.m[1]);
      occ = g.lattice == 1 ? (((x + y + z) & 1) == 0)
                           : ((x & 1) == (y & 1) && (y & 1) == (z & 1));
      if (occ) {
        const int addr = padded_addr_xyz(g, x, y, z);
        const int sub = sub_of(g.lattice, x & 1, y & 1);
        const int sp = lat[addr];
        v = shell_sum_lane(lat, g, off, vex, addr, sub, sp, 0);
        for (int n = 1; n < g.R; ++n) v = v + shell_sum_lane(lat, g, off, vex, addr, sub, sp, n);
      }
    }
    const unsigned long long mask = __ballot(occ);
    for (int l = 0; l < 64; ++l)
      if ((mask >> l) & 1ull) e = e + readlane_d(v, l);
  }
  return 0.5 * e;
}

}  // namespace bwl
