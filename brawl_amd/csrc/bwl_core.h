// bwl_core.h -- shared host/device definitions for the gfx950 lattice kernels.
//
// Data layout (HBM and LDS)
//   The reference stores occupations on a doubled cubic grid
//   config(1, 2*n1, 2*n2, 2*n3), of which only 1/4 of the cells are sites
//   (shared_data.f90:30; fcc occupied iff x+y+z even, bcc iff x=y=z mod 2,
//   0-based doubled coordinates).  Here every lattice is stored per
//   sublattice: a site at doubled coordinates (x,y,z) has parity bits
//   (x&1, y&1, z&1) -> sublattice s, and cell (x>>1, y>>1, z>>1).
//     fcc: 4 sublattices (000, 011, 101, 110), s = 2*(x&1) + (y&1)
//     bcc: 2 sublattices (000, 111),           s = x&1
//   HBM ("compact"): int8 [s][cz][cy][cx], n1*n2*n3 bytes per sublattice.
//   LDS ("padded"):  int8 [s][cz+h][cy+h][cx+h] with a periodic halo of h
//   cells on every face, so every neighbour of an interior site is at a
//   fixed signed byte offset from it (no wrap arithmetic in the hot loop).
#pragma once

#include <stdint.h>

#include "bwl_tables.h"

namespace bwl {

constexpr int kMaxShells = 10;
constexpr int kMaxSpecies = 8;
constexpr int kMaxNbr = 168;        // bcc, 10 shells
constexpr int kLdsBytes = 160 * 1024;  // gfx950 LDS per workgroup

// Geometry of one lattice, computed on the host, passed by value.
struct Geo {
  int32_t lattice;          // 1 fcc, 2 bcc
  int32_t n[3];             // cells per axis
  int32_t m[3];             // doubled-grid extent 2*n
  int32_t h;                // halo width (cells)
  int32_t P[3];             // padded extents n + 2h
  int32_t nsub;             // 4 (fcc) or 2 (bcc)
  int32_t sub_stride;       // P0*P1*P2
  int32_t lds_bytes;        // nsub*sub_stride rounded up to 16
  int32_t compact_bytes;    // nsub*n0*n1*n2
  int32_t S, R;             // species, shells
  int32_t z;                // neighbours over shells 1..R
  int32_t rcut;             // largest |offset component| over shells 1..R
  int32_t n_atoms;
  int32_t shell_start[kMaxShells + 1];
};

// The straight-line variant of the sequential walkers' ordered pair-energy
// sum for a lattice: lattice*16 + R for fcc and bcc up to 6 shells, 0 for
// the mask-driven loop (bwl_seq.h).
__host__ __device__ constexpr int pair_sum_key(int lattice, int R) {
  return ((lattice == 1 || lattice == 2) && R >= 1 && R <= 6) ? lattice * 16 + R : 0;
}

// Parity bits of each sublattice.
__host__ __device__ inline void sub_bits(int lattice, int s, int &bx, int &by, int &bz) {
  if (lattice == 1) {
    // 000, 011, 101, 110
    bx = s >> 1;
    by = s & 1;
    bz = bx ^ by;
  } else {
    bx = by = bz = s;
  }
}

__host__ __device__ inline int sub_of(int lattice, int bx, int by) {
  return lattice == 1 ? 2 * bx + by : bx;
}

// One kind of round of the parallel sampler.  Active sites form two
// interleaved grids: along axis `axis` the layers sit at
//   o + (j/2)*L + (j odd ? u : 0),  j = 0..nL1+nL2-1   (even j: grid 1,
// odd j: grid 2), and along the other two axes both grids are stride-s
// grids with independent random origins.  Every two active sites are at
// least rcut+1 apart along some axis (periodically), and an odd u moves
// grid 2 to a different sublattice, so swaps cross sublattices.
struct RoundVar {
  int32_t axis, u;
  int32_t nL1, nL2;      // layers of grid 1 / grid 2
  int32_t n1, n2;        // sites of grid 1 / grid 2 (n2 <= n1), pairs = n2
  int32_t mul_off, n_mul;  // multipliers coprime with nL1
};
constexpr int kMaxRoundVars = 6;

// "Natural" LDS layout of a cubic supercell (n1 = n2 = n3 = n): lattice
// coordinates in which every neighbour shell is a fixed set of integer
// vectors, so neighbour offsets do not depend on the site.
//   fcc: (a,b,c) = (x, y, (z-x-y)/2 mod n), box (2n, 2n, n)
//   bcc: (a,b,c) = (z, (x-z)/2 mod n, (y-z)/2 mod n), box (2n, n, n)
// Stored [c+h2][b+h1][a+h0] with periodic halos h.
struct NatGeo {
  int32_t D[3], h[3], P[3];
  int32_t lds_bytes;
  int32_t lane_order;    // bit A set: on round axis A the first free axis
                         // varies fastest across a wavefront's lanes
  int32_t off[kMaxNbr];  // linear neighbour offsets in shell/list order
};

// Lattice-coordinate delta of the reference neighbour offset t.
__host__ __device__ constexpr void nat_delta(int lattice, int t, int &da, int &db, int &dc) {
  if (lattice == 1) {
    const int dx = kFccOffsets[t][0], dy = kFccOffsets[t][1], dz = kFccOffsets[t][2];
    da = dx;
    db = dy;
    dc = (dz - dx - dy) / 2;
  } else {
    const int dx = kBccOffsets[t][0], dy = kBccOffsets[t][1], dz = kBccOffsets[t][2];
    da = dz;
    db = (dx - dz) / 2;
    dc = (dy - dz) / 2;
  }
}

constexpr int kNatBudget = 150 * 1024;  // LDS bytes the natural lattice may take

// Natural layout of a cubic supercell with n cells per axis and R shells,
// padded so that the byte gathers of a round's wavefront (lanes on a 16 x 4
// block of the stride-s grid over the two free axes) fall in distinct LDS
// banks as far as the lattice allows:
//  * bcc with grid stride 4 (rcut 2 or 3) and 16 grid positions per axis
//    (n = 32): P0 = 4 mod 8 and P1 = 8 mod 16, so the 4-steps along x, y
//    and z move (2 mod 4, 16 mod 32, odd) dwords; the lanes then take the
//    second free axis fastest on round axes x and y and the first on z (two
//    of three round axes conflict-free on 32 banks per half-wavefront, all
//    three on 64 banks; other sizes wrap the 16-lane rows and gain nothing);
//  * otherwise P0 = 2 mod 4 and P1 odd (2-cell steps are odd numbers of
//    dwords), lanes with the first free axis fastest.
// Returns false when an axis is shorter than twice its halo.
// Halo of the natural layout (largest |delta| per axis over shells 1..R),
// the neighbour count z and the interaction range rcut (doubled units).
__host__ __device__ constexpr void nat_halo(int lattice, int R, int (&h)[3], int &z, int &rcut) {
  z = 0;
  rcut = 0;
  for (int s = 0; s < R; ++s) z += lattice == 1 ? kFccShellSize[s] : kBccShellSize[s];
  h[0] = h[1] = h[2] = 0;
  for (int t = 0; t < z; ++t) {
    int d[3] = {0, 0, 0};
    nat_delta(lattice, t, d[0], d[1], d[2]);
    for (int a = 0; a < 3; ++a) {
      const int v = d[a] < 0 ? -d[a] : d[a];
      if (v > h[a]) h[a] = v;
      const int w = lattice == 1 ? kFccOffsets[t][a] : kBccOffsets[t][a];
      rcut = w > rcut ? w : (-w > rcut ? -w : rcut);
    }
  }
}

__host__ __device__ constexpr void nat_offsets(int lattice, int z, NatGeo &G) {
  for (int t = 0; t < kMaxNbr; ++t) G.off[t] = 0;
  for (int t = 0; t < z; ++t) {
    int da = 0, db = 0, dc = 0;
    nat_delta(lattice, t, da, db, dc);
    G.off[t] = (dc * G.P[1] + db) * G.P[0] + da;
  }
}

__host__ __device__ constexpr bool nat_dims(int lattice, int R, int n, NatGeo &G) {
  int z = 0, rcut = 0, h[3] = {0, 0, 0};
  nat_halo(lattice, R, h, z, rcut);
  const int need = rcut + 1, stride = need % 2 ? need + 1 : need;
  G.D[0] = 2 * n;
  G.D[1] = lattice == 1 ? 2 * n : n;
  G.D[2] = n;
  for (int a = 0; a < 3; ++a) {
    G.h[a] = h[a];
    if (G.D[a] < 2 * h[a]) return false;
    G.P[a] = G.D[a] + 2 * h[a];
  }
  int p0 = G.P[0], p1 = G.P[1];
  const bool bcc4 = lattice == 2 && stride == 4 && 2 * n == 16 * stride;
  if (bcc4) {
    while (p0 % 8 != 4) ++p0;
    while (p1 % 16 != 8) ++p1;
  } else {
    if (p0 % 4 == 0) p0 += 2;
    if (p1 % 2 == 0) p1 += 1;
  }
  G.lane_order = 7;
  if (p0 * p1 * G.P[2] <= kNatBudget) {
    G.P[0] = p0;
    G.P[1] = p1;
    if (bcc4) G.lane_order = 4;
  }
  G.lds_bytes = G.P[0] * G.P[1] * G.P[2];
  nat_offsets(lattice, z, G);
  return true;
}

// Fixed-pitch natural layout: one row length P0 and plane count P1 for every
// cubic supercell of a lattice and shell count up to nmax cells per axis
// (the planes used are n + 2h), so that the neighbour offsets are the same
// for all those sizes (compile-time constants of one kernel).  The pitches
// come from a model of the LDS banks hit by a round's byte gathers (lanes on
// the stride-s grid of the free axes), averaged over n = 8..28 (fcc) and
// 8..32 (bcc): the rule of nat_dims (P0 = 2 mod 4, P1 odd) gives e.g. fcc
// R3 (66, 65), where the 4-step along x moves a multiple of 32 dwords and
// the gathers conflict 5-fold; (66, 69) conflicts 2-fold on average.
// nmax is the largest n that fits the pitches and the LDS budget.
__host__ __device__ constexpr void nat_fixed_pitch(int lattice, int R, int &p0, int &p1) {
  constexpr int kFcc[4][2] = {{68, 63}, {66, 69}, {66, 69}, {66, 69}};
  constexpr int kBcc[4][2] = {{84, 47}, {86, 43}, {82, 43}, {82, 41}};
  const int r = R < 1 ? 0 : (R > 4 ? 3 : R - 1);
  p0 = lattice == 1 ? kFcc[r][0] : kBcc[r][0];
  p1 = lattice == 1 ? kFcc[r][1] : kBcc[r][1];
}

__host__ __device__ constexpr int nat_fixed_nmax(int lattice, int R) {
  int h[3] = {0, 0, 0}, z = 0, rcut = 0;
  nat_halo(lattice, R, h, z, rcut);
  int p0 = 0, p1 = 0;
  nat_fixed_pitch(lattice, R, p0, p1);
  int best = 0;
  for (int n = 1; n <= 64; ++n) {
    const int d1 = lattice == 1 ? 2 * n : n;
    if (2 * n + 2 * h[0] <= p0 && d1 + 2 * h[1] <= p1 && (n + 2 * h[2]) * p0 * p1 <= kNatBudget)
      best = n;
  }
  return best;
}

__host__ __device__ constexpr bool nat_dims_fixed(int lattice, int R, int n, NatGeo &G) {
  int h[3] = {0, 0, 0}, z = 0, rcut = 0;
  nat_halo(lattice, R, h, z, rcut);
  const int nmax = nat_fixed_nmax(lattice, R);
  G.D[0] = 2 * n;
  G.D[1] = lattice == 1 ? 2 * n : n;
  G.D[2] = n;
  for (int a = 0; a < 3; ++a) {
    G.h[a] = h[a];
    if (G.D[a] < 2 * h[a]) return false;
  }
  if (n > nmax) return false;
  int p0 = 0, p1 = 0;
  nat_fixed_pitch(lattice, R, p0, p1);
  G.P[0] = p0;
  G.P[1] = p1;
  G.P[2] = n + 2 * h[2];
  G.lds_bytes = G.P[0] * G.P[1] * G.P[2];
  G.lane_order = 7;
  nat_offsets(lattice, z, G);
  return true;
}

// The scalars of a geometry that the sequential samplers' trial loops use,
// without the shell table.
struct SeqGeo {
  int32_t lattice;
  int32_t n[3], m[3], P[3];
  int32_t h, sub_stride, S, R, z;
};

// Padded LDS address of the interior cell (s, cx, cy, cz).  GEO is Geo or
// SeqGeo.
template <class GEO>
__host__ __device__ inline int padded_addr(const GEO &g, int s, int cx, int cy, int cz) {
  return s * g.sub_stride + ((cz + g.h) * g.P[1] + (cy + g.h)) * g.P[0] + (cx + g.h);
}

// Padded address of the site at 0-based doubled coordinates (x,y,z).
template <class GEO>
__host__ __device__ inline int padded_addr_xyz(const GEO &g, int x, int y, int z) {
  return padded_addr(g, sub_of(g.lattice, x & 1, y & 1), x >> 1, y >> 1, z >> 1);
}

__host__ __device__ inline int compact_addr(const Geo &g, int s, int cx, int cy, int cz) {
  return ((s * g.n[2] + cz) * g.n[1] + cy) * g.n[0] + cx;
}

// ------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based RNG,
// written from the published round function and constants.
// ------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in [0,1) from two words
__host__ __device__ inline double u01_53(uint32_t a, uint32_t b) {
  uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)v * (1.0 / 9007199254740992.0);
}

// uniform integer in [0, m) from a 32-bit word (m < 2^31)
__host__ __device__ inline uint32_t below(uint32_t r, uint32_t m) {
  return (uint32_t)(((uint64_t)r * m) >> 32);
}

}  // namespace bwl
