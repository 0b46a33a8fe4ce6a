// bwl_metro_kernels.h -- parallel Metropolis sampler on an LDS-resident lattice
// (one workgroup per chain; many chains per launch).  Kernel templates with
// internal linkage: bwl_metro_par.hip instantiates the per-size and
// per-sublattice variants, bwl_metro_fixed.hip the fixed-pitch ones (two
// translation units that compile in parallel).
//
// Rounds.  A round picks a kind (RoundVar: axis A and layer offset u) and
// two random origins; its active sites are two interleaved grids (grid 1 on
// layers o_A + k*L, grid 2 on layers o_A + u + k*L; stride-s grids with
// their own origins on the other axes).  Any two active sites are at least
// rcut+1 apart along some axis (periodically), so no active site lies in
// another's neighbour shells and the pair trials of a round are independent
// Metropolis moves on disjoint, non-interacting pairs.  Grid-2 site
// (layer j, b, c) is paired with grid-1 site (ml*j+al mod nL1, mb*b+ab mod
// KB, mc*c+ac mod KC), a bijection drawn per round.  Round kind, origins and
// pairing never depend on the state, so each pair swap is a symmetric
// proposal; the round kinds connect every pair of sublattices (checked on
// the host), so the chain is irreducible over the fixed-composition
// ensemble.
//
// LDS layout ("natural" coordinates, cubic supercells).  fcc and bcc sites
// are the points of a lattice, so in lattice coordinates every neighbour
// shell is a fixed set of integer vectors:
//   fcc: (a,b,c) = (x, y, (z-x-y)/2 mod n)     box 2n x 2n x n
//   bcc: (a,b,c) = (z, (x-z)/2 mod n, (y-z)/2 mod n)   box 2n x n x n
// For n1 = n2 = n3 these boxes are plainly periodic, so with a halo every
// neighbour of every site is at one fixed linear offset: compile-time
// immediates of the loads for single-shape kernels and for the fixed-pitch
// layout (one pitch for every size up to a bound, see nat_dims_fixed),
// kernel constants held in registers otherwise; the hot loop is one LDS
// byte load and one shift-add per neighbour.  Other shapes (or lattices too
// large for LDS with this halo) use the per-sublattice layout with
// per-sublattice offset tables (k_metro_cmp).
//
// Energy change.  LDS holds bit positions FIELD*(species-1); a site's
// shell-n neighbour counts are packed FIELD-bit fields.  dE is contracted
// in FP64 from the count differences:
//   dE = sum_n sum_s (c_a[n][s] - c_b[n][s]) * (V(beta,s,n) - V(alpha,s,n))
// (exact because a and b are not neighbours).  The energy is recounted from
// the lattice at the end of every temperature step (and before the first)
// and carried through
// the sweeps by summing the accepted dE.
#pragma once

#include <hip/hip_runtime.h>

#include "bwl_device.h"
#include "bwl_kernels.h"
#include "bwl_tables.h"

namespace bwl {

namespace {

__device__ const double kKbInRyPar = 8.167333262e-5 / 13.605693122990;

template <int LAT>
constexpr int shell_size(int n) {
  return LAT == 1 ? kFccShellSize[n] : kBccShellSize[n];
}

template <int LAT>
constexpr int shell_begin(int n) {
  int s = 0;
  for (int k = 0; k < n; ++k) s += shell_size<LAT>(k);
  return s;
}

// x / m and x % m for x < 2^22, m <= 2^12: float reciprocal, one correction.
__device__ __forceinline__ void divmod_small(uint32_t x, uint32_t m, float inv_m, uint32_t &q,
                                             uint32_t &r) {
  uint32_t qq = (uint32_t)((float)x * inv_m);
  int rr = (int)x - (int)(qq * m);
  if (rr < 0) {
    qq -= 1;
    rr += (int)m;
  } else if (rr >= (int)m) {
    qq += 1;
    rr -= (int)m;
  }
  q = qq;
  r = (uint32_t)rr;
}

__device__ __forceinline__ uint32_t mod_small(uint32_t x, uint32_t m, float inv_m) {
  uint32_t q, r;
  divmod_small(x, m, inv_m, q, r);
  return r;
}

// Map (value on round axis, first free axis, second free axis) to (x,y,z);
// the free axes of round axis A are (A+1)%3 and (A+2)%3.
__device__ __forceinline__ void place(int ax, int A, int B, int Cc, int &x, int &y, int &z) {
  x = ax == 0 ? A : (ax == 1 ? Cc : B);
  y = ax == 0 ? B : (ax == 1 ? A : Cc);
  z = ax == 0 ? Cc : (ax == 1 ? B : A);
}

__device__ __forceinline__ int pick3(int ax, const int32_t *v) {
  return ax == 0 ? v[0] : (ax == 1 ? v[1] : v[2]);
}

struct Round {
  int ax;
  int mA, mB, mC;        // extents along round axis / free axes
  int oA1, oB1, oC1;     // grid 1 origin
  int oA2, oB2, oC2;     // grid 2 origin
  int sub1, sub2;
  int KB, KC, nL1, n2;
  float invKB, invKC, invL1, invKBC;
  uint32_t mb, ab, mc, ac, ml, al;
};

__device__ inline Round make_round(const ParArgs &C, unsigned long long phase) {
  const Geo &g = C.g;
  const uint32_t plo = (uint32_t)phase, phi = (uint32_t)(phase >> 32);
  U4 rp = philox4x32(U4{plo, phi, 0xffffffffu, C.chain_id}, C.seed_lo, C.seed_hi);
  U4 rq = philox4x32(U4{plo, phi, 0xfffffffeu, C.chain_id}, C.seed_lo, C.seed_hi);
  U4 rs = philox4x32(U4{plo, phi, 0xfffffffdu, C.chain_id}, C.seed_lo, C.seed_hi);
  Round R;
  const RoundVar &v = C.var[below(rp.w, (uint32_t)C.n_var)];
  R.ax = v.axis;
  const int bx = R.ax == 2 ? 0 : R.ax + 1, cx = R.ax == 0 ? 2 : R.ax - 1;
  R.mA = pick3(R.ax, g.m);
  R.mB = pick3(bx, g.m);
  R.mC = pick3(cx, g.m);
  const int nB = pick3(bx, g.n), nC = pick3(cx, g.n);
  const bool fcc = g.lattice == 1;
  // grid 1 origin: a random occupied site
  R.oA1 = (int)below(rp.x, (uint32_t)R.mA);
  R.oB1 = fcc ? (int)below(rp.y, (uint32_t)R.mB) : 2 * (int)below(rp.y, (uint32_t)nB) + (R.oA1 & 1);
  R.oC1 = 2 * (int)below(rp.z, (uint32_t)nC) + (fcc ? ((R.oA1 + R.oB1) & 1) : (R.oA1 & 1));
  // grid 2 origin: u further along the round axis, free elsewhere
  const int a2 = R.oA1 + v.u;
  R.oB2 = fcc ? (int)below(rq.x, (uint32_t)R.mB) : 2 * (int)below(rq.x, (uint32_t)nB) + (a2 & 1);
  R.oC2 = 2 * (int)below(rq.y, (uint32_t)nC) + (fcc ? ((a2 + R.oB2) & 1) : (a2 & 1));
  R.oA2 = a2 >= R.mA ? a2 - R.mA : a2;
  int x, y, z;
  place(R.ax, R.oA1, R.oB1, R.oC1, x, y, z);
  R.sub1 = sub_of(g.lattice, x & 1, y & 1);
  place(R.ax, R.oA2, R.oB2, R.oC2, x, y, z);
  R.sub2 = sub_of(g.lattice, x & 1, y & 1);
  R.KB = pick3(bx, C.K);
  R.KC = pick3(cx, C.K);
  R.nL1 = v.nL1;
  R.n2 = v.n2;
  R.invKB = 1.0f / (float)R.KB;
  R.invKC = 1.0f / (float)R.KC;
  R.invL1 = 1.0f / (float)R.nL1;
  R.invKBC = 1.0f / (float)(R.KB * R.KC);
  const uint32_t nub = (uint32_t)pick3(bx, C.n_units), nuc = (uint32_t)pick3(cx, C.n_units);
  R.mb = C.units[bx][below(rq.z, nub)];
  R.mc = C.units[cx][below(rq.w, nuc)];
  R.ml = C.muls[v.mul_off + below(rs.x, (uint32_t)v.n_mul)];
  R.ab = below(rs.y, (uint32_t)R.KB);
  R.ac = below(rs.z, (uint32_t)R.KC);
  R.al = below(rs.w, (uint32_t)R.nL1);
  return R;
}

// Doubled coordinates of a grid site (layer jl, free-axis indices kb, kc).
__device__ __forceinline__ void grid_xyz(const ParArgs &C, const Round &R, int oA, int oB, int oC,
                                         int jl, int kb, int kc, int &x, int &y, int &z) {
  int pa = oA + jl * C.L, pb = oB + kb * C.stride, pc = oC + kc * C.stride;
  if (pa >= R.mA) pa -= R.mA;
  if (pb >= R.mB) pb -= R.mB;
  if (pc >= R.mC) pc -= R.mC;
  place(R.ax, pa, pb, pc, x, y, z);
}

// ---- natural layout --------------------------------------------------
template <int LAT>
__device__ __forceinline__ void to_natural(int x, int y, int z, int n, int &a, int &b, int &c) {
  const int m = 2 * n;
  if (LAT == 1) {
    a = x;
    b = y;
    int t = (z - x - y + 2 * m) >> 1;  // in (0, 3n)
    if (t >= n) t -= n;
    if (t >= n) t -= n;
    c = t;
  } else {
    a = z;
    int tb = (x - z + m) >> 1, tc = (y - z + m) >> 1;  // in (0, 2n)
    b = tb >= n ? tb - n : tb;
    c = tc >= n ? tc - n : tc;
  }
}

// Extents of the natural box: compile-time constants for the specialised
// shapes, kernel arguments otherwise.
struct ND {
  int D0, D1, D2, h0, h1, h2, P0, P1, P2;
};

__host__ __device__ constexpr ND nd_of(const NatGeo &N) {
  return ND{N.D[0], N.D[1], N.D[2], N.h[0], N.h[1], N.h[2], N.P[0], N.P[1], N.P[2]};
}

// Compile-time natural layout: NC > 0 is the layout of the n = NC cubic
// supercell; NC = kFixedPitch the fixed-pitch layout shared by all sizes up
// to nat_fixed_nmax (only its row length, plane count and offsets are used).
constexpr int kFixedPitch = -1;

template <int LAT, int R, int NC>
__host__ __device__ constexpr NatGeo nat_geo_c() {
  NatGeo G{};
  if (NC == kFixedPitch)
    nat_dims_fixed(LAT, R, nat_fixed_nmax(LAT, R), G);
  else
    nat_dims(LAT, R, NC, G);
  return G;
}

__device__ __forceinline__ int nat_addr(const ND &N, int a, int b, int c) {
  // 24-bit multiplies (full rate); LDS addresses are far below 2^24
  return (int)__umul24(__umul24((uint32_t)(c + N.h2), (uint32_t)N.P1) + (uint32_t)(b + N.h1),
                       (uint32_t)N.P0) + (a + N.h0);
}

// Store v at the interior cell (LDS address addr, natural coordinates a, b,
// c) and its periodic images: with D >= 2h a cell has at most one image
// position per axis, at a fixed address step, so the (up to 8) stores are
// straight-line code under per-lane predicates.
__device__ __forceinline__ void nat_put(int8_t *lat, const ND &N, int addr, int a, int b, int c,
                                        int8_t v) {
  const int row = N.P0, plane = (int)__umul24((uint32_t)N.P0, (uint32_t)N.P1);
  const int da = a < N.h0 ? N.D0 : (a >= N.D0 - N.h0 ? -N.D0 : 0);
  const int db = b < N.h1 ? N.D1 * row : (b >= N.D1 - N.h1 ? -N.D1 * row : 0);
  const int dc = c < N.h2 ? N.D2 * plane : (c >= N.D2 - N.h2 ? -N.D2 * plane : 0);
  lat[addr] = v;
  if (da) lat[addr + da] = v;
  if (db) {
    lat[addr + db] = v;
    if (da) lat[addr + db + da] = v;
  }
  if (dc) {
    lat[addr + dc] = v;
    if (da) lat[addr + dc + da] = v;
    if (db) {
      lat[addr + dc + db] = v;
      if (da) lat[addr + dc + db + da] = v;
    }
  }
}

template <int LAT>
__device__ void nat_stage(int8_t *lat, const int8_t *__restrict__ compact, const Geo &g,
                          const ND &N, int shift) {
  const int n = g.n[0], m = 2 * n;
  const int total = N.P0 * N.P1 * N.P2;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int pc = idx / (N.P0 * N.P1);
    int r = idx - pc * N.P0 * N.P1;
    int pb = r / N.P0;
    int pa = r - pb * N.P0;
    int a = pmod(pa - N.h0, N.D0), b = pmod(pb - N.h1, N.D1), c = pmod(pc - N.h2, N.D2);
    int x, y, z;
    if (LAT == 1) {
      x = a; y = b; z = (a + b + 2 * c) % m;
    } else {
      z = a; x = (a + 2 * b) % m; y = (a + 2 * c) % m;
    }
    int v = compact[compact_addr(g, sub_of(g.lattice, x & 1, y & 1), x >> 1, y >> 1, z >> 1)];
    lat[idx] = (int8_t)(shift * (v - 1));
  }
}

template <int LAT>
__device__ void nat_unstage(const int8_t *lat, int8_t *__restrict__ compact, const Geo &g,
                            const ND &N, int shift) {
  const int n = g.n[0], m = 2 * n;
  const int total = N.D0 * N.D1 * N.D2;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int c = idx / (N.D0 * N.D1);
    int r = idx - c * N.D0 * N.D1;
    int b = r / N.D0;
    int a = r - b * N.D0;
    int x, y, z;
    if (LAT == 1) {
      x = a; y = b; z = (a + b + 2 * c) % m;
    } else {
      z = a; x = (a + 2 * b) % m; y = (a + 2 * c) % m;
    }
    compact[compact_addr(g, sub_of(g.lattice, x & 1, y & 1), x >> 1, y >> 1, z >> 1)] =
        (int8_t)(lat[nat_addr(N, a, b, c)] / shift + 1);
  }
}

// Packed shell counts with offsets fixed at compile time: every neighbour
// load is one ds_read_u8 with an immediate offset from a per-site base.
template <int LAT, int R, typename CW, int NC>
__device__ __forceinline__ void count_c(const int8_t *lat, int addr, CW (&c)[R]) {
  constexpr NatGeo K = nat_geo_c<LAT, R, NC>();
  constexpr int Z = shell_begin<LAT>(R);
  constexpr int MIN = [] {
    int mn = 0;
    for (int t = 0; t < Z; ++t) mn = K.off[t] < mn ? K.off[t] : mn;
    return mn;
  }();
  // hide base = addr + MIN from the optimiser so it cannot re-split the
  // constant; every offset below is then a non-negative immediate
  int b = addr + MIN;
  asm volatile("" : "+v"(b));
  const int8_t *base = lat + b;
#pragma unroll
  for (int n = 0; n < R; ++n) {
    CW acc = 0;
#pragma unroll
    for (int t = shell_begin<LAT>(n); t < shell_begin<LAT>(n + 1); ++t)
      acc += (CW)1 << base[K.off[t] - MIN];
    c[n] = acc;
  }
}

// Packed shell counts with kernel-constant neighbour offsets.
template <int LAT, int R, typename CW, int ZN>
__device__ __forceinline__ void count_nat(const int8_t *lat, int addr, const int (&off)[ZN],
                                          CW (&c)[R]) {
#pragma unroll
  for (int n = 0; n < R; ++n) {
    CW acc = 0;
#pragma unroll
    for (int t = shell_begin<LAT>(n); t < shell_begin<LAT>(n + 1); ++t)
      acc += (CW)1 << lat[addr + off[t]];
    c[n] = acc;
  }
}

// accept with probability min(1, exp(x)), x = -beta*dE <= 0: a float
// exponential decides unless u falls within 1e-5 (relative) of it, then FP64.
__device__ __forceinline__ bool metropolis_accept(double u, double x) {
  const float ef = __expf((float)x);
  const double e = (double)ef;
  if (u < e * (1.0 - 1e-5)) return true;
  if (u > e * (1.0 + 1e-5) + 1e-300) return false;
  return u < exp(x);
}

// Parameters of one round, computed ahead (state-independent) into LDS.
struct RoundP {
  int32_t ax, mA, mB, mC;
  int32_t oA1, oB1, oC1, oA2, oB2, oC2;
  int32_t KB, KC, nL1, n2;
  int32_t al, ab, ac, pad;
};
constexpr int kRoundBatch = 64;

__device__ inline RoundP round_params(const ParArgs &C, unsigned long long phase) {
  const Round R = make_round(C, phase);
  RoundP P;
  P.ax = R.ax;
  P.mA = R.mA; P.mB = R.mB; P.mC = R.mC;
  P.oA1 = R.oA1; P.oB1 = R.oB1; P.oC1 = R.oC1;
  P.oA2 = R.oA2; P.oB2 = R.oB2; P.oC2 = R.oC2;
  P.KB = R.KB; P.KC = R.KC; P.nL1 = R.nL1; P.n2 = R.n2;
  P.al = (int32_t)R.al; P.ab = (int32_t)R.ab; P.ac = (int32_t)R.ac;
  P.pad = 0;
  return P;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Natural coordinates of the grid sites of a round as affine functions of
// the grid indices (j along the round axis, u and v along its free axes):
//   coordinate = base + j*cj + u*cu + v*cv,  then reduced into [0, D).
// The doubled coordinates are x, y, z = origin + j*L | u*s | v*s (one index
// each, by round axis), and the natural ones linear in those (bcc: a = z,
// b = (x-z)/2, c = (y-z)/2; fcc: a = x, b = y, c = (z-x-y)/2), so the
// coefficients are uniform per round.  a has one non-negative term and needs
// one conditional subtract; the b and c bases carry a bias of 8D so the sums
// stay positive and are reduced with a float reciprocal.
struct NatAffine {
  int a1, b1, c1, a2, b2, c2;  // biased bases of grids 1 and 2
  int aj, au, av, bj, bu, bv, cj, cu, cv;
};

template <int LAT>
__device__ __forceinline__ NatAffine nat_affine(int ax, int L, int s, int n, int oA1, int oB1,
                                                int oC1, int oA2, int oB2, int oC2) {
  // (origin grid 1, origin grid 2, j, u, v) of x, y and z
  int X[5], Y[5], Z[5];
  auto set = [](int *t, int o1, int o2, int j, int u, int v) {
    t[0] = o1; t[1] = o2; t[2] = j; t[3] = u; t[4] = v;
  };
  if (ax == 0) {
    set(X, oA1, oA2, L, 0, 0); set(Y, oB1, oB2, 0, s, 0); set(Z, oC1, oC2, 0, 0, s);
  } else if (ax == 1) {
    set(X, oC1, oC2, 0, 0, s); set(Y, oA1, oA2, L, 0, 0); set(Z, oB1, oB2, 0, s, 0);
  } else {
    set(X, oB1, oB2, 0, s, 0); set(Y, oC1, oC2, 0, 0, s); set(Z, oA1, oA2, L, 0, 0);
  }
  int A[5], B[5], Cc[5];
  for (int i = 0; i < 5; ++i) {
    if (LAT == 1) {
      A[i] = X[i];
      B[i] = Y[i];
      Cc[i] = (Z[i] - X[i] - Y[i]) / 2;  // even differences
    } else {
      A[i] = Z[i];
      B[i] = (X[i] - Z[i]) / 2;
      Cc[i] = (Y[i] - Z[i]) / 2;
    }
  }
  const int D1 = LAT == 1 ? 2 * n : n;
  NatAffine F;
  F.a1 = A[0]; F.a2 = A[1];
  F.b1 = B[0] + 8 * D1; F.b2 = B[1] + 8 * D1;
  F.c1 = Cc[0] + 8 * n; F.c2 = Cc[1] + 8 * n;
  F.aj = A[2]; F.au = A[3]; F.av = A[4];
  F.bj = B[2]; F.bu = B[3]; F.bv = B[4];
  F.cj = Cc[2]; F.cu = Cc[3]; F.cv = Cc[4];
  return F;
}

// x mod d for 0 <= x < 2^20, d <= 2^12: a mask for powers of two, else a
// float reciprocal (the quotient is exact or one low), then one conditional
// subtract.
__device__ __forceinline__ int mod_pos(int x, int d, float inv_d) {
  if ((d & (d - 1)) == 0) {  // folds when d is a constant
    // opaque to the demanded-bits rewrite, which would otherwise turn the
    // 24-bit multiplies feeding x into quarter-rate 32-bit ones
    asm volatile("" : "+v"(x));
    return x & (d - 1);
  }
  const uint32_t q = (uint32_t)((float)x * inv_d);
  const uint32_t r = (uint32_t)x - __umul24(q, (uint32_t)d);
  return (int)min(r, r - (uint32_t)d);
}

__device__ __forceinline__ void nat_site(const NatAffine &F, int ba, int bb, int bc, int j, int u,
                                         int v, const ND &N, float inv1, float inv2, int &a,
                                         int &b, int &c) {
  const uint32_t va = (uint32_t)(ba + __mul24(j, F.aj) + __mul24(u, F.au) + __mul24(v, F.av));
  a = (int)min(va, va - (uint32_t)N.D0);
  b = mod_pos(bb + __mul24(j, F.bj) + __mul24(u, F.bu) + __mul24(v, F.bv), N.D1, inv1);
  c = mod_pos(bc + __mul24(j, F.cj) + __mul24(u, F.cu) + __mul24(v, F.cv), N.D2, inv2);
}

// (2^(F-1) + c_a - c_b) per F-bit field of a packed count word (no field
// borrows: |c_a - c_b| <= shell size < 2^(F-1))
template <int F>
constexpr uint32_t field_bias() {
  uint32_t b = 0;
  for (int s = 0; s + F <= 32; s += F) b |= (uint32_t)1 << (s + F - 1);
  return b;
}
template <int F>
__device__ __forceinline__ uint32_t biased_diff(uint32_t a, uint32_t b) {
  return a + field_bias<F>() - b;
}

// One radial-density sample of the lattice: H[a][b][n] += number of
// shell-n neighbours of species b around sites of species a.  Each lane
// widens its sites' packed FIELD-bit counts to 16-bit fields and sums them
// per (species, shell); the lanes of a wavefront are then summed with
// shuffles and lane 0 adds the totals to H (LDS, 64-bit).  With at most 4
// species in 8-bit fields the per-lane sums stay in registers over all of
// the lane's sites (flushed before a 16-bit field could overflow);
// otherwise every 64 sites.
template <int LAT, int R, typename CW, int NC, int FIELD>
__device__ __forceinline__ void asro_sample(const int8_t *lat, const ND N, const int32_t *offl, int S,
                                         int tid, int nthr, bool active, unsigned long long *hacc) {
  constexpr uint64_t kMask = (1ull << FIELD) - 1;
  constexpr int Z = shell_begin<LAT>(R);
  int off[NC ? 1 : Z];
#pragma unroll
  for (int t = 0; t < (NC ? 1 : Z); ++t) off[t] = NC ? 0 : offl[t];
  constexpr bool kRegs = sizeof(CW) == 4 && FIELD == 8 && (R <= 3 || NC != 0);
  constexpr int XS = kRegs ? 4 : 8;
  constexpr int kMaxField = shell_begin<LAT>(R);  // largest count a site can add
  const int nsites = N.D0 * N.D1 * N.D2;
  const int lane = tid & 63;
  uint64_t acc[kRegs ? 4 : 1][kRegs ? R : 1];
  int pending = 0;
  auto flush = [&]() {
    if constexpr (kRegs) {
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int n = 0; n < R; ++n) {
          uint64_t v = acc[x][n];
#pragma unroll
          for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
          if (lane == 0 && x < S)
            for (int s2 = 0; s2 < S; ++s2) {
              const unsigned long long f = (v >> (16 * s2)) & 0xffffull;
              if (f) atomicAdd(&hacc[(x * S + s2) * R + n], f);
            }
          acc[x][n] = 0;
        }
    }
  };
  if constexpr (kRegs) {
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int n = 0; n < R; ++n) acc[x][n] = 0;
  }
  for (int i0 = active ? (tid & ~63) : nsites; i0 < nsites; i0 += nthr) {
    const int i = i0 + lane;
    const bool valid = i < nsites;
    int sp = -1;
    CW cnt[R];
#pragma unroll
    for (int n = 0; n < R; ++n) cnt[n] = 0;
    if (valid) {
      int c = i / (N.D0 * N.D1);
      int r2 = i - c * N.D0 * N.D1;
      int b = r2 / N.D0;
      int a = r2 - b * N.D0;
      int ad = nat_addr(N, a, b, c);
      sp = lat[ad] / FIELD;
      if constexpr (NC != 0)
        count_c<LAT, R, CW, NC>(lat, ad, cnt);
      else
        count_nat<LAT, R, CW>(lat, ad, off, cnt);
    }
    if constexpr (kRegs) {
#pragma unroll
      for (int n = 0; n < R; ++n) {
        const uint64_t w = (uint64_t)cnt[n];
        const uint64_t e = (w & 0xffull) | ((w & 0xff00ull) << 8) | ((w & 0xff0000ull) << 16) |
                           ((w & 0xff000000ull) << 24);
#pragma unroll
        for (int x = 0; x < 4; ++x) acc[x][n] += sp == x ? e : 0ull;
      }
      // a lane's 16-bit field holds at most 65535; the wavefront sum of
      // fields must too: flush before 64 lanes x pending x kMaxField could
      if (++pending * 64 * kMaxField > 60000) {
        flush();
        pending = 0;
      }
    } else {
      for (int x = 0; x < S && x < XS; ++x) {
#pragma unroll
        for (int n = 0; n < R; ++n) {
          const uint64_t w = sp == x ? (uint64_t)cnt[n] : 0ull;
          uint64_t lo = 0, hi = 0;
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            lo |= ((w >> (FIELD * f)) & kMask) << (16 * f);
            if (FIELD * (f + 5) <= 8 * (int)sizeof(CW))
              hi |= ((w >> (FIELD * (f + 4))) & kMask) << (16 * f);
          }
#pragma unroll
          for (int m = 32; m > 0; m >>= 1) {
            lo += __shfl_xor(lo, m);
            hi += __shfl_xor(hi, m);
          }
          if (lane == 0)
            for (int s2 = 0; s2 < S; ++s2) {
              const uint64_t v = s2 < 4 ? (lo >> (16 * s2)) & 0xffffull : (hi >> (16 * (s2 - 4))) & 0xffffull;
              if (v) atomicAdd(&hacc[(x * S + s2) * R + n], (unsigned long long)v);
            }
        }
      }
    }
  }
  if constexpr (kRegs) {
    if (active) flush();
  }
}

template <int LAT, int R, typename CW, int NC, int FIELD>
__global__ __launch_bounds__(NC != 0 && sizeof(CW) == 4 ? kMaxNatThreads : 512) void k_metro_nat(
    const ParArgs *__restrict__ chains) {
  extern __shared__ __align__(16) uint8_t smem[];
  static_assert(FIELD == 8 || (FIELD == 6 && sizeof(CW) == 4), "count fields: 8 bits, or 6 in 32-bit words");
  const ParArgs &C = chains[blockIdx.x];
  const Geo &g = C.g;
  constexpr NatGeo KG = nat_geo_c<LAT, R, NC ? NC : 8>();
  // extents: all compile-time for one shape (NC > 0); for the fixed pitch
  // the box is the chain's but the pitches and halos are constants
  ND N = NC > 0 ? nd_of(KG) : nd_of(C.nat);
  if constexpr (NC == kFixedPitch) {
    N.P0 = KG.P[0];
    N.P1 = KG.P[1];
    N.h0 = KG.h[0];
    N.h1 = KG.h[1];
    N.h2 = KG.h[2];
  }
  constexpr int Z = shell_begin<LAT>(R);
  const int tid = threadIdx.x;
  // the chain's own thread count; a batched launch may give the block more
  // threads, which then only join barriers (results do not depend on the
  // block size a chain is launched with)
  const int nthr = C.threads;
  const bool active = tid < nthr;
  const int S = g.S;
  // dtab[alpha][beta][n][s<S1P]: (V(beta,s,n)-V(alpha,s,n)) - (same for
  // s=S-1) for s < S-1, zero beyond; the S-1 independent count differences
  // suffice since each shell is full.  Rows are padded to S1P (the most
  // species the count word holds, less one) so the contraction below is
  // unrolled at compile time (the zero terms leave every sum unchanged), and
  // the row stride is odd in doubles so the lanes' row gathers spread over
  // the banks.
  constexpr int S1P = 8 * (int)sizeof(CW) / FIELD - 1;
  constexpr int RSTR = (R * S1P) | 1;
  int8_t *lat = (int8_t *)smem;
  size_t o = ((N.P0 * N.P1 * N.P2) + 15) & ~15;
  double *dtab = (double *)(smem + o);
  o += sizeof(double) * S * S * RSTR;
  double *vrow = (double *)(smem + o);  // [x][n][s]: V(x,s,n)
  o += sizeof(double) * S * R * S;
  RoundP *rps = (RoundP *)(smem + o);
  o += sizeof(RoundP) * kRoundBatch;
  int32_t *offl = (int32_t *)(smem + o);  // neighbour offsets, read into VGPRs
  o += sizeof(int32_t) * ((Z + 3) & ~3);
  double *red = (double *)(smem + o);
  unsigned long long *redu = (unsigned long long *)(smem + o);
  o += sizeof(double) * (blockDim.x / 64 + 1);
  // shell pair counts H[a][b][n] summed over the radial-density samples of
  // the current temperature step
  unsigned long long *hacc = (unsigned long long *)(smem + o);

  for (int i = tid; i < Z; i += blockDim.x) offl[i] = C.nat.off[i];
  for (int i = tid; i < S * S * R; i += blockDim.x) hacc[i] = 0ull;
  for (int i = tid; i < S * S * RSTR; i += blockDim.x) {
    const int row = i / RSTR, k = i - row * RSTR;
    const int s = k % S1P, n = k / S1P, be = row % S, al = row / S;
    const double *V = C.vex;
    double v = 0.0;
    if (n < R && s < S - 1) {
      double d_s = V[be + S * (s + S * n)] - V[al + S * (s + S * n)];
      double d_l = V[be + S * ((S - 1) + S * n)] - V[al + S * ((S - 1) + S * n)];
      v = d_s - d_l;
    }
    dtab[i] = v;
  }
  for (int i = tid; i < S * R * S; i += blockDim.x) {
    int s = i % S, n = (i / S) % R, x = i / (S * R);
    vrow[i] = C.vex[x + S * (s + S * n)];
  }
  nat_stage<LAT>(lat, C.compact, g, N, FIELD);
  __syncthreads();
  int off[NC ? 1 : Z];
#pragma unroll
  for (int t = 0; t < (NC ? 1 : Z); ++t) off[t] = offl[t];

  const int n_cells = g.n[0];
  const int L = C.L, stride = C.stride;
  const int lane_order = NC ? KG.lane_order : C.nat.lane_order;
  // lane -> (layer, slow, fast) decomposition of a round's pair index q in
  // row-major order; the natural layout is cubic, so every free axis has
  // the same K grid positions and the decomposition is the same every round
  const int Kq = C.K[0];
  uint32_t jl0, ks0, kf0, djl, dks, dkf;
  {
    uint32_t rem;
    const float invKK = 1.0f / (float)(Kq * Kq), invK = 1.0f / (float)Kq;
    divmod_small((uint32_t)tid, (uint32_t)(Kq * Kq), invKK, jl0, rem);
    divmod_small(rem, (uint32_t)Kq, invK, ks0, kf0);
    divmod_small((uint32_t)nthr, (uint32_t)(Kq * Kq), invKK, djl, rem);
    divmod_small(rem, (uint32_t)Kq, invK, dks, dkf);
  }
  const float invD1 = 1.0f / (float)N.D1, invD2 = 1.0f / (float)N.D2;
  unsigned long long phase = C.phase0;
  long long acc_total = 0, trials_total = 0;

  // Energy of the lattice recounted from its shell counts (all threads).
  auto recount = [&]() -> double {
    double part = 0.0;
    const int nsites = N.D0 * N.D1 * N.D2;
    for (int i = active ? tid : nsites; i < nsites; i += nthr) {
      int c = i / (N.D0 * N.D1);
      int r = i - c * N.D0 * N.D1;
      int b = r / N.D0;
      int a = r - b * N.D0;
      int ad = nat_addr(N, a, b, c);
      int sp = lat[ad] / FIELD;
      CW cnt[R];
      if constexpr (NC != 0)
        count_c<LAT, R, CW, NC>(lat, ad, cnt);
      else
        count_nat<LAT, R, CW>(lat, ad, off, cnt);
#pragma unroll
      for (int n = 0; n < R; ++n) {
        const double *vr = vrow + (sp * R + n) * S;
        for (int s = 0; s < S; ++s) part += (double)((cnt[n] >> (FIELD * s)) & ((1u << FIELD) - 1)) * vr[s];
      }
    }
    return 0.5 * block_sum(part, red);
  };
  // The energy is recounted before the first temperature step and at the end
  // of every step; between recounts it is carried as the sum of the accepted
  // dE, and the carried value's difference from the recount is reported.
  double E = recount();

  for (int t = 0; t < C.T_steps; ++t) {
    const double temp = C.T0 + (double)t * C.dT;
    const double sim_temp = temp * kKbInRyPar;
    const double beta = 1.0 / sim_temp;
    double step_E = 0.0, step_Esq = 0.0;
    unsigned long long acc_local = 0;
    long long trials_sampled = 0;
    const int nsw = C.burn_sweeps + C.n_sweeps;
    for (int sw = 0; sw < nsw; ++sw) {
      const bool sampled = sw >= C.burn_sweeps;
      double dE_sum = 0.0;
      long long done = 0;
      while (done < C.min_trials_per_sweep) {
        // parameters of the next kRoundBatch rounds, one thread each
        if (tid < kRoundBatch) rps[tid] = round_params(C, phase + (unsigned long long)tid);
        __syncthreads();
        // Each thread draws its acceptance uniforms in order over the whole
        // batch of rounds: its d-th draw is word d % 4 of Philox at counter
        // (batch phase, tid | (d / 4) << 12, chain). One call serves four
        // pairs even when a thread has only one or two pairs per round.
        uint32_t draw = 0;
        U4 rnd = {0, 0, 0, 0};
        int r = 0;
        for (; r < kRoundBatch && done < C.min_trials_per_sweep; ++r) {
          const RoundP &P = rps[r];
          const int ax = uni(P.ax);
          const int nL1 = uni(P.nL1), n2 = uni(P.n2);
          const int al = uni(P.al), ab = uni(P.ab), ac = uni(P.ac);
          const NatAffine F = nat_affine<LAT>(ax, L, stride, n_cells, uni(P.oA1), uni(P.oB1),
                                              uni(P.oC1), uni(P.oA2), uni(P.oB2), uni(P.oC2));
          // the fast free axis of the lanes is chosen per round axis by the
          // layout (bank spread)
          const bool bfast = (lane_order >> ax) & 1;
          uint32_t jl = jl0, ks = ks0, kf = kf0;
          for (int q = active ? tid : n2; q < n2; q += nthr, ++draw) {
            if ((draw & 3) == 0)
              rnd = philox4x32(U4{(uint32_t)phase, (uint32_t)(phase >> 32),
                                  (uint32_t)tid | ((draw >> 2) << 12), C.chain_id},
                               C.seed_lo, C.seed_hi);
            const uint32_t ru = (draw & 3) == 0 ? rnd.x : (draw & 3) == 1 ? rnd.y : (draw & 3) == 2 ? rnd.z : rnd.w;
            // grid-2 site, and its grid-1 partner (translation by (al, ab,
            // ac) in grid indices)
            const int kb = (int)(bfast ? kf : ks), kc = (int)(bfast ? ks : kf);
            int aB, bB, cB, aA, bA, cA;
            nat_site(F, F.a2, F.b2, F.c2, (int)jl, kb, kc, N, invD1, invD2, aB, bB, cB);
            int j1 = (int)jl + al, k1 = kb + ab, c1 = kc + ac;
            if (j1 >= nL1) j1 -= nL1;
            if (k1 >= Kq) k1 -= Kq;
            if (c1 >= Kq) c1 -= Kq;
            nat_site(F, F.a1, F.b1, F.c1, j1, k1, c1, N, invD1, invD2, aA, bA, cA);
            // advance to this thread's next pair
            kf += dkf;
            uint32_t cy = kf >= (uint32_t)Kq;
            kf -= cy * Kq;
            ks += dks + cy;
            cy = ks >= (uint32_t)Kq;
            ks -= cy * Kq;
            jl += djl + cy;

            // Both neighbourhoods are read before the species are compared,
            // so all loads of the pair are in flight together; for equal
            // species every dtab entry is zero, dE is zero and the swap is
            // a counted no-op.
            const int addrA = nat_addr(N, aA, bA, cA), addrB = nat_addr(N, aB, bB, cB);
            const int codeA = lat[addrA], codeB = lat[addrB];
            CW ca[R], cb[R];
            if constexpr (NC != 0) {
              count_c<LAT, R, CW, NC>(lat, addrA, ca);
              count_c<LAT, R, CW, NC>(lat, addrB, cb);
            } else {
              count_nat<LAT, R, CW>(lat, addrA, off, ca);
              count_nat<LAT, R, CW>(lat, addrB, off, cb);
            }
            // keep the contraction below out of the gather: interleaving it
            // makes the scheduler trade the gather's loads in flight for
            // registers (one load at a time for the last shells)
            __builtin_amdgcn_sched_barrier(0);
            // codes are FIELD*(species-1) >= 0: unsigned division is a shift
            // (or a multiply-shift) without a sign fix-up
            const int alpha = (int)((uint32_t)codeA / FIELD), betasp = (int)((uint32_t)codeB / FIELD);
            const double *dt = dtab + __umul24((uint32_t)(alpha * S + betasp), (uint32_t)RSTR);
            // a short table row is read into registers in one burst (the
            // loads then overlap instead of one round trip per term)
            constexpr bool kRowRegs = R * S1P <= 16;
            double dv[kRowRegs ? R * S1P : 1];
            if constexpr (kRowRegs) {
#pragma unroll
              for (int i = 0; i < R * S1P; ++i) dv[i] = dt[i];
              __builtin_amdgcn_sched_barrier(0);
            }
            auto dtv = [&](int i) { return kRowRegs ? dv[kRowRegs ? i : 0] : dt[i]; };
            // fused multiply-adds: each product enters the sum unrounded (one
            // rounding per term instead of two, one instruction instead of two)
            double dE = 0.0;
#pragma unroll
            for (int n = 0; n < R; ++n) {
              constexpr uint32_t kM = (1u << FIELD) - 1;
              constexpr int kB = 1 << (FIELD - 1), kLo = sizeof(CW) == 4 ? S1P : 4;
              const uint32_t d0 = biased_diff<FIELD>((uint32_t)ca[n], (uint32_t)cb[n]);
#pragma unroll
              for (int s = 0; s < kLo; ++s)
                dE = __builtin_fma((double)((int)((d0 >> (FIELD * s)) & kM) - kB), dtv(n * S1P + s), dE);
              if constexpr (sizeof(CW) > 4) {
                const uint32_t d1 = biased_diff<FIELD>((uint32_t)((uint64_t)ca[n] >> 32),
                                                       (uint32_t)((uint64_t)cb[n] >> 32));
#pragma unroll
                for (int s = 4; s < S1P; ++s)
                  dE = __builtin_fma((double)((int)((d1 >> (FIELD * (s - 4))) & kM) - kB), dtv(n * S1P + s), dE);
              }
            }
            const bool same = codeA == codeB;
            bool acc = same || dE < 0.0;
            if (!acc) acc = metropolis_accept((double)ru * (1.0 / 4294967296.0), -(beta * dE));
            if (acc) {
              if (!same) {
                nat_put(lat, N, addrA, aA, bA, cA, (int8_t)codeB);
                nat_put(lat, N, addrB, aB, bB, cB, (int8_t)codeA);
                dE_sum += dE;
              }
              if (sampled) acc_local++;
            }
          }
          done += n2;
          __syncthreads();
        }
        phase += (unsigned long long)r;
      }
      trials_total += done;
      E += block_sum(dE_sum, red);
      if (sampled) {
        trials_sampled += done;
        if (tid == 0) {
          step_E = step_E + E;
          step_Esq = step_Esq + E * E;
          if (C.e_trace) C.e_trace[(long long)t * C.n_sweeps + (sw - C.burn_sweeps)] = E;
        }
        if (C.asro_every > 0 && (sw - C.burn_sweeps + 1) % C.asro_every == 0)
          asro_sample<LAT, R, CW, NC, FIELD>(lat, N, offl, S, tid, nthr, active, hacc);
      }
    }
    unsigned long long acc_sum = block_sum_u64(acc_local, redu);
    if (C.asro_every > 0) {
      // rho(a, b, l) = mean over samples of H(a, b, l) / N_a, shell 0 on-site
      // (the reference's radial_densities layout, wc_range = R + 1)
      const int ns_asro = C.n_sweeps / C.asro_every;
      for (int i = tid; i < S * S * (R + 1); i += blockDim.x) {
        const int a = i % S, b = (i / S) % S, l = i / (S * S);
        unsigned long long tot = 0;
        for (int b2 = 0; b2 < S; ++b2) tot += hacc[(a * S + b2) * R + 0];
        const double na = (double)(tot / (unsigned long long)(shell_size<LAT>(0) * (ns_asro > 0 ? ns_asro : 1)));
        double v;
        if (l == 0)
          v = (a == b ? na : 0.0) / na;
        else
          v = ((double)hacc[(a * S + b) * R + (l - 1)] / na) / (double)ns_asro;
        C.rho_of_T[(size_t)t * S * S * (R + 1) + i] = v;
      }
      __syncthreads();
      for (int i = tid; i < S * S * R; i += blockDim.x) hacc[i] = 0ull;
    }
    if (tid == 0) {
      const double ns = (double)C.n_sweeps;
      const double na = (double)g.n_atoms;
      C.energies_of_T[t] = step_E / ns / na;
      double Cv = (step_Esq / ns - (step_E / ns) * (step_E / ns)) / (sim_temp * temp) / na;
      if (Cv < 0.0) Cv = 0.0;
      if (temp <= 0.0) Cv = 0.0;
      C.C_of_T[t] = Cv;
      C.acc_of_T[t] = (double)acc_sum / (double)trials_sampled;
      acc_total += (long long)acc_sum;
    }
    const double Er = recount();
    if (tid == 0) {
      C.e_drift[t] = E - Er;
      C.e_drift[C.T_steps + t] = Er;
    }
    E = Er;
  }
  __syncthreads();
  nat_unstage<LAT>(lat, C.compact, g, N, FIELD);
  if (tid == 0) {
    C.counters[0] = (long long)phase;
    C.counters[1] += acc_total;
    C.counters[2] += trials_total;
  }
}

// ---- per-sublattice (compact) layout fallback ------------------------------
template <int LAT, int R, typename CW>
__device__ __forceinline__ void count_cmp(const int8_t *lat, int addr, const int32_t *__restrict__ off,
                                          CW (&c)[R]) {
#pragma unroll
  for (int n = 0; n < R; ++n) {
    CW acc = 0;
    for (int t = shell_begin<LAT>(n); t < shell_begin<LAT>(n + 1); ++t)
      acc += (CW)1 << lat[addr + off[t]];
    c[n] = acc;
  }
}

template <int LAT, int R, typename CW, int FIELD>
__global__ __launch_bounds__(512) void k_metro_cmp(const ParArgs *__restrict__ chains) {
  extern __shared__ __align__(16) uint8_t smem[];
  const ParArgs &C = chains[blockIdx.x];
  const Geo &g = C.g;
  constexpr int Z = shell_begin<LAT>(R);
  const int tid = threadIdx.x;
  const int nthr = C.threads;  // the chain's own thread count (see k_metro_nat)
  const bool active = tid < nthr;
  const int S = g.S;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  double *vt = (double *)(smem + o);  // vt[(n*S + x)*S + s] = V_ex(x+1, s+1, n+1)
  o += sizeof(double) * S * S * R;
  double *red = (double *)(smem + o);
  unsigned long long *redu = (unsigned long long *)(smem + o);
  int32_t *offs = (int32_t *)(red + (blockDim.x / 64 + 1));  // [nsub][Z]

  for (int i = tid; i < S * S * R; i += blockDim.x) {
    int n = i / (S * S), x = (i / S) % S, s = i % S;
    vt[i] = C.vex[x + S * (s + S * n)];
  }
  for (int i = tid; i < g.nsub * Z; i += blockDim.x) offs[i] = C.off[i];
  stage_lattice(lat, C.compact, g, FIELD);
  __syncthreads();

  unsigned long long phase = C.phase0;
  long long acc_total = 0, trials_total = 0;
  const int cells = g.n[0] * g.n[1] * g.n[2];

  // Energy of the lattice recounted from its shell counts (all threads).
  auto recount = [&]() -> double {
    double part = 0.0;
    for (int i = active ? tid : g.nsub * cells; i < g.nsub * cells; i += nthr) {
      int s = i / cells;
      int r = i - s * cells;
      int cz = r / (g.n[0] * g.n[1]);
      r -= cz * g.n[0] * g.n[1];
      int cy = r / g.n[0];
      int cx = r - cy * g.n[0];
      int ad = padded_addr(g, s, cx, cy, cz);
      int a = lat[ad] / FIELD;
      CW cnt[R];
      count_cmp<LAT, R, CW>(lat, ad, offs + s * Z, cnt);
#pragma unroll
      for (int n = 0; n < R; ++n) {
        const double *va = vt + (n * S + a) * S;
        for (int b = 0; b < S; ++b)
          part += (double)((cnt[n] >> (FIELD * b)) & ((1 << FIELD) - 1)) * va[b];
      }
    }
    return 0.5 * block_sum(part, red);
  };
  // recounted before the first step and at the end of every step (the
  // carried value's difference from the recount is reported)
  double E = recount();

  for (int t = 0; t < C.T_steps; ++t) {
    const double temp = C.T0 + (double)t * C.dT;
    const double sim_temp = temp * kKbInRyPar;
    const double beta = 1.0 / sim_temp;
    double step_E = 0.0, step_Esq = 0.0;
    unsigned long long acc_local = 0;
    long long trials_sampled = 0;
    const int nsw = C.burn_sweeps + C.n_sweeps;
    for (int sw = 0; sw < nsw; ++sw) {
      const bool sampled = sw >= C.burn_sweeps;
      double dE_sum = 0.0;
      long long done = 0;
      while (done < C.min_trials_per_sweep) {
        const Round RG = make_round(C, phase);
        const int32_t *off1 = offs + RG.sub1 * Z;
        const int32_t *off2 = offs + RG.sub2 * Z;
        uint32_t jl, rem, kb, kc, djl, drem, dkb, dkc;
        divmod_small((uint32_t)tid, (uint32_t)(RG.KB * RG.KC), RG.invKBC, jl, rem);
        divmod_small(rem, (uint32_t)RG.KB, RG.invKB, kc, kb);
        divmod_small((uint32_t)nthr, (uint32_t)(RG.KB * RG.KC), RG.invKBC, djl, drem);
        divmod_small(drem, (uint32_t)RG.KB, RG.invKB, dkc, dkb);
        for (int q = active ? tid : RG.n2; q < RG.n2; q += nthr) {
          int x, y, z, cellA[3], cellB[3];
          grid_xyz(C, RG, RG.oA2, RG.oB2, RG.oC2, (int)jl, (int)kb, (int)kc, x, y, z);
          cellB[0] = x >> 1; cellB[1] = y >> 1; cellB[2] = z >> 1;
          const uint32_t jl1 = mod_small(RG.ml * jl + RG.al, (uint32_t)RG.nL1, RG.invL1);
          const uint32_t kb1 = mod_small(RG.mb * kb + RG.ab, (uint32_t)RG.KB, RG.invKB);
          const uint32_t kc1 = mod_small(RG.mc * kc + RG.ac, (uint32_t)RG.KC, RG.invKC);
          grid_xyz(C, RG, RG.oA1, RG.oB1, RG.oC1, (int)jl1, (int)kb1, (int)kc1, x, y, z);
          cellA[0] = x >> 1; cellA[1] = y >> 1; cellA[2] = z >> 1;
          kb += dkb;
          uint32_t cy = kb >= (uint32_t)RG.KB;
          kb -= cy * RG.KB;
          kc += dkc + cy;
          cy = kc >= (uint32_t)RG.KC;
          kc -= cy * RG.KC;
          jl += djl + cy;
          const int addrA = padded_addr(g, RG.sub1, cellA[0], cellA[1], cellA[2]);
          const int addrB = padded_addr(g, RG.sub2, cellB[0], cellB[1], cellB[2]);
          const int codeA = lat[addrA], codeB = lat[addrB];
          if (codeA == codeB) {
            if (sampled) acc_local++;
            continue;
          }
          CW ca[R], cb[R];
          count_cmp<LAT, R, CW>(lat, addrA, off1, ca);
          count_cmp<LAT, R, CW>(lat, addrB, off2, cb);
          const int alpha = codeA / FIELD, betasp = codeB / FIELD;
          double dE = 0.0;
#pragma unroll
          for (int n = 0; n < R; ++n) {
            const double *va = vt + (n * S + alpha) * S;
            const double *vb = vt + (n * S + betasp) * S;
            for (int s = 0; s < S; ++s) {
              int d = (int)((ca[n] >> (FIELD * s)) & ((1 << FIELD) - 1)) -
                      (int)((cb[n] >> (FIELD * s)) & ((1 << FIELD) - 1));
              if (d) dE += (double)d * (vb[s] - va[s]);
            }
          }
          bool acc = dE < 0.0;
          if (!acc) {
            U4 r = philox4x32(U4{(uint32_t)phase, (uint32_t)(phase >> 32), (uint32_t)q, C.chain_id},
                              C.seed_lo, C.seed_hi);
            acc = metropolis_accept(u01_53(r.x, r.y), -(beta * dE));
          }
          if (acc) {
            write_site(lat, g, RG.sub1, cellA[0], cellA[1], cellA[2], (int8_t)codeB, 0, 1);
            write_site(lat, g, RG.sub2, cellB[0], cellB[1], cellB[2], (int8_t)codeA, 0, 1);
            dE_sum += dE;
            if (sampled) acc_local++;
          }
        }
        done += RG.n2;
        ++phase;
        __syncthreads();
      }
      trials_total += done;
      E += block_sum(dE_sum, red);
      if (sampled) {
        trials_sampled += done;
        if (tid == 0) {
          step_E = step_E + E;
          step_Esq = step_Esq + E * E;
          if (C.e_trace) C.e_trace[(long long)t * C.n_sweeps + (sw - C.burn_sweeps)] = E;
        }
      }
    }
    unsigned long long acc_sum = block_sum_u64(acc_local, redu);
    if (tid == 0) {
      const double ns = (double)C.n_sweeps;
      const double na = (double)g.n_atoms;
      C.energies_of_T[t] = step_E / ns / na;
      double Cv = (step_Esq / ns - (step_E / ns) * (step_E / ns)) / (sim_temp * temp) / na;
      if (Cv < 0.0) Cv = 0.0;
      if (temp <= 0.0) Cv = 0.0;
      C.C_of_T[t] = Cv;
      C.acc_of_T[t] = (double)acc_sum / (double)trials_sampled;
      acc_total += (long long)acc_sum;
    }
    const double Er = recount();
    if (tid == 0) {
      C.e_drift[t] = E - Er;
      C.e_drift[C.T_steps + t] = Er;
    }
    E = Er;
  }
  __syncthreads();
  unstage_lattice(lat, C.compact, g, FIELD);
  if (tid == 0) {
    C.counters[0] = (long long)phase;
    C.counters[1] += acc_total;
    C.counters[2] += trials_total;
  }
}

}  // namespace

}  // namespace bwl
