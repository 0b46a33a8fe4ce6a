// bwl_kernels.hip -- gfx950 kernels of the lattice-sampling path.
//
//   k_pack / k_unpack      reference doubled-grid layout <-> compact
//                          per-sublattice layout (HBM)
//   k_histogram            integer pair-count histogram, one workgroup per
//                          lattice, lattice staged whole in LDS
//   k_site_energy          one site's energy in the reference's FP64 order
//   k_metro_seq            bit-exact sequential Metropolis driven by the
//                          reference's MT19937 stream (one wavefront)
//   k_metro_par            parallel Metropolis: rounds of simultaneous,
//                          non-interacting pair swaps on an LDS-resident
//                          lattice, one workgroup per chain
#include <hip/hip_runtime.h>

#include "bwl_device.h"
#include "bwl_kernels.h"

namespace bwl {

// The reference's Boltzmann constant (constants.f90:41-46, typo kept).
__device__ const double kKbInRy = 8.167333262e-5 / 13.605693122990;

// ------------------------------------------------------------------
// layout conversion
// ------------------------------------------------------------------
__global__ void k_pack(const int8_t *__restrict__ ref, int8_t *__restrict__ compact, Geo g) {
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int total = g.nsub * cells;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += gridDim.x * blockDim.x) {
    int s = idx / cells;
    int r = idx - s * cells;
    int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    int cy = r / g.n[0];
    int cx = r - cy * g.n[0];
    int bx, by, bz;
    sub_bits(g.lattice, s, bx, by, bz);
    int x = 2 * cx + bx, y = 2 * cy + by, z = 2 * cz + bz;
    compact[idx] = ref[x + g.m[0] * (y + g.m[1] * z)];
  }
}

__global__ void k_unpack(const int8_t *__restrict__ compact, int8_t *__restrict__ ref, Geo g) {
  const int total = g.m[0] * g.m[1] * g.m[2];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += gridDim.x * blockDim.x) {
    int x = idx % g.m[0];
    int y = (idx / g.m[0]) % g.m[1];
    int z = idx / (g.m[0] * g.m[1]);
    bool occ = g.lattice == 1 ? (((x + y + z) & 1) == 0)
                              : ((x & 1) == (y & 1) && (y & 1) == (z & 1));
    int8_t v = 0;
    if (occ)
      v = compact[compact_addr(g, sub_of(g.lattice, x & 1, y & 1), x >> 1, y >> 1, z >> 1)];
    ref[idx] = v;
  }
}

// ------------------------------------------------------------------
// pair-count histogram: H[n][b][a] += number of shell-n neighbours of
// species b around sites of species a (Fortran H(a,b,n) layout).
// Per-wave private copies in LDS, summed in a fixed order at the end.
// ------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_histogram(const HistArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const HistArgs &A = args[blockIdx.x];
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  const int hsize = g.S * g.S * g.R;
  const int nw = blockDim.x >> 6;
  uint32_t *hw = (uint32_t *)(smem + ((g.lds_bytes + 15) & ~15));
  for (int i = threadIdx.x; i < hsize * nw; i += blockDim.x) hw[i] = 0;
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  uint32_t *mine = hw + (threadIdx.x >> 6) * hsize;
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int nsites = g.nsub * cells;
  for (int i = threadIdx.x; i < nsites; i += blockDim.x) {
    int s = i / cells;
    int r = i - s * cells;
    int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    int cy = r / g.n[0];
    int cx = r - cy * g.n[0];
    int addr = padded_addr(g, s, cx, cy, cz);
    int a = lat[addr];
    if (a <= 0) continue;
    const int32_t *off = A.off + s * g.z;
    for (int n = 0; n < g.R; ++n) {
      // 8-bit packed counts per species (a shell has at most 32 entries)
      uint64_t c = 0;
      for (int t = g.shell_start[n]; t < g.shell_start[n + 1]; ++t)
        c += (uint64_t)1 << (8 * (lat[addr + off[t]] - 1));
      for (int b = 0; b < g.S; ++b) {
        uint32_t k = (uint32_t)(c >> (8 * b)) & 0xffu;
        if (k) atomicAdd(&mine[(a - 1) + g.S * (b + g.S * n)], k);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < hsize; i += blockDim.x) {
    unsigned long long t = 0;
    for (int w = 0; w < nw; ++w) t += hw[w * hsize + i];
    A.hist[i] = (long long)t;
  }
}

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

__global__ __launch_bounds__(64) void k_site_energy(SiteArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  double *vex = (double *)(smem + ((g.lds_bytes + 15) & ~15));
  for (int i = threadIdx.x; i < g.S * g.S * g.R; i += blockDim.x) vex[i] = A.vex[i];
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  double eA = 0.0;
  int addrA = padded_addr_xyz(g, A.x[0], A.y[0], A.z[0]);
  int addrB = padded_addr_xyz(g, A.x[1], A.y[1], A.z[1]);
  int subA = sub_of(g.lattice, A.x[0] & 1, A.y[0] & 1);
  int subB = sub_of(g.lattice, A.x[1] & 1, A.y[1] & 1);
  double e = pair_energy_lanes(lat, g, A.off, vex, addrA, subA, addrB, subB, &eA);
  if (threadIdx.x == 0) {
    A.out[0] = eA;
    A.out[1] = e;
  }
}

// ------------------------------------------------------------------
// Bit-exact sequential Metropolis (metropolis.F90:751-891), one wavefront.
// All lanes run the same control flow; the MT19937 state is in LDS.
// ------------------------------------------------------------------
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
__constant__ int8_t kRsBcc[8][3] = {{1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
                                    {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
__constant__ int8_t kRsFcc[12][3] = {{0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1},
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

__global__ __launch_bounds__(64) void k_metro_seq(SeqArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const Geo &g = A.g;
  const int lane = threadIdx.x;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  uint32_t *st = (uint32_t *)(smem + o);
  o += 624 * 4;
  double *vex = (double *)(smem + o);
  o += sizeof(double) * g.S * g.S * g.R;
  double *red = (double *)(smem + o);

  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  for (int i = lane; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  int mti = A.mt_index[0];
  const double beta = A.beta;
  long long accepted = 0;
  long long sample_k = 0;

  for (long long trial = 1; trial <= A.n_trials; ++trial) {
    SiteXYZ s1 = random_site_dev(g, st, mti);
    SiteXYZ s2 = A.nbr_swap ? random_nbr_dev(g, s1, st, mti) : random_site_dev(g, st, mti);
    int addr1 = padded_addr_xyz(g, s1.x, s1.y, s1.z);
    int addr2 = padded_addr_xyz(g, s2.x, s2.y, s2.z);
    int sub1 = sub_of(g.lattice, s1.x & 1, s1.y & 1);
    int sub2 = sub_of(g.lattice, s2.x & 1, s2.y & 1);
    int sp1 = lat[addr1], sp2 = lat[addr2];
    bool acc;
    if (sp1 == sp2) {
      acc = true;  // metropolis.F90:774-777: counted as accepted, no draw
    } else {
      double e_un = pair_energy_lanes(lat, g, A.off, vex, addr1, sub1, addr2, sub2, nullptr);
      __syncthreads();
      write_site(lat, g, sub1, s1.x >> 1, s1.y >> 1, s1.z >> 1, (int8_t)sp2, lane, 64);
      write_site(lat, g, sub2, s2.x >> 1, s2.y >> 1, s2.z >> 1, (int8_t)sp1, lane, 64);
      __syncthreads();
      double e_sw = pair_energy_lanes(lat, g, A.off, vex, addr1, sub1, addr2, sub2, nullptr);
      double de = e_sw - e_un;
      if (de < 0.0) {
        acc = true;
      } else {
        uint32_t r = mt_next(st, mti);
        double u = (double)r * (1.0 / 4294967296.0);
        acc = lt_exp(u, -(beta * de));
      }
      if (!acc) {
        __syncthreads();
        write_site(lat, g, sub1, s1.x >> 1, s1.y >> 1, s1.z >> 1, (int8_t)sp1, lane, 64);
        write_site(lat, g, sub2, s2.x >> 1, s2.y >> 1, s2.z >> 1, (int8_t)sp2, lane, 64);
      }
      __syncthreads();
    }
    if (acc) accepted++;
    if (A.energies && A.sample_every > 0 && trial % A.sample_every == 0) {
      double e = wave_total_energy(lat, g, A.off, vex, red);
      if (lane == 0) A.energies[sample_k] = e;
      sample_k++;
    }
  }
  __syncthreads();
  unstage_lattice(lat, A.compact, g, 0);
  for (int i = lane; i < 624; i += 64) A.mt[i] = st[i];
  if (lane == 0) {
    A.mt_index[0] = mti;
    A.accepted[0] = accepted;
  }
}

// ------------------------------------------------------------------
// Parallel Metropolis sampler.
//
// A round picks a kind (RoundVar: axis A and layer offset u) and two random
// origins; its active sites are two interleaved grids (grid 1 on layers
// o_A + k*L, grid 2 on layers o_A + u + k*L; stride-s grids with their own
// origins on the other axes).  Any two active sites are at least rcut+1
// apart along some axis (periodically), so no active site lies in another's
// neighbour shells and the pair trials of a round are independent
// Metropolis moves on disjoint, non-interacting pairs.  Grid-2 site q is
// paired with grid-1 site (mul*q + add) mod n1.  Round kind, origins and
// pairing never depend on the state, each pair swap is a symmetric
// proposal, and with u odd or even the two grids lie on different or equal
// sublattices, so every pair of sites is connected by some sequence of
// rounds (the chain is irreducible over the fixed-composition ensemble).
//
// LDS holds the lattice as bit positions FIELD*(species-1); a site's
// shell-n neighbour species counts are accumulated as packed FIELD-bit
// fields with one shift-add per neighbour.  dE is contracted in FP64:
//   dE = sum_n sum_s (c_a[n][s] - c_b[n][s]) * (V(beta,s,n) - V(alpha,s,n))
// (exact because a and b are not neighbours of each other).
// ------------------------------------------------------------------
struct RoundGeo {
  int ax, bx, cx;        // permutation: round axis, then the two free axes
  int o1[3], o2[3];      // origins of grid 1 / grid 2 (doubled coords)
  int sub1, sub2;
  uint32_t mul, add;
  RoundVar v;
};

__device__ inline int occupied_coord(int lattice, int pa, int pb, uint32_t r, int n) {
  // third coordinate of an occupied site given two others
  return 2 * (int)below(r, (uint32_t)n) + (lattice == 1 ? ((pa + pb) & 1) : (pa & 1));
}

__device__ inline RoundGeo make_round(const ParArgs &C, unsigned long long phase) {
  const Geo &g = C.g;
  U4 rp = philox4x32(U4{(uint32_t)phase, (uint32_t)(phase >> 32), 0xffffffffu, C.chain_id},
                     C.seed_lo, C.seed_hi);
  U4 rq = philox4x32(U4{(uint32_t)phase, (uint32_t)(phase >> 32), 0xfffffffeu, C.chain_id},
                     C.seed_lo, C.seed_hi);
  RoundGeo R;
  R.v = C.var[below(rp.w, (uint32_t)C.n_var)];
  R.ax = R.v.axis;
  R.bx = (R.ax + 1) % 3;
  R.cx = (R.ax + 2) % 3;
  // grid 1 origin: random occupied site
  R.o1[R.ax] = (int)below(rp.x, (uint32_t)g.m[R.ax]);
  if (g.lattice == 1)
    R.o1[R.bx] = (int)below(rp.y, (uint32_t)g.m[R.bx]);
  else
    R.o1[R.bx] = 2 * (int)below(rp.y, (uint32_t)g.n[R.bx]) + (R.o1[R.ax] & 1);
  R.o1[R.cx] = occupied_coord(g.lattice, R.o1[R.ax], R.o1[R.bx], rp.z, g.n[R.cx]);
  // grid 2 origin: shifted by u along the round axis, free elsewhere
  R.o2[R.ax] = R.o1[R.ax] + R.v.u;
  if (g.lattice == 1)
    R.o2[R.bx] = (int)below(rq.x, (uint32_t)g.m[R.bx]);
  else
    R.o2[R.bx] = 2 * (int)below(rq.x, (uint32_t)g.n[R.bx]) + (R.o2[R.ax] & 1);
  R.o2[R.cx] = occupied_coord(g.lattice, R.o2[R.ax], R.o2[R.bx], rq.y, g.n[R.cx]);
  R.o2[R.ax] %= g.m[R.ax];
  R.sub1 = sub_of(g.lattice, R.o1[0] & 1, R.o1[1] & 1);
  R.sub2 = sub_of(g.lattice, R.o2[0] & 1, R.o2[1] & 1);
  R.mul = C.muls[R.v.mul_off + below(rq.z, (uint32_t)R.v.n_mul)];
  R.add = below(rq.w, (uint32_t)R.v.n1);
  return R;
}

// Site of grid-`which` with linear index q: (layer, kb, kc), layer slowest.
__device__ inline void round_site(const ParArgs &C, const RoundGeo &R, int which, uint32_t q,
                                  int &addr, int cell[3]) {
  const Geo &g = C.g;
  const int KB = C.K[R.bx], KC = C.K[R.cx];
  const int kb = q % KB;
  const int kc = (q / KB) % KC;
  const int jl = q / (KB * KC);
  const int *o = which ? R.o2 : R.o1;
  int p[3];
  p[R.ax] = o[R.ax] + jl * C.L;
  p[R.bx] = o[R.bx] + kb * C.stride;
  p[R.cx] = o[R.cx] + kc * C.stride;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (p[a] >= g.m[a]) p[a] -= g.m[a];
    cell[a] = p[a] >> 1;
  }
  addr = padded_addr(g, which ? R.sub2 : R.sub1, cell[0], cell[1], cell[2]);
}

template <typename CW, int FIELD>
__global__ __launch_bounds__(1024) void k_metro_par(const ParArgs *__restrict__ chains) {
  extern __shared__ __align__(16) uint8_t smem[];
  const ParArgs &C = chains[blockIdx.x];
  const Geo &g = C.g;
  const int tid = threadIdx.x;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  double *vt = (double *)(smem + o);  // vt[(n*S + x)*S + s] = V_ex(x+1, s+1, n+1)
  o += sizeof(double) * g.S * g.S * g.R;
  double *red = (double *)(smem + o);
  unsigned long long *redu = (unsigned long long *)(smem + o);

  for (int i = tid; i < g.S * g.S * g.R; i += blockDim.x) {
    int n = i / (g.S * g.S), x = (i / g.S) % g.S, s = i % g.S;
    vt[i] = C.vex[x + g.S * (s + g.S * n)];
  }
  stage_lattice(lat, C.compact, g, FIELD);
  __syncthreads();

  const int S = g.S, R = g.R;
  unsigned long long phase = C.phase0;
  long long acc_total = 0, trials_total = 0;

  for (int t = 0; t < C.T_steps; ++t) {
    const double temp = C.T0 + (double)t * C.dT;
    const double sim_temp = temp * kKbInRy;
    const double beta = 1.0 / sim_temp;
    double step_E = 0.0, step_Esq = 0.0;
    unsigned long long acc_local = 0;
    long long trials_sampled = 0;
    const int nsw = C.burn_sweeps + C.n_sweeps;
    for (int sw = 0; sw < nsw; ++sw) {
      const bool sampled = sw >= C.burn_sweeps;
      long long done = 0;
      while (done < C.min_trials_per_sweep) {
        const RoundGeo RG = make_round(C, phase);
        const int32_t *off1 = C.off + RG.sub1 * g.z;
        const int32_t *off2 = C.off + RG.sub2 * g.z;
        const int npairs = RG.v.n2;
        for (int q = tid; q < npairs; q += blockDim.x) {
          int addrA, addrB, cellA[3], cellB[3];
          round_site(C, RG, 1, (uint32_t)q, addrB, cellB);
          uint32_t ia = (uint32_t)(((uint64_t)RG.mul * (uint32_t)q + RG.add) % (uint32_t)RG.v.n1);
          round_site(C, RG, 0, ia, addrA, cellA);
          const int codeA = lat[addrA], codeB = lat[addrB];
          if (codeA == codeB) {
            if (sampled) acc_local++;
            continue;
          }
          CW ca[kMaxShells], cb[kMaxShells];
#pragma unroll
          for (int n = 0; n < kMaxShells; ++n) {
            CW sa = 0, sb = 0;
            if (n < R) {
              for (int tt = g.shell_start[n]; tt < g.shell_start[n + 1]; ++tt) {
                sa += (CW)1 << lat[addrA + off1[tt]];
                sb += (CW)1 << lat[addrB + off2[tt]];
              }
            }
            ca[n] = sa;
            cb[n] = sb;
          }
          const int alpha = codeA / FIELD, betasp = codeB / FIELD;
          double dE = 0.0;
#pragma unroll
          for (int n = 0; n < kMaxShells; ++n) {
            if (n < R) {
              const double *va = vt + (n * S + alpha) * S;
              const double *vb = vt + (n * S + betasp) * S;
              for (int s = 0; s < S; ++s) {
                int d = (int)((ca[n] >> (FIELD * s)) & ((1 << FIELD) - 1)) -
                        (int)((cb[n] >> (FIELD * s)) & ((1 << FIELD) - 1));
                if (d) dE += (double)d * (vb[s] - va[s]);
              }
            }
          }
          bool acc = dE < 0.0;
          if (!acc) {
            U4 r = philox4x32(U4{(uint32_t)phase, (uint32_t)(phase >> 32), (uint32_t)q, C.chain_id},
                              C.seed_lo, C.seed_hi);
            acc = u01_53(r.x, r.y) < exp(-(beta * dE));
          }
          if (acc) {
            write_site(lat, g, RG.sub1, cellA[0], cellA[1], cellA[2], (int8_t)codeB, 0, 1);
            write_site(lat, g, RG.sub2, cellB[0], cellB[1], cellB[2], (int8_t)codeA, 0, 1);
            if (sampled) acc_local++;
          }
        }
        done += npairs;
        ++phase;
        __syncthreads();
      }
      trials_total += done;
      if (sampled) {
        trials_sampled += done;
        // energy recount after the sweep
        const int cells = g.n[0] * g.n[1] * g.n[2];
        const int nsites = g.nsub * cells;
        double part = 0.0;
        for (int i = tid; i < nsites; i += blockDim.x) {
          int s = i / cells;
          int r = i - s * cells;
          int cz = r / (g.n[0] * g.n[1]);
          r -= cz * g.n[0] * g.n[1];
          int cy = r / g.n[0];
          int cx = r - cy * g.n[0];
          int ad = padded_addr(g, s, cx, cy, cz);
          int a = lat[ad] / FIELD;
          const int32_t *oo = C.off + s * g.z;
#pragma unroll
          for (int n = 0; n < kMaxShells; ++n) {
            if (n < R) {
              CW c = 0;
              for (int tt = g.shell_start[n]; tt < g.shell_start[n + 1]; ++tt)
                c += (CW)1 << lat[ad + oo[tt]];
              const double *va = vt + (n * S + a) * S;
              for (int s2 = 0; s2 < S; ++s2)
                part += (double)((c >> (FIELD * s2)) & ((1 << FIELD) - 1)) * va[s2];
            }
          }
        }
        double E = 0.5 * block_sum(part, red);
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
  }
  __syncthreads();
  unstage_lattice(lat, C.compact, g, FIELD);
  if (tid == 0) {
    C.counters[0] = (long long)phase;
    C.counters[1] += acc_total;
    C.counters[2] += trials_total;
  }
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
hipError_t launch_pack(const int8_t *ref, int8_t *compact, const Geo &g, hipStream_t s) {
  int total = g.compact_bytes;
  int blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(256), 0, s, ref, compact, g);
  return hipGetLastError();
}

hipError_t launch_unpack(const int8_t *compact, int8_t *ref, const Geo &g, hipStream_t s) {
  int total = g.m[0] * g.m[1] * g.m[2];
  int blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_unpack, dim3(blocks), dim3(256), 0, s, compact, ref, g);
  return hipGetLastError();
}

size_t hist_lds_bytes(const Geo &g, int threads) {
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(uint32_t) * g.S * g.S * g.R * (threads / 64);
}

hipError_t launch_histogram(const HistArgs *args_dev, int n, const Geo &g_max, int threads,
                            hipStream_t s) {
  size_t lds = hist_lds_bytes(g_max, threads);
  hipError_t e = hipFuncSetAttribute((const void *)k_histogram,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_histogram, dim3(n), dim3(threads), lds, s, args_dev);
  return hipGetLastError();
}

size_t site_lds_bytes(const Geo &g) {
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * g.R;
}

hipError_t launch_site_energy(const SiteArgs &a, hipStream_t s) {
  size_t lds = site_lds_bytes(a.g);
  hipError_t e = hipFuncSetAttribute((const void *)k_site_energy,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_site_energy, dim3(1), dim3(64), lds, s, a);
  return hipGetLastError();
}

size_t seq_lds_bytes(const Geo &g) {
  return (size_t)((g.lds_bytes + 15) & ~15) + 624 * 4 + sizeof(double) * g.S * g.S * g.R + 64;
}

hipError_t launch_metro_seq(const SeqArgs &a, hipStream_t s) {
  size_t lds = seq_lds_bytes(a.g);
  hipError_t e = hipFuncSetAttribute((const void *)k_metro_seq,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_metro_seq, dim3(1), dim3(64), lds, s, a);
  return hipGetLastError();
}

size_t par_lds_bytes(const Geo &g, int threads) {
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * g.R +
         sizeof(double) * (threads / 64);
}

hipError_t launch_metro_par(const ParArgs *args_dev, int n_chains, size_t lds, int threads,
                            bool wide, hipStream_t s) {
  const void *fn = wide ? (const void *)k_metro_par<uint64_t, 8>
                        : (const void *)k_metro_par<uint32_t, 6>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (wide)
    hipLaunchKernelGGL((k_metro_par<uint64_t, 8>), dim3(n_chains), dim3(threads), lds, s, args_dev);
  else
    hipLaunchKernelGGL((k_metro_par<uint32_t, 6>), dim3(n_chains), dim3(threads), lds, s, args_dev);
  return hipGetLastError();
}

}  // namespace bwl
