// bwl_kernels.hip -- gfx950 kernels of the lattice-sampling path.
//
//   k_pack / k_unpack      reference doubled-grid layout <-> compact
//                          per-sublattice layout (HBM)
//   k_histogram            integer pair-count histogram, one workgroup per
//                          lattice, lattice staged whole in LDS
//   k_site_energy          one site's energy in the reference's FP64 order
//   k_metro_seq            bit-exact sequential Metropolis driven by the
//                          reference's MT19937 stream (one wavefront)
// (the parallel Metropolis sampler is in bwl_metro_kernels.h)
#include <hip/hip_runtime.h>

#include "bwl_kernels.h"
#include "bwl_seq.h"

namespace bwl {


// ------------------------------------------------------------------
// layout conversion
// ------------------------------------------------------------------
__global__ void k_pack(const int8_t *__restrict__ ref, int8_t *__restrict__ compact, Geo g) {
  // blockIdx.y selects one of several lattices laid out back to back
  ref += (size_t)blockIdx.y * g.m[0] * g.m[1] * g.m[2];
  compact += (size_t)blockIdx.y * g.compact_bytes;
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
// Block (x, y) stages lattice y whole in LDS and counts the x-th of
// gridDim.x equal slices of its sites into per-wave private copies in LDS;
// the block's sums are added to the (zeroed) output with 64-bit atomics.
// Integer sums: the result does not depend on the slicing.
// ------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_histogram(const HistArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const HistArgs &A = args[blockIdx.y];
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  const int hsize = g.S * g.S * g.R;
  const int nw = blockDim.x >> 6;
  uint32_t *hw = (uint32_t *)(smem + ((g.lds_bytes + 15) & ~15));
  int32_t *offl = (int32_t *)(hw + hsize * nw);
  for (int i = threadIdx.x; i < hsize * nw; i += blockDim.x) hw[i] = 0;
  stage_offsets(offl, A.off, g);
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  uint32_t *mine = hw + (threadIdx.x >> 6) * hsize;
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int nsites = g.nsub * cells;
  const int per = (nsites + gridDim.x - 1) / gridDim.x;
  const int i0 = blockIdx.x * per;
  const int i1 = i0 + per < nsites ? i0 + per : nsites;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    int s = i / cells;
    int r = i - s * cells;
    int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    int cy = r / g.n[0];
    int cx = r - cy * g.n[0];
    int addr = padded_addr(g, s, cx, cy, cz);
    int a = lat[addr];
    if (a <= 0) continue;
    const int32_t *off = offl + s * g.z;
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
    if (t) atomicAdd((unsigned long long *)&A.hist[i], t);
  }
}

__global__ __launch_bounds__(64) void k_site_energy(SiteArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  double *vex = (double *)(smem + ((g.lds_bytes + 15) & ~15));
  int32_t *offl = (int32_t *)(vex + g.S * g.S * g.R);
  for (int i = threadIdx.x; i < g.S * g.S * g.R; i += blockDim.x) vex[i] = A.vex[i];
  stage_offsets(offl, A.off, g);
  PairLanes pl;
  init_pair_lanes(g, pl);
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  double eA = 0.0;
  int addrA = padded_addr_xyz(g, A.x[0], A.y[0], A.z[0]);
  int addrB = padded_addr_xyz(g, A.x[1], A.y[1], A.z[1]);
  int subA = sub_of(g.lattice, A.x[0] & 1, A.y[0] & 1);
  int subB = sub_of(g.lattice, A.x[1] & 1, A.y[1] & 1);
  double e = pair_energy_lanes(lat, g, offl, vex, pl, addrA, subA, addrB, subB, &eA);
  if (threadIdx.x == 0) {
    A.out[0] = eA;
    A.out[1] = e;
  }
}

// ------------------------------------------------------------------
// Bit-exact sequential Metropolis (metropolis.F90:751-891), one wavefront.
// All lanes run the same control flow; the MT19937 state is in LDS.
// ------------------------------------------------------------------
template <int KEY>
__device__ __forceinline__ void metro_seq_run(const SeqArgs &A, uint8_t *smem) {
  const Geo &g = A.g;
  const int lane = threadIdx.x;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  uint32_t *st = (uint32_t *)(smem + o);
  o += 624 * 4;
  double *sumbuf = (double *)(smem + o);
  clear_sum_buf(sumbuf);
  o += kSumBufBytes;
  double *vex = (double *)(smem + o);
  o += sizeof(double) * g.S * g.S * g.R;
  int32_t *offl = (int32_t *)(smem + o);
  o += sizeof(int32_t) * g.nsub * g.z;
  // the trial loop's geometry and swap mode (see put_seq_geo)
  int32_t *gw = (int32_t *)(smem + ((o + 15) & ~(size_t)15));

  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  for (int i = lane; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  put_seq_geo(gw, g);
  if (lane == 0) {
    gw[kSeqGeoWords] = A.nbr_swap;
    gw[kSeqGeoWords + 1] = A.energies != nullptr && A.sample_every > 0;
  }
  put_loop64(gw, 2, (unsigned long long)A.n_trials);
  put_loop64(gw, 4, (unsigned long long)A.sample_every);
  put_loop64(gw, 6, (unsigned long long)__double_as_longlong(A.beta));
  stage_offsets(offl, A.off, g);
  PairLanes pl;
  init_pair_lanes(g, pl);
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  int mti = A.mt_index[0];
  MtWin win{0u, kMtWinEmpty};
  long long accepted = 0;
  long long sample_k = 0;
  const SeqGeo sg = get_seq_geo(gw);
  const int nbr_swap = gw[kSeqGeoWords];
  const bool sampling = gw[kSeqGeoWords + 1] != 0;
  const long long n_trials = (long long)get_loop64(gw, 2), sample_every = (long long)get_loop64(gw, 4);
  const double beta = __longlong_as_double((long long)get_loop64(gw, 6));

  long long to_sample = sample_every;  // trial % sample_every == 0, counted down
  for (long long trial = 1; trial <= n_trials; ++trial) {
    SiteXYZ s1 = random_site_dev(sg, st, mti, win);
    SiteXYZ s2 = nbr_swap ? random_nbr_dev(sg, s1, st, mti, win) : random_site_dev(sg, st, mti, win);
    int addr1 = padded_addr_xyz(sg, s1.x, s1.y, s1.z);
    int addr2 = padded_addr_xyz(sg, s2.x, s2.y, s2.z);
    int sub1 = sub_of(sg.lattice, s1.x & 1, s1.y & 1);
    int sub2 = sub_of(sg.lattice, s2.x & 1, s2.y & 1);
    int sp1 = lat[addr1], sp2 = lat[addr2];
    bool acc;
    if (sp1 == sp2) {
      acc = true;  // metropolis.F90:774-777: counted as accepted, no draw
    } else {
      // energies before and after the swap from one gather of the unswapped lattice
      double e_un, e_sw;
      pair_energy_swap_lanes<KEY>(lat, sg, offl, vex, pl, addr1, sub1, s1, addr2, sub2, s2, sp1, sp2, sumbuf, e_un,
                                  e_sw);
      double de = e_sw - e_un;
      if (de < 0.0) {
        acc = true;
      } else {
        uint32_t r = mt_next(st, mti, win);
        double u = (double)r * (1.0 / 4294967296.0);
        acc = lt_exp_dev(u, -(beta * de));
      }
      if (acc) put_pair(lat, sg, s1, sub1, (int8_t)sp2, s2, sub2, (int8_t)sp1);
    }
    if (acc) accepted++;
    if (sampling && --to_sample == 0) {
      to_sample = sample_every;
      const double e = ref_total_energy(lat, g, offl, vex);
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
// Proposal draws (rdm_site, rdm_nbr, mc_select) from the reference stream,
// one wavefront (the MT19937 twist is wavefront-wide); lane 0 writes.
// ------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_draw_sites(DrawArgs A) {
  __shared__ uint32_t st[624];
  const int lane = threadIdx.x;
  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  __syncthreads();
  int mti = A.mt_index[0];
  const Geo &g = A.g;
  auto put = [&](int32_t *o, SiteXYZ s) {
    if (lane == 0) {
      o[0] = 1;
      o[1] = s.x + 1;
      o[2] = s.y + 1;
      o[3] = s.z + 1;
    }
  };
  if (A.mode == 1) {
    SiteXYZ s;
    s.x = A.site[1] - 1;
    s.y = A.site[2] - 1;
    s.z = A.site[3] - 1;
    put(A.out, random_nbr_dev(g, s, st, mti));
  } else {
    for (int q = 0; q < A.n; ++q) {
      SiteXYZ s1 = random_site_dev(g, st, mti);
      if (A.mode == 0) {
        put(A.out + 4 * q, s1);
        continue;
      }
      SiteXYZ s2 = A.mode == 3 ? random_nbr_dev(g, s1, st, mti) : random_site_dev(g, st, mti);
      put(A.out + 8 * q, s1);
      put(A.out + 8 * q + 4, s2);
    }
  }
  __syncthreads();
  for (int i = lane; i < 624; i += 64) A.mt[i] = st[i];
  if (lane == 0) A.mt_index[0] = mti;
}

hipError_t launch_draw_sites(const DrawArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_draw_sites, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_zero_i64(long long *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t launch_zero_i64(long long *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
hipError_t launch_pack(const int8_t *ref, int8_t *compact, const Geo &g, hipStream_t s) {
  return launch_pack_many(ref, compact, g, 1, s);
}

hipError_t launch_pack_many(const int8_t *ref, int8_t *compact, const Geo &g, int count,
                            hipStream_t s) {
  int total = g.compact_bytes;
  int blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  for (int first = 0; first < count; first += 65535) {
    const int n = count - first < 65535 ? count - first : 65535;
    hipLaunchKernelGGL(k_pack, dim3(blocks, n), dim3(256), 0, s,
                       ref + (size_t)first * g.m[0] * g.m[1] * g.m[2],
                       compact + (size_t)first * g.compact_bytes, g);
  }
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
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(uint32_t) * g.S * g.S * g.R * (threads / 64) +
         sizeof(int32_t) * g.nsub * g.z;
}

hipError_t launch_histogram(const HistArgs *args_dev, int n, const Geo &g_max, int threads,
                            hipStream_t s) {
  size_t lds = hist_lds_bytes(g_max, threads);
  hipError_t e = hipFuncSetAttribute((const void *)k_histogram,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // slices of about 2048 sites (each block restages the lattice, from L2)
  const int nsites = g_max.nsub * g_max.n[0] * g_max.n[1] * g_max.n[2];
  int parts = nsites / 2048;
  parts = parts < 1 ? 1 : (parts > 256 ? 256 : parts);
  hipLaunchKernelGGL(k_histogram, dim3(parts, n), dim3(threads), lds, s, args_dev);
  return hipGetLastError();
}

size_t site_lds_bytes(const Geo &g) {
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * g.R +
         sizeof(int32_t) * g.nsub * g.z;
}

hipError_t launch_site_energy(const SiteArgs &a, hipStream_t s) {
  size_t lds = site_lds_bytes(a.g);
  hipError_t e = hipFuncSetAttribute((const void *)k_site_energy,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_site_energy, dim3(1), dim3(64), lds, s, a);
  return hipGetLastError();
}

// One chain (the context's), its arguments in the kernel arguments.
template <int KEY>
__global__ __launch_bounds__(64) void k_metro_seq(SeqArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  metro_seq_run<KEY>(A, smem);
}

// Many independent chains, one wavefront each (block b runs args[b]): the
// reference's replica ensemble, every chain on its own MT19937 stream.
template <int KEY>
__global__ __launch_bounds__(64) void k_metro_seq_batch(const SeqArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  metro_seq_run<KEY>(args[blockIdx.x], smem);
}

size_t seq_lds_bytes(const Geo &g) {
  return seq_geo_region((size_t)((g.lds_bytes + 15) & ~15) + 624 * 4 + kSumBufBytes +
                        sizeof(double) * g.S * g.S * g.R + sizeof(int32_t) * g.nsub * g.z) +
         64;
}

namespace {
const void *metro_seq_kernel(int key) { BWL_KERNEL_FOR_PAIR_SUM_KEY(k_metro_seq, key) }
const void *metro_seq_batch_kernel(int key) { BWL_KERNEL_FOR_PAIR_SUM_KEY(k_metro_seq_batch, key) }
}  // namespace

hipError_t launch_metro_seq(const SeqArgs &a, hipStream_t s) {
  size_t lds = seq_lds_bytes(a.g);
  const void *fn = metro_seq_kernel(pair_sum_key(a.g.lattice, a.g.R));
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  SeqArgs arg = a;
  void *kargs[] = {(void *)&arg};
  return hipLaunchKernel(fn, dim3(1), dim3(64), kargs, lds, s);
}

hipError_t launch_metro_seq_batch(const SeqArgs *args_dev, int n, size_t lds, hipStream_t s,
                                  int key) {
  const void *fn = metro_seq_batch_kernel(key);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void *kargs[] = {(void *)&args_dev};
  return hipLaunchKernel(fn, dim3(n), dim3(64), kargs, lds, s);
}

// ------------------------------------------------------------------
// LDS read bandwidth probe (measurement only): each block fills 64 KiB of
// LDS, then every lane streams 16-byte reads, a wavefront reading 1 KiB of
// consecutive addresses per step (no bank conflicts).  The index is made
// opaque to the compiler each step so no read can be merged away; the
// folded values are written only if they hit one improbable pattern.
// ------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_lds_bandwidth(int iters, uint32_t *sink) {
  __shared__ uint4 buf[4096];
  const int t = threadIdx.x;
  for (int i = t; i < 4096; i += blockDim.x) buf[i] = make_uint4(i, 3u * i, 5u * i, 7u * i);
  __syncthreads();
  uint32_t acc = 0;
  int idx = t & 4095;
  for (int k = 0; k < iters; ++k) {
    const uint4 v = buf[idx];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
    idx = (idx + 1024) & 4095;
    asm volatile("" : "+v"(idx));
  }
  if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc;
}

hipError_t launch_lds_bandwidth(int blocks, int iters, uint32_t *sink, hipStream_t s) {
  hipLaunchKernelGGL(k_lds_bandwidth, dim3(blocks), dim3(1024), 0, s, iters, sink);
  return hipGetLastError();
}

}  // namespace bwl
