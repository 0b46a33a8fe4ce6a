// bwl_walkers.hip -- sequential walkers on the device: nested-sampling runs,
// Wang-Landau walker sweeps and window entry, and the windowed pair-count
// histogram behind the radial densities.
//
// Each nested-sampling run and each Wang-Landau walker is one wavefront
// that consumes its own MT19937 stream exactly as the reference's serial
// code does (one MPI rank per walker there), with site and pair energies in
// the reference's FP64 order, so every walker's trajectory is bit-identical
// to the reference's.  Throughput comes from running many walkers (one per
// CU) in one launch.
#include <hip/hip_runtime.h>

#include "bwl_kernels.h"
#include "bwl_seq.h"

namespace bwl {

namespace {

// bin_index (wang-landau.F90:515-523): truncation toward zero, +1.
__device__ inline int wl_bin_index(double e, const double *edges, int bins) {
  const double range = edges[bins] - edges[0];
  return (int)(((e - edges[0]) / range) * (double)bins) + 1;
}

// bin_index with its constants in registers, for an energy that every lane
// holds alike (the result is wavefront-uniform).  x = ((e - e0) * (1/range)) *
// bins differs from the reference's ((e - e0) / range) * bins by at most a
// few units in the last place (under 1e-15 |x|); when x is further than
// 1e-9 max(1, |x|) from every integer both truncate to the same integer, and
// otherwise the reference's expression is evaluated.  The result is always
// that of wl_bin_index.
struct BinMap {
  double e0, range, rinv, nb;
};

__device__ inline BinMap bin_map(const double *edges, int bins) {
  BinMap b;
  b.e0 = edges[0];
  b.range = edges[bins] - edges[0];
  b.rinv = 1.0 / b.range;
  b.nb = (double)bins;
  return b;
}

__device__ inline int wl_bin_index(double e, const BinMap &b) {
  const double a = e - b.e0;
  const double x = (a * b.rinv) * b.nb;
  const double t = trunc(x);
  const double f = fabs(x - t);
  const double tol = 1e-9 * fmax(1.0, fabs(x));
  if (wave_uniform(f > tol && 1.0 - f > tol)) return wave_uniform((int)t + 1);
  return wave_uniform((int)((a / b.range) * b.nb) + 1);
}

__device__ inline double genrand_dev(uint32_t *st, int &mti) {
  return (double)mt_next(st, mti) * (1.0 / 4294967296.0);
}

__device__ inline double genrand_dev(uint32_t *st, int &mti, MtWin &win) {
  return (double)mt_next(st, mti, win) * (1.0 / 4294967296.0);
}

}  // namespace

// ------------------------------------------------------------------
// Initial walker energies of nested-sampling runs, all walkers of all runs
// at once (block (w, run)): total_energy + genrand()*1e-8
// (nested_sampling.f90:95; 1e-8 is a REAL(4) literal there), the total in
// the reference's summation order.
// ------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_ns_energies(const NsArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const NsArgs &A = args[blockIdx.y];
  const int w = blockIdx.x;
  if (w >= A.n_walkers) return;
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  double *vex = (double *)(smem + o);
  o += sizeof(double) * g.S * g.S * g.R;
  int32_t *offl = (int32_t *)(smem + o);
  for (int i = threadIdx.x; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  stage_offsets(offl, A.off, g);
  stage_lattice(lat, A.walkers + (size_t)g.compact_bytes * w, g, 0);
  __syncthreads();
  const double e = ref_total_energy(lat, g, offl, vex);
  if (threadIdx.x == 0) A.E[w] = e + A.rnde[w] * (double)1e-8f;
}

// ------------------------------------------------------------------
// Nested sampling: the whole serial loop of nested_sampling.f90:78-200 for
// one run.  Walker lattices stay in HBM; the walker being walked is staged
// in LDS.  Walker energies live in LDS.
// ------------------------------------------------------------------
template <int KEY>
__global__ __launch_bounds__(64) void k_ns(const NsArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const NsArgs &A = args[blockIdx.x];
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
  double *E = (double *)(smem + o);
  o += sizeof(double) * A.n_walkers;
  int32_t *offl = (int32_t *)(smem + o);
  o += sizeof(int32_t) * g.nsub * g.z;
  // the trial loop's geometry (see put_seq_geo)
  int32_t *gw = (int32_t *)(smem + ((o + 15) & ~(size_t)15));

  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  for (int i = lane; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  put_seq_geo(gw, g);
  if (lane == 0) gw[kSeqGeoWords] = A.n_steps;
  put_loop64(gw, 2, (unsigned long long)A.max_redraws);
  stage_offsets(offl, A.off, g);
  PairLanes pl;
  init_pair_lanes(g, pl);
  int mti = A.mt_index[0];
  MtWin win{0u, kMtWinEmpty};
  const int nw = A.n_walkers;
  const size_t cb = (size_t)g.compact_bytes;

  // walker energies, computed beforehand by k_ns_energies
  for (int w = lane; w < nw; w += 64) E[w] = A.E[w];
  __syncthreads();

  long long trials = 0, accepted = 0;
  int status = 0;
  int extra = 0, n_acc = 0;
  const int half = (int)((float)nw / 2.0f);
  const SeqGeo sg = get_seq_geo(gw);
  const ImageLanes il = image_lanes(sg);
  const int n_steps = gw[kSeqGeoWords];
  const long long max_redraws = (long long)get_loop64(gw, 2);
  for (int it = 1; it <= A.n_iter && status == 0; ++it) {
    // maxloc: the first maximum (nested_sampling.f90:117)
    double best = -INFINITY;
    int bi = 0x7fffffff;  // no walker seen
    for (int w = lane; w < nw; w += 64)
      if (bi == 0x7fffffff || E[w] > best) {
        best = E[w];
        bi = w;
      }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      const double ob = __shfl_xor(best, m);
      const int oi = __shfl_xor(bi, m);
      if (oi != 0x7fffffff && (bi == 0x7fffffff || ob > best || (ob == best && oi < bi))) {
        best = ob;
        bi = oi;
      }
    }
    const int imax = bi;
    const double limit = best;
    if (lane == 0) A.e_limits[it - 1] = limit;
    // walk-length adaptation (nested_sampling.f90:129-144)
    if (it % half == 0) {
      if (((float)n_acc < (float)A.n_at * 0.05f) && (extra < n_steps * 100)) extra += n_steps;
    }
    // clone a random walker over the discarded one (nested_sampling.f90:149-152)
    const double rnd = genrand_dev(st, mti, win);
    int irnd = (int)ceil(rnd * (double)nw);
    if (irnd < 1) irnd = 1;
    const int src = irnd - 1;
    __syncthreads();
    stage_lattice_walk(lat, A.walkers + cb * src, g);
    __syncthreads();
    double ecur = E[src];
    n_acc = 0;
    // the walk (nested_sampling.f90:157-192)
    for (int stp = 1; stp <= n_steps + extra; ++stp) {
      // rdm2 is redrawn while its species equals rdm1's; the first rdm2 is
      // drawn before either species is read, so both reads go out together
      const SiteXYZ r1 = random_site_dev(sg, st, mti, win);
      SiteXYZ r2 = random_site_dev(sg, st, mti, win);
      const int a1 = padded_addr_xyz(sg, r1.x, r1.y, r1.z);
      int a2 = padded_addr_xyz(sg, r2.x, r2.y, r2.z);
      const int s1 = lat[a1];
      int s2 = lat[a2];
      long long redraws = 0;
      while (s1 == s2 && ++redraws < max_redraws) {
        r2 = random_site_dev(sg, st, mti, win);
        a2 = padded_addr_xyz(sg, r2.x, r2.y, r2.z);
        s2 = lat[a2];
      }
      if (s1 == s2) {
        status = 1;
        break;
      }
      const int sub1 = sub_of(sg.lattice, r1.x & 1, r1.y & 1);
      const int sub2 = sub_of(sg.lattice, r2.x & 1, r2.y & 1);
      // energies before and after the swap from one gather of the unswapped lattice
      double eu, es;
      pair_energy_swap_lanes<KEY>(lat, sg, offl, vex, pl, a1, sub1, r1, a2, sub2, r2, s1, s2, sumbuf, eu, es);
      const double de = es - eu;
      ++trials;
      const bool acc = ecur + de < limit;
      if (acc) {
        ecur = ecur + de;
        ++n_acc;
        ++accepted;
      }
      if (acc) put_pair(lat, sg, il, r1, sub1, (int8_t)s2, r2, sub2, (int8_t)s1);
    }
    __syncthreads();
    if (lane == 0) E[imax] = ecur;
    unstage_lattice_walk(lat, A.walkers + cb * imax, g);
    __syncthreads();
  }
  for (int w = lane; w < nw; w += 64) A.E[w] = E[w];
  for (int i = lane; i < 624; i += 64) A.mt[i] = st[i];
  if (lane == 0) {
    A.mt_index[0] = mti;
    A.counters[0] = trials;
    A.counters[1] = accepted;
    A.counters[2] = status;
  }
}

// ------------------------------------------------------------------
// Wang-Landau `sweeps` (wang-landau.F90:539-632) for one walker, without the
// window average (a separate kernel does that after every walker's call).
// The walker's ln g and visit histogram are staged in LDS.  KEY is the
// pair-sum variant of every walker of the launch (pair_sum_key), or -1 when
// the launch mixes lattices: a kernel for one variant carries only that
// sum's code, which leaves the registers to the rest of the trial.
// ------------------------------------------------------------------
template <int KEY>
__global__ __launch_bounds__(64) void k_wl_sweeps(const WlArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const WlArgs &A = args[blockIdx.x];
  const Geo &g = A.g;
  const int lane = threadIdx.x;
  const int hb = A.end_idx - A.start_idx + 1;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  uint32_t *st = (uint32_t *)(smem + o);
  o += 624 * 4;
  double *sumbuf = (double *)(smem + o);
  clear_sum_buf(sumbuf);
  o += kSumBufBytes;
  double *vex = (double *)(smem + o);
  o += sizeof(double) * g.S * g.S * g.R;
  double *lng = (double *)(smem + o);
  o += sizeof(double) * A.bins;
  double *edges = (double *)(smem + o);
  o += sizeof(double) * (A.bins + 1);
  double *hist = (double *)(smem + o);
  o += sizeof(double) * hb;
  int32_t *offl = (int32_t *)(smem + o);
  o += sizeof(int32_t) * g.nsub * g.z;
  // the trial loop's geometry and window (see put_seq_geo)
  int32_t *gw = (int32_t *)(smem + ((o + 15) & ~(size_t)15));

  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  for (int i = lane; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  put_seq_geo(gw, g);
  if (lane == 0) {
    gw[kSeqGeoWords + 0] = A.bins;
    gw[kSeqGeoWords + 1] = A.start_idx;
    gw[kSeqGeoWords + 2] = A.end_idx;
    gw[kSeqGeoWords + 3] = A.nbr_swap;
  }
  stage_offsets(offl, A.off, g);
  PairLanes pl;
  init_pair_lanes(g, pl);
  for (int i = lane; i < A.bins; i += 64) lng[i] = A.lng[i];
  for (int i = lane; i <= A.bins; i += 64) edges[i] = A.bin_edges[i];
  for (int i = lane; i < hb; i += 64) hist[i] = A.hist[i];
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  int mti = A.mt_index[0];
  MtWin win{0u, kMtWinEmpty};
  const SeqGeo sg = get_seq_geo(gw);
  const ImageLanes il = image_lanes(sg);
  const int bins = gw[kSeqGeoWords + 0], start_idx = gw[kSeqGeoWords + 1];
  const int end_idx = gw[kSeqGeoWords + 2], nbr_swap = gw[kSeqGeoWords + 3];
  const double wl_f = A.wl_f;
  const long long cadence = (long long)(int)(0.02 * (double)(float)g.n_atoms);
  double e_un = ref_total_energy(lat, g, offl, vex);
  const BinMap bmap = bin_map(edges, bins);
  int ibin = wl_bin_index(e_un, edges, bins);  // bin of e_un, kept with it
  // ln g of ibin (0 outside 1..bins), kept with it: lane 0 is the only
  // writer of lng and writes this value back after every trial
  double lcur = (ibin >= 1 && ibin <= bins) ? lng[ibin - 1] : 0.0;
  long long accepted = 0;
  const long long total = (long long)A.mc_sweeps * g.n_atoms;
  // trials until the next histogram update: i % cadence == 0, counted down
  // (a 64-bit remainder per trial costs more than the rest of the decision)
  long long to_hit = cadence;
  for (long long i = 1; i <= total; ++i) {
    const SiteXYZ r1 = random_site_dev(sg, st, mti, win);
    const SiteXYZ r2 =
        nbr_swap ? random_nbr_dev(sg, r1, st, mti, win) : random_site_dev(sg, st, mti, win);
    const int a1 = padded_addr_xyz(sg, r1.x, r1.y, r1.z);
    const int a2 = padded_addr_xyz(sg, r2.x, r2.y, r2.z);
    const int sub1 = sub_of(sg.lattice, r1.x & 1, r1.y & 1);
    const int sub2 = sub_of(sg.lattice, r2.x & 1, r2.y & 1);
    const int s1 = wave_uniform(lat[a1]), s2 = wave_uniform(lat[a2]);
    // The reference also evaluates pair_energy and swaps when the species
    // are equal; neither changes the state, so both are skipped here.
    double e_sw = e_un;
    if (s1 != s2) {
      // energies before and after the swap from one gather of the unswapped lattice
      double pu, ps;
      pair_energy_swap_lanes<KEY>(lat, sg, offl, vex, pl, a1, sub1, r1, a2, sub2, r2, s1, s2, sumbuf, pu, ps);
      e_sw = e_un - pu + ps;
    }
    bool acc = false;
    // e_sw == e_un when the species are equal (no swap): same bin
    int jbin = s1 != s2 ? wl_bin_index(e_sw, bmap) : ibin;
    const bool hit = cadence > 0 && --to_hit == 0;
    if (hit) to_hit = cadence;
    if (jbin > start_idx - 1 && jbin < end_idx + 1) {
      const double lj = lng[jbin - 1];
      const double u = genrand_dev(st, mti, win);
      if (log_lt_dev_uniform(u, lcur - lj)) {
        acc = true;
        ++accepted;
        e_un = e_sw;
        ibin = jbin;
        lcur = lj;
      } else {
        jbin = ibin;
      }
    } else {
      jbin = ibin;
    }
    if (s1 != s2 && acc) put_pair(lat, sg, il, r1, sub1, (int8_t)s2, r2, sub2, (int8_t)s1);
    // a walker outside its window would index outside the reference's
    // arrays; such updates are dropped
    const bool in_range = jbin >= 1 && jbin <= bins;
    if (in_range) lcur = lcur + wl_f;
    if (lane == 0) {
      if (hit && jbin >= start_idx && jbin <= end_idx) hist[jbin - start_idx] += 1.0;
      if (in_range) lng[jbin - 1] = lcur;
    }
    wave_lds_order();  // one wavefront: the next trial's LDS reads follow these writes
  }
  __syncthreads();
  unstage_lattice(lat, A.compact, g, 0);
  for (int i = lane; i < A.bins; i += 64) A.lng[i] = lng[i];
  for (int i = lane; i < hb; i += 64) A.hist[i] = hist[i];
  for (int i = lane; i < 624; i += 64) A.mt[i] = st[i];
  if (lane == 0) {
    A.mt_index[0] = mti;
    A.accepted[0] = accepted;
  }
}

// Window average after `sweeps` (wang-landau.F90:628-631): ln g and the
// visit histogram summed over a window's walkers in walker order, divided
// by the number of walkers, and written back to every walker of the window.
// One thread per (window, bin).
__global__ void k_wl_average(double *lng, double *hist, int n_windows, int wpw, int bins,
                             int hist_stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = bins + hist_stride;
  if (t >= n_windows * per) return;
  const int w = t / per;
  const int b = t - w * per;
  double *base;
  int stride;
  if (b < bins) {
    base = lng + (size_t)w * wpw * bins + b;
    stride = bins;
  } else {
    base = hist + (size_t)w * wpw * hist_stride + (b - bins);
    stride = hist_stride;
  }
  double s = base[0];
  for (int k = 1; k < wpw; ++k) s = s + base[(size_t)k * stride];
  s = s / (double)wpw;
  for (int k = 0; k < wpw; ++k) base[(size_t)k * stride] = s;
}

// ------------------------------------------------------------------
// replica_exchange (wang-landau.F90:1392-1501).
// k_wl_energies: e_unswapped = full_energy(config) of every walker
// (bw_hamiltonian.f90:58-81 order), one wavefront per walker.
// ------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_wl_energies(const WlArgs *__restrict__ args, double *E) {
  extern __shared__ __align__(16) uint8_t smem[];
  const WlArgs &A = args[blockIdx.x];
  const Geo &g = A.g;
  int8_t *lat = (int8_t *)smem;
  size_t o = (g.lds_bytes + 15) & ~15;
  double *vex = (double *)(smem + o);
  o += sizeof(double) * g.S * g.S * g.R;
  int32_t *offl = (int32_t *)(smem + o);
  for (int i = threadIdx.x; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  stage_offsets(offl, A.off, g);
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  const double e = ref_total_energy(lat, g, offl, vex);
  if (threadIdx.x == 0) E[blockIdx.x] = e;
}

// The exchange decisions, one wavefront for the whole walker set (a few
// hundred random numbers at most).  Each walker's overlap location comes
// from the bin of its energy (:1406-1431).  For each pair of neighbouring
// windows (i, i+1), rank 0's stream (walker 0) shuffles the candidate rows
// of both windows (shuffle_rows, :1513-1524), the rows are matched greedily
// in shuffled order (:1447-1468), and each matched pair (a in window i, b in
// window i+1), in order, draws genrand() from a's stream and exchanges iff
// it is below exp(lng_a(bin_a) - lng_a(bin_b)) (:1476-1497).  All lanes
// follow the same control flow (the MT19937 helpers are whole-wavefront);
// lane 0 writes the shared rows and the outputs.
__global__ __launch_bounds__(64) void k_wl_exchange(WlExchangeArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int lane = threadIdx.x;
  const int W = A.W, wpw = A.wpw, nwin = A.num_windows;
  uint32_t *st0 = (uint32_t *)smem;       // walker 0's stream
  uint32_t *st1 = st0 + 624;              // the deciding walker's stream
  int32_t *ibin = (int32_t *)(st1 + 624); // W
  int32_t *loc = ibin + W;                // W
  int32_t *lo = loc + W;                  // rows (walker, location) of window i
  int32_t *up = lo + 2 * wpw;             // rows of window i+1
  int32_t *pr = up + 2 * wpw;             // matched pairs, then their count
  for (int k = lane; k < W; k += 64) {
    const int b = wl_bin_index(A.E[k], A.bin_edges, A.bins);
    const int m = k / wpw + 1;  // the rank's window, 1-based
    bool lower = false, upper = false;
    if (m > 1) lower = b < A.windows[2 * (m - 2) + 1] + 1 && b > A.windows[2 * (m - 1)] - 1;
    if (m < nwin) upper = b > A.windows[2 * m] - 1 && b < A.windows[2 * (m - 1) + 1] + 1;
    ibin[k] = b;
    loc[k] = upper ? m : (lower ? m - 1 : 0);
    A.partner[k] = -1;
    A.accepted[k] = 0;
  }
  for (int i = lane; i < 624; i += 64) st0[i] = A.mt[i];
  int mti0 = A.mt_index[0];
  int nsw = 0;
  __syncthreads();
  for (int i = 1; i < nwin; ++i) {
    for (int j = lane; j < wpw; j += 64) {
      lo[2 * j] = (i - 1) * wpw + j;
      lo[2 * j + 1] = loc[(i - 1) * wpw + j];
      up[2 * j] = i * wpw + j;
      up[2 * j + 1] = loc[i * wpw + j];
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
      int32_t *rows = side ? up : lo;
      for (int t = wpw; t >= 2; --t) {
        const double u = genrand_dev(st0, mti0);
        int r = 1 + (int)(u * (double)t);
        if (r > wpw) r = wpw;
        if (lane == 0) {
          const int32_t a0 = rows[2 * (t - 1)], a1 = rows[2 * (t - 1) + 1];
          rows[2 * (t - 1)] = rows[2 * (r - 1)];
          rows[2 * (t - 1) + 1] = rows[2 * (r - 1) + 1];
          rows[2 * (r - 1)] = a0;
          rows[2 * (r - 1) + 1] = a1;
        }
        __syncthreads();
      }
    }
    if (lane == 0) {
      int np = 0;
      for (int j = 0; j < wpw; ++j) {
        if (lo[2 * j + 1] == 0) continue;
        for (int k = 0; k < wpw; ++k) {
          if (up[2 * k + 1] == 0) continue;
          if (lo[2 * j + 1] == up[2 * k + 1]) {
            pr[2 * np] = lo[2 * j];
            pr[2 * np + 1] = up[2 * k];
            ++np;
            lo[2 * j] = lo[2 * j + 1] = 0;
            up[2 * k] = up[2 * k + 1] = 0;
          }
        }
      }
      pr[2 * wpw] = np;
    }
    __syncthreads();
    const int np = pr[2 * wpw];
    for (int j = 0; j < np; ++j) {
      const int a = pr[2 * j], b = pr[2 * j + 1];
      double u;
      const int mti_a = a == 0 ? 0 : A.mt_index[a];
      if (a == 0) {
        u = genrand_dev(st0, mti0);
      } else if (mti_a < 624) {
        // no twist due: the draw is the next stored word, tempered
        u = (double)mt_temper(A.mt[(size_t)624 * a + mti_a]) * (1.0 / 4294967296.0);
        __syncthreads();  // every lane has read the index before lane 0 moves it
        if (lane == 0) A.mt_index[a] = mti_a + 1;
      } else {
        for (int q = lane; q < 624; q += 64) st1[q] = A.mt[(size_t)624 * a + q];
        int m1 = A.mt_index[a];
        __syncthreads();
        u = genrand_dev(st1, m1);
        __syncthreads();
        for (int q = lane; q < 624; q += 64) A.mt[(size_t)624 * a + q] = st1[q];
        if (lane == 0) A.mt_index[a] = m1;
        __syncthreads();
      }
      // ibin of a matched walker lies inside both windows' bins
      const double x = A.lng[(size_t)A.bins * a + ibin[a] - 1] - A.lng[(size_t)A.bins * a + ibin[b] - 1];
      const bool acc = x >= 0.0 || lt_exp_dev(u, x);  // u < 1 <= exp(x) when x >= 0
      if (lane == 0) {
        A.partner[a] = b;
        A.partner[b] = a;
        A.accepted[a] = A.accepted[b] = acc ? 1 : 0;
        if (acc) {
          A.swaps[2 * nsw] = a;
          A.swaps[2 * nsw + 1] = b;
        }
      }
      nsw += acc ? 1 : 0;
    }
    __syncthreads();
  }
  for (int i = lane; i < 624; i += 64) A.mt[i] = st0[i];
  if (lane == 0) {
    A.mt_index[0] = mti0;
    A.n_swaps[0] = nsw;
  }
}

// Exchange the configurations of the accepted pairs: block j swaps pair j.
__global__ void k_wl_swap_configs(int8_t *walkers, long long bytes, const int32_t *swaps,
                                  const int32_t *n_swaps) {
  const int j = blockIdx.x;
  if (j >= n_swaps[0]) return;
  int8_t *a = walkers + bytes * swaps[2 * j];
  int8_t *b = walkers + bytes * swaps[2 * j + 1];
  for (long long i = threadIdx.x; i < bytes; i += blockDim.x) {
    const int8_t t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}

// ------------------------------------------------------------------
// enter_energy_window (wang-landau.F90:643-741) for one walker.
// ------------------------------------------------------------------
template <int KEY>
__global__ __launch_bounds__(64) void k_wl_enter(const WlEnterArgs *__restrict__ args) {
  extern __shared__ __align__(16) uint8_t smem[];
  const WlEnterArgs &A = args[blockIdx.x];
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
  // the trial loop's geometry (see put_seq_geo)
  int32_t *gw = (int32_t *)(smem + ((o + 15) & ~(size_t)15));

  for (int i = lane; i < 624; i += 64) st[i] = A.mt[i];
  for (int i = lane; i < g.S * g.S * g.R; i += 64) vex[i] = A.vex[i];
  put_seq_geo(gw, g);
  put_loop64(gw, 0, (unsigned long long)__double_as_longlong(A.min_e));
  put_loop64(gw, 2, (unsigned long long)__double_as_longlong(A.max_e));
  put_loop64(gw, 4, (unsigned long long)__double_as_longlong(A.denom));
  put_loop64(gw, 6, (unsigned long long)A.period);
  put_loop64(gw, 8, (unsigned long long)A.budget);
  stage_offsets(offl, A.off, g);
  PairLanes pl;
  init_pair_lanes(g, pl);
  stage_lattice(lat, A.compact, g, 0);
  __syncthreads();
  int mti = A.mt_index[0];
  MtWin win{0u, kMtWinEmpty};
  const double min_e = __longlong_as_double((long long)get_loop64(gw, 0));
  const double max_e = __longlong_as_double((long long)get_loop64(gw, 2));
  const double denom = __longlong_as_double((long long)get_loop64(gw, 4));
  const long long period = (long long)get_loop64(gw, 6), budget = (long long)get_loop64(gw, 8);
  const double target = (min_e + max_e) / 2.0;
  const double cond = fabs(max_e - min_e) * 0.1;
  double e_un = A.init_e ? ref_total_energy(lat, g, offl, vex) : A.e_unswapped[0];
  long long isteps = A.i_steps[0];
  long long trials = A.trials[0];
  long long made = 0;
  bool skip = A.resume[0] != 0;
  int status = 0;
  const SeqGeo sg = get_seq_geo(gw);
  const ImageLanes il = image_lanes(sg);
  // iterations until isteps is next a multiple of period, counted down
  long long to_period = period - isteps % period;
  for (;;) {
    if (!skip) {
      ++isteps;
      if (--to_period == 0) {
        isteps = 0;
        status = 1;  // the host re-runs initial_setup, then resumes here
        break;
      }
    }
    skip = false;
    if (e_un < max_e - cond && e_un > min_e + cond) {
      e_un = ref_total_energy(lat, g, offl, vex);
      if (e_un < max_e - cond && e_un > min_e + cond) {
        status = 0;
        break;
      }
      continue;
    }
    if (made >= budget) {
      status = 2;
      skip = true;  // resume before this iteration's trial
      break;
    }
    ++made;
    ++trials;
    const SiteXYZ r1 = random_site_dev(sg, st, mti, win);
    const SiteXYZ r2 = random_site_dev(sg, st, mti, win);
    const int a1 = padded_addr_xyz(sg, r1.x, r1.y, r1.z);
    const int a2 = padded_addr_xyz(sg, r2.x, r2.y, r2.z);
    const int s1 = lat[a1], s2 = lat[a2];
    if (s1 != s2) {
      const int sub1 = sub_of(sg.lattice, r1.x & 1, r1.y & 1);
      const int sub2 = sub_of(sg.lattice, r2.x & 1, r2.y & 1);
      // energies before and after the swap from one gather of the unswapped lattice
      double pu, ps;
      pair_energy_swap_lanes<KEY>(lat, sg, offl, vex, pl, a1, sub1, r1, a2, sub2, r2, s1, s2, sumbuf, pu, ps);
      const double e_sw = e_un - pu + ps;
      const double de = ((e_sw - target) * (e_sw - target) - (e_un - target) * (e_un - target)) /
                        denom;
      const double u = genrand_dev(st, mti, win);
      const bool acc = log_lt_dev(u, -de);
      if (acc) e_un = e_sw;
      if (acc) put_pair(lat, sg, il, r1, sub1, (int8_t)s2, r2, sub2, (int8_t)s1);
    }
  }
  __syncthreads();
  unstage_lattice(lat, A.compact, g, 0);
  for (int i = lane; i < 624; i += 64) A.mt[i] = st[i];
  if (lane == 0) {
    A.mt_index[0] = mti;
    A.e_unswapped[0] = e_un;
    A.i_steps[0] = isteps;
    A.trials[0] = trials;
    A.resume[0] = (status == 1 || status == 2) ? 1 : 0;
    A.status[0] = status;
  }
}

// ------------------------------------------------------------------
// Windowed pair counts of radial_densities: for every occupied site of
// species a and every window offset whose minimum-image distance falls on
// shell l, the species b at that offset: H(a, b, l) += 1.  The lattice is
// read from HBM (it is small and stays in L2); counts are packed 8 bits per
// species per shell and flushed to a per-block LDS histogram, then added to
// the global one with 64-bit atomics (integer sums, so the result does not
// depend on the order).
// ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_asro(AsroArgs A) {
  __shared__ unsigned int lh[kMaxSpecies * kMaxSpecies * kMaxAsroShells];
  const Geo &g = A.g;
  const int S = g.S;
  const int hsize = S * S * A.wc;
  for (int i = threadIdx.x; i < hsize; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int nsites = g.nsub * cells;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nsites;
       idx += gridDim.x * blockDim.x) {
    const int s = idx / cells;
    int r = idx - s * cells;
    const int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    const int cy = r / g.n[0];
    const int cx = r - cy * g.n[0];
    int bx, by, bz;
    sub_bits(g.lattice, s, bx, by, bz);
    const int x = 2 * cx + bx, y = 2 * cy + by, z = 2 * cz + bz;
    const int a = A.compact[idx];
    if (a <= 0) continue;
    for (int l = 0; l < A.wc; ++l) {
      uint64_t c = 0;
      int k = 0;
      for (int t = A.shell_start[l]; t < A.shell_start[l + 1]; ++t) {
        int xx = x + A.off[3 * t], yy = y + A.off[3 * t + 1], zz = z + A.off[3 * t + 2];
        xx = xx < 0 ? xx + g.m[0] : (xx >= g.m[0] ? xx - g.m[0] : xx);
        yy = yy < 0 ? yy + g.m[1] : (yy >= g.m[1] ? yy - g.m[1] : yy);
        zz = zz < 0 ? zz + g.m[2] : (zz >= g.m[2] ? zz - g.m[2] : zz);
        const int b = A.compact[compact_addr(g, sub_of(g.lattice, xx & 1, yy & 1), xx >> 1, yy >> 1,
                                             zz >> 1)];
        c += (uint64_t)1 << (8 * (b - 1));
        if (++k == 255 || t + 1 == A.shell_start[l + 1]) {
          for (int bb = 0; bb < S; ++bb) {
            const unsigned q = (unsigned)(c >> (8 * bb)) & 0xffu;
            if (q) atomicAdd(&lh[(a - 1) + S * (bb + S * l)], q);
          }
          c = 0;
          k = 0;
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < hsize; i += blockDim.x)
    if (lh[i]) atomicAdd(&A.hist[i], (unsigned long long)lh[i]);
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
namespace {
const void *wl_sweeps_kernel(int key) { BWL_KERNEL_FOR_PAIR_SUM_KEY(k_wl_sweeps, key) }
const void *ns_kernel(int key) { BWL_KERNEL_FOR_PAIR_SUM_KEY(k_ns, key) }
const void *wl_enter_kernel(int key) { BWL_KERNEL_FOR_PAIR_SUM_KEY(k_wl_enter, key) }
}  // namespace

size_t ns_lds_bytes(const Geo &g, int n_walkers) {
  return seq_geo_region((size_t)((g.lds_bytes + 15) & ~15) + 624 * 4 + kSumBufBytes +
                        sizeof(double) * g.S * g.S * g.R + sizeof(double) * (size_t)n_walkers +
                        sizeof(int32_t) * g.nsub * g.z);
}

hipError_t launch_ns(const NsArgs *args_dev, int n, size_t lds, int max_walkers, size_t lds_energies,
                     hipStream_t s, int key) {
  hipError_t e = hipFuncSetAttribute((const void *)k_ns_energies,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_energies);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ns_energies, dim3(max_walkers, n), dim3(64), lds_energies, s, args_dev);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const void *fn = ns_kernel(key);
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void *kargs[] = {(void *)&args_dev};
  return hipLaunchKernel(fn, dim3(n), dim3(64), kargs, lds, s);
}

size_t ns_energies_lds_bytes(const Geo &g) {
  return (size_t)((g.lds_bytes + 15) & ~15) + sizeof(double) * g.S * g.S * g.R +
         sizeof(int32_t) * g.nsub * g.z;
}

size_t wl_lds_bytes(const Geo &g, int bins, int window_bins) {
  return seq_geo_region((size_t)((g.lds_bytes + 15) & ~15) + 624 * 4 + kSumBufBytes +
                        sizeof(double) * g.S * g.S * g.R +
                        sizeof(double) * (size_t)(2 * bins + 1 + window_bins) + sizeof(int32_t) * g.nsub * g.z);
}

hipError_t launch_wl_sweeps(const WlArgs *args_dev, int n, size_t lds, hipStream_t s, int key) {
  const void *fn = wl_sweeps_kernel(key);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void *kargs[] = {(void *)&args_dev};
  return hipLaunchKernel(fn, dim3(n), dim3(64), kargs, lds, s);
}

hipError_t launch_wl_average(double *lng, double *hist, int n_windows, int wpw, int bins,
                             int hist_stride, hipStream_t s) {
  const int total = n_windows * (bins + hist_stride);
  hipLaunchKernelGGL(k_wl_average, dim3((total + 255) / 256), dim3(256), 0, s, lng, hist,
                     n_windows, wpw, bins, hist_stride);
  return hipGetLastError();
}

hipError_t launch_wl_energies(const WlArgs *args_dev, int n, double *E, size_t lds, hipStream_t s) {
  hipError_t e = hipFuncSetAttribute((const void *)k_wl_energies,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_wl_energies, dim3(n), dim3(64), lds, s, args_dev, E);
  return hipGetLastError();
}

size_t wl_exchange_lds_bytes(int W, int wpw) {
  return 2 * 624 * 4 + sizeof(int32_t) * (size_t)(2 * W + 6 * wpw + 1);
}

hipError_t launch_wl_exchange(const WlExchangeArgs &a, size_t lds, hipStream_t s) {
  hipError_t e = hipFuncSetAttribute((const void *)k_wl_exchange,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_wl_exchange, dim3(1), dim3(64), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_wl_swap_configs(int8_t *walkers, long long bytes, const int32_t *swaps,
                                  const int32_t *n_swaps, int max_pairs, hipStream_t s) {
  hipLaunchKernelGGL(k_wl_swap_configs, dim3(max_pairs > 0 ? max_pairs : 1), dim3(256), 0, s,
                     walkers, bytes, swaps, n_swaps);
  return hipGetLastError();
}

size_t wl_enter_lds_bytes(const Geo &g) {
  return seq_geo_region((size_t)((g.lds_bytes + 15) & ~15) + 624 * 4 + kSumBufBytes +
                        sizeof(double) * g.S * g.S * g.R + sizeof(int32_t) * g.nsub * g.z);
}

hipError_t launch_wl_enter(const WlEnterArgs *args_dev, int n, size_t lds, hipStream_t s, int key) {
  const void *fn = wl_enter_kernel(key);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  void *kargs[] = {(void *)&args_dev};
  return hipLaunchKernel(fn, dim3(n), dim3(64), kargs, lds, s);
}

hipError_t launch_asro(const AsroArgs &a, hipStream_t s) {
  const int nsites = a.g.nsub * a.g.n[0] * a.g.n[1] * a.g.n[2];
  int blocks = (nsites + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_asro, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace bwl
