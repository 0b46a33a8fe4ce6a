// bwl_host.cpp -- host-side logic (geometry, MT19937, initial configuration,
// parallel round plan).  Plain C++, no device calls.
#include "bwl_host.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>

#include "bwl_tables.h"

namespace bwl {

namespace {

int floor_div2(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

}  // namespace

int make_geo(const bwl_lattice_desc &d, Geo &g, std::vector<int32_t> &off, std::string &err) {
  memset(&g, 0, sizeof(g));
  if (d.lattice != BWL_FCC && d.lattice != BWL_BCC) {
    err = "lattice must be BWL_FCC or BWL_BCC (simple_cubic is not supported)";
    return d.lattice == 0 ? BWL_ERR_ARG : BWL_ERR_UNSUPPORTED;
  }
  if (d.n1 < 1 || d.n2 < 1 || d.n3 < 1 || d.n1 > 512 || d.n2 > 512 || d.n3 > 512) {
    err = "cell counts n1, n2, n3 must be in 1..512";
    return BWL_ERR_ARG;
  }
  if (d.n_species < 1 || d.n_species > kMaxSpecies) {
    err = "n_species must be in 1..8";
    return BWL_ERR_UNSUPPORTED;
  }
  const int max_shells = d.lattice == BWL_FCC ? 6 : 10;
  if (d.n_shells < 1 || d.n_shells > max_shells) {
    err = "unsupported number of shells";  // initialise.F90:168, 206, 237
    return BWL_ERR_UNSUPPORTED;
  }
  g.lattice = d.lattice;
  g.n[0] = d.n1; g.n[1] = d.n2; g.n[2] = d.n3;
  for (int a = 0; a < 3; ++a) g.m[a] = 2 * g.n[a];
  g.S = d.n_species;
  g.R = d.n_shells;
  g.nsub = d.lattice == BWL_FCC ? 4 : 2;
  g.n_atoms = g.nsub * d.n1 * d.n2 * d.n3;
  const int8_t(*tab)[3] = d.lattice == BWL_FCC ? kFccOffsets : kBccOffsets;
  const int *sz = d.lattice == BWL_FCC ? kFccShellSize : kBccShellSize;
  int z = 0;
  g.shell_start[0] = 0;
  for (int n = 0; n < g.R; ++n) {
    z += sz[n];
    g.shell_start[n + 1] = z;
  }
  g.z = z;
  int rcut = 0, h = 0;
  for (int t = 0; t < z; ++t)
    for (int a = 0; a < 3; ++a) {
      rcut = std::max(rcut, (int)std::abs(tab[t][a]));
      for (int b = 0; b <= 1; ++b) h = std::max(h, std::abs(floor_div2(b + tab[t][a])));
    }
  g.rcut = rcut;
  g.h = h;
  for (int a = 0; a < 3; ++a) g.P[a] = g.n[a] + 2 * h;
  g.sub_stride = g.P[0] * g.P[1] * g.P[2];
  long long lds = (long long)g.nsub * g.sub_stride;
  g.lds_bytes = (int)((lds + 15) & ~15LL);
  g.compact_bytes = g.nsub * d.n1 * d.n2 * d.n3;
  off.assign((size_t)g.nsub * z, 0);
  for (int s = 0; s < g.nsub; ++s) {
    int bx, by, bz;
    sub_bits(g.lattice, s, bx, by, bz);
    for (int t = 0; t < z; ++t) {
      int ex = bx + tab[t][0], ey = by + tab[t][1], ez = bz + tab[t][2];
      int s2 = sub_of(g.lattice, ex & 1, ey & 1);
      int dcx = floor_div2(ex), dcy = floor_div2(ey), dcz = floor_div2(ez);
      off[(size_t)s * z + t] =
          (s2 - s) * g.sub_stride + (dcz * g.P[1] + dcy) * g.P[0] + dcx;
    }
  }
  return BWL_OK;
}

// ---------------------------------------------------------------------------
void MT19937::seed(uint32_t s) {
  mt[0] = s;
  for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  mti = 624;
}

uint32_t MT19937::next() {
  if (mti >= 624) {
    if (mti == 625) seed(5489u);
    for (int k = 0; k < 624; ++k) {
      uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
      mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    mti = 0;
  }
  uint32_t y = mt[mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// ---------------------------------------------------------------------------
// initial_setup: species counts from numbers (REAL(4) concentrations,
// initialise.F90:468-472) or from concentrations (nint + round-robin
// correction, :480-504); then rejection sampling site by site, in the
// reference's loop order, against inclusive cumulative bounds.
int initial_setup_host(const Geo &g, const int64_t *species_numbers, const double *concentrations,
                       MT19937 &mt, int8_t *config, std::string &err) {
  const int S = g.S;
  const int n_sites = g.n_atoms;
  std::vector<int64_t> count(S, 0), placed(S, 0);
  std::vector<double> conc(S + 1, 0.0);
  int64_t total_numbers = 0;
  if (species_numbers)
    for (int i = 0; i < S; ++i) total_numbers += species_numbers[i];
  if (species_numbers && total_numbers == n_sites) {
    for (int i = 0; i < S; ++i) {
      count[i] = species_numbers[i];
      conc[i + 1] = (double)((float)count[i] / (float)n_sites);
    }
  } else {
    if (!concentrations) {
      err = "species numbers do not sum to the number of sites and no concentrations given";
      return BWL_ERR_ARG;
    }
    int64_t total = 0;
    for (int i = 0; i < S; ++i) {
      conc[i + 1] = concentrations[i];
      count[i] = (int64_t)llround((double)(float)n_sites * concentrations[i]);
      total += count[i];
    }
    int64_t pending = total - n_sites;
    const int64_t step = pending > 0 ? -1 : 1;
    int guard = 0;
    while (pending != 0) {
      bool moved = false;
      for (int j = 0; j < S && pending != 0; ++j) {
        if (count[j] == 0) continue;
        count[j] += step;
        pending += step;
        moved = true;
      }
      if (!moved || ++guard > (1 << 24)) {
        err = "cannot correct species counts";
        return BWL_ERR_ARG;
      }
    }
  }
  int64_t sum = std::accumulate(count.begin(), count.end(), (int64_t)0);
  if (sum != n_sites) {
    err = "Species counts do not sum to the number of lattice sites";
    return BWL_ERR_ARG;
  }
  for (int i = 0; i < S; ++i)
    if (count[i] < 0) {
      err = "negative species count";
      return BWL_ERR_ARG;
    }
  // bounds sum(conc(0:l)), accumulated from element 0 each time
  std::vector<double> bound(S + 1);
  for (int l = 0; l <= S; ++l) {
    double acc = 0.0;
    for (int q = 0; q <= l; ++q) acc = acc + conc[q];
    bound[l] = acc;
  }
  const size_t cells = (size_t)g.m[0] * g.m[1] * g.m[2];
  memset(config, 0, cells);
  auto fill = [&](size_t idx) -> bool {
    long long draws = 0;
    while (config[idx] == 0) {
      double r = mt.genrand();
      for (int l = 1; l <= S; ++l) {
        if (r >= bound[l - 1] && r <= bound[l] && placed[l - 1] < count[l - 1]) {
          config[idx] = (int8_t)l;
          placed[l - 1]++;
        }
      }
      if (++draws > (1LL << 40)) return false;
    }
    return true;
  };
  const int m0 = g.m[0], m1 = g.m[1], m2 = g.m[2];
  for (int k = 1; k <= m2; ++k) {
    if (g.lattice == BWL_BCC) {
      // sites (2i - mod(k,2), 2j - mod(k,2), k), initialise.F90:540-570
      const int sh = k & 1;
      for (int j = 1; j <= m1 / 2; ++j)
        for (int i = 1; i <= m0 / 2; ++i) {
          size_t idx = (size_t)(2 * i - sh - 1) + (size_t)m0 * ((2 * j - sh - 1) + (size_t)m1 * (k - 1));
          if (!fill(idx)) { err = "initial_setup did not terminate"; return BWL_ERR_ARG; }
        }
    } else {
      // fcc, initialise.F90:572-611
      for (int j = 1; j <= m1; ++j)
        for (int i = 1; i <= m0 / 2; ++i) {
          int x = 2 * i - (k & 1) * (j & 1) - ((k + 1) & 1) * ((j + 1) & 1);
          size_t idx = (size_t)(x - 1) + (size_t)m0 * ((j - 1) + (size_t)m1 * (k - 1));
          if (!fill(idx)) { err = "initial_setup did not terminate"; return BWL_ERR_ARG; }
        }
    }
  }
  return BWL_OK;
}

// ---------------------------------------------------------------------------
int plan_parallel(const Geo &g, ParPlan &p, std::string &err) {
  const int r = g.rcut;
  const int need = r + 1;
  p.stride = (need % 2) ? need + 1 : need;  // smallest even >= rcut+1
  p.L = 2 * r + 4;                           // even; u and L-u both >= rcut+1
  for (int a = 0; a < 3; ++a) {
    const int m = g.m[a];
    if (m < need) {
      err = "lattice too small for simultaneous swaps (extent below interaction range)";
      return BWL_ERR_UNSUPPORTED;
    }
    p.K[a] = (m % p.stride == 0) ? m / p.stride : (m - need) / p.stride + 1;
  }
  p.vars.clear();
  p.muls.clear();
  p.fast_halo = true;
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] < 2 * g.h) p.fast_halo = false;
    int got = 0;
    for (uint32_t c = 1; c <= (uint32_t)p.K[a] && got < 16; ++c)
      if ((c < (uint32_t)p.K[a] || p.K[a] == 1) && std::gcd(c, (uint32_t)p.K[a]) == 1)
        p.units[a][got++] = c;
    p.n_units[a] = got;
  }
  for (int axis = 0; axis < 3; ++axis)
    for (int par = 0; par < 2; ++par) {
      RoundVar v;
      v.axis = axis;
      v.u = ((r + 1) % 2 == par) ? r + 1 : r + 2;
      // layers at (j/2)*L + (j&1)*u while the wrap gap stays >= rcut+1
      int nL = 0;
      for (int j = 0;; ++j) {
        int pos = (j >> 1) * p.L + (j & 1) * v.u;
        if (pos > g.m[axis] - need) break;
        ++nL;
      }
      v.nL1 = (nL + 1) / 2;
      v.nL2 = nL / 2;
      if (v.nL2 < 1) continue;
      const int kk = p.K[(axis + 1) % 3] * p.K[(axis + 2) % 3];
      v.n1 = v.nL1 * kk;
      v.n2 = v.nL2 * kk;
      v.mul_off = (int32_t)p.muls.size();
      int got = 0;
      for (uint32_t c = 1; c <= (uint32_t)v.nL1 && got < 16; ++c)
        if (c < (uint32_t)v.nL1 || v.nL1 == 1)
          if (std::gcd(c, (uint32_t)v.nL1) == 1) {
            p.muls.push_back(c);
            ++got;
          }
      v.n_mul = got;
      p.vars.push_back(v);
    }
  if (p.vars.empty()) {
    err = "lattice too small for simultaneous swaps (fewer than two active layers)";
    return BWL_ERR_UNSUPPORTED;
  }
  // The rounds must be able to exchange atoms between every pair of
  // sublattices, or the sublattice compositions would be frozen: union the
  // sublattices each round kind can pair and require one component.
  int parent[4] = {0, 1, 2, 3};
  auto find = [&](int x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (const RoundVar &v : p.vars)
    for (int s1 = 0; s1 < g.nsub; ++s1) {
      int b1[3];
      sub_bits(g.lattice, s1, b1[0], b1[1], b1[2]);
      const int A = v.axis, B = (A + 1) % 3, Cx = (A + 2) % 3;
      for (int fb = 0; fb <= 1; ++fb) {
        int b2[3];
        b2[A] = b1[A] ^ (v.u & 1);
        b2[B] = g.lattice == BWL_FCC ? fb : b2[A];
        b2[Cx] = g.lattice == BWL_FCC ? (b2[A] ^ b2[B]) : b2[A];
        int s2 = sub_of(g.lattice, b2[0], b2[1]);
        parent[find(s1)] = find(s2);
      }
    }
  for (int s = 1; s < g.nsub; ++s)
    if (find(s) != find(0)) {
      err = "lattice too small for simultaneous swaps across sublattices";
      return BWL_ERR_UNSUPPORTED;
    }
  return BWL_OK;
}

bool make_nat(const Geo &g, NatGeo &nat) {
  memset(&nat, 0, sizeof(nat));
  if (g.n[0] != g.n[1] || g.n[1] != g.n[2]) return false;
  return nat_dims(g.lattice, g.R, g.n[0], nat);
}

bool make_nat_fixed(const Geo &g, NatGeo &nat) {
  memset(&nat, 0, sizeof(nat));
  if (g.n[0] != g.n[1] || g.n[1] != g.n[2]) return false;
  return nat_dims_fixed(g.lattice, g.R, g.n[0], nat);
}

int check_config(const Geo &g, const int8_t *config, std::string &err) {
  const int m0 = g.m[0], m1 = g.m[1], m2 = g.m[2];
  for (int z = 0; z < m2; ++z)
    for (int y = 0; y < m1; ++y)
      for (int x = 0; x < m0; ++x) {
        int v = config[x + (size_t)m0 * (y + (size_t)m1 * z)];
        bool occ = g.lattice == BWL_FCC ? (((x + y + z) & 1) == 0)
                                        : ((x & 1) == (y & 1) && (y & 1) == (z & 1));
        if (occ ? (v < 1 || v > g.S) : (v != 0)) {
          err = "config holds an invalid occupation at (" + std::to_string(x + 1) + "," +
                std::to_string(y + 1) + "," + std::to_string(z + 1) + ")";
          return BWL_ERR_ARG;
        }
      }
  return BWL_OK;
}

}  // namespace bwl

namespace bwl {

void lattice_shells_host(const Geo &g, int wc, double *shells) {
  // Distances are sqrt of integers, so the distinct values are those of the
  // distinct squared distances; 0 is always present (site (1,1,1) is
  // occupied, and the reference pads its sorted list with zeros).
  const int m0 = g.m[0], m1 = g.m[1], m2 = g.m[2];
  const long long dmax = (long long)(m0 - 1) * (m0 - 1) + (long long)(m1 - 1) * (m1 - 1) +
                         (long long)(m2 - 1) * (m2 - 1);
  std::vector<uint8_t> seen((size_t)dmax + 1, 0);
  seen[0] = 1;
  for (int k = 0; k < m2; ++k)
    for (int j = 0; j < m1; ++j)
      for (int i = 0; i < m0; ++i) {
        const bool occ = g.lattice == BWL_FCC ? (((i + j + k) & 1) == 0)
                                              : ((i & 1) == (j & 1) && (j & 1) == (k & 1));
        if (occ) seen[(size_t)(i * i + j * j + k * k)] = 1;
      }
  std::vector<double> vals;
  for (long long d = 0; d <= dmax; ++d)
    if (seen[(size_t)d]) {
      // REAL(4) sum of the squares, then REAL(4) sqrt (analytics.f90:247-249)
      vals.push_back((double)sqrtf((float)d));
    }
  for (int s = 0; s < wc; ++s) shells[s] = 0.0;
  // a shell ends where the next distance differs by 1e-3 or more
  int cnt = 0;
  for (size_t i = 0; i + 1 < vals.size() && cnt < wc; ++i) {
    if (fabs(vals[i] - vals[i + 1]) < 1e-3) continue;
    shells[cnt++] = vals[i];
  }
}

void asro_offsets_host(const Geo &g, int wc, const double *shells, std::vector<int32_t> &off3,
                       int32_t *shell_start) {
  const int w[3] = {std::min(g.n[0], 5), std::min(g.n[1], 5), std::min(g.n[2], 5)};
  std::vector<std::vector<int32_t>> per((size_t)wc);
  for (int dz = -w[2]; dz <= w[2]; ++dz)
    for (int dy = -w[1]; dy <= w[1]; ++dy)
      for (int dx = -w[0]; dx <= w[0]; ++dx) {
        const bool lat = g.lattice == BWL_FCC
                             ? (((dx + dy + dz) & 1) == 0)
                             : (((dx ^ dy) & 1) == 0 && ((dy ^ dz) & 1) == 0);
        if (!lat) continue;
        // minimum image of the index difference, as the reference computes
        // it from site (1,1,1): REAL(4) differences, NINT of the ratio
        const int d[3] = {dx, dy, dz};
        double r[3];
        for (int a = 0; a < 3; ++a) {
          const int m = g.m[a];
          const int j = ((d[a] % m) + m) % m + 1;
          double v = (double)(float)(1 - j);
          v = v - (double)(float)m * (double)lround(v / (double)(float)m);
          r[a] = v;
        }
        const double dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int l = 0; l < wc; ++l)
          if (fabs(dist - shells[l]) < 1e-3) {
            per[(size_t)l].push_back(dx);
            per[(size_t)l].push_back(dy);
            per[(size_t)l].push_back(dz);
          }
      }
  off3.clear();
  shell_start[0] = 0;
  for (int l = 0; l < wc; ++l) {
    off3.insert(off3.end(), per[(size_t)l].begin(), per[(size_t)l].end());
    shell_start[l + 1] = (int32_t)(off3.size() / 3);
  }
}

void wl_setup_host(const Geo &g, int bins, int num_windows, float bin_overlap, float energy_min,
                   float energy_max, int32_t *windows, double *bin_edges) {
  // divide_range: window i ends at floor(factor * i + 1) with
  // factor = (bins - 1) / ((n + 1)^1 - 1); the next starts one bin later
  std::vector<int32_t> iv((size_t)2 * num_windows);
  const double factor = ((double)bins - 1.0) / (pow((double)num_windows + 1.0, 1.0) - 1.0);
  iv[0] = 1;
  iv[(size_t)2 * (num_windows - 1) + 1] = bins;
  for (int i = 2; i <= num_windows; ++i) {
    iv[(size_t)2 * (i - 2) + 1] = (int32_t)floor(factor * pow((double)(i - 1), 1.0) + 1.0);
    iv[(size_t)2 * (i - 1)] = iv[(size_t)2 * (i - 2) + 1] + 1;
  }
  for (int q = 0; q < 2 * num_windows; ++q) windows[q] = iv[(size_t)q];
  // create_overlap: every window after the first starts ceil(overlap *
  // width of the previous window) bins earlier (at least 2)
  if (num_windows > 1) {
    for (int i = 2; i <= num_windows - 1; ++i) {
      const int nb = windows[2 * (i - 2) + 1] - windows[2 * (i - 2)] + 1;
      int ov = (int)ceilf(bin_overlap * (float)nb);
      ov = std::max(ov, 2);
      windows[2 * (i - 1)] = iv[(size_t)2 * (i - 1)] - ov;
      windows[2 * (i - 1) + 1] = iv[(size_t)2 * (i - 1) + 1];
    }
    const int nw = num_windows;
    const int nb = windows[2 * (nw - 2) + 1] - windows[2 * (nw - 2)];
    int ov = (int)ceilf(bin_overlap * (float)nb);
    ov = std::max(ov, 2);
    windows[2 * (nw - 1)] = iv[(size_t)2 * (nw - 1)] - ov;
    windows[2 * (nw - 1) + 1] = iv[(size_t)2 * (nw - 1) + 1];
  }
  // create_energy_bins: meV/atom (REAL(4)) to Ry for the whole lattice
  const double to_ry = (double)g.n_atoms / (13.605693122 * 1000.0);
  const double width = (double)((energy_max - energy_min) / (float)bins) * to_ry;
  for (int i = 1; i <= bins + 1; ++i)
    bin_edges[i - 1] = (double)energy_min * to_ry + (double)(i - 1) * width;
}

}  // namespace bwl
