// bwl_device.h -- device building blocks shared by the gfx950 kernels:
// LDS staging of a lattice with periodic halos, halo-consistent site writes,
// block reductions, the reference's MT19937 stream, and correctly-rounded
// exp/log comparisons used by the bit-exact sequential mode.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "bwl_core.h"

namespace bwl {

// Keeps the compiler from moving memory accesses across this point, without
// a barrier or a wait: in a workgroup of one wavefront the LDS accesses of
// the wavefront complete in program order (an LDS read issued after an LDS
// write sees it, whichever lanes made them), so ordering the instructions
// orders the accesses.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_wave_barrier();
  __asm__ __volatile__("" ::: "memory");
}

// v as a wavefront-uniform value (read from the first active lane), for a
// value that every lane computes alike: the compiler keeps it in a scalar
// register and branches on it without exec-mask bookkeeping.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Positive modulo for a small positive divisor.
__device__ __forceinline__ int pmod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// A cell index of a box of extents (E0, E1, E2) x sublattices, held as its
// digits (s, z, y, x) and advanced by a fixed step with carries: the digits
// of a thread's strided walk through the box without a division per cell.
struct BoxWalk {
  int x, y, z, s;
  int dx, dy, dz, ds;  // the step's digits
};

__device__ inline BoxWalk box_walk(int start, int step, int e0, int e1, int e2) {
  const int plane = e0 * e1, vol = plane * e2;
  BoxWalk w;
  w.s = start / vol;
  int r = start - w.s * vol;
  w.z = r / plane;
  r -= w.z * plane;
  w.y = r / e0;
  w.x = r - w.y * e0;
  w.ds = step / vol;
  r = step - w.ds * vol;
  w.dz = r / plane;
  r -= w.dz * plane;
  w.dy = r / e0;
  w.dx = r - w.dy * e0;
  return w;
}

__device__ __forceinline__ void box_step(BoxWalk &w, int e0, int e1, int e2) {
  w.x += w.dx;
  int c = w.x >= e0;
  w.x -= c ? e0 : 0;
  w.y += w.dy + c;
  c = w.y >= e1;
  w.y -= c ? e1 : 0;
  w.z += w.dz + c;
  c = w.z >= e2;
  w.z -= c ? e2 : 0;
  w.s += w.ds + c;
}

// (p - h) mod n for a padded coordinate p in [0, n + 2h), h <= n.
__device__ __forceinline__ int unpad(int p, int h, int n) {
  const int c = p - h;
  return c < 0 ? c + n : (c >= n ? c - n : c);
}

// Stage a compact lattice (HBM) into the padded LDS layout, halos included.
// `shift` > 0 stores shift*(species-1) (a bit position for packed species
// counting) instead of the species itself.
__device__ inline void stage_lattice(int8_t *lat, const int8_t *__restrict__ compact,
                                     const Geo &g, int shift) {
  const int total = g.nsub * g.sub_stride;
  const int plane = g.P[0] * g.P[1];
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int s = idx / g.sub_stride;
    int r = idx - s * g.sub_stride;
    int pz = r / plane;
    r -= pz * plane;
    int py = r / g.P[0];
    int px = r - py * g.P[0];
    int cx = pmod(px - g.h, g.n[0]);
    int cy = pmod(py - g.h, g.n[1]);
    int cz = pmod(pz - g.h, g.n[2]);
    int v = compact[compact_addr(g, s, cx, cy, cz)];
    lat[idx] = (int8_t)(shift ? shift * (v - 1) : v);
  }
}

// Write interior cells back to the compact HBM lattice.
__device__ inline void unstage_lattice(const int8_t *lat, int8_t *__restrict__ compact,
                                       const Geo &g, int shift) {
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int total = g.nsub * cells;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    int s = idx / cells;
    int r = idx - s * cells;
    int cz = r / (g.n[0] * g.n[1]);
    r -= cz * g.n[0] * g.n[1];
    int cy = r / g.n[0];
    int cx = r - cy * g.n[0];
    int v = lat[padded_addr(g, s, cx, cy, cz)];
    compact[idx] = (int8_t)(shift ? v / shift + 1 : v);
  }
}

// stage_lattice for a species lattice staged many times in one kernel (the
// nested-sampling clone at every iteration): the same cells, each thread's
// cell digits carried along its strided walk instead of divided out per
// cell.  (Kernels that stage once keep stage_lattice: there the walk's code
// measured slower in the trial loop around it.)
__device__ inline void stage_lattice_walk(int8_t *lat, const int8_t *__restrict__ compact,
                                          const Geo &g) {
  const int total = g.nsub * g.sub_stride;
  if (g.h <= g.n[0] && g.h <= g.n[1] && g.h <= g.n[2]) {
    BoxWalk w = box_walk(threadIdx.x, blockDim.x, g.P[0], g.P[1], g.P[2]);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int cx = unpad(w.x, g.h, g.n[0]), cy = unpad(w.y, g.h, g.n[1]), cz = unpad(w.z, g.h, g.n[2]);
      const int v = compact[compact_addr(g, w.s, cx, cy, cz)];
      lat[idx] = (int8_t)v;
      box_step(w, g.P[0], g.P[1], g.P[2]);
    }
    return;
  }
  stage_lattice(lat, compact, g, 0);
}

// unstage_lattice by the same carried walk.
__device__ inline void unstage_lattice_walk(const int8_t *lat, int8_t *__restrict__ compact,
                                            const Geo &g) {
  const int cells = g.n[0] * g.n[1] * g.n[2];
  const int total = g.nsub * cells;
  BoxWalk w = box_walk(threadIdx.x, blockDim.x, g.n[0], g.n[1], g.n[2]);
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int v = lat[padded_addr(g, w.s, w.x, w.y, w.z)];
    compact[idx] = (int8_t)v;
    box_step(w, g.n[0], g.n[1], g.n[2]);
  }
}

// Store `v` at interior cell (s,cx,cy,cz) and at every periodic image of it
// inside the padded box.  Images are split over `nl` cooperating lanes
// (lane `l`); a single thread passes l = 0, nl = 1.
template <class GEO>
__device__ inline void write_site(int8_t *lat, const GEO &g, int s, int cx, int cy, int cz,
                                  int8_t v, int l, int nl) {
  if (g.h <= g.n[0] && g.h <= g.n[1] && g.h <= g.n[2]) {
    // halo no wider than the box (every lattice but the smallest): the first
    // image is (c + h) mod n by one conditional subtract, there are at most
    // 3 images per axis, and image q splits into (ix, iy, iz) by compares
    // -- the same cells as the general loop below, without divisions
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int p0x = cx + g.h >= n0 ? cx + g.h - n0 : cx + g.h;
    const int p0y = cy + g.h >= n1 ? cy + g.h - n1 : cy + g.h;
    const int p0z = cz + g.h >= n2 ? cz + g.h - n2 : cz + g.h;
    const int kx = 1 + (p0x + n0 < g.P[0]) + (p0x + 2 * n0 < g.P[0]);
    const int ky = 1 + (p0y + n1 < g.P[1]) + (p0y + 2 * n1 < g.P[1]);
    const int kz = 1 + (p0z + n2 < g.P[2]) + (p0z + 2 * n2 < g.P[2]);
    const int kxy = kx * ky, total = kxy * kz;
    for (int q = l; q < total; q += nl) {
      const int iz = (q >= kxy) + (q >= 2 * kxy);
      const int r = q - iz * kxy;
      const int iy = (r >= kx) + (r >= 2 * kx);
      const int ix = r - iy * kx;
      lat[s * g.sub_stride + ((p0z + iz * n2) * g.P[1] + (p0y + iy * n1)) * g.P[0] + p0x + ix * n0] = v;
    }
    return;
  }
  int p0x = (cx + g.h) % g.n[0], p0y = (cy + g.h) % g.n[1], p0z = (cz + g.h) % g.n[2];
  int kx = (g.P[0] - 1 - p0x) / g.n[0] + 1;
  int ky = (g.P[1] - 1 - p0y) / g.n[1] + 1;
  int kz = (g.P[2] - 1 - p0z) / g.n[2] + 1;
  int total = kx * ky * kz;
  for (int q = l; q < total; q += nl) {
    int ix = q % kx;
    int iy = (q / kx) % ky;
    int iz = q / (kx * ky);
    int px = p0x + ix * g.n[0], py = p0y + iy * g.n[1], pz = p0z + iz * g.n[2];
    lat[s * g.sub_stride + (pz * g.P[1] + py) * g.P[0] + px] = v;
  }
}

// Deterministic block sum (fixed reduction order for a fixed block size).
__device__ inline double block_sum(double v, double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  __syncthreads();
  return t;
}

__device__ inline unsigned long long block_sum_u64(unsigned long long v, unsigned long long *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  unsigned long long t = 0;
  for (int w = 0; w < nw; ++w) t += red[w];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------------
// MT19937 on the device, state in LDS, driven by one whole wavefront
// (every lane calls with the same mti).  The twist runs in 64-wide chunks
// whose reads all precede their writes, which reproduces the serial
// recurrence of the generator exactly.
// ------------------------------------------------------------------
__device__ inline void mt_twist(uint32_t *st) {
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < 623; base += 64) {
    int k = base + lane;
    uint32_t v = 0;
    if (k < 623) {
      uint32_t y = (st[k] & 0x80000000u) | (st[k + 1] & 0x7fffffffu);
      uint32_t other = k < 227 ? st[k + 397] : st[k - 227];
      v = other ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    __syncthreads();
    if (k < 623) st[k] = v;
    __syncthreads();
  }
  if (lane == 0) {
    uint32_t y = (st[623] & 0x80000000u) | (st[0] & 0x7fffffffu);
    st[623] = st[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  __syncthreads();
}

__device__ inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// Three consecutive draws: when no twist falls between them the three state
// words are read together (one round trip instead of three); the values are
// those of three mt_next calls.
__device__ inline void mt_next3(uint32_t *st, int &mti, uint32_t &r1, uint32_t &r2, uint32_t &r3);

__device__ inline uint32_t mt_next(uint32_t *st, int &mti) {
  if (mti >= 624) {
    mt_twist(st);
    mti = 0;
  }
  const uint32_t y = st[mti];
  mti++;
  return mt_temper(y);
}

__device__ inline void mt_next3(uint32_t *st, int &mti, uint32_t &r1, uint32_t &r2, uint32_t &r3) {
  if (mti <= 621) {
    const uint32_t y1 = st[mti], y2 = st[mti + 1], y3 = st[mti + 2];
    mti += 3;
    r1 = mt_temper(y1);
    r2 = mt_temper(y2);
    r3 = mt_temper(y3);
    return;
  }
  r1 = mt_next(st, mti);
  r2 = mt_next(st, mti);
  r3 = mt_next(st, mti);
}

// A window of 64 consecutive tempered draws held across the wavefront's
// lanes (lane i: state word base + i), so a draw is one cross-lane read
// instead of an LDS round trip; the window is refilled (one LDS read per
// lane) when the draws leave it and after every twist.  The draws and the
// state are those of mt_next.  A kernel that uses a window takes every draw
// of that state through it (a twist elsewhere would leave it stale).
struct MtWin {
  uint32_t w;
  int base;  // state index of lane 0's word; kMtWinEmpty before the first fill
};
constexpr int kMtWinEmpty = -4096;

__device__ inline void mt_win_fill(const uint32_t *st, int from, MtWin &win) {
  const int k = from + (int)(threadIdx.x & 63);
  win.w = mt_temper(st[k < 624 ? k : 623]);  // lanes past the end are never read
  win.base = from;
}

__device__ inline uint32_t mt_next(uint32_t *st, int &mti, MtWin &win) {
  if (mti >= 624) {
    mt_twist(st);
    mti = 0;
    mt_win_fill(st, 0, win);
  } else if ((unsigned)(mti - win.base) >= 64u) {
    mt_win_fill(st, mti, win);
  }
  const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)win.w, mti - win.base);
  ++mti;
  return r;
}

__device__ inline void mt_next3(uint32_t *st, int &mti, MtWin &win, uint32_t &r1, uint32_t &r2,
                                uint32_t &r3) {
  if (mti <= 621 && (unsigned)(mti - win.base) <= 61u) {
    const int i = mti - win.base;
    r1 = (uint32_t)__builtin_amdgcn_readlane((int)win.w, i);
    r2 = (uint32_t)__builtin_amdgcn_readlane((int)win.w, i + 1);
    r3 = (uint32_t)__builtin_amdgcn_readlane((int)win.w, i + 2);
    mti += 3;
    return;
  }
  r1 = mt_next(st, mti, win);
  r2 = mt_next(st, mti, win);
  r3 = mt_next(st, mti, win);
}

// ------------------------------------------------------------------
// Double-double arithmetic (Dekker/Knuth error-free transforms) for a
// correctly-rounded exp and log, used to decide u < exp(x) and
// log(u) < t exactly as a correctly-rounded libm would.
// ------------------------------------------------------------------
struct DD {
  double hi, lo;
};

__host__ __device__ inline DD two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__host__ __device__ inline DD quick_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}

__host__ __device__ inline DD two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}

__host__ __device__ inline DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  DD t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ inline DD dd_mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}

__host__ __device__ inline DD dd_mul_d(DD a, double b) {
  DD p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return quick_two_sum(p.hi, p.lo);
}

// exp(r) for a reduced argument |r| <= 0.35 in double-double.
__host__ __device__ inline DD dd_exp_reduced(DD r) {
  // exp(r) for |r| <= 0.35: scale by 2^-10, Taylor to degree 12, square 10x.
  DD s = {ldexp(r.hi, -10), ldexp(r.lo, -10)};
  // expm1(s) = s + s^2/2! + ... (Horner in double-double)
  DD acc = {1.0 / 479001600.0, 0.0};  // 1/12!
  const double inv_fact[12] = {1.0,
                               1.0 / 2.0,
                               1.0 / 6.0,
                               1.0 / 24.0,
                               1.0 / 120.0,
                               1.0 / 720.0,
                               1.0 / 5040.0,
                               1.0 / 40320.0,
                               1.0 / 362880.0,
                               1.0 / 3628800.0,
                               1.0 / 39916800.0,
                               1.0 / 479001600.0};
  // 1/k! as double-double: hi from the double, lo from the exact remainder
  // is not needed at this scale because |s|^k/k! shrinks below 2^-60 fast
  // for k >= 3; the k = 1, 2 coefficients (1, 1/2) are exact.
  for (int k = 11; k >= 1; --k) {
    acc = dd_mul(acc, s);
    acc = dd_add(acc, DD{inv_fact[k - 1], 0.0});
  }
  DD em1 = dd_mul(acc, s);  // s * (1 + s/2 + ...) = expm1(s)
  // (1+e)^2 - 1 = 2e + e^2
  for (int i = 0; i < 10; ++i) em1 = dd_add(dd_mul_d(em1, 2.0), dd_mul(em1, em1));
  return dd_add(DD{1.0, 0.0}, em1);
}

__host__ __device__ inline DD dd_exp(double x) {
  const double LN2_HI = 6.93147180369123816490e-01;   // 0x3FE62E42FEE00000
  const double LN2_MID = 1.90821492927058770002e-10;  // 0x3DEA39EF35793C76
  const double LN2_LO = 1.1612227229362532e-26;       // ln2 - HI - MID
  // ln 2 = HI + MID + LO; HI has 21 trailing zero bits, so k*HI is exact
  // for |k| < 2^21.
  double k = rint(x * 1.44269504088896338700e+00);
  DD r = two_sum(x, -k * LN2_HI);
  r = dd_add(r, two_prod(-k, LN2_MID));
  r = dd_add(r, DD{-k * LN2_LO, 0.0});
  DD e = dd_exp_reduced(r);
  int ki = (int)k;
  return {ldexp(e.hi, ki), ldexp(e.lo, ki)};
}

// True iff u < exp(x) with exp correctly rounded to FP64.
__host__ __device__ inline bool lt_exp(double u, double x) {
  double e = exp(x);
  double tol = e * 8.9e-16;  // 4 ulp
  if (u < e - tol) return true;
  if (u > e + tol) return false;
  DD d = dd_exp(x);
  double rn = d.hi + d.lo;  // correctly rounded value of the double-double
  return u < rn;
}

// lt_exp on the device with a float exponential deciding first whenever u
// is clear of it by 1e-4 (relative; the float's error is below 1e-6), which
// leaves the FP64 and double-double paths to the rare close calls.
__device__ inline bool lt_exp_dev(double u, double x) {
  if (x > -80.0 && x < 80.0) {
    const double ef = (double)__expf((float)x);
    if (u < ef * (1.0 - 1e-4)) return true;
    if (u > ef * (1.0 + 1e-4)) return false;
  }
  return lt_exp(u, x);
}

// log(u) for u in (0,1) as double-double: Newton step on exp.
__host__ __device__ inline DD dd_log(double u) {
  double y0 = log(u);
  // y1 = y0 + (u - exp(y0)) / exp(y0)  (one Newton step doubles precision)
  DD e = dd_exp(y0);
  DD diff = dd_add(DD{u, 0.0}, DD{-e.hi, -e.lo});
  double corr = (diff.hi + diff.lo) / (e.hi + e.lo);
  return two_sum(y0, corr);
}

// True iff log(u) < t with log correctly rounded to FP64 (u > 0).
__host__ __device__ inline bool log_lt(double u, double t) {
  if (u <= 0.0) return t > -INFINITY;  // log(0) = -inf
  double l = log(u);
  double tol = fabs(l) * 8.9e-16 + 1e-300;
  if (l < t - tol) return true;
  if (l > t + tol) return false;
  DD d = dd_log(u);
  double rn = d.hi + d.lo;
  return rn < t;
}

// log_lt on the device with a float logarithm deciding first whenever it is
// clear of t by 1e-4 (for u >= 2^-32 its error is below 1e-5).
__device__ inline bool log_lt_dev(double u, double t) {
  if (u >= 0x1p-32 && t > -1e30 && t < 1e30) {
    const double lf = (double)__logf((float)u);
    if (lf < t - 1e-4) return true;
    if (lf > t + 1e-4) return false;
  }
  return log_lt(u, t);
}

// log_lt_dev for a u and t that every lane holds alike.
__device__ inline bool log_lt_dev_uniform(double u, double t) {
  if (wave_uniform(u >= 0x1p-32 && t > -1e30 && t < 1e30)) {
    const double lf = (double)__logf((float)u);
    if (wave_uniform(lf < t - 1e-4)) return true;
    if (wave_uniform(lf > t + 1e-4)) return false;
  }
  return wave_uniform(log_lt(u, t));
}


}  // namespace bwl
