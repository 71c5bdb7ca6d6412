// MI355X (gfx950) implementation of the BVE Biot-Savart hot path and the
// C ABI of include/lpm_bve.h.  Kernel design notes: lpm_bve_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "lpm_bve.h"
#include "lpm_bve_kernels.hpp"
#include "lpm_mesh.hpp"
#include "lpm_meshio.hpp"
#include "lpm_status.hpp"

namespace lpm {
namespace dev {

constexpr double kPi = 3.1415926535897932384626433832795027975;  // src/LpmDefs.hpp:82
constexpr double kLn2 = 0.69314718055994530941723212145817657;

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------

// Reciprocal of a positive double in the FP32 normal range.  r = v_rcp_f32
// estimate (relative error |e| < 2^-23, issued to the transcendental unit
// beside the VALU); with e = 1 - a r (one FMA, nearly exact),
// 1/a = r/(1 - e) = r (1 + e + e^2) + O(e^3 r), and e^3 < 2^-69, so two more
// FMAs give the reciprocal to within the final rounding.
__device__ __forceinline__ double rcp64(double a) {
  const double r = (double)__builtin_amdgcn_rcpf((float)a);
  const double e = fma(-a, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// 1/a, 1/b, 1/c, 1/d of four nonzero doubles from one reciprocal:
// p = 1/(ab cd) from the v_rcp_f64 estimate r refined to second order
// (e = 1 - P r, p = r (1 + e + e^2), as in rcp64), then 1/a = b (cd p),
// 1/b = a (cd p), 1/c = d (ab p), 1/d = c (ab p).  12 FP64 operations and one
// v_rcp_f64 replace four rcp64 calls (12 FP64 operations, 8 conversions and
// 4 v_rcp_f32); each result carries about five roundings.  Signs pass
// through (the literal 1 - x.y of an off-sphere stage argument can be
// negative).  The product must lie in the FP64 normal range, which four
// pair denominators of magnitude >= 1e-70 or padding records satisfy.  A
// zero denominator makes all four results non-finite.
__device__ __forceinline__ void rcp4_64(double a, double b, double c, double d, double& ia, double& ib,
                                        double& ic, double& id) {
  const double ab = a * b, cd = c * d, P = ab * cd;
  const double r = __builtin_amdgcn_rcp(P);
  const double e = fma(-P, r, 1.0);
  const double p = fma(r, fma(e, e, e), r);
  const double iab = cd * p, icd = ab * p;
  ia = b * iab;
  ib = a * iab;
  ic = d * icd;
  id = c * icd;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// source preparation
// ---------------------------------------------------------------------------

// FP64 source record (64 bytes).  Dot form (k_prep_f64_dot):
// a = {y, q}, b = {q y, 0} with q = (-zeta A)/(4 pi).  Tile-local |y - x|^2/2
// form (k_prep_f64_tiles): a = {d, |d|^2}, b = {q y.n, q y.e1, q y.e2, q}
// with d = y - c the offset from the tile centre c, q = (-zeta A)/(2 pi) and
// y written in the tile's orthonormal frame (n, e1, e2; tile_frame).
struct Src64 {
  double4 a, b;
};

// Dot-form records.  Source j is target smap[j]; its area is area[smap[j]].
// Padding records sit at the origin with q = 0 (denominator 1).  ns_dev
// (may be null) holds the source count on the device, at most ns.
__device__ __forceinline__ void prep_f64_dot_block(int blk, Src64* __restrict__ src, Pos3 pos,
                                                   const double* __restrict__ zeta, const double* __restrict__ area,
                                                   const int* __restrict__ smap, int ns, int ns_pad,
                                                   const int* __restrict__ ns_dev) {
  const int j = blk * kBlock + threadIdx.x;
  if (j >= ns_pad) return;
  if (ns_dev) ns = min(ns, *ns_dev);
  Src64 r;
  if (j < ns) {
    const int t = smap[j];
    const double q = (-zeta[t] * area[t]) / (4 * kPi);
    const double y0 = pos.at(t, 0), y1 = pos.at(t, 1), y2 = pos.at(t, 2);
    r.a = make_double4(y0, y1, y2, q);
    r.b = make_double4(q * y0, q * y1, q * y2, 0.0);
  } else {
    r.a = make_double4(0.0, 0.0, 0.0, 0.0);
    r.b = make_double4(0.0, 0.0, 0.0, 0.0);
  }
  src[j] = r;
}

__global__ __launch_bounds__(kBlock) void k_prep_f64_dot(Src64* __restrict__ src, Pos3 pos,
                                                         const double* __restrict__ zeta,
                                                         const double* __restrict__ area,
                                                         const int* __restrict__ smap, int ns, int ns_pad,
                                                         const int* __restrict__ ns_dev) {
  prep_f64_dot_block(blockIdx.x, src, pos, zeta, area, smap, ns, ns_pad, ns_dev);
}

// FP32C source record: a = {dx, dy, dz, q y.n}, b = {q y.e1, q y.e2, q, |d|^2}
// with d = fl32(y - c) the offset from the tile centre c,
// q = fl32((-zeta A)/(2 pi)) and y = c + d written in the tile's orthonormal
// frame (n along c, e1, e2; FP64, rounded once).  The far forms sum
// {a.w, b.x, b.y}/|d|^2, i.e. q y/|d|^2 in frame components (the velocity
// form reads a and b.xy: one 16-byte and one 8-byte LDS read), and the tile's frame
// maps the three sums back in FP64: the component along x that this leaves
// out of q (y - x)/|d|^2 is cancelled by the cross product with x, and the
// large coefficient (along n) is rounded relative to itself, so its error
// stays in proportion to the term's own size after the cross product.
// b.w = |d|^2 (FP32, from the stored offsets) is not read by the pair
// kernels, which form the far denominator (1 - x.c) - x.d and the near one
// from d; it fills the record to 32 bytes.
struct Src32 {
  float4 a, b;
};

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Orthonormal frame of a tile: n = c/|c| (e_z when |c| is tiny), e1 = the
// normalised n x (the axis of n's smallest component), e2 = n x e1.
__device__ __forceinline__ void tile_frame(const double (&c)[4], double (&n)[3], double (&e1)[3], double (&e2)[3]) {
  const double cn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (cn > 1e-3) {
    for (int k = 0; k < 3; ++k) n[k] = c[k] / cn;
  } else {
    n[0] = 0.0;
    n[1] = 0.0;
    n[2] = 1.0;
  }
  const double a0 = fabs(n[0]), a1 = fabs(n[1]), a2 = fabs(n[2]);
  double v[3];
  if (a0 <= a1 && a0 <= a2) {
    v[0] = 0.0, v[1] = n[2], v[2] = -n[1];
  } else if (a1 <= a2) {
    v[0] = -n[2], v[1] = 0.0, v[2] = n[0];
  } else {
    v[0] = n[1], v[1] = -n[0], v[2] = 0.0;
  }
  const double vn = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) e1[k] = v[k] * vn;
  e2[0] = n[1] * e1[2] - n[2] * e1[1];
  e2[1] = n[2] * e1[0] - n[0] * e1[2];
  e2[2] = n[0] * e1[1] - n[1] * e1[0];
}

// One workgroup per tile: tile centre (FP64 mean of its sources), the
// tile's FP64 sum of q, its radius max |d| (rounded up), its frame (at
// ctr[ntiles + 3t + {0,1,2}] = n, e1, e2), and the records.  Padding records
// sit at the origin with q = 0.
__device__ __forceinline__ void prep_f32c_block(int t, int ntiles, Src32* __restrict__ src,
                                                double4* __restrict__ ctr, float* __restrict__ trad,
                                                double* __restrict__ srcq, Pos3 pos, const double* __restrict__ zeta,
                                                const double* __restrict__ area, const int* __restrict__ smap,
                                                int ns) {
  __shared__ double red[5][kTile / 64];
  const int j = t * kTile + threadIdx.x;
  const bool valid = j < ns;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, q = 0.0;
  if (valid) {
    const int i = smap[j];
    p0 = pos.at(i, 0);
    p1 = pos.at(i, 1);
    p2 = pos.at(i, 2);
    q = (-zeta[i] * area[i]) / (2 * kPi);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s0 = wave_sum(p0), s1 = wave_sum(p1), s2 = wave_sum(p2), sq = wave_sum(q);
  if (lane == 0) {
    red[0][wave] = s0;
    red[1][wave] = s1;
    red[2][wave] = s2;
    red[3][wave] = sq;
  }
  __syncthreads();
  const int nvalid = min(kTile, ns - t * kTile);
  double c[4];
  for (int k = 0; k < 4; ++k) c[k] = ((red[k][0] + red[k][1]) + (red[k][2] + red[k][3]));
  const double inv = 1.0 / nvalid;
  c[0] *= inv;
  c[1] *= inv;
  c[2] *= inv;
  // every lane forms the same frame from the same sums
  double n[3], e1[3], e2[3];
  tile_frame(c, n, e1, e2);
  Src32 r;
  double rad = 0.0;
  if (valid) {
    const float dx = (float)(p0 - c[0]), dy = (float)(p1 - c[1]), dz = (float)(p2 - c[2]);
    const float qf = (float)q;
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    // y = c + d as the kernels see it, in frame components
    const double y0 = c[0] + (double)dx, y1 = c[1] + (double)dy, y2 = c[2] + (double)dz, qd = qf;
    r.a = make_float4(dx, dy, dz, (float)(qd * (y0 * n[0] + y1 * n[1] + y2 * n[2])));
    r.b = make_float4((float)(qd * (y0 * e1[0] + y1 * e1[1] + y2 * e1[2])),
                      (float)(qd * (y0 * e2[0] + y1 * e2[1] + y2 * e2[2])), qf, d2);
    rad = sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
  } else {
    // at the origin (offset -c): 1 - x.y = 1 for every target x, so no
    // padding denominator can vanish and spoil a shared reciprocal
    const float dx = (float)(-c[0]), dy = (float)(-c[1]), dz = (float)(-c[2]);
    r.a = make_float4(dx, dy, dz, 0.0f);
    r.b = make_float4(0.0f, 0.0f, 0.0f, fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
  }
  src[j] = r;
  srcq[j] = q;
  rad = wave_max(rad);
  __syncthreads();
  if (lane == 0) red[4][wave] = rad;
  __syncthreads();
  if (threadIdx.x == 0) {
    ctr[t] = make_double4(c[0], c[1], c[2], c[3]);
    double4* fr = ctr + ntiles + 3 * (long)t;
    fr[0] = make_double4(n[0], n[1], n[2], 0.0);
    fr[1] = make_double4(e1[0], e1[1], e1[2], 0.0);
    fr[2] = make_double4(e2[0], e2[1], e2[2], 0.0);
    const double rr = fmax(fmax(red[4][0], red[4][1]), fmax(red[4][2], red[4][3]));
    trad[t] = (float)(rr * (1.0 + 1e-6));
  }
}

__global__ __launch_bounds__(kTile) void k_prep_f32c(Src32* __restrict__ src, double4* __restrict__ ctr,
                                                     float* __restrict__ trad, double* __restrict__ srcq, Pos3 pos,
                                                     const double* __restrict__ zeta,
                                                     const double* __restrict__ area,
                                                     const int* __restrict__ smap, int ns) {
  prep_f32c_block(blockIdx.x, gridDim.x, src, ctr, trad, srcq, pos, zeta, area, smap, ns);
}

// One workgroup per tile: centre (mean of its sources) and radius max |d|
// (rounded up) in ctr[t], frame (at ctr[ntiles + 3t + {0,1,2}] = n, e1, e2)
// and FP64 records; padding records sit at the origin (offset -c) with
// q = 0, where 1 - x.y = 1 for every target x.
__device__ __forceinline__ void prep_f64_tiles_block(int t, int ntiles, Src64* __restrict__ src,
                                                     double4* __restrict__ ctr, Pos3 pos,
                                                     const double* __restrict__ zeta,
                                                     const double* __restrict__ area,
                                                     const int* __restrict__ smap, int ns) {
  __shared__ double red[4][kTile / 64];
  const int j = t * kTile + threadIdx.x;
  const bool valid = j < ns;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, q = 0.0;
  if (valid) {
    const int i = smap[j];
    p0 = pos.at(i, 0);
    p1 = pos.at(i, 1);
    p2 = pos.at(i, 2);
    q = (-zeta[i] * area[i]) / (2 * kPi);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double s0 = wave_sum(p0), s1 = wave_sum(p1), s2 = wave_sum(p2);
  if (lane == 0) {
    red[0][wave] = s0;
    red[1][wave] = s1;
    red[2][wave] = s2;
  }
  __syncthreads();
  const double inv = 1.0 / min(kTile, ns - t * kTile);
  double c[4];
  for (int k = 0; k < 3; ++k) c[k] = ((red[k][0] + red[k][1]) + (red[k][2] + red[k][3])) * inv;
  c[3] = 0.0;
  double n[3], e1[3], e2[3];
  tile_frame(c, n, e1, e2);
  Src64 r;
  double rad = 0.0;
  if (valid) {
    const double dx = p0 - c[0], dy = p1 - c[1], dz = p2 - c[2];
    const double d2 = fma(dx, dx, fma(dy, dy, dz * dz));
    r.a = make_double4(dx, dy, dz, d2);
    r.b = make_double4(q * (p0 * n[0] + p1 * n[1] + p2 * n[2]), q * (p0 * e1[0] + p1 * e1[1] + p2 * e1[2]),
                       q * (p0 * e2[0] + p1 * e2[1] + p2 * e2[2]), q);
    rad = sqrt(d2);
  } else {
    r.a = make_double4(-c[0], -c[1], -c[2], fma(c[0], c[0], fma(c[1], c[1], c[2] * c[2])));
    r.b = make_double4(0.0, 0.0, 0.0, 0.0);
  }
  src[j] = r;
  rad = wave_max(rad);
  __syncthreads();
  if (lane == 0) red[3][wave] = rad;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rr = fmax(fmax(red[3][0], red[3][1]), fmax(red[3][2], red[3][3]));
    ctr[t] = make_double4(c[0], c[1], c[2], rr * (1.0 + 1e-6));
    double4* fr = ctr + ntiles + 3 * (long)t;
    fr[0] = make_double4(n[0], n[1], n[2], 0.0);
    fr[1] = make_double4(e1[0], e1[1], e1[2], 0.0);
    fr[2] = make_double4(e2[0], e2[1], e2[2], 0.0);
  }
}

__global__ __launch_bounds__(kTile) void k_prep_f64_tiles(Src64* __restrict__ src, double4* __restrict__ ctr,
                                                          Pos3 pos, const double* __restrict__ zeta,
                                                          const double* __restrict__ area,
                                                          const int* __restrict__ smap, int ns) {
  prep_f64_tiles_block(blockIdx.x, gridDim.x, src, ctr, pos, zeta, area, smap, ns);
}

// Sum of per-tile q sums in tile order (one workgroup; deterministic).
__global__ __launch_bounds__(kBlock) void k_qsum(const double4* __restrict__ ctr, int ntiles,
                                                 double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (int t = threadIdx.x; t < ntiles; t += kBlock) s += ctr[t].w;
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// pair kernels
// ---------------------------------------------------------------------------

struct PairArgs {
  Pos3 tgt;              // target positions
  const int* self_slot;  // source slot of each target, or -1; may be null
  int nt;                // targets
  int ntiles;            // source tiles
  int tiles_per_split;   // tiles per grid.y slice
  double* part;          // [splits][kNC][nt_pad]
  int nt_pad;
  float near_factor;  // FP32C and FP64 tile form: near-form threshold in tile radii
  // FP64 kernels: the source count on the device (may be null); tiles past
  // it hold only padding records and are skipped
  const int* ns_dev;
};

template <int R>
__device__ __forceinline__ void load_targets(const PairArgs& a, int bx, double (&x)[R][3], int (&self)[R],
                                             int (&idx)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = (bx * R + r) * kBlock + threadIdx.x;
    idx[r] = i;
    const int ii = min(i, a.nt - 1);
    x[r][0] = a.tgt.at(ii, 0);
    x[r][1] = a.tgt.at(ii, 1);
    x[r][2] = a.tgt.at(ii, 2);
    self[r] = (a.self_slot && i < a.nt) ? a.self_slot[ii] : -1;
  }
}

template <int R>
__device__ __forceinline__ bool wave_has_self(const int (&self)[R], int base) {
  int hit = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) hit |= (unsigned)(self[r] - base) < (unsigned)kTile;
  return __ballot(hit) != 0;
}

template <int R>
__device__ __forceinline__ void store_partials(const PairArgs& a, int by, const int (&idx)[R],
                                               const double (&S)[R][kNC]) {
  double* out = a.part + (long)by * kNC * a.nt_pad;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (idx[r] < a.nt) {
#pragma unroll
      for (int c = 0; c < kNC; ++c) out[(long)c * a.nt_pad + idx[r]] = S[r][c];
    }
  }
}

// Natural logarithm of a pair denominator for the stream-function sums, by
// the table-driven reduction of P. T. P. Tang ("Table-driven implementation
// of the logarithm function in IEEE floating-point arithmetic", ACM TOMS 16,
// 1990) with a 129-entry table: x = 2^k m, m in [1, 2); F = 1 + j/128 with j
// the nearest multiple of 1/128 to m - 1 (j = 0..128), f = m - F (exact, in
// [-2^-8, 2^-8]);
//   log x = (k ln 2 + log F) + log1p(f/F),
// log1p(r) by its Taylor series to r^7 (remainder below r^8/8 < 7e-21); at
// j = 128 (x just below a power of two) k ln 2 + log 2 cancels exactly.  The
// table {log F, 1/F} is built per workgroup with the library's log
// (log_table_fill).  About 20 operations per call against some 75 for the
// library's log; measured against long-double logarithms (20M arguments
// over [1e-11, 16)): within 1.5 units in the last place where |log x| > 1,
// and within 1.9e-16 absolute where |log x| <= 1 (near x = 1 from below the
// sum k ln 2 + log F keeps one rounding of ln 2's size).  Zero, negative,
// subnormal and non-finite arguments go to the library (the reference's
// literal 1 - x.y of an off-sphere stage argument can be <= 0).
constexpr int kLogTable = 129;

__device__ __forceinline__ void log_table_fill(double2* tab) {
  if (threadIdx.x < kLogTable) {
    const double F = 1.0 + threadIdx.x * (1.0 / 128);
    tab[threadIdx.x] = make_double2(log(F), 1.0 / F);
  }
}

// The library's log for plog's rare arguments, out of line so that the pair
// loops do not carry a copy of it per pair.
__device__ __attribute__((noinline)) double log_special(double x) { return log(x); }

__device__ __forceinline__ double plog(double x, const double2* __restrict__ tab) {
  // anything but a positive normal finite number (class bits: +normal = 256)
  if (!__builtin_amdgcn_class(x, 256)) return log_special(x);
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned hi = (unsigned)(b >> 32);
  const int k = (int)(hi >> 20) - 1023;
  const int j = (int)(((hi >> 12) & 255) + 1) >> 1;
  const double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
  // F = 1 + j/128 (j = 128: 2) from its bits
  const double F = __longlong_as_double((long long)((unsigned long long)(0x3FF00000u + ((unsigned)j << 13)) << 32));
  const double2 t = tab[j];
  const double r = (m - F) * t.y;
  double p = 1.0 / 7;
  p = fma(p, r, -1.0 / 6);
  p = fma(p, r, 1.0 / 5);
  p = fma(p, r, -1.0 / 4);
  p = fma(p, r, 1.0 / 3);
  p = fma(p, r, -1.0 / 2);
  p = fma(p, r, 1.0);
  return fma(p, r, fma((double)k, kLn2, t.x));
}

// Literal denominator 1 - x.y of src/LpmBVEKernels.hpp:21-44 as three FMAs
// from 1 (each rounding once, as the dot product's own FMAs do before the
// exact subtraction from 1 in the four-operation form).
__device__ __forceinline__ double dot_den(const double (&x)[3], const Src64& s) {
  return fma(-x[2], s.a.z, fma(-x[1], s.a.y, fma(-x[0], s.a.x, 1.0)));
}

// Dot-form pair term (biotSavart and greensFn, src/LpmBVEKernels.hpp:21-44): den = 1 - x.y,
// S += (q y)/den; psi: S3 += q ln(den).  `me` drops the target's own source.
template <bool PSI, bool CHECK>
__device__ __forceinline__ void dot_term(const double (&x)[3], const Src64& s, bool me, double (&S)[kNC],
                                         const double2* lt) {
  double den = dot_den(x, s);
  double q = s.a.w;
  if (CHECK && me) {
    den = 1.0;
    q = 0.0;
  }
  double inv = rcp64(den);
  if (CHECK && me) inv = 0.0;
  S[0] = fma(inv, s.b.x, S[0]);
  S[1] = fma(inv, s.b.y, S[1]);
  S[2] = fma(inv, s.b.z, S[2]);
  if (PSI) S[3] = fma(q, plog(den, lt), S[3]);
}

// dot_term's arithmetic for one target and four sources with one shared
// reciprocal (rcp4_64).  The four pairs belong to one target, so a degenerate
// pair (a target on a source that is not its own, as a divided face's centre
// on its middle child's) leaves every other target's sums as they are.

template <bool PSI>
__device__ __forceinline__ void dot_acc(double inv, double den, const Src64& s, double (&S)[kNC],
                                        const double2* lt) {
  S[0] = fma(inv, s.b.x, S[0]);
  S[1] = fma(inv, s.b.y, S[1]);
  S[2] = fma(inv, s.b.z, S[2]);
  if (PSI) S[3] = fma(s.a.w, plog(den, lt), S[3]);
}

template <bool PSI>
__device__ __forceinline__ void dot_terms4(const double (&x)[3], const Src64& s0, const Src64& s1,
                                           const Src64& s2, const Src64& s3, double (&S)[kNC], const double2* lt) {
  const double a = dot_den(x, s0), b = dot_den(x, s1), c = dot_den(x, s2), d = dot_den(x, s3);
  double ia, ib, ic, id;
  rcp4_64(a, b, c, d, ia, ib, ic, id);
  dot_acc<PSI>(ia, a, s0, S, lt);
  dot_acc<PSI>(ib, b, s1, S, lt);
  dot_acc<PSI>(ic, c, s2, S, lt);
  dot_acc<PSI>(id, d, s3, S, lt);
}

// Far form of the |y - x|^2/2 kernel (tile-local records, see pair_block_f64)
// for one target and four sources with one shared reciprocal.  On the unit
// sphere |y - x|^2/2 = 1 - x.y, and with y = c + d,
//   den = A - x.d,   A = 1 - x.c  (per target and tile),
// three FMAs per pair.  On a tile whose centre lies beyond near_factor tile
// radii of the target, |x.d| <= A (1 + ...) is bounded by the tile radius, so
// the cancellation in den is at most ((f + 1)/(f - 1))^2 (2.8 at f = 4);
// nearer tiles form d = D + dx explicitly.  F[0..2] += q y/den in frame
// components (twice q y/|y - x|^2), and with PSI F[3] += q ln(den).
__device__ __forceinline__ double far_den64(const double (&x)[3], double A, const Src64& s) {
  return fma(-x[2], s.a.z, fma(-x[1], s.a.y, fma(-x[0], s.a.x, A)));
}

template <bool PSI>
__device__ __forceinline__ void far_acc64(double inv, double den, const Src64& s, double (&F)[kNC],
                                          const double2* lt) {
  F[0] = fma(inv, s.b.x, F[0]);
  F[1] = fma(inv, s.b.y, F[1]);
  F[2] = fma(inv, s.b.z, F[2]);
  if (PSI) F[3] = fma(s.b.w, plog(den, lt), F[3]);
}

template <bool PSI>
__device__ __forceinline__ void far_terms4_64(const double (&x)[3], double A, const Src64& s0, const Src64& s1,
                                              const Src64& s2, const Src64& s3, double (&F)[kNC],
                                              const double2* lt) {
  const double a = far_den64(x, A, s0), b = far_den64(x, A, s1);
  const double c = far_den64(x, A, s2), d = far_den64(x, A, s3);
  double ia, ib, ic, id;
  rcp4_64(a, b, c, d, ia, ib, ic, id);
  far_acc64<PSI>(ia, a, s0, F, lt);
  far_acc64<PSI>(ib, b, s1, F, lt);
  far_acc64<PSI>(ic, c, s2, F, lt);
  far_acc64<PSI>(id, d, s3, F, lt);
}

// Near form: d = D + dx formed explicitly (D = c - x), S += q d/|d|^2 in
// global components, with PSI S3 += q ln(|d|^2/2).  `me` drops the target's
// own source.
template <bool PSI, bool CHECK>
__device__ __forceinline__ void near_term64(const double (&D)[3], const Src64& s, bool me, double (&S)[kNC],
                                            const double2* lt) {
  const double e0 = D[0] + s.a.x, e1 = D[1] + s.a.y, e2 = D[2] + s.a.z;
  double r2 = fma(e0, e0, fma(e1, e1, e2 * e2));
  double q = s.b.w;
  if (CHECK && me) {
    r2 = 1.0;
    q = 0.0;
  }
  const double f = q * rcp64(r2);
  S[0] = fma(f, e0, S[0]);
  S[1] = fma(f, e1, S[1]);
  S[2] = fma(f, e2, S[2]);
  if (PSI) S[3] = fma(q, plog(0.5 * r2, lt), S[3]);
}

// FP64 pair sums of one target block (R targets per lane) over one source
// split.  LDS: source tiles are staged through LDS (double-buffered, one
// coalesced 16 KB copy per tile, tile t+1 loaded while tile t is summed);
// otherwise every lane reads the same record, which the compiler issues as
// scalar loads.
//   FORM_DOT: dot_term for every pair.
//   FORM_CHORD (|y - x|^2/2 form, tile-local records): with D = c - x per
//     target and tile, tiles holding a target's own source or lying within
//     near_factor tile radii of a target of the wave form d = D + dx
//     explicitly (the self pair dropped) and sum S += q d/|d|^2; other tiles
//     form den = (1 - x.c) - x.dx and sum F = q y/den in the tile frame,
//     S += (n F0 + e1 F1 + e2 F2)/2.  The two differ by the component
//     -x sum q/|d|^2, which the finalising cross product with x cancels.
template <int R, int FORM, bool PSI, bool LDS>
__device__ __forceinline__ void pair_block_f64(const Src64* __restrict__ src, const double4* __restrict__ ctr,
                                               const PairArgs& a, int bx, int by, Src64 (*sh)[kTile],
                                               const double2* lt = nullptr) {
  double x[R][3];
  int self[R], idx[R];
  load_targets<R>(a, bx, x, self, idx);
  double S[R][kNC];
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < kNC; ++c) S[r][c] = 0.0;
  const int ntiles = a.ns_dev ? min(a.ntiles, (*a.ns_dev + kTile - 1) / kTile) : a.ntiles;
  const int t0 = by * a.tiles_per_split;
  const int t1 = min(ntiles, t0 + a.tiles_per_split);
  if constexpr (LDS) {
    if (t0 < t1) sh[0][threadIdx.x] = src[(long)t0 * kTile + threadIdx.x];
    __syncthreads();
  }
  for (int t = t0; t < t1; ++t) {
    const int base = t * kTile;
    const Src64* ts;
    if constexpr (LDS) {
      const int buf = (t - t0) & 1;
      if (t + 1 < t1) sh[buf ^ 1][threadIdx.x] = src[(long)(t + 1) * kTile + threadIdx.x];
      ts = sh[buf];
    } else {
      ts = src + base;
    }
    const bool has_self = wave_has_self<R>(self, base);
    if constexpr (FORM == FORM_DOT) {
      if (!has_self) {
        for (int j = 0; j < kTile; j += 4) {
          const Src64 s0 = ts[j], s1 = ts[j + 1], s2 = ts[j + 2], s3 = ts[j + 3];
#pragma unroll
          for (int r = 0; r < R; ++r) dot_terms4<PSI>(x[r], s0, s1, s2, s3, S[r], lt);
        }
      } else {
        for (int j = 0; j < kTile; ++j) {
          const Src64 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) dot_term<PSI, true>(x[r], sj, base + j == self[r], S[r], lt);
        }
      }
    } else {
      // c.w: the tile radius (k_prep_f64_tiles)
      const double4 c = ctr[t];
      const double rn = (double)a.near_factor * c.w;
      double D[R][3];
      int near = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        D[r][0] = c.x - x[r][0];
        D[r][1] = c.y - x[r][1];
        D[r][2] = c.z - x[r][2];
        near |= fma(D[r][0], D[r][0], fma(D[r][1], D[r][1], D[r][2] * D[r][2])) < rn * rn;
      }
      if (has_self) {
        for (int j = 0; j < kTile; ++j) {
          const Src64 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) near_term64<PSI, true>(D[r], sj, base + j == self[r], S[r], lt);
        }
      } else if (__ballot(near) != 0) {
#pragma unroll 2
        for (int j = 0; j < kTile; ++j) {
          const Src64 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) near_term64<PSI, false>(D[r], sj, false, S[r], lt);
        }
      } else {
        double A[R], F[R][kNC];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          A[r] = 1.0 - (x[r][0] * c.x + x[r][1] * c.y + x[r][2] * c.z);
          for (int k = 0; k < kNC; ++k) F[r][k] = 0.0;
        }
        for (int j = 0; j < kTile; j += 4) {
          const Src64 s0 = ts[j], s1 = ts[j + 1], s2 = ts[j + 2], s3 = ts[j + 3];
#pragma unroll
          for (int r = 0; r < R; ++r) far_terms4_64<PSI>(x[r], A[r], s0, s1, s2, s3, F[r], lt);
        }
        const double4* fr = ctr + a.ntiles + 3 * (long)t;
        const double4 fn = fr[0], f1 = fr[1], f2 = fr[2];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double p0 = 0.5 * F[r][0], p1 = 0.5 * F[r][1], p2 = 0.5 * F[r][2];
          S[r][0] += fma(fn.x, p0, fma(f1.x, p1, f2.x * p2));
          S[r][1] += fma(fn.y, p0, fma(f1.y, p1, f2.y * p2));
          S[r][2] += fma(fn.z, p0, fma(f1.z, p1, f2.z * p2));
          S[r][3] += F[r][3];
        }
      }
    }
    if constexpr (LDS) __syncthreads();
  }
  store_partials<R>(a, by, idx, S);
}

template <int R, int FORM, bool PSI, bool LDS>
__global__ __launch_bounds__(kBlock) void k_pair_f64(const Src64* __restrict__ src, const double4* __restrict__ ctr,
                                                     PairArgs a) {
  // with PSI, the logarithm table of plog (2 KB, filled before the first
  // barrier of the source loop)
  __shared__ double2 lt[PSI ? kLogTable : 1];
  if constexpr (PSI) {
    log_table_fill(lt);
    __syncthreads();
  }
  if constexpr (LDS) {
    __shared__ Src64 sh[2][kTile];
    pair_block_f64<R, FORM, PSI, true>(src, ctr, a, blockIdx.x, blockIdx.y, sh, lt);
  } else {
    pair_block_f64<R, FORM, PSI, false>(src, ctr, a, blockIdx.x, blockIdx.y, nullptr, lt);
  }
}

// Near form: d = D + dx formed explicitly, f = q / |d|^2, S += f d.
// P accumulates q log2(|d|^2).  `me` drops the target's own source.
template <bool PSI, bool CHECK>
__device__ __forceinline__ void pair32_near(const float (&D)[3], const Src32& s, bool me, float (&T)[kNC]) {
  const float e0 = D[0] + s.a.x, e1 = D[1] + s.a.y, e2 = D[2] + s.a.z;
  float r2 = fmaf(e0, e0, fmaf(e1, e1, e2 * e2));
  float q = s.b.z;
  if (CHECK && me) {
    r2 = 1.0f;
    q = 0.0f;
  }
  const float f = q * __builtin_amdgcn_rcpf(r2);
  T[0] = fmaf(f, e0, T[0]);
  T[1] = fmaf(f, e1, T[1]);
  T[2] = fmaf(f, e2, T[2]);
  if (PSI) T[3] = fmaf(q, __builtin_amdgcn_logf(r2), T[3]);
}

// S += scale (n F0 + e1 F1 + e2 F2) in FP64 with the tile frame fr[0..2].
__device__ __forceinline__ void add_frame(double (&S)[kNC], const double4* fr, float F0, float F1, float F2,
                                          double scale) {
  const double4 n = fr[0], e1 = fr[1], e2 = fr[2];
  const double p0 = scale * F0, p1 = scale * F1, p2 = scale * F2;
  S[0] += fma(n.x, p0, fma(e1.x, p1, e2.x * p2));
  S[1] += fma(n.y, p0, fma(e1.y, p1, e2.y * p2));
  S[2] += fma(n.z, p0, fma(e1.z, p1, e2.z * p2));
}

// Far form of the step kernel (velocity only).  On the unit sphere
// |y - x|^2 / 2 = 1 - x.y; in tile-local coordinates y = c + dx this is
//   den = (1 - x.c) - x.dx,
// where A = 1 - x.c is formed per target and tile in FP64 (no cancellation
// there) and rounded once, and x.dx is small: on a tile beyond 4 tile radii
// the cancellation in den is at most ~4, as for |D|^2 + |dx|^2 + 2 D.dx.
// 3 FMAs per pair; the three sums of b.{y,z,w}/den, halved and mapped by
// the tile frame, give the tile's sum of q d/|d|^2 up to a component along
// x (see Src32).
__device__ __forceinline__ float far_den(const float (&NX)[3], float A, const Src32& s) {
  float v = fmaf(NX[0], s.a.x, A);
  v = fmaf(NX[1], s.a.y, v);
  return fmaf(NX[2], s.a.z, v);
}

__device__ __forceinline__ void far_acc(float inv, const Src32& s, float (&F)[3]) {
  F[0] = fmaf(inv, s.a.w, F[0]);
  F[1] = fmaf(inv, s.b.x, F[1]);
  F[2] = fmaf(inv, s.b.y, F[2]);
}

// Four pairs sharing one reciprocal: with a, b, c, d the four denominators,
// p = 1/(ab cd), 1/a = b (cd p), 1/b = a (cd p), 1/c = d (ab p),
// 1/d = c (ab p).  9 multiplies and one v_rcp_f32 replace four v_rcp_f32,
// which issue at a quarter of the FMA rate and do not overlap it.  Far-form
// denominators lie in [((near_factor - 1) r_tile)^2 / 2, 2] (1 for padding
// records), so the product of four stays in the FP32 normal range; each
// reciprocal carries about 5 roundings instead of 1.
__device__ __forceinline__ void pair32_far4(const float (&NX)[3], float A, const Src32& s0, const Src32& s1,
                                            const Src32& s2, const Src32& s3, float (&F)[3]) {
  const float a = far_den(NX, A, s0), b = far_den(NX, A, s1);
  const float c = far_den(NX, A, s2), d = far_den(NX, A, s3);
  const float ab = a * b, cd = c * d;
  const float p = __builtin_amdgcn_rcpf(ab * cd);
  const float iab = cd * p, icd = ab * p;
  far_acc(b * iab, s0, F);
  far_acc(a * iab, s1, F);
  far_acc(d * icd, s2, F);
  far_acc(c * icd, s3, F);
}

// The same for two targets (A, B) and two sources.
__device__ __forceinline__ void pair32_far22(const float (&NXA)[3], float AA, const float (&NXB)[3], float AB,
                                             const Src32& s0, const Src32& s1, float (&FA)[3],
                                             float (&FB)[3]) {
  const float a = far_den(NXA, AA, s0), b = far_den(NXA, AA, s1);
  const float c = far_den(NXB, AB, s0), d = far_den(NXB, AB, s1);
  const float ab = a * b, cd = c * d;
  const float p = __builtin_amdgcn_rcpf(ab * cd);
  const float iab = cd * p, icd = ab * p;
  far_acc(b * iab, s0, FA);
  far_acc(a * iab, s1, FA);
  far_acc(d * icd, s0, FB);
  far_acc(c * icd, s1, FB);
}

__device__ __forceinline__ Src32 load_src(const Src32* p) { return *p; }

// Far-form sums of one tile for R targets (velocity only), one reciprocal per
// four pairs: target pairs x source pairs, and four sources for the last
// target when R is odd.
template <int R>
__device__ __forceinline__ void far_tile(const Src32* ts, const float (&NX)[R][3], const float (&A)[R],
                                         float (&F)[R][3]) {
  if constexpr (R == 1) {
    for (int j = 0; j < kTile; j += 4) {
      const Src32 s0 = load_src(ts + j), s1 = load_src(ts + j + 1), s2 = load_src(ts + j + 2),
                  s3 = load_src(ts + j + 3);
      pair32_far4(NX[0], A[0], s0, s1, s2, s3, F[0]);
    }
  } else if constexpr (R % 2 == 0) {
#pragma unroll 2
    for (int j = 0; j < kTile; j += 2) {
      const Src32 s0 = load_src(ts + j), s1 = load_src(ts + j + 1);
#pragma unroll
      for (int r = 0; r < R; r += 2) pair32_far22(NX[r], A[r], NX[r + 1], A[r + 1], s0, s1, F[r], F[r + 1]);
    }
  } else {
    for (int j = 0; j < kTile; j += 4) {
      const Src32 s0 = load_src(ts + j), s1 = load_src(ts + j + 1), s2 = load_src(ts + j + 2),
                  s3 = load_src(ts + j + 3);
#pragma unroll
      for (int r = 0; r + 1 < R; r += 2) {
        pair32_far22(NX[r], A[r], NX[r + 1], A[r + 1], s0, s1, F[r], F[r + 1]);
        pair32_far22(NX[r], A[r], NX[r + 1], A[r + 1], s2, s3, F[r], F[r + 1]);
      }
      pair32_far4(NX[R - 1], A[R - 1], s0, s1, s2, s3, F[R - 1]);
    }
  }
}

// A tile's far-form contribution for R targets (velocity-only kernels),
// added to S: A = 1 - x.c in FP64, the sums, then the frame in FP64.
template <int R>
__device__ __forceinline__ void far_tile_sum(const Src32* ts, const double (&x)[R][3], const double4& c,
                                             const double4* fr, const float (&NX)[R][3], double (&S)[R][kNC]) {
  float A[R], F[R][3];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    A[r] = (float)(1.0 - (x[r][0] * c.x + x[r][1] * c.y + x[r][2] * c.z));
    for (int k = 0; k < 3; ++k) F[r][k] = 0.0f;
  }
  far_tile<R>(ts, NX, A, F);
#pragma unroll
  for (int r = 0; r < R; ++r) add_frame(S[r], fr, F[r][0], F[r][1], F[r][2], 0.5);
}

// A tile's far-form contribution with psi for R targets, added to S: the
// velocity kernel's denominators den = (1 - x.c) - x.d = |d|^2/2 and
// reciprocals (one per four sources of a target; far denominators are
// bounded below, padding ones are 1), and P = sum q log2(den).  The
// stream-function sums use q log2 |d|^2 = q log2(den) + q, so the tile adds
// P plus its FP64 sum of q (ctr[t].w).
template <int R>
__device__ __forceinline__ void far_tile_psi(const Src32* ts, const double (&x)[R][3], const double4& c,
                                             const double4* fr, const float (&NX)[R][3], double (&S)[R][kNC]) {
  float A[R], F[R][3], P[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    A[r] = (float)(1.0 - (x[r][0] * c.x + x[r][1] * c.y + x[r][2] * c.z));
    for (int k = 0; k < 3; ++k) F[r][k] = 0.0f;
    P[r] = 0.0f;
  }
  for (int j = 0; j < kTile; j += 4) {
    const Src32 s0 = ts[j], s1 = ts[j + 1], s2 = ts[j + 2], s3 = ts[j + 3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float a = far_den(NX[r], A[r], s0), b = far_den(NX[r], A[r], s1);
      const float cc = far_den(NX[r], A[r], s2), d = far_den(NX[r], A[r], s3);
      const float ab = a * b, cd = cc * d;
      const float p = __builtin_amdgcn_rcpf(ab * cd);
      const float iab = cd * p, icd = ab * p;
      far_acc(b * iab, s0, F[r]);
      far_acc(a * iab, s1, F[r]);
      far_acc(d * icd, s2, F[r]);
      far_acc(cc * icd, s3, F[r]);
      P[r] = fmaf(s0.b.z, __builtin_amdgcn_logf(a), P[r]);
      P[r] = fmaf(s1.b.z, __builtin_amdgcn_logf(b), P[r]);
      P[r] = fmaf(s2.b.z, __builtin_amdgcn_logf(cc), P[r]);
      P[r] = fmaf(s3.b.z, __builtin_amdgcn_logf(d), P[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    add_frame(S[r], fr, F[r][0], F[r][1], F[r][2], 0.5);
    S[r][3] += (double)P[r] + c.w;
  }
}

// Tiles whose centre lies within `near_factor` tile radii of a target of
// the wave use the near form, and so does the tile holding a target's own
// source; the far form's cancellation in |d|^2 is bounded by
// ((near_factor + 1) / (near_factor - 1))^2.  Default 4.
constexpr float kNearDefault = 4.0f;

template <int R, bool PSI>
__global__ __launch_bounds__(kBlock) void k_pair_f32c(const Src32* __restrict__ src,
                                                      const double4* __restrict__ ctr,
                                                      const float* __restrict__ trad, PairArgs a) {
  double x[R][3];
  int self[R], idx[R];
  load_targets<R>(a, blockIdx.x, x, self, idx);
  float NX[R][3];  // -x in FP32 (far form of the velocity-only kernels)
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int k = 0; k < 3; ++k) NX[r][k] = -(float)x[r][k];
  double S[R][kNC];
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < kNC; ++c) S[r][c] = 0.0;
  const int t0 = blockIdx.y * a.tiles_per_split;
  const int t1 = min(a.ntiles, t0 + a.tiles_per_split);
  for (int t = t0; t < t1; ++t) {
    const int base = t * kTile;
    const double4 c = ctr[t];
    const float rn = a.near_factor * trad[t];
    float D[R][3], E[R];
    int near = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      D[r][0] = (float)(c.x - x[r][0]);
      D[r][1] = (float)(c.y - x[r][1]);
      D[r][2] = (float)(c.z - x[r][2]);
      E[r] = fmaf(D[r][0], D[r][0], fmaf(D[r][1], D[r][1], D[r][2] * D[r][2]));
      near |= E[r] < rn * rn;
    }
    const Src32* ts = src + base;
    const double4* fr = ctr + a.ntiles + 3 * (long)t;
    const bool has_self = wave_has_self<R>(self, base);
    if (has_self || __ballot(near) != 0) {
      float T[R][kNC];
#pragma unroll
      for (int r = 0; r < R; ++r)
        for (int k = 0; k < kNC; ++k) T[r][k] = 0.0f;
      if (!has_self) {
#pragma unroll 4
        for (int j = 0; j < kTile; ++j) {
          const Src32 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) pair32_near<PSI, false>(D[r], sj, false, T[r]);
        }
      } else {
        for (int j = 0; j < kTile; ++j) {
          const Src32 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) pair32_near<PSI, true>(D[r], sj, base + j == self[r], T[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        for (int k = 0; k < (PSI ? kNC : 3); ++k) S[r][k] += (double)T[r][k];
    } else if constexpr (PSI) {
      far_tile_psi<R>(ts, x, c, fr, NX, S);
    } else {
      far_tile_sum<R>(ts, x, c, fr, NX, S);
    }
  }
  store_partials<R>(a, blockIdx.y, idx, S);
}

// Same arithmetic as k_pair_f32c with the source tiles staged through LDS:
// the workgroup copies tile t+1 (coalesced, 8 KB) into the second buffer as
// it starts tile t, and waves read records with uniform-address LDS loads
// into VGPRs (LDS returns in order, so the compiler prefetches the next
// records while the current ones are in use).
// With xs (R >= 3), each lane parks its targets' FP64 positions in its own
// LDS slots ([3R][kBlock], 8 KB per target per lane) and reloads them once
// per tile, which keeps the register count at four waves per SIMD.
template <int R, bool PSI>
__device__ __forceinline__ void pair_block_f32c_lds(const Src32* __restrict__ src, const double4* __restrict__ ctr,
                                                    const float* __restrict__ trad, const PairArgs& a, int bx,
                                                    int by, Src32 (*sh)[kTile], double* xs) {
  constexpr bool kXL = R >= 3;
  double x[R][3];
  int self[R], idx[R];
  load_targets<R>(a, bx, x, self, idx);
  if constexpr (kXL) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      for (int k = 0; k < 3; ++k) xs[(3 * r + k) * kBlock + threadIdx.x] = x[r][k];
  }
  double S[R][kNC];
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < kNC; ++c) S[r][c] = 0.0;
  const int t0 = by * a.tiles_per_split;
  const int t1 = min(a.ntiles, t0 + a.tiles_per_split);
  if (t0 < t1) sh[0][threadIdx.x] = src[(long)t0 * kTile + threadIdx.x];
  __syncthreads();
  for (int t = t0; t < t1; ++t) {
    const int buf = (t - t0) & 1;
    // Stage tile t+1 into the other buffer.  Its last readers (tile t-1)
    // finished before the barrier that ended the previous iteration.
    if (t + 1 < t1) sh[buf ^ 1][threadIdx.x] = src[(long)(t + 1) * kTile + threadIdx.x];
    const int base = t * kTile;
    const double4 c = ctr[t];
    const float rn = a.near_factor * trad[t];
    if constexpr (kXL) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        for (int k = 0; k < 3; ++k) x[r][k] = xs[(3 * r + k) * kBlock + threadIdx.x];
    }
    float D[R][3], E[R], NX[R][3];  // NX = -x in FP32 (far form of the velocity-only kernels)
    int near = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      D[r][0] = (float)(c.x - x[r][0]);
      D[r][1] = (float)(c.y - x[r][1]);
      D[r][2] = (float)(c.z - x[r][2]);
      E[r] = fmaf(D[r][0], D[r][0], fmaf(D[r][1], D[r][1], D[r][2] * D[r][2]));
      near |= E[r] < rn * rn;
      for (int k = 0; k < 3; ++k) NX[r][k] = -(float)x[r][k];
    }
    const Src32* ts = sh[buf];
    const double4* fr = ctr + a.ntiles + 3 * (long)t;
    const bool has_self = wave_has_self<R>(self, base);
    if (has_self || __ballot(near) != 0) {
      float T[R][kNC];
#pragma unroll
      for (int r = 0; r < R; ++r)
        for (int k = 0; k < kNC; ++k) T[r][k] = 0.0f;
      if (!has_self) {
#pragma unroll 4
        for (int j = 0; j < kTile; ++j) {
          const Src32 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) pair32_near<PSI, false>(D[r], sj, false, T[r]);
        }
      } else {
        for (int j = 0; j < kTile; ++j) {
          const Src32 sj = ts[j];
#pragma unroll
          for (int r = 0; r < R; ++r) pair32_near<PSI, true>(D[r], sj, base + j == self[r], T[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
        for (int k = 0; k < (PSI ? kNC : 3); ++k) S[r][k] += (double)T[r][k];
    } else if constexpr (PSI) {
      far_tile_psi<R>(ts, x, c, fr, NX, S);
    } else {
      far_tile_sum<R>(ts, x, c, fr, NX, S);
    }
    __syncthreads();
  }
  store_partials<R>(a, by, idx, S);
}

// The explicit waves-per-EU range (at least four, the occupancy the R = 4
// kernel has anyway: 109 VGPRs) changes only the register assignment of the
// far loop -- same instructions in the same order -- and that assignment
// runs the level-8 launch 1.3 % faster (404.4 against 409.6 ms, one box,
// alternating processes; the R = 1, 2 kernels are unchanged in speed).
template <int R, bool PSI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pair_f32c_lds(
    const Src32* __restrict__ src, const double4* __restrict__ ctr, const float* __restrict__ trad, PairArgs a) {
  __shared__ Src32 sh[2][kTile];
  if constexpr (R >= 3) {
    __shared__ double xs[3 * R * kBlock];
    pair_block_f32c_lds<R, PSI>(src, ctr, trad, a, blockIdx.x, blockIdx.y, sh, xs);
  } else {
    pair_block_f32c_lds<R, PSI>(src, ctr, trad, a, blockIdx.x, blockIdx.y, sh, nullptr);
  }
}

// ---------------------------------------------------------------------------
// ensemble: one persistent launch over a flattened work list
// ---------------------------------------------------------------------------

// Per-member arguments of a pair evaluation.
struct EnsMember {
  PairArgs a;
  const Src32* src32;
  const double4* ctr;
  const float* trad;
  const Src64* src64;
};

// One work item: target block bx and source split by of member m.  Items
// are processed in list order by whichever workgroup takes them next; each
// item writes its own partial slots, so the result does not depend on which
// workgroup ran it or on the other members in the list.
struct EnsItem {
  int member, bx, by;
};

__global__ void k_reset_counter(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// Completion marker: lane 0 writes `seq` to host-visible (pinned) memory after
// everything queued before it on the stream.
__global__ void k_mark(volatile unsigned long long* flag, unsigned long long seq) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    *flag = seq;
  }
}

__device__ __forceinline__ int next_item(int* counter, int* slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
  __syncthreads();
  const int it = *slot;
  __syncthreads();
  return it;
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_pair_f32c_ens(const EnsMember* __restrict__ mem,
                                                          const EnsItem* __restrict__ items, int nitems,
                                                          int* counter) {
  // With R = 4 the two tile buffers, the parked positions and the item slot
  // take 40,968 bytes, just over a quarter of the CU's LDS: three workgroups
  // per CU, where the compiler gives the loop 141 VGPRs.  Measured on the
  // 64-member batching configuration, that runs 5 % faster than four
  // workgroups per CU (the slot moved into a tile buffer: 118 VGPRs).
  __shared__ Src32 sh[2][kTile];
  __shared__ double xs[R >= 3 ? 3 * R * kBlock : 1];
  __shared__ int slot;
  for (;;) {
    const int it = next_item(counter, &slot);
    if (it >= nitems) break;
    const EnsItem w = items[it];
    const EnsMember& m = mem[w.member];
    pair_block_f32c_lds<R, false>(m.src32, m.ctr, m.trad, m.a, w.bx, w.by, sh, xs);
  }
}

template <int R, int FORM>
__global__ __launch_bounds__(kBlock) void k_pair_f64_ens(const EnsMember* __restrict__ mem,
                                                         const EnsItem* __restrict__ items, int nitems,
                                                         int* counter) {
  __shared__ Src64 sh[2][kTile];
  __shared__ int slot;
  for (;;) {
    const int it = next_item(counter, &slot);
    if (it >= nitems) break;
    const EnsItem w = items[it];
    const EnsMember& m = mem[w.member];
    pair_block_f64<R, FORM, false, true>(m.src64, m.ctr, m.a, w.bx, w.by, sh);
  }
}

// ---------------------------------------------------------------------------
// finalize: sum split partials in order, u = x x S, stream function
// ---------------------------------------------------------------------------
enum FinMode { FIN_DOT = 0, FIN_CHORD = 1, FIN_F32C = 2 };

struct FinArgs {
  const double* part;
  int splits, nt, nt_pad;
  Pos3 tgt;
  MutPos3 u;             // velocity output
  double* psi;           // may be null
  long psi_stride;
  const double* qtot;    // FP32C: total q over sources
  const double* srcq;    // FP32C: FP64 q per source slot
  const int* self_slot;  // may be null
};

template <int MODE>
__device__ __forceinline__ void finalize_at(const FinArgs& a, int i) {
  double S[kNC] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < a.splits; ++k)
    for (int c = 0; c < kNC; ++c) S[c] += a.part[((long)k * kNC + c) * a.nt_pad + i];
  const double x0 = a.tgt.at(i, 0), x1 = a.tgt.at(i, 1), x2 = a.tgt.at(i, 2);
  a.u.at(i, 0) = x1 * S[2] - x2 * S[1];
  a.u.at(i, 1) = x2 * S[0] - x0 * S[2];
  a.u.at(i, 2) = x0 * S[1] - x1 * S[0];
  if (a.psi) {
    double p;
    if (MODE == FIN_DOT) {
      p = S[3];
    } else if (MODE == FIN_CHORD) {
      p = 0.5 * S[3];
    } else {
      // psi = sum_{j != i} (q_j / 2) (ln |d|^2 - ln 2)
      double qx = *a.qtot;
      const int s = a.self_slot ? a.self_slot[i] : -1;
      if (s >= 0) qx -= a.srcq[s];
      p = 0.5 * kLn2 * (S[3] - qx);
    }
    a.psi[i * a.psi_stride] = p;
  }
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_finalize(FinArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < a.nt) finalize_at<MODE>(a, i);
}

// ---------------------------------------------------------------------------
// RK4 stage arithmetic (src/LpmBVERK4_Impl.hpp:11-121)
// ---------------------------------------------------------------------------
struct StageArgs {
  int nt, nv, stage;  // stage 1..4
  int lagr;
  double dt, omega;
  const double *u, *v, *w;  // velocity at the stage argument
  double *x, *y, *z, *zeta;  // state (updated at stage 4)
  const double* absv;        // conserved absolute vorticity (lagr)
  double* kx[4];             // position increments, SoA triples [3][nt]
  double* kz[4];             // vorticity increments (ref)
  double *wx, *wy, *wz, *wzeta;  // next stage argument
};

__device__ __forceinline__ void normalize3(double& a, double& b, double& c) {
  const double s = 1.0 / sqrt(a * a + b * b + c * c);
  a *= s;
  b *= s;
  c *= s;
}

__device__ __forceinline__ void stage_at(const StageArgs& a, int i) {
  const int s = a.stage - 1;
  const long nt = a.nt;
  // k_s = dt * u; dzeta_s = -2 Omega w dt (BVEVorticityTendency, src/LpmBVEKernels.hpp:317-319)
  const double k0 = a.dt * a.u[i], k1 = a.dt * a.v[i], k2 = a.dt * a.w[i];
  a.kx[s][i] = k0;
  a.kx[s][nt + i] = k1;
  a.kx[s][2 * nt + i] = k2;
  const double kzs = -2.0 * a.omega * a.w[i] * a.dt;
  if (!a.lagr) a.kz[s][i] = kzs;
  if (a.stage < 4) {
    const double c = a.stage == 3 ? 1.0 : 0.5;
    double wx = a.x[i] + c * k0, wy = a.y[i] + c * k1, wz = a.z[i] + c * k2;
    double wzeta;
    if (a.lagr) {
      normalize3(wx, wy, wz);
      wzeta = a.absv[i] - 2.0 * a.omega * wz;
    } else {
      wzeta = a.zeta[i] + c * kzs;
    }
    a.wx[i] = wx;
    a.wy[i] = wy;
    a.wz[i] = wz;
    a.wzeta[i] = wzeta;
    return;
  }
  // BVERK4Update: x += (k1 + k4)/6 + (k2 + k3)/3
  const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  double* X[3] = {a.x, a.y, a.z};
  double nx[3];
  for (int c = 0; c < 3; ++c) {
    const long o = c * nt + i;
    nx[c] = X[c][i] + (sixth * (a.kx[0][o] + a.kx[3][o]) + third * (a.kx[1][o] + a.kx[2][o]));
  }
  double nz;
  if (a.lagr) {
    normalize3(nx[0], nx[1], nx[2]);
    nz = a.absv[i] - 2.0 * a.omega * nx[2];
  } else {
    // faces combine vorticity stages (1, 2, 4, 4) (src/LpmBVERK4_Impl.hpp:120-121)
    const double k3z = i < a.nv ? a.kz[2][i] : a.kz[3][i];
    nz = a.zeta[i] + (sixth * (a.kz[0][i] + a.kz[3][i]) + third * (a.kz[1][i] + k3z));
  }
  for (int c = 0; c < 3; ++c) X[c][i] = nx[c];
  a.zeta[i] = nz;
}

__global__ __launch_bounds__(kBlock) void k_stage(StageArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < a.nt) stage_at(a, i);
}

// ---------------------------------------------------------------------------
// ensemble-wide stage, prep and finalize: one launch for all members
// ---------------------------------------------------------------------------

// Per-member descriptors of one evaluation (the member's own buffers): the
// source records' inputs and outputs, and the finalize arguments.
struct EnsEval {
  Src32* src32;
  Src64* src64;
  double4* ctr;
  float* trad;
  double* srcq;
  Pos3 srcpos;
  const double* zeta;
  const double* area;
  const int* smap;
  int ns, ntiles;
  FinArgs fin;
};

enum PrepKind { PREP_F32C = 0, PREP_F64_TILES = 1, PREP_F64_DOT = 2 };

// One workgroup per (member, source tile) of the list `blk`: the member's
// records exactly as its own prep kernel writes them.
template <int KIND>
__global__ __launch_bounds__(kTile) void k_prep_ens(const EnsEval* __restrict__ ev, const int2* __restrict__ blk) {
  const int2 b = blk[blockIdx.x];
  const EnsEval& m = ev[b.x];
  if constexpr (KIND == PREP_F32C)
    prep_f32c_block(b.y, m.ntiles, m.src32, m.ctr, m.trad, m.srcq, m.srcpos, m.zeta, m.area, m.smap, m.ns);
  else if constexpr (KIND == PREP_F64_TILES)
    prep_f64_tiles_block(b.y, m.ntiles, m.src64, m.ctr, m.srcpos, m.zeta, m.area, m.smap, m.ns);
  else
    prep_f64_dot_block(b.y, m.src64, m.srcpos, m.zeta, m.area, m.smap, m.ns, m.ntiles * kTile, nullptr);
}

// One workgroup per (member, target block) of the list `blk`.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_finalize_ens(const EnsEval* __restrict__ ev, const int2* __restrict__ blk) {
  const int2 b = blk[blockIdx.x];
  const FinArgs a = ev[b.x].fin;
  const int i = b.y * kBlock + threadIdx.x;
  if (i < a.nt) finalize_at<MODE>(a, i);
}

// RK4 stage `stage` of every member: one workgroup per (member, target
// block) of the list `blk`; st[m] holds the member's arrays.
__global__ __launch_bounds__(kBlock) void k_stage_ens(const StageArgs* __restrict__ st, const int2* __restrict__ blk,
                                                      int stage, int lagr, double dt, double omega) {
  const int2 b = blk[blockIdx.x];
  StageArgs a = st[b.x];
  a.stage = stage;
  a.lagr = lagr;
  a.dt = dt;
  a.omega = omega;
  const int i = b.y * kBlock + threadIdx.x;
  if (i < a.nt) stage_at(a, i);
}

// absv = zeta + 2 Omega z
__global__ __launch_bounds__(kBlock) void k_absv(int nt, double omega, const double* __restrict__ zeta,
                                                 const double* __restrict__ z, double* __restrict__ absv) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nt) absv[i] = zeta[i] + 2.0 * omega * z[i];
}

// sum over sources of zeta * area, in slot order per lane then a fixed tree
__global__ __launch_bounds__(kBlock) void k_circulation(const double* __restrict__ zeta,
                                                        const double* __restrict__ area,
                                                        const int* __restrict__ smap, int ns,
                                                        double* __restrict__ out) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (int j = threadIdx.x; j < ns; j += kBlock) {
    const int t = smap[j];
    s += zeta[t] * area[t];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// Deterministic single-workgroup stream compaction of unmasked faces:
// smap[slot] = offset + face, self_slot[face] = slot (or -1).
__global__ __launch_bounds__(1024) void k_compact(const uint8_t* __restrict__ mask, int nf, int offset,
                                                  int* __restrict__ smap, int* __restrict__ self_slot,
                                                  int* __restrict__ count) {
  __shared__ int cnt[1024];
  const int chunk = (nf + 1023) / 1024;
  const int lo = min(nf, (int)threadIdx.x * chunk), hi = min(nf, lo + chunk);
  int c = 0;
  for (int f = lo; f < hi; ++f) c += mask[f] == 0;
  cnt[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < 1024; ++k) {
      const int v = cnt[k];
      cnt[k] = run;
      run += v;
    }
    *count = run;
  }
  __syncthreads();
  int slot = cnt[threadIdx.x];
  for (int f = lo; f < hi; ++f) {
    if (mask[f] == 0) {
      smap[slot] = offset + f;
      if (self_slot) self_slot[f] = slot;
      ++slot;
    } else if (self_slot) {
      self_slot[f] = -1;
    }
  }
}

// ---------------------------------------------------------------------------
// initial conditions on the device (the composite field of lpm_ic_eval:
// solid-body rotation, Rossby-Haurwitz wave, Gaussian vortices, summed in
// that order; BVESphere::init_vorticity, src/LpmBVESphere.cpp:126-149)
// ---------------------------------------------------------------------------
struct IcArgs {
  double sbr;
  int rh_n;
  double rh_w, rh_K;
  int nvort;
  const double4* vort;  // {centre xyz, amplitude} per vortex
  double s2;            // sigma^2
};

// atan4 (src/LpmUtilities.hpp:14-46): longitude in [0, 2 pi)
__device__ double atan4_dev(double y, double x) {
  if (x == 0.0) return y > 0.0 ? 0.5 * kPi : (y < 0.0 ? 1.5 * kPi : 0.0);
  if (y == 0.0) return x > 0.0 ? 0.0 : kPi;
  const double th = atan2(fabs(y), fabs(x));
  if (x > 0.0 && y > 0.0) return th;
  if (x < 0.0 && y > 0.0) return kPi - th;
  if (x < 0.0 && y < 0.0) return kPi + th;
  return 2.0 * kPi - th;
}

__global__ __launch_bounds__(kBlock) void k_ic_eval(Pos3 pos, int n, IcArgs a, double* __restrict__ zeta) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double x = pos.at(i, 0), y = pos.at(i, 1), z = pos.at(i, 2);
  double zt = 0.0;
  if (a.sbr != 0.0) zt += a.sbr * (2 * (2 * kPi) * z);
  if (a.rh_n > 0) {
    const double lat = atan2(z, sqrt(x * x + y * y)), lon = atan4_dev(y, x);
    const double cl = cos(lat), sl = sin(lat);
    zt += 2 * a.rh_w * sl - a.rh_K * (a.rh_n + 1) * (a.rh_n + 2) * pow(cl, (double)a.rh_n) * sl * cos(a.rh_n * lon);
  }
  for (int k = 0; k < a.nvort; ++k) {
    const double4 v = a.vort[k];
    zt += v.w * exp(-(1.0 - (x * v.x + y * v.y + z * v.z)) / a.s2);
  }
  zeta[i] = zt;
}

// Area-weighted moments of zeta in a fixed order: workgroup b sums the
// contiguous chunk b of the targets (lane-strided, then a fixed tree).
constexpr int kMomentBlocks = 256;
__global__ __launch_bounds__(kBlock) void k_zeta_moments(const double* __restrict__ zeta,
                                                         const double* __restrict__ area, int n,
                                                         double2* __restrict__ part) {
  __shared__ double r0[kBlock / 64], r1[kBlock / 64];
  const int chunk = (n + kMomentBlocks - 1) / kMomentBlocks;
  const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double za = 0.0, aa = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kBlock) {
    za += zeta[i] * area[i];
    aa += area[i];
  }
  za = wave_sum(za);
  aa = wave_sum(aa);
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = za;
    r1[threadIdx.x >> 6] = aa;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = make_double2((r0[0] + r0[1]) + (r0[2] + r0[3]), (r1[0] + r1[1]) + (r1[2] + r1[3]));
}

// mean = sum(zeta A) / sum(A) over the partials in order
__global__ void k_zeta_mean(const double2* __restrict__ part, double* __restrict__ mean) {
  if (threadIdx.x != 0) return;
  double za = 0.0, aa = 0.0;
  for (int b = 0; b < kMomentBlocks; ++b) {
    za += part[b].x;
    aa += part[b].y;
  }
  *mean = za / aa;
}

__global__ __launch_bounds__(kBlock) void k_sub_mean(double* __restrict__ zeta, int n,
                                                     const double* __restrict__ mean) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) zeta[i] -= *mean;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// device-resident leaf meshes, one per (device, seed, level), shared by contexts
// ---------------------------------------------------------------------------
struct MeshEntry {
  int device = 0, kind = 0, level = 0;
  int64_t nv = 0, nf = 0;
  int ns = 0, nt = 0;
  std::vector<int64_t> fmap;        // active face -> tree face
  std::vector<double> face_x_host;  // every tree face [nf][3]
  std::vector<uint8_t> fmask;       // every tree face: 1 = divided
  double* pos = nullptr;            // [3][nt]: vertices, then active faces in tree order
  double* area = nullptr;           // [nt], 0 for vertices
  int* smap = nullptr;              // source slot -> target index
  int* self_slot = nullptr;         // target -> source slot or -1
  int users = 0;
  size_t bytes = 0;
};

// ===========================================================================
// host side
// ===========================================================================
namespace {

using namespace dev;

#define LPM_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail(LPM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Launch shape of one evaluation.
struct Shape {
  float near_factor = kNearDefault;
  bool lds = true;  // stage source tiles through LDS (else scalar-cache reads)
  int R = 2;
  int splits = 1;
  int ntiles = 0;
  int tiles_per_split = 0;
};

// Launch shapes chosen by the autotuner, keyed on the number of targets N.
// A context whose launch is automatic takes the entry for its N, so a
// problem run alone and the same problem inside an ensemble use the same
// decomposition (and get the same bits).
struct ShapeRule {
  int R, splits;
};
std::mutex g_rule_mu;
std::map<int, ShapeRule> g_rules;
// Persistent-grid rules of ensembles: total target count N of the members ->
// workgroups (of kBlock lanes, 4 waves) per CU, from lpm_ens_autotune.
std::map<int64_t, int> g_ens_rules;

// short_items: contexts split the sources finely (short workgroups); the
// functor entry points keep coarser splits, which bound their scratch.
Shape choose_shape(int nt, int ns, int R_req, int splits_req, bool short_items = true) {
  if (R_req == 0) {
    std::lock_guard<std::mutex> g(g_rule_mu);
    auto it = g_rules.find(nt);
    if (it != g_rules.end()) {
      R_req = it->second.R;
      if (splits_req == 0) splits_req = it->second.splits;
    }
  }
  Shape sh;
  // default without a rule: four targets per lane on large problems (half the
  // LDS reads per pair of two; 1,921 target blocks x 18 source splits at
  // 1.97M targets), two below a million targets, where four would leave too
  // few workgroups
  sh.R = (R_req >= 1 && R_req <= 4) ? R_req : (nt >= (1 << 20) ? 4 : 2);
  sh.ntiles = std::max(1, cdiv(ns, kTile));
  const int wg_x = cdiv(nt, (long)kBlock * sh.R);
  int splits = splits_req;
  if (splits <= 0) {
    // at least 2048 workgroups (8 per CU on 256 CUs) with >= 4 tiles per
    // split; then short workgroups: about 131072 of them with >= 4 tiles
    // each, at most 72 splits and about 4.6 GB of split partials (32 bytes
    // per target and split).  Workgroup durations vary, so the last ones to
    // finish leave CUs idle for about one workgroup's duration, and another
    // stream's kernels (a serving lane's) wait for workgroup slots: with 1.31M
    // sources and 1.97M targets the FP32C launch took 457 ms with 2 splits
    // (workgroups of ~116 ms), 400 ms with 18 and 396-397 ms with 36 or 72,
    // and level-5 requests beside it finished 33 % sooner with 72; a level-7
    // FP32C step took 118 ms with 20 splits and 111 ms with 72, a level-6 one
    // 9.2 ms with 9 and 7.7 ms with 72.  Functor launches keep the earlier
    // rule: about 32768 workgroups of >= 64 tiles, at most 32 splits.
    const int fill = std::min(cdiv(2048, wg_x), std::max(1, sh.ntiles / 4));
    const int budget = (int)std::max(1.0, 4.6e9 / (32.0 * (double)nt));
    const int tail = short_items ? std::min({cdiv(131072, wg_x), sh.ntiles / 4, 72, budget})
                                 : std::min({cdiv(32768, wg_x), sh.ntiles / 64, 32});
    splits = std::max({1, fill, tail});
  }
  splits = std::min(splits, sh.ntiles);
  sh.tiles_per_split = cdiv(sh.ntiles, splits);
  sh.splits = cdiv(sh.ntiles, sh.tiles_per_split);
  return sh;
}

// Device buffers of one evaluation: sources, partials, FP32C extras.
struct EvalBuffers {
  Src64* src64 = nullptr;  // dot-form or tile-local FP64 records
  Src32* src32 = nullptr;
  double4* ctr = nullptr;
  float* trad = nullptr;
  double* srcq = nullptr;
  double* qtot = nullptr;
  double* part = nullptr;
  size_t part_elems = 0;
};

// One source sweep over targets `tgt` from sources smap[ns] of the
// target-indexed arrays (pos, zeta, area).  Writes u (and psi).
struct EvalRequest {
  Pos3 tgt;
  int nt;
  Pos3 srcpos;
  const double* zeta;
  const double* area;
  const int* smap;
  int ns;
  const int* self_slot;
  MutPos3 u;
  double* psi;
  long psi_stride;
  int prec;  // lpm_prec
  int form;  // Form (FP64 only)
  bool want_psi;
  Shape shape;
  // FP64 dot form: the source count on the device (at most ns), or null
  const int* ns_dev = nullptr;
};

template <int R, bool LDS>
void launch_pair_f64(const EvalRequest& q, const EvalBuffers& b, const PairArgs& pa, dim3 grid, hipStream_t st) {
  if (q.form == FORM_DOT) {
    if (q.want_psi)
      hipLaunchKernelGGL((k_pair_f64<R, FORM_DOT, true, LDS>), grid, dim3(kBlock), 0, st, b.src64, b.ctr, pa);
    else
      hipLaunchKernelGGL((k_pair_f64<R, FORM_DOT, false, LDS>), grid, dim3(kBlock), 0, st, b.src64, b.ctr, pa);
  } else {
    if (q.want_psi)
      hipLaunchKernelGGL((k_pair_f64<R, FORM_CHORD, true, LDS>), grid, dim3(kBlock), 0, st, b.src64, b.ctr, pa);
    else
      hipLaunchKernelGGL((k_pair_f64<R, FORM_CHORD, false, LDS>), grid, dim3(kBlock), 0, st, b.src64, b.ctr, pa);
  }
}

template <int R>
void launch_pair(const EvalRequest& q, const EvalBuffers& b, const PairArgs& pa, dim3 grid, hipStream_t st) {
  if (q.prec == LPM_PREC_FP32C) {
    if (q.shape.lds) {
      if (q.want_psi)
        hipLaunchKernelGGL((k_pair_f32c_lds<R, true>), grid, dim3(kBlock), 0, st, b.src32, b.ctr, b.trad, pa);
      else
        hipLaunchKernelGGL((k_pair_f32c_lds<R, false>), grid, dim3(kBlock), 0, st, b.src32, b.ctr, b.trad, pa);
    } else if (q.want_psi) {
      hipLaunchKernelGGL((k_pair_f32c<R, true>), grid, dim3(kBlock), 0, st, b.src32, b.ctr, b.trad, pa);
    } else {
      hipLaunchKernelGGL((k_pair_f32c<R, false>), grid, dim3(kBlock), 0, st, b.src32, b.ctr, b.trad, pa);
    }
  } else if (q.shape.lds) {
    launch_pair_f64<R, true>(q, b, pa, grid, st);
  } else {
    launch_pair_f64<R, false>(q, b, pa, grid, st);
  }
}

PairArgs make_pair_args(const EvalRequest& q, const EvalBuffers& b) {
  PairArgs pa;
  pa.tgt = q.tgt;
  pa.self_slot = q.self_slot;
  pa.nt = q.nt;
  pa.ntiles = q.shape.ntiles;
  pa.tiles_per_split = q.shape.tiles_per_split;
  pa.part = b.part;
  pa.nt_pad = q.nt;
  pa.near_factor = q.shape.near_factor;
  pa.ns_dev = q.ns_dev;
  return pa;
}

// Source records for one evaluation.
int enqueue_prep(const EvalRequest& q, const EvalBuffers& b, hipStream_t st) {
  const Shape& sh = q.shape;
  const int ns_pad = sh.ntiles * kTile;
  if (q.prec == LPM_PREC_FP32C) {
    hipLaunchKernelGGL(k_prep_f32c, dim3(sh.ntiles), dim3(kTile), 0, st, b.src32, b.ctr, b.trad, b.srcq, q.srcpos,
                       q.zeta, q.area, q.smap, q.ns);
    if (q.want_psi) hipLaunchKernelGGL(k_qsum, dim3(1), dim3(kBlock), 0, st, b.ctr, sh.ntiles, b.qtot);
  } else if (q.form == FORM_DOT) {
    hipLaunchKernelGGL(k_prep_f64_dot, dim3(cdiv(ns_pad, kBlock)), dim3(kBlock), 0, st, b.src64, q.srcpos, q.zeta,
                       q.area, q.smap, q.ns, ns_pad, q.ns_dev);
  } else {
    hipLaunchKernelGGL(k_prep_f64_tiles, dim3(sh.ntiles), dim3(kTile), 0, st, b.src64, b.ctr, q.srcpos, q.zeta,
                       q.area, q.smap, q.ns);
  }
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

// The pair kernel over the request's own grid; ev0/ev1 (optional) bracket it.
int enqueue_pair(const EvalRequest& q, const EvalBuffers& b, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  const PairArgs pa = make_pair_args(q, b);
  // The stream-function kernels hold four more sums and a logarithm per pair:
  // with at most two targets per lane they keep more waves resident.
  // Measured at 1.97M targets (one box): FP64 2,564-2,582 ms at two per lane
  // against 2,939 at four; FP32C 626 ms at two against 707 at four and 819
  // at one.  The split decomposition, and so the partial buffers, stay the
  // shape's.
  int R = q.shape.R;
  if (q.want_psi) R = std::min(R, 2);
  const dim3 grid(cdiv(q.nt, (long)kBlock * R), q.shape.splits);
  if (ev0) LPM_HIP(hipEventRecord(ev0, st));
  switch (R) {
    case 1: launch_pair<1>(q, b, pa, grid, st); break;
    case 3: launch_pair<3>(q, b, pa, grid, st); break;
    case 4: launch_pair<4>(q, b, pa, grid, st); break;
    default: launch_pair<2>(q, b, pa, grid, st); break;
  }
  LPM_HIP(hipGetLastError());
  if (ev1) LPM_HIP(hipEventRecord(ev1, st));
  return LPM_OK;
}

FinArgs make_fin_args(const EvalRequest& q, const EvalBuffers& b) {
  FinArgs fa;
  fa.part = b.part;
  fa.splits = q.shape.splits;
  fa.nt = q.nt;
  fa.nt_pad = q.nt;
  fa.tgt = q.tgt;
  fa.u = q.u;
  fa.psi = q.want_psi ? q.psi : nullptr;
  fa.psi_stride = q.psi_stride;
  fa.qtot = b.qtot;
  fa.srcq = b.srcq;
  fa.self_slot = q.self_slot;
  return fa;
}

// Sum split partials, cross product, stream function.
int enqueue_finalize(const EvalRequest& q, const EvalBuffers& b, hipStream_t st) {
  const FinArgs fa = make_fin_args(q, b);
  const dim3 fg(cdiv(q.nt, kBlock));
  if (q.prec == LPM_PREC_FP32C)
    hipLaunchKernelGGL((k_finalize<FIN_F32C>), fg, dim3(kBlock), 0, st, fa);
  else if (q.form == FORM_DOT)
    hipLaunchKernelGGL((k_finalize<FIN_DOT>), fg, dim3(kBlock), 0, st, fa);
  else
    hipLaunchKernelGGL((k_finalize<FIN_CHORD>), fg, dim3(kBlock), 0, st, fa);
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

// Queue prep + pair + finalize on `st`.
int enqueue_eval(const EvalRequest& q, const EvalBuffers& b, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  int rc;
  if ((rc = enqueue_prep(q, b, st))) return rc;
  if ((rc = enqueue_pair(q, b, st, ev0, ev1))) return rc;
  return enqueue_finalize(q, b, st);
}

// Sizes (bytes) of the evaluation buffers for ns sources, nt targets.
struct BufferPlan {
  size_t src64, src32, ctr, trad, srcq, part;
};
BufferPlan plan_buffers(int nt, const Shape& sh) {
  const size_t ns_pad = (size_t)sh.ntiles * kTile;
  BufferPlan p;
  p.src64 = ns_pad * sizeof(Src64);
  p.src32 = ns_pad * sizeof(Src32);
  p.ctr = 4 * (size_t)sh.ntiles * sizeof(double4) + 256;  // centres, then FP32C tile frames
  p.trad = (size_t)sh.ntiles * sizeof(float) + 256;
  p.srcq = ns_pad * sizeof(double);
  p.part = (size_t)sh.splits * kNC * nt * sizeof(double);
  return p;
}

// ---------------------------------------------------------------------------
// functor entry points: scratch layout, per-device workspace, static scratch
// ---------------------------------------------------------------------------
inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Scratch layout of one functor evaluation, as byte offsets from a 256-byte
// aligned base: source records and split partials planned for nf sources
// (the active count is known only on the device, after compaction), the
// compacted source map, self slots and count, and a velocity buffer when the
// caller asks for the stream function only.
struct FunctorPlan {
  Shape shape;
  size_t src64 = 0, part = 0, smap = 0, self = 0, count = 0, utmp = 0, bytes = 0;
};

FunctorPlan functor_plan(int64_t nt, int64_t nf, bool need_u) {
  FunctorPlan p;
  p.shape = choose_shape((int)nt, (int)std::max<int64_t>(nf, 1), 2, 0, false);
  const BufferPlan bp = plan_buffers((int)nt, p.shape);
  size_t off = 0;
  auto take = [&](size_t n) {
    const size_t r = off;
    off += align256(n);
    return r;
  };
  p.src64 = take(bp.src64);
  p.part = take(bp.part);
  p.smap = take(sizeof(int) * (size_t)(nf + 1));
  p.self = take(sizeof(int) * (size_t)(nf + 1));
  p.count = take(sizeof(int));
  p.utmp = need_u ? take(sizeof(double) * 3 * (size_t)nt) : 0;
  p.bytes = off;
  return p;
}

int functor_check(double* u, double* psi, const double* tx, int64_t nt, const double* fx, const double* fzeta,
                  const double* farea, const uint8_t* fmask, int64_t nf, bool collocated) {
  if (nt < 0 || nf < 0 || nt > INT32_MAX / 4 || nf > INT32_MAX / 4) return fail(LPM_ERR_ARG, "functor: bad sizes");
  if (collocated && nt != nf) return fail(LPM_ERR_ARG, "functor: face targets must be the nf faces");
  if (nt > 0 && (!tx || !fx || !fzeta || !farea || !fmask || (!u && !psi)))
    return fail(LPM_ERR_ARG, "functor: null pointer");
  return LPM_OK;
}

// Queue one functor evaluation on `st` with its scratch at `base`: compact
// the mask on the device, then one FP64 dot-form evaluation whose kernels read
// the active count from the device (tiles past it are skipped, so no value
// crosses to the host and nothing waits).  No active source: every sum is 0.
int functor_enqueue(const FunctorPlan& fp, char* base, double* u, double* psi, const double* tx, int64_t nt,
                    const double* fx, const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf,
                    bool collocated, hipStream_t st) {
  EvalBuffers b;
  b.src64 = (Src64*)(base + fp.src64);
  b.part = (double*)(base + fp.part);
  int* smap = (int*)(base + fp.smap);
  int* self = (int*)(base + fp.self);
  int* count = (int*)(base + fp.count);
  hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, st, fmask, (int)nf, 0, smap, collocated ? self : nullptr,
                     count);
  LPM_HIP(hipGetLastError());
  EvalRequest q;
  q.tgt = Pos3{tx, 3, 1};
  q.nt = (int)nt;
  q.srcpos = Pos3{fx, 3, 1};
  q.zeta = fzeta;
  q.area = farea;
  q.smap = smap;
  q.ns = (int)nf;  // upper bound; the kernels take the count from ns_dev
  q.ns_dev = count;
  q.self_slot = collocated ? self : nullptr;
  q.u = MutPos3{u ? u : (double*)(base + fp.utmp), 3, 1};
  q.psi = psi;
  q.psi_stride = 1;
  q.prec = LPM_PREC_FP64;
  q.form = FORM_DOT;
  q.want_psi = psi != nullptr;
  q.shape = fp.shape;
  return enqueue_eval(q, b, st, nullptr, nullptr);
}

// Per-device workspace of the allocating functor entry points.  It grows on
// demand and is shared by every call on the device: each call waits on the
// device for the previous call's kernels (an event, whatever their stream)
// before it reuses the workspace.
struct Workspace {
  std::mutex mu;
  char* base = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded behind the last call's kernels
};
std::mutex g_ws_mu;
std::map<int, Workspace*> g_ws;

Workspace* workspace_for(int dev) {
  std::lock_guard<std::mutex> g(g_ws_mu);
  auto& w = g_ws[dev];
  if (!w) w = new Workspace();
  return w;
}

// Functor evaluation on caller-owned device arrays, asynchronous on `stream`.
int functor_eval(double* u, double* psi, const double* tx, int64_t nt, const double* fx, const double* fzeta,
                 const double* farea, const uint8_t* fmask, int64_t nf, bool collocated, void* stream) {
  int rc = functor_check(u, psi, tx, nt, fx, fzeta, farea, fmask, nf, collocated);
  if (rc) return rc;
  if (nt == 0) return ok();
  int dev = 0;
  LPM_HIP(hipGetDevice(&dev));
  hipStream_t st = (hipStream_t)stream;
  Workspace* ws = workspace_for(dev);
  std::lock_guard<std::mutex> g(ws->mu);
  const FunctorPlan fp = functor_plan(nt, nf, u == nullptr);
  if (!ws->done) LPM_HIP(hipEventCreateWithFlags(&ws->done, hipEventDisableTiming));
  if (ws->bytes < fp.bytes) {
    if (ws->base) {
      // the previous call's kernels still read the old workspace
      LPM_HIP(hipEventSynchronize(ws->done));
      LPM_HIP(hipFree(ws->base));
      ws->base = nullptr;
      ws->bytes = 0;
    }
    if (hipMalloc(&ws->base, fp.bytes) != hipSuccess) return fail(LPM_ERR_NOMEM, "functor: workspace allocation failed");
    ws->bytes = fp.bytes;
  } else {
    LPM_HIP(hipStreamWaitEvent(st, ws->done, 0));
  }
  if ((rc = functor_enqueue(fp, ws->base, u, psi, tx, nt, fx, fzeta, farea, fmask, nf, collocated, st))) return rc;
  LPM_HIP(hipEventRecord(ws->done, st));
  return ok();
}

// Per-device address of the static functor scratch (`scratch` below), looked
// up once per device with that device current.
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;

int static_scratch_for(int dev, char** out);

}  // namespace
}  // namespace lpm

// Device scratch of lpm_bve_functor_static: a fixed array in every device's
// copy of the code object, held until the process ends.  Host code never
// passes this name to a kernel or a copy (on the host it is only a stand-in
// for the symbol): it uses the address hipGetSymbolAddress returns for the
// current device.
constexpr size_t kStaticScratchBytes = LPM_FUNCTOR_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace lpm {
namespace {

int static_scratch_for(int dev, char** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return LPM_OK;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(LPM_ERR_HIP, std::string("lpm_bve_functor_static: scratch symbol address: ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = (char*)p;
  *out = (char*)p;
  return LPM_OK;
}

}  // namespace
}  // namespace lpm

// ===========================================================================
// solver context
// ===========================================================================
struct lpm_bve_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t nv = 0, nf = 0;
  int nt = 0, ns = 0;  // targets = vertices then active faces; sources = active faces
  // tree-face index of each active face; host copies of every face's
  // position and vorticity (divided faces are not evaluated and keep theirs)
  std::vector<int64_t> fmap;
  std::vector<double> face_x_host, face_zeta_host;
  std::vector<uint8_t> fmask_host;  // every tree face's mask (state set from arrays)
  // device arrays (SoA, target-indexed: vertices then tree faces)
  double *x = nullptr, *y = nullptr, *z = nullptr, *zeta = nullptr;
  double *u = nullptr, *v = nullptr, *w = nullptr, *psi = nullptr;
  double *area = nullptr, *absv = nullptr;
  double *wx = nullptr, *wy = nullptr, *wz = nullptr, *wzeta = nullptr;
  double* kx[4] = {nullptr, nullptr, nullptr, nullptr};
  double* kz[4] = {nullptr, nullptr, nullptr, nullptr};
  int *smap = nullptr, *self_slot = nullptr;
  double* circ = nullptr;
  lpm::EvalBuffers eb;
  std::vector<void*> allocations;
  // launch shape (R_req 0: automatic -- autotuned rule for this N, else 2)
  int R_req = 0, splits_req = 0;
  // bumped whenever device buffers are (re)allocated
  uint64_t generation = 0;
  float near_factor = lpm::dev::kNearDefault;
  bool lds = true;
  // velocity validity: form and precision that produced u
  bool vel_valid = false;
  int vel_prec = -1, vel_form = -1;
  bool psi_valid = false;
  // absv is valid for this omega
  bool absv_valid = false;
  double absv_omega = 0.0;
  // captured step graphs
  struct GraphKey {
    double dt, omega;
    int scheme, prec, R, splits;
    float near_factor;
    bool lds;
    bool operator<(const GraphKey& o) const {
      return std::tie(dt, omega, scheme, prec, R, splits, near_factor, lds) <
             std::tie(o.dt, o.omega, o.scheme, o.prec, o.R, o.splits, o.near_factor, o.lds);
    }
  };
  std::map<GraphKey, hipGraphExec_t> graphs;
  // pair-kernel timing
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timing_events;
  bool timing = false;
  // state built from a cached device mesh (lpm_bve_set_state_level): the
  // mesh's read-only arrays are shared, and the divided faces' vorticity is
  // evaluated on the host from the field parameters when first read back
  lpm::MeshEntry* mesh = nullptr;
  bool own_stream = true;
  // completion marker behind the last asynchronous work queued for the
  // context (lpm_bve_solve, lpm_bve_advance, ensemble steps): a one-lane
  // kernel writes done_seq into pinned host memory; lpm_bve_query compares
  unsigned long long* done_flag = nullptr;
  unsigned long long done_seq = 0;
  // the marker of the ensemble whose stream the context shares (its steps)
  const unsigned long long* done_flag_shared = nullptr;
  unsigned long long done_seq_shared = 0;
  bool face_zeta_pending = false;
  double ic_sbr = 0.0, ic_rh_w = 0.0, ic_rh_K = 0.0, ic_sigma = 1.0;
  int ic_rh_n = 0;
  bool ic_centered = false;
  double* ic_mean = nullptr;  // device scalar: subtracted mean
  std::vector<double> ic_centers, ic_amps;
};

namespace lpm {
namespace {

std::mutex g_mesh_mu;
std::map<std::tuple<int, int, int>, MeshEntry*> g_meshes;  // (device, seed kind, level)

void release_mesh(lpm_bve_ctx* c) {
  c->face_zeta_pending = false;
  c->ic_mean = nullptr;
  if (!c->mesh) return;
  std::lock_guard<std::mutex> g(g_mesh_mu);
  --c->mesh->users;
  c->mesh = nullptr;
}

void free_ctx_memory(lpm_bve_ctx* c);

// Queue a completion marker behind the work queued so far on `st`.
int queue_mark(unsigned long long** flag, unsigned long long* seq, hipStream_t st) {
  if (!*flag) {
    if (hipHostMalloc((void**)flag, sizeof(unsigned long long), hipHostMallocCoherent) != hipSuccess)
      return fail(LPM_ERR_NOMEM, "completion marker: pinned allocation failed");
    **flag = 0;
  }
  ++*seq;
  hipLaunchKernelGGL(k_mark, dim3(1), dim3(64), 0, st, *flag, *seq);
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

bool mark_reached(const unsigned long long* flag, unsigned long long seq) {
  return !flag || *(const volatile unsigned long long*)flag >= seq;
}

int ctx_mark_done(lpm_bve_ctx* c) { return queue_mark(&c->done_flag, &c->done_seq, c->stream); }

// Device arrays of a context's state for c->nt targets and c->ns sources.
// own_static: also the read-only area, source map and self slots (else they
// come from a cached mesh).
int ctx_alloc_state(lpm_bve_ctx* c, bool own_static) {
  const int nt = c->nt;
  auto alloc = [&](void** p, size_t n) -> int {
    if (hipMalloc(p, std::max<size_t>(n, 256)) != hipSuccess) return fail(LPM_ERR_NOMEM, "device allocation failed");
    c->allocations.push_back(*p);
    return LPM_OK;
  };
  const size_t d = sizeof(double) * (size_t)nt;
  int rc = 0;
  // position, stage-argument and velocity triples are contiguous [3][nt]
  // blocks so kernels address them as one SoA array
  double** triples[] = {&c->x, &c->wx, &c->u};
  for (double** p : triples)
    if ((rc = alloc((void**)p, 3 * d))) return rc;
  c->y = c->x + nt;
  c->z = c->x + 2 * (size_t)nt;
  c->wy = c->wx + nt;
  c->wz = c->wx + 2 * (size_t)nt;
  c->v = c->u + nt;
  c->w = c->u + 2 * (size_t)nt;
  double** singles[] = {&c->zeta, &c->psi, &c->absv, &c->wzeta};
  for (double** p : singles)
    if ((rc = alloc((void**)p, d))) return rc;
  for (int s = 0; s < 4; ++s) {
    if ((rc = alloc((void**)&c->kx[s], 3 * d))) return rc;
    if ((rc = alloc((void**)&c->kz[s], d))) return rc;
  }
  if (own_static) {
    if ((rc = alloc((void**)&c->area, d))) return rc;
    if ((rc = alloc((void**)&c->smap, sizeof(int) * c->ns))) return rc;
    if ((rc = alloc((void**)&c->self_slot, sizeof(int) * nt))) return rc;
  }
  if ((rc = alloc((void**)&c->circ, 256))) return rc;
  return LPM_OK;
}

// The cached level-`level` tree of seed `kind` on `device`, built and
// uploaded on first use; counts one more user.
int get_mesh(int device, int kind, int level, MeshEntry** out) {
  std::lock_guard<std::mutex> g(g_mesh_mu);
  const auto key = std::make_tuple(device, kind, level);
  auto it = g_meshes.find(key);
  if (it != g_meshes.end()) {
    ++it->second->users;
    *out = it->second;
    return LPM_OK;
  }
  MeshEntry* m = new MeshEntry();
  m->device = device;
  m->kind = kind;
  m->level = level;
  try {
    const TreeMesh t = build_tree(kind, level);
    m->nv = t.nverts();
    m->nf = t.nfaces();
    m->fmask = t.face_mask;
    for (int64_t f = 0; f < m->nf; ++f)
      if (!t.face_mask[f]) m->fmap.push_back(f);
    m->ns = (int)m->fmap.size();
    m->nt = (int)(m->nv + m->ns);
    m->face_x_host.resize(3 * (size_t)m->nf);
    std::memcpy(m->face_x_host.data(), t.face_phys.data(), sizeof(double) * 3 * m->nf);
    const int nt = m->nt;
    std::vector<double> hp(3 * (size_t)nt), ha(nt, 0.0);
    std::vector<int> hs(m->ns), hself(nt, -1);
    for (int64_t i = 0; i < m->nv; ++i)
      for (int k = 0; k < 3; ++k) hp[k * (size_t)nt + i] = t.vert_phys[i][k];
    for (int j = 0; j < m->ns; ++j) {
      const int64_t f = m->fmap[j], i = m->nv + j;
      for (int k = 0; k < 3; ++k) hp[k * (size_t)nt + i] = t.face_phys[f][k];
      ha[i] = t.face_area[f];
      hs[j] = (int)i;
      hself[i] = j;
    }
    const size_t d = sizeof(double) * (size_t)nt;
    m->bytes = 4 * d + sizeof(int) * ((size_t)m->ns + nt);
    bool okm = hipMalloc(&m->pos, 3 * d) == hipSuccess;
    okm = okm && hipMalloc(&m->area, d) == hipSuccess;
    okm = okm && hipMalloc(&m->smap, sizeof(int) * m->ns) == hipSuccess;
    okm = okm && hipMalloc(&m->self_slot, sizeof(int) * nt) == hipSuccess;
    if (!okm) {
      (void)hipFree(m->pos);
      (void)hipFree(m->area);
      (void)hipFree(m->smap);
      (void)hipFree(m->self_slot);
      delete m;
      return fail(LPM_ERR_NOMEM, "mesh cache: device allocation failed");
    }
    LPM_HIP(hipMemcpy(m->pos, hp.data(), 3 * d, hipMemcpyHostToDevice));
    LPM_HIP(hipMemcpy(m->area, ha.data(), d, hipMemcpyHostToDevice));
    LPM_HIP(hipMemcpy(m->smap, hs.data(), sizeof(int) * m->ns, hipMemcpyHostToDevice));
    LPM_HIP(hipMemcpy(m->self_slot, hself.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
  } catch (const std::exception& e) {
    delete m;
    return fail(LPM_ERR_NOMEM, std::string("mesh cache: ") + e.what());
  }
  m->users = 1;
  g_meshes[key] = m;
  *out = m;
  return LPM_OK;
}

const std::vector<int64_t>& ctx_fmap(const lpm_bve_ctx* c) { return c->mesh ? c->mesh->fmap : c->fmap; }

// Host copy of every tree face's vorticity: the divided faces' values (the
// field at their positions, less the subtracted mean) for a cached-mesh state.
int ctx_face_zeta_host(lpm_bve_ctx* c) {
  if (!c->face_zeta_pending) return LPM_OK;
  double mean = 0.0;
  if (c->ic_centered) {
    LPM_HIP(hipMemcpyAsync(&mean, c->ic_mean, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(hipStreamSynchronize(c->stream));
  }
  const MeshEntry* m = c->mesh;
  std::vector<int64_t> pf;
  std::vector<double> px;
  for (int64_t f = 0; f < c->nf; ++f)
    if (m->fmask[f]) {
      pf.push_back(f);
      for (int k = 0; k < 3; ++k) px.push_back(m->face_x_host[3 * f + k]);
    }
  std::vector<double> pz(pf.size());
  const int nvort = (int)c->ic_amps.size();
  if (!pf.empty()) {
    const int rc = lpm_ic_eval(px.data(), (int64_t)pf.size(), pz.data(), c->ic_sbr, c->ic_rh_n, c->ic_rh_w,
                               c->ic_rh_K, nvort, nvort ? c->ic_centers.data() : nullptr,
                               nvort ? c->ic_amps.data() : nullptr, c->ic_sigma);
    if (rc) return rc;
  }
  // active faces are filled from the device state by the caller
  c->face_zeta_host.assign(c->nf, 0.0);
  for (size_t p = 0; p < pf.size(); ++p) c->face_zeta_host[pf[p]] = pz[p] - mean;
  c->face_zeta_pending = false;
  return LPM_OK;
}

void free_ctx_memory(lpm_bve_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  for (void* p : c->allocations) (void)hipFree(p);
  c->allocations.clear();
  release_mesh(c);
  for (auto& e : c->timing_events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  c->timing_events.clear();
}

Shape ctx_shape(const lpm_bve_ctx* c) {
  Shape sh = choose_shape(c->nt, std::max(c->ns, 1), c->R_req, c->splits_req);
  sh.near_factor = c->near_factor;
  sh.lds = c->lds;
  return sh;
}

// Make sure the evaluation buffers fit the current launch shape.
int ensure_eval_buffers(lpm_bve_ctx* c, const Shape& sh) {
  const BufferPlan bp = plan_buffers(c->nt, sh);
  const size_t part_elems = bp.part / sizeof(double);
  if (c->eb.src64 && c->eb.part_elems >= part_elems) return LPM_OK;
  if (c->eb.part) {
    // grow the partial buffer only
    LPM_HIP(hipStreamSynchronize(c->stream));
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    auto it = std::find(c->allocations.begin(), c->allocations.end(), (void*)c->eb.part);
    if (it != c->allocations.end()) c->allocations.erase(it);
    LPM_HIP(hipFree(c->eb.part));
    c->eb.part = nullptr;
    if (hipMalloc(&c->eb.part, bp.part) != hipSuccess) return fail(LPM_ERR_NOMEM, "partial buffer allocation failed");
    c->allocations.push_back(c->eb.part);
    c->eb.part_elems = part_elems;
    ++c->generation;
    return LPM_OK;
  }
  auto alloc = [&](void** p, size_t n) -> int {
    if (hipMalloc(p, std::max<size_t>(n, 256)) != hipSuccess) return fail(LPM_ERR_NOMEM, "device allocation failed");
    c->allocations.push_back(*p);
    return LPM_OK;
  };
  int rc;
  if ((rc = alloc((void**)&c->eb.src64, bp.src64))) return rc;
  if ((rc = alloc((void**)&c->eb.src32, bp.src32))) return rc;
  if ((rc = alloc((void**)&c->eb.ctr, bp.ctr))) return rc;
  if ((rc = alloc((void**)&c->eb.trad, bp.trad))) return rc;
  if ((rc = alloc((void**)&c->eb.srcq, bp.srcq))) return rc;
  if ((rc = alloc((void**)&c->eb.qtot, 256))) return rc;
  if ((rc = alloc((void**)&c->eb.part, bp.part))) return rc;
  c->eb.part_elems = part_elems;
  ++c->generation;
  return LPM_OK;
}

// Request for an evaluation at positions px ([3][nt] SoA) with vorticity
// pzeta; makes sure the evaluation buffers fit the launch shape.
int ctx_request(lpm_bve_ctx* c, const double* px, const double* pzeta, int prec, int form, bool want_psi,
                EvalRequest& q) {
  const Shape sh = ctx_shape(c);
  int rc = ensure_eval_buffers(c, sh);
  if (rc) return rc;
  q.tgt = Pos3{px, 1, (long)c->nt};
  q.nt = c->nt;
  q.srcpos = q.tgt;
  q.zeta = pzeta;
  q.area = c->area;
  q.smap = c->smap;
  q.ns = c->ns;
  q.self_slot = c->self_slot;
  q.u = MutPos3{c->u, 1, (long)c->nt};
  q.psi = c->psi;
  q.psi_stride = 1;
  q.prec = prec;
  q.form = form;
  q.want_psi = want_psi;
  q.shape = sh;
  return LPM_OK;
}

// Evaluate velocity (and psi) at positions px with vorticity pzeta on `st`.
int ctx_eval(lpm_bve_ctx* c, const double* px, const double* pzeta, int prec, int form, bool want_psi,
             hipStream_t st) {
  EvalRequest q;
  int rc = ctx_request(c, px, pzeta, prec, form, want_psi, q);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    LPM_HIP(hipEventCreate(&e0));
    LPM_HIP(hipEventCreate(&e1));
    c->timing_events.push_back({e0, e1});
  }
  return enqueue_eval(q, c->eb, st, e0, e1);
}

// The context's arrays for the stage kernel (stage, scheme, dt and Omega
// left at defaults).
StageArgs ctx_stage_args(const lpm_bve_ctx* c) {
  StageArgs a;
  a.nt = c->nt;
  a.nv = (int)c->nv;
  a.stage = 1;
  a.lagr = 0;
  a.dt = 0.0;
  a.omega = 0.0;
  a.u = c->u;
  a.v = c->v;
  a.w = c->w;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.zeta = c->zeta;
  a.absv = c->absv;
  for (int s = 0; s < 4; ++s) {
    a.kx[s] = c->kx[s];
    a.kz[s] = c->kz[s];
  }
  a.wx = c->wx;
  a.wy = c->wy;
  a.wz = c->wz;
  a.wzeta = c->wzeta;
  return a;
}

int ctx_stage(lpm_bve_ctx* c, int stage, double dt, double omega, int scheme, hipStream_t st) {
  StageArgs a = ctx_stage_args(c);
  a.stage = stage;
  a.lagr = scheme == LPM_SCHEME_LAGR;
  a.dt = dt;
  a.omega = omega;
  hipLaunchKernelGGL(k_stage, dim3(cdiv(c->nt, kBlock)), dim3(kBlock), 0, st, a);
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

// The absolute vorticity zeta + 2 Omega z of the state at the first step with
// rate `omega` (BVESphere::init_vorticity, src/LpmBVESphere.cpp:133,145):
// conserved along trajectories, so computed once per rate and kept.  The
// Lagrangian scheme takes zeta from it; both schemes report it.
int ctx_ensure_absv(lpm_bve_ctx* c, double omega, hipStream_t st) {
  if (c->absv_valid && c->absv_omega == omega) return LPM_OK;
  hipLaunchKernelGGL(k_absv, dim3(cdiv(c->nt, kBlock)), dim3(kBlock), 0, st, c->nt, omega, c->zeta, c->z, c->absv);
  LPM_HIP(hipGetLastError());
  c->absv_valid = true;
  c->absv_omega = omega;
  return LPM_OK;
}

// Before stepping: absolute vorticity and a valid entry velocity for
// (prec, form), queued on `st`.
int ctx_prepare_step(lpm_bve_ctx* c, double omega, int scheme, int prec, hipStream_t st) {
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  int rc = ctx_ensure_absv(c, omega, st);
  if (rc) return rc;
  if (!c->vel_valid || (c->vel_prec >= 0 && (c->vel_prec != prec || c->vel_form != form))) {
    if ((rc = ctx_eval(c, c->x, c->zeta, prec, form, false, st))) return rc;
    c->vel_valid = true;
    c->vel_prec = prec;
    c->vel_form = form;
  }
  return ensure_eval_buffers(c, ctx_shape(c));
}

// One advance_timestep (src/LpmBVERK4_Impl.hpp:42-130) queued on `st`.
int ctx_enqueue_step(lpm_bve_ctx* c, double dt, double omega, int scheme, int prec, hipStream_t st) {
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  int rc;
  for (int stage = 1; stage <= 3; ++stage) {
    if ((rc = ctx_stage(c, stage, dt, omega, scheme, st))) return rc;
    if ((rc = ctx_eval(c, c->wx, c->wzeta, prec, form, false, st))) return rc;
  }
  if ((rc = ctx_stage(c, 4, dt, omega, scheme, st))) return rc;
  return ctx_eval(c, c->x, c->zeta, prec, form, false, st);
}

}  // namespace
}  // namespace lpm

using namespace lpm;

extern "C" {

int lpm_device_count(int* count) {
  LPM_HIP(hipGetDeviceCount(count));
  return ok();
}

int lpm_bve_vertex_velocity(double* u, const double* vx, int64_t nv, const double* fx, const double* fzeta,
                            const double* farea, const uint8_t* fmask, int64_t nf, void* stream) {
  return functor_eval(u, nullptr, vx, nv, fx, fzeta, farea, fmask, nf, false, stream);
}
int lpm_bve_face_velocity(double* u, const double* fx, const double* fzeta, const double* farea,
                          const uint8_t* fmask, int64_t nf, void* stream) {
  return functor_eval(u, nullptr, fx, nf, fx, fzeta, farea, fmask, nf, true, stream);
}
int lpm_bve_vertex_streamfn(double* psi, const double* vx, int64_t nv, const double* fx, const double* fzeta,
                            const double* farea, const uint8_t* fmask, int64_t nf, void* stream) {
  return functor_eval(nullptr, psi, vx, nv, fx, fzeta, farea, fmask, nf, false, stream);
}
int lpm_bve_face_streamfn(double* psi, const double* fx, const double* fzeta, const double* farea,
                          const uint8_t* fmask, int64_t nf, void* stream) {
  return functor_eval(nullptr, psi, fx, nf, fx, fzeta, farea, fmask, nf, true, stream);
}
int lpm_bve_vertex_solve(double* psi, double* u, const double* vx, int64_t nv, const double* fx,
                         const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, void* stream) {
  if (!psi || !u) return fail(LPM_ERR_ARG, "lpm_bve_vertex_solve: null output");
  return functor_eval(u, psi, vx, nv, fx, fzeta, farea, fmask, nf, false, stream);
}
int lpm_bve_face_solve(double* psi, double* u, const double* fx, const double* fzeta, const double* farea,
                       const uint8_t* fmask, int64_t nf, void* stream) {
  if (!psi || !u) return fail(LPM_ERR_ARG, "lpm_bve_face_solve: null output");
  return functor_eval(u, psi, fx, nf, fx, fzeta, farea, fmask, nf, true, stream);
}

int lpm_bve_functor_static(double* psi, double* u, const double* tx, int64_t nt, const double* fx,
                           const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, int collocated,
                           void* stream) {
  const bool col = collocated != 0;
  int rc = functor_check(u, psi, tx, nt, fx, fzeta, farea, fmask, nf, col);
  if (rc) return rc;
  if (nt == 0) return ok();
  const FunctorPlan fp = functor_plan(nt, nf, u == nullptr);
  if (fp.bytes > kStaticScratchBytes)
    return fail(LPM_ERR_ARG, "lpm_bve_functor_static: needs " + std::to_string(fp.bytes) +
                                 " bytes of scratch for these sizes, above its fixed " +
                                 std::to_string(kStaticScratchBytes) + " (use the allocating functor entry points)");
  int dev = 0;
  LPM_HIP(hipGetDevice(&dev));
  char* base = nullptr;
  if ((rc = static_scratch_for(dev, &base))) return rc;
  if ((rc = functor_enqueue(fp, base, u, psi, tx, nt, fx, fzeta, farea, fmask, nf, col, (hipStream_t)stream)))
    return rc;
  return ok();
}

int lpm_bve_create(int device, lpm_bve_ctx** out) {
  if (!out) return fail(LPM_ERR_ARG, "lpm_bve_create: null output");
  int n = 0;
  LPM_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(LPM_ERR_ARG, "lpm_bve_create: no such device");
  LPM_HIP(hipSetDevice(device));
  lpm_bve_ctx* c = new lpm_bve_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(LPM_ERR_HIP, "lpm_bve_create: stream creation failed");
  }
  *out = c;
  return ok();
}

int lpm_bve_create_on_stream(int device, void* stream, lpm_bve_ctx** out) {
  if (!out || !stream) return fail(LPM_ERR_ARG, "lpm_bve_create_on_stream: null argument");
  int n = 0;
  LPM_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(LPM_ERR_ARG, "lpm_bve_create_on_stream: no such device");
  lpm_bve_ctx* c = new lpm_bve_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  *out = c;
  return ok();
}

int lpm_bve_destroy(lpm_bve_ctx* c) {
  if (!c) return ok();
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_ctx_memory(c);
  if (c->done_flag) (void)hipHostFree(c->done_flag);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ok();
}

int lpm_bve_set_state(lpm_bve_ctx* c, int64_t nv, const double* vx, const double* vzeta, int64_t nf,
                      const double* fx, const double* fzeta, const double* farea, const uint8_t* fmask) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_set_state: null context");
  if (nv < 0 || nf < 1 || nv > (int64_t)INT32_MAX || nf > (int64_t)INT32_MAX)
    return fail(LPM_ERR_ARG, "lpm_bve_set_state: sizes out of range");
  if ((nv && (!vx || !vzeta)) || !fx || !fzeta || !farea || !fmask)
    return fail(LPM_ERR_ARG, "lpm_bve_set_state: null array");
  // Divided faces are sources of nothing (mask) and, as targets, sit on
  // their centre child's centroid, where the reference's collocated sum
  // divides by 1 - x.y = 0; they are kept on the host and not evaluated.
  std::vector<int64_t> fmap;
  for (int64_t f = 0; f < nf; ++f)
    if (!fmask[f]) fmap.push_back(f);
  const int64_t ns = (int64_t)fmap.size();
  if (ns == 0) return fail(LPM_ERR_ARG, "lpm_bve_set_state: no active (unmasked) face");
  if (nv + ns > (int64_t)INT32_MAX / 8)
    return fail(LPM_ERR_ARG, "lpm_bve_set_state: vertices + active faces out of range (the limit of lpm_bve_state_bytes)");
  LPM_HIP(hipSetDevice(c->device));
  LPM_HIP(hipStreamSynchronize(c->stream));
  // Same sizes and mask as the resident state (the reference-side copy-in of
  // every advance_timestep): update the arrays in place, keeping every buffer
  // and captured step graph, with no allocation and no free.
  const bool in_place = c->x && !c->mesh && c->nv == nv && c->nf == nf && (int64_t)c->fmask_host.size() == nf &&
                        std::equal(c->fmask_host.begin(), c->fmask_host.end(), fmask);
  if (!in_place) {
    free_ctx_memory(c);
    c->eb = EvalBuffers();
    c->nv = nv;
    c->nf = nf;
    c->fmap = std::move(fmap);
    c->ns = (int)ns;
    c->nt = (int)(nv + ns);
    c->fmask_host.assign(fmask, fmask + nf);
  }
  c->face_x_host.assign(fx, fx + 3 * nf);
  c->face_zeta_host.assign(fzeta, fzeta + nf);
  const int nt = c->nt;
  // host staging: SoA, target-indexed (vertices, then active faces in tree order)
  std::vector<double> h(5 * (size_t)nt);
  double* harea = h.data() + 4 * (size_t)nt;
  for (int64_t i = 0; i < nv; ++i) {
    for (int k = 0; k < 3; ++k) h[k * (size_t)nt + i] = vx[3 * i + k];
    h[3 * (size_t)nt + i] = vzeta[i];
    harea[i] = 0.0;
  }
  for (int j = 0; j < c->ns; ++j) {
    const int64_t f = c->fmap[j], i = nv + j;
    for (int k = 0; k < 3; ++k) h[k * (size_t)nt + i] = fx[3 * f + k];
    h[3 * (size_t)nt + i] = fzeta[f];
    harea[i] = farea[f];
  }
  const size_t d = sizeof(double) * (size_t)nt;
  if (!in_place) {
    std::vector<int> hsmap(c->ns), hself(nt, -1);
    for (int j = 0; j < c->ns; ++j) {
      hsmap[j] = (int)(nv + j);
      hself[nv + j] = j;
    }
    int rc = ctx_alloc_state(c, true);
    if (rc) return rc;
    LPM_HIP(hipMemcpyAsync(c->smap, hsmap.data(), sizeof(int) * c->ns, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(hipMemcpyAsync(c->self_slot, hself.data(), sizeof(int) * nt, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(hipStreamSynchronize(c->stream));
  }
  // positions (one [3][nt] block), vorticity, areas
  LPM_HIP(hipMemcpyAsync(c->x, h.data(), 3 * d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipMemcpyAsync(c->zeta, h.data() + 3 * (size_t)nt, d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipMemcpyAsync(c->area, harea, d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  c->vel_valid = false;
  c->psi_valid = false;
  c->absv_valid = false;
  return ok();
}

static_assert(LPM_SEED_ICOS_TRI == lpm::kSeedIcosTri && LPM_SEED_CUBED_SPHERE == lpm::kSeedCubedSphere,
              "seed kinds of the C ABI and the mesh builder agree");

int lpm_bve_set_state_level(lpm_bve_ctx* c, int level, const lpm_ic_params* ic) {
  return lpm_bve_set_state_seed_level(c, LPM_SEED_ICOS_TRI, level, ic);
}

int lpm_bve_set_state_seed_level(lpm_bve_ctx* c, int seed_kind, int level, const lpm_ic_params* ic) {
  if (!c || !ic) return fail(LPM_ERR_ARG, "lpm_bve_set_state_level: null argument");
  if (seed_kind != LPM_SEED_ICOS_TRI && seed_kind != LPM_SEED_CUBED_SPHERE)
    return fail(LPM_ERR_ARG, "lpm_bve_set_state_level: unknown seed kind");
  if (level < 0 || level > 10) return fail(LPM_ERR_ARG, "lpm_bve_set_state_level: level must be in [0, 10]");
  if (ic->nvort < 0 || (ic->nvort > 0 && (!ic->centers || !ic->amps)) || !(ic->sigma > 0.0))
    return fail(LPM_ERR_ARG, "lpm_bve_set_state_level: bad field parameters");
  LPM_HIP(hipSetDevice(c->device));
  LPM_HIP(hipStreamSynchronize(c->stream));
  free_ctx_memory(c);
  c->eb = EvalBuffers();
  MeshEntry* m = nullptr;
  int rc = get_mesh(c->device, seed_kind, level, &m);
  if (rc) return rc;
  c->mesh = m;
  c->nv = m->nv;
  c->nf = m->nf;
  c->ns = m->ns;
  c->nt = m->nt;
  c->fmap.clear();
  c->fmask_host.clear();
  c->face_x_host.clear();
  c->face_zeta_host.clear();
  if ((rc = ctx_alloc_state(c, false))) return rc;
  c->area = m->area;
  c->smap = m->smap;
  c->self_slot = m->self_slot;
  const int nt = c->nt;
  LPM_HIP(hipMemcpyAsync(c->x, m->pos, 3 * sizeof(double) * (size_t)nt, hipMemcpyDeviceToDevice, c->stream));
  // field parameters: vortices as {centre, amplitude}; the subtracted mean
  const int nv = ic->nvort;
  std::vector<double4> hv(std::max(nv, 1));
  for (int k = 0; k < nv; ++k)
    hv[k] = make_double4(ic->centers[3 * k], ic->centers[3 * k + 1], ic->centers[3 * k + 2], ic->amps[k]);
  double4* dv = nullptr;
  double* ic_scratch = nullptr;
  if (hipMalloc(&dv, sizeof(double4) * hv.size()) != hipSuccess) return fail(LPM_ERR_NOMEM, "device allocation failed");
  c->allocations.push_back(dv);
  const size_t scratch_bytes = sizeof(double2) * dev::kMomentBlocks + 256;
  if (hipMalloc(&ic_scratch, scratch_bytes) != hipSuccess) return fail(LPM_ERR_NOMEM, "device allocation failed");
  c->allocations.push_back(ic_scratch);
  LPM_HIP(hipMemcpyAsync(dv, hv.data(), sizeof(double4) * hv.size(), hipMemcpyHostToDevice, c->stream));
  IcArgs a;
  a.sbr = ic->sbr;
  a.rh_n = ic->rh_n;
  a.rh_w = ic->rh_w;
  a.rh_K = ic->rh_K;
  a.nvort = nv;
  a.vort = dv;
  a.s2 = ic->sigma * ic->sigma;
  const dim3 grid(cdiv(nt, kBlock));
  hipLaunchKernelGGL(k_ic_eval, grid, dim3(kBlock), 0, c->stream, Pos3{c->x, 1, (long)nt}, nt, a, c->zeta);
  double* mean = ic_scratch;
  double2* part = reinterpret_cast<double2*>(ic_scratch + 32);
  c->ic_centered = ic->subtract_mean != 0;
  if (c->ic_centered) {
    hipLaunchKernelGGL(k_zeta_moments, dim3(kMomentBlocks), dim3(kBlock), 0, c->stream, c->zeta, c->area, nt, part);
    hipLaunchKernelGGL(k_zeta_mean, dim3(1), dim3(64), 0, c->stream, part, mean);
    hipLaunchKernelGGL(k_sub_mean, grid, dim3(kBlock), 0, c->stream, c->zeta, nt, mean);
  }
  LPM_HIP(hipGetLastError());
  LPM_HIP(hipStreamSynchronize(c->stream));
  c->ic_mean = mean;
  c->ic_sbr = ic->sbr;
  c->ic_rh_n = ic->rh_n;
  c->ic_rh_w = ic->rh_w;
  c->ic_rh_K = ic->rh_K;
  c->ic_sigma = ic->sigma;
  c->ic_centers.assign(ic->centers, ic->centers + 3 * (size_t)nv);
  c->ic_amps.assign(ic->amps, ic->amps + nv);
  c->face_zeta_pending = true;
  c->vel_valid = false;
  c->psi_valid = false;
  c->absv_valid = false;
  return ok();
}

int lpm_mesh_cache_clear(int device, int* kept) {
  std::lock_guard<std::mutex> g(g_mesh_mu);
  int n_kept = 0;
  for (auto it = g_meshes.begin(); it != g_meshes.end();) {
    MeshEntry* m = it->second;
    if ((device >= 0 && m->device != device) || m->users > 0) {
      n_kept += m->users > 0;
      ++it;
      continue;
    }
    (void)hipSetDevice(m->device);
    (void)hipFree(m->pos);
    (void)hipFree(m->area);
    (void)hipFree(m->smap);
    (void)hipFree(m->self_slot);
    delete m;
    it = g_meshes.erase(it);
  }
  if (kept) *kept = n_kept;
  return ok();
}

int lpm_bve_set_velocity(lpm_bve_ctx* c, const double* vu, const double* fu) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_set_velocity: no state");
  if ((c->nv && !vu) || !fu) return fail(LPM_ERR_ARG, "lpm_bve_set_velocity: null array");
  LPM_HIP(hipSetDevice(c->device));
  const int nt = c->nt;
  std::vector<double> h(3 * (size_t)nt);
  for (int64_t i = 0; i < c->nv; ++i)
    for (int k = 0; k < 3; ++k) h[k * (size_t)nt + i] = vu[3 * i + k];
  const std::vector<int64_t>& fmap = ctx_fmap(c);
  for (int j = 0; j < c->ns; ++j)
    for (int k = 0; k < 3; ++k) h[k * (size_t)nt + c->nv + j] = fu[3 * fmap[j] + k];
  const size_t d = sizeof(double) * nt;
  LPM_HIP(hipMemcpyAsync(c->u, h.data(), d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipMemcpyAsync(c->v, h.data() + nt, d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipMemcpyAsync(c->w, h.data() + 2 * (size_t)nt, d, hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  c->vel_valid = true;
  c->vel_prec = -1;  // caller-provided: accepted for any scheme
  c->vel_form = -1;
  return ok();
}

int lpm_bve_streamfn(lpm_bve_ctx* c, int prec, int scheme) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_streamfn: no state");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C) return fail(LPM_ERR_ARG, "lpm_bve_streamfn: bad precision");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR) return fail(LPM_ERR_ARG, "lpm_bve_streamfn: bad scheme");
  LPM_HIP(hipSetDevice(c->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  EvalRequest q;
  int rc = ctx_request(c, c->x, c->zeta, prec, form, true, q);
  if (rc) return rc;
  // the pair kernels sum velocity and stream function together; the
  // velocity goes to the fourth position increment (free between steps), so
  // the entry velocity of the next step stays the one the last step left
  q.u = MutPos3{c->kx[3], 1, (long)c->nt};
  if ((rc = enqueue_eval(q, c->eb, c->stream, nullptr, nullptr))) return rc;
  c->psi_valid = true;
  if ((rc = ctx_mark_done(c))) return rc;
  return ok();
}

int lpm_bve_solve(lpm_bve_ctx* c, int prec, int scheme, int want_psi) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_solve: no state");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C) return fail(LPM_ERR_ARG, "lpm_bve_solve: bad precision");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR) return fail(LPM_ERR_ARG, "lpm_bve_solve: bad scheme");
  LPM_HIP(hipSetDevice(c->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  int rc = ctx_eval(c, c->x, c->zeta, prec, form, want_psi != 0, c->stream);
  if (rc) return rc;
  c->vel_valid = true;
  c->vel_prec = prec;
  c->vel_form = form;
  c->psi_valid = want_psi != 0;
  if ((rc = ctx_mark_done(c))) return rc;
  return ok();
}

int lpm_bve_advance(lpm_bve_ctx* c, double dt, double omega, int scheme, int prec, int nsteps, int use_graph) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_advance: no state");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR) return fail(LPM_ERR_ARG, "lpm_bve_advance: bad scheme");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C) return fail(LPM_ERR_ARG, "lpm_bve_advance: bad precision");
  if (prec == LPM_PREC_FP32C && scheme != LPM_SCHEME_LAGR)
    return fail(LPM_ERR_ARG, "lpm_bve_advance: FP32C pair sums require the renormalised (LAGR) scheme");
  if (nsteps < 0) return fail(LPM_ERR_ARG, "lpm_bve_advance: nsteps < 0");
  if (!(dt == dt) || !(omega == omega)) return fail(LPM_ERR_ARG, "lpm_bve_advance: dt/omega is NaN");
  LPM_HIP(hipSetDevice(c->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  int rc;
  if ((rc = ctx_prepare_step(c, omega, scheme, prec, c->stream))) return rc;
  const Shape sh = ctx_shape(c);
  if (use_graph && !c->timing && nsteps > 0) {
    lpm_bve_ctx::GraphKey key{dt, omega, scheme, prec, sh.R, sh.splits, sh.near_factor, sh.lds};
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
      hipGraph_t g;
      LPM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = ctx_enqueue_step(c, dt, omega, scheme, prec, c->stream);
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return fail(LPM_ERR_HIP, std::string("step capture failed: ") + hipGetErrorString(ce));
      hipGraphExec_t ge;
      hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(LPM_ERR_HIP, std::string("graph instantiate failed: ") + hipGetErrorString(ie));
      it = c->graphs.emplace(key, ge).first;
    }
    for (int s = 0; s < nsteps; ++s) LPM_HIP(hipGraphLaunch(it->second, c->stream));
  } else {
    for (int s = 0; s < nsteps; ++s)
      if ((rc = ctx_enqueue_step(c, dt, omega, scheme, prec, c->stream))) return rc;
  }
  c->vel_prec = prec;
  c->vel_form = form;
  c->psi_valid = false;
  if ((rc = ctx_mark_done(c))) return rc;
  return ok();
}

int lpm_bve_get_state(lpm_bve_ctx* c, double* vx, double* vzeta, double* vu, double* vpsi, double* fx,
                      double* fzeta, double* fu, double* fpsi) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_get_state: no state");
  LPM_HIP(hipSetDevice(c->device));
  const int nt = c->nt;
  const size_t d = sizeof(double) * nt;
  std::vector<double> h(8 * (size_t)nt);
  double* srcs[8] = {c->x, c->y, c->z, c->zeta, c->u, c->v, c->w, c->psi};
  for (int k = 0; k < 8; ++k) LPM_HIP(hipMemcpyAsync(h.data() + k * (size_t)nt, srcs[k], d, hipMemcpyDeviceToHost, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  auto col = [&](int k, int64_t i) { return h[k * (size_t)nt + i]; };
  for (int64_t i = 0; i < c->nv; ++i) {
    for (int k = 0; k < 3; ++k) {
      if (vx) vx[3 * i + k] = col(k, i);
      if (vu) vu[3 * i + k] = col(4 + k, i);
    }
    if (vzeta) vzeta[i] = col(3, i);
    if (vpsi) vpsi[i] = col(7, i);
  }
  // divided faces: stored position and vorticity, zero velocity and psi
  if (fzeta) {
    const int rc = ctx_face_zeta_host(c);
    if (rc) return rc;
  }
  const std::vector<double>& fxh = c->mesh ? c->mesh->face_x_host : c->face_x_host;
  if (fx) std::memcpy(fx, fxh.data(), sizeof(double) * 3 * c->nf);
  if (fzeta) std::memcpy(fzeta, c->face_zeta_host.data(), sizeof(double) * c->nf);
  if (fu) std::memset(fu, 0, sizeof(double) * 3 * c->nf);
  if (fpsi) std::memset(fpsi, 0, sizeof(double) * c->nf);
  const std::vector<int64_t>& fmap = ctx_fmap(c);
  for (int j = 0; j < c->ns; ++j) {
    const int64_t f = fmap[j], i = c->nv + j;
    for (int k = 0; k < 3; ++k) {
      if (fx) fx[3 * f + k] = col(k, i);
      if (fu) fu[3 * f + k] = col(4 + k, i);
    }
    if (fzeta) fzeta[f] = col(3, i);
    if (fpsi) fpsi[f] = col(7, i);
  }
  return ok();
}

}  // extern "C"

namespace lpm {
namespace {

// The context's state on the host for the file writers: positions,
// vorticity, velocity and psi (zero when not valid), absolute vorticity for
// `omega` (computed as the first step at that rate would), areas and mask.
int ctx_download(lpm_bve_ctx* c, double omega, BveFields& s) {
  const int64_t nv = c->nv, nf = c->nf;
  const int nt = c->nt;
  s = BveFields();
  s.nv = nv;
  s.nf = nf;
  s.vx.resize(3 * nv);
  s.vzeta.resize(nv);
  s.vu.resize(3 * nv);
  s.vpsi.resize(nv);
  s.fx.resize(3 * nf);
  s.fzeta.resize(nf);
  s.fu.resize(3 * nf);
  s.fpsi.resize(nf);
  int rc = lpm_bve_get_state(c, s.vx.data(), s.vzeta.data(), s.vu.data(), s.vpsi.data(), s.fx.data(), s.fzeta.data(),
                             s.fu.data(), s.fpsi.data());
  if (rc) return rc;
  if ((rc = ctx_ensure_absv(c, omega, c->stream))) return rc;
  std::vector<double> absv(nt), area(nt);
  LPM_HIP(hipMemcpyAsync(absv.data(), c->absv, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream));
  LPM_HIP(hipMemcpyAsync(area.data(), c->area, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  s.vabsv.assign(absv.begin(), absv.begin() + nv);
  // divided faces: not evaluated, masked, area 0, absolute vorticity from
  // their stored vorticity and position
  s.fabsv.resize(nf);
  s.farea.assign(nf, 0.0);
  s.fmask.assign(nf, 1);
  for (int64_t f = 0; f < nf; ++f) s.fabsv[f] = s.fzeta[f] + 2.0 * omega * s.fx[3 * f + 2];
  const std::vector<int64_t>& fmap = ctx_fmap(c);
  for (int j = 0; j < c->ns; ++j) {
    const int64_t f = fmap[j];
    s.fmask[f] = 0;
    s.farea[f] = area[nv + j];
    s.fabsv[f] = absv[nv + j];
  }
  if (!c->psi_valid) {
    std::fill(s.vpsi.begin(), s.vpsi.end(), 0.0);
    std::fill(s.fpsi.begin(), s.fpsi.end(), 0.0);
  }
  if (!c->vel_valid) {
    std::fill(s.vu.begin(), s.vu.end(), 0.0);
    std::fill(s.fu.begin(), s.fu.end(), 0.0);
  }
  s.omega = omega;
  s.vel_prec = c->vel_valid ? c->vel_prec : -2;
  s.vel_form = c->vel_valid ? c->vel_form : -1;
  s.psi_valid = c->psi_valid;
  s.absv_valid = true;
  s.absv_omega = omega;
  return LPM_OK;
}

// The tree the state was built on: `level` >= 0 names it (checked against the
// state's sizes), -1 takes the cached mesh's level, or none (null) for a
// state set from arrays.
int ctx_tree(const lpm_bve_ctx* c, int level, bool required, std::unique_ptr<TreeMesh>& out, const char* who) {
  out.reset();
  if (level < 0 && c->mesh) level = c->mesh->level;
  if (level < 0) {
    if (required)
      return fail(LPM_ERR_ARG, std::string(who) + ": face connectivity unknown: pass the tree level the state was built on");
    return LPM_OK;
  }
  if (level > 10) return fail(LPM_ERR_ARG, std::string(who) + ": level must be in [0, 10]");
  // the seed whose level-`level` tree has the state's sizes (the icosahedral
  // and cubed-sphere trees never share a size: 10 4^a != 6 4^b)
  int kind = -1;
  for (int k : {kSeedIcosTri, kSeedCubedSphere}) {
    int64_t nvl = 0, nel = 0, nfl = 0;
    seed_max_allocations(k, level, nvl, nel, nfl);
    if (nvl == c->nv && nfl == c->nf) kind = k;
  }
  if (kind < 0)
    return fail(LPM_ERR_ARG, std::string(who) + ": the state's sizes are not those of tree level " + std::to_string(level));
  try {
    out.reset(new TreeMesh(build_tree(kind, level)));
  } catch (const std::exception& e) {
    return fail(LPM_ERR_NOMEM, std::string(who) + ": " + e.what());
  }
  return LPM_OK;
}

// Upload vertex and tree-face values into a target-indexed device array.
int ctx_upload_targets(lpm_bve_ctx* c, double* dst, int ncomp, const double* vvals, const double* fvals) {
  const int nt = c->nt;
  std::vector<double> h((size_t)ncomp * nt);
  for (int k = 0; k < ncomp; ++k) {
    for (int64_t i = 0; i < c->nv; ++i) h[k * (size_t)nt + i] = vvals[ncomp * i + k];
    const std::vector<int64_t>& fmap = ctx_fmap(c);
    for (int j = 0; j < c->ns; ++j) h[k * (size_t)nt + c->nv + j] = fvals[ncomp * fmap[j] + k];
  }
  LPM_HIP(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  return LPM_OK;
}

}  // namespace
}  // namespace lpm

extern "C" {

int lpm_bve_checkpoint(lpm_bve_ctx* c, const char* path, int level, double time, double omega) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_checkpoint: no state");
  if (!path) return fail(LPM_ERR_ARG, "lpm_bve_checkpoint: null path");
  LPM_HIP(hipSetDevice(c->device));
  std::unique_ptr<TreeMesh> tree;
  int rc = ctx_tree(c, level, false, tree, "lpm_bve_checkpoint");
  if (rc) return rc;
  BveFields s;
  if ((rc = ctx_download(c, omega, s))) return rc;
  s.time = time;
  std::string err;
  if (!write_bve_nc(path, s, tree.get(), err)) return fail(LPM_ERR_IO, "lpm_bve_checkpoint: " + err);
  return ok();
}

int lpm_bve_restore(lpm_bve_ctx* c, const char* path, double* time, double* omega) {
  if (!c || !path) return fail(LPM_ERR_ARG, "lpm_bve_restore: null argument");
  BveFields s;
  std::string err;
  if (!read_bve_nc(path, s, err)) return fail(LPM_ERR_IO, "lpm_bve_restore: " + err);
  int rc = lpm_bve_set_state(c, s.nv, s.vx.data(), s.vzeta.data(), s.nf, s.fx.data(), s.fzeta.data(), s.farea.data(),
                             s.fmask.data());
  if (rc) return rc;
  if (s.vel_prec > -2) {
    if ((rc = ctx_upload_targets(c, c->u, 3, s.vu.data(), s.fu.data()))) return rc;
    c->vel_valid = true;
    c->vel_prec = s.vel_prec;
    c->vel_form = s.vel_form;
  }
  if (s.psi_valid) {
    if ((rc = ctx_upload_targets(c, c->psi, 1, s.vpsi.data(), s.fpsi.data()))) return rc;
    c->psi_valid = true;
  }
  if (s.absv_valid) {
    if ((rc = ctx_upload_targets(c, c->absv, 1, s.vabsv.data(), s.fabsv.data()))) return rc;
    c->absv_valid = true;
    c->absv_omega = s.absv_omega;
  }
  if (time) *time = s.time;
  if (omega) *omega = s.omega;
  return ok();
}

int lpm_bve_write_vtk(lpm_bve_ctx* c, const char* path, int level, double omega, int format) {
  return lpm_bve_write_vtk_arrays(c, path, level, omega, format, nullptr, 0);
}

int lpm_bve_write_vtk_arrays(lpm_bve_ctx* c, const char* path, int level, double omega, int format,
                             const lpm_vtk_array* arrays, int narrays) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_write_vtk: no state");
  if (!path) return fail(LPM_ERR_ARG, "lpm_bve_write_vtk: null path");
  if (format < LPM_VTK_LEGACY_ASCII || format > LPM_VTK_XML_BINARY)
    return fail(LPM_ERR_ARG, "lpm_bve_write_vtk: unknown format");
  std::vector<VtkExtra> extra;
  std::string err;
  if (!vtk_extras(arrays, narrays, extra, err)) return fail(LPM_ERR_ARG, "lpm_bve_write_vtk: " + err);
  LPM_HIP(hipSetDevice(c->device));
  std::unique_ptr<TreeMesh> tree;
  int rc = ctx_tree(c, level, true, tree, "lpm_bve_write_vtk");
  if (rc) return rc;
  BveFields s;
  if ((rc = ctx_download(c, omega, s))) return rc;
  if (!write_bve_vtk(path, s, *tree, format, err, extra)) return fail(LPM_ERR_IO, "lpm_bve_write_vtk: " + err);
  return ok();
}

int lpm_bve_circulation(lpm_bve_ctx* c, double* circ) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_circulation: no state");
  if (!circ) return fail(LPM_ERR_ARG, "lpm_bve_circulation: null output");
  LPM_HIP(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_circulation, dim3(1), dim3(kBlock), 0, c->stream, c->zeta, c->area, c->smap, c->ns, c->circ);
  LPM_HIP(hipGetLastError());
  LPM_HIP(hipMemcpyAsync(circ, c->circ, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  LPM_HIP(hipStreamSynchronize(c->stream));
  return ok();
}

int lpm_bve_synchronize(lpm_bve_ctx* c) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_synchronize: null context");
  LPM_HIP(hipSetDevice(c->device));
  LPM_HIP(hipStreamSynchronize(c->stream));
  return ok();
}

int lpm_bve_sizes(lpm_bve_ctx* c, int64_t* s) {
  if (!c || !s) return fail(LPM_ERR_ARG, "lpm_bve_sizes: null argument");
  s[0] = c->nv;
  s[1] = c->nf;
  s[2] = c->ns;
  s[3] = c->nt;
  return ok();
}

int lpm_bve_set_launch(lpm_bve_ctx* c, int targets_per_lane, int source_splits) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_set_launch: null context");
  if (targets_per_lane < 0 || targets_per_lane > 4)
    return fail(LPM_ERR_ARG, "lpm_bve_set_launch: targets_per_lane must be 0 (automatic) or 1..4");
  if (source_splits < 0) return fail(LPM_ERR_ARG, "lpm_bve_set_launch: source_splits < 0");
  c->R_req = targets_per_lane;
  c->splits_req = source_splits;
  return ok();
}

int lpm_bve_set_source_staging(lpm_bve_ctx* c, int lds) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_set_source_staging: null context");
  c->lds = lds != 0;
  return ok();
}

int lpm_bve_set_near_factor(lpm_bve_ctx* c, double factor) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_set_near_factor: null context");
  if (!(factor >= 0.0)) return fail(LPM_ERR_ARG, "lpm_bve_set_near_factor: factor must be >= 0");
  c->near_factor = (float)std::min(factor, 1e18);
  return ok();
}

int lpm_bve_stream(lpm_bve_ctx* c, void** stream, int* device) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_stream: null context");
  if (stream) *stream = (void*)c->stream;
  if (device) *device = c->device;
  return ok();
}

int lpm_bve_last_pair_timing(lpm_bve_ctx* c, int* launches, double* mean_ms) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_last_pair_timing: null context");
  LPM_HIP(hipSetDevice(c->device));
  LPM_HIP(hipStreamSynchronize(c->stream));
  double total = 0.0;
  int n = 0;
  for (auto& e : c->timing_events) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
      total += ms;
      ++n;
    }
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  c->timing_events.clear();
  if (launches) *launches = n;
  if (mean_ms) *mean_ms = n ? total / n : 0.0;
  return ok();
}

// Enable (1) or disable (0) per-launch timing events around pair kernels.
// Timed steps run without graph replay.
int lpm_bve_enable_timing(lpm_bve_ctx* c, int on) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_enable_timing: null context");
  c->timing = on != 0;
  return ok();
}

}  // extern "C"

// ===========================================================================
// ensemble: independent problems advanced together, one persistent pair
// kernel per evaluation over all members' (target block, source split) items
// ===========================================================================
struct lpm_bve_ensemble {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<lpm_bve_ctx*> members;
  int blocks_per_cu = 0;  // persistent grid: 0 = occupancy limit
  // plan (rebuilt when membership, generations or the step kind change)
  bool plan_valid = false;
  std::vector<uint64_t> plan_generations;
  // each member's launch shape when the plan was built (R, splits, tiles per
  // split, near factor): a change (lpm_bve_set_launch, lpm_bve_set_near_factor,
  // a new autotuned rule) rebuilds the plan and its graphs
  std::vector<std::tuple<int, int, int, float>> plan_shapes;
  int plan_prec = -1, plan_form = -1;
  int plan_per_cu = 0;                     // workgroups per CU the plan's grids were sized for
  std::vector<int> plan_R;                 // distinct R values, one launch each
  std::vector<int> plan_nitems;            // items per R launch
  std::vector<int> plan_grid;              // persistent workgroups per R launch
  std::vector<lpm::dev::EnsItem*> d_items;  // per R launch
  // Plan buffers grow on demand and are kept (freed at destroy): a hipFree
  // waits for the whole device, so a membership change while other work
  // runs on the GPU must not free.
  lpm::dev::EnsMember* d_mem_w = nullptr;  // descriptors: stage arguments as targets
  lpm::dev::EnsMember* d_mem_x = nullptr;  // descriptors: state as targets
  lpm::dev::EnsEval* d_eval_w = nullptr;   // prep / finalize descriptors, the same two kinds
  lpm::dev::EnsEval* d_eval_x = nullptr;
  lpm::dev::StageArgs* d_stage = nullptr;  // each member's arrays for the stage kernel
  size_t cap_mem = 0;
  // (member, target block) and (member, source tile) lists of the
  // ensemble-wide stage, finalize and prep launches
  int2* d_tblk = nullptr;
  int2* d_sblk = nullptr;
  size_t cap_tblk = 0, cap_sblk = 0;
  int n_tblk = 0, n_sblk = 0;
  std::vector<lpm::dev::EnsItem*> item_bufs;  // by launch index
  std::vector<size_t> item_caps;
  int* d_counters = nullptr;  // four per persistent launch in a step
  std::vector<void*> plan_allocs;  // every plan buffer ever allocated
  std::map<std::tuple<double, double, int, int>, hipGraphExec_t> graphs;
  hipEvent_t ev = nullptr;
  unsigned long long* done_flag = nullptr;  // completion marker (lpm_ens_query)
  unsigned long long done_seq = 0;
};

namespace lpm {
namespace {

void ens_drop_plan(lpm_bve_ensemble* e) {
  for (auto& kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
  e->graphs.clear();
  e->d_items.clear();
  e->plan_R.clear();
  e->plan_nitems.clear();
  e->plan_grid.clear();
  e->plan_valid = false;
}

void ens_free_buffers(lpm_bve_ensemble* e) {
  for (void* p : e->plan_allocs) (void)hipFree(p);
  e->plan_allocs.clear();
  e->d_mem_w = e->d_mem_x = nullptr;
  e->d_eval_w = e->d_eval_x = nullptr;
  e->d_stage = nullptr;
  e->cap_mem = 0;
  e->d_tblk = e->d_sblk = nullptr;
  e->cap_tblk = e->cap_sblk = 0;
  e->n_tblk = e->n_sblk = 0;
  e->item_bufs.clear();
  e->item_caps.clear();
  e->d_counters = nullptr;
}

std::tuple<int, int, int, float> shape_key(const lpm_bve_ctx* c) {
  const Shape sh = ctx_shape(c);
  return std::make_tuple(sh.R, sh.splits, sh.tiles_per_split, sh.near_factor);
}

int ens_blocks_per_cu(const lpm_bve_ensemble* e);

bool ens_plan_current(const lpm_bve_ensemble* e, int prec, int form) {
  if (!e->plan_valid || e->plan_prec != prec || e->plan_form != form) return false;
  if (e->plan_per_cu != ens_blocks_per_cu(e)) return false;
  if (e->plan_generations.size() != e->members.size() || e->plan_shapes.size() != e->members.size()) return false;
  for (size_t m = 0; m < e->members.size(); ++m)
    if (e->plan_generations[m] != e->members[m]->generation || e->plan_shapes[m] != shape_key(e->members[m]))
      return false;
  return true;
}

int ens_grid(lpm_bve_ensemble* e, int R, int prec, int form, int nitems);

// Build the flattened work lists and member descriptors for one kind of
// evaluation (precision and kernel form).  Items are ordered by decreasing
// cost so that the persistent workgroups take the long items first.
int ens_build_plan(lpm_bve_ensemble* e, int prec, int form) {
  ens_drop_plan(e);
  e->plan_per_cu = ens_blocks_per_cu(e);
  const size_t M = e->members.size();
  std::vector<EnsMember> hw(M), hx(M);
  std::vector<EnsEval> vw(M), vx(M);
  std::vector<StageArgs> hs(M);
  std::vector<int2> tblk, sblk;
  std::map<int, std::vector<std::pair<long, EnsItem>>> byR;
  for (size_t m = 0; m < M; ++m) {
    lpm_bve_ctx* c = e->members[m];
    hs[m] = ctx_stage_args(c);
    for (int which = 0; which < 2; ++which) {
      EvalRequest q;
      int rc = ctx_request(c, which ? c->x : c->wx, which ? c->zeta : c->wzeta, prec, form, false, q);
      if (rc) return rc;
      EnsMember& d = which ? hx[m] : hw[m];
      d.a = make_pair_args(q, c->eb);
      d.src32 = c->eb.src32;
      d.ctr = c->eb.ctr;
      d.trad = c->eb.trad;
      d.src64 = c->eb.src64;
      EnsEval& v = which ? vx[m] : vw[m];
      v.src32 = c->eb.src32;
      v.src64 = c->eb.src64;
      v.ctr = c->eb.ctr;
      v.trad = c->eb.trad;
      v.srcq = c->eb.srcq;
      v.srcpos = q.srcpos;
      v.zeta = q.zeta;
      v.area = q.area;
      v.smap = q.smap;
      v.ns = q.ns;
      v.ntiles = q.shape.ntiles;
      v.fin = make_fin_args(q, c->eb);
      if (which) {
        for (int k = 0; k < cdiv(c->nt, kBlock); ++k) tblk.push_back(make_int2((int)m, k));
        for (int t = 0; t < q.shape.ntiles; ++t) sblk.push_back(make_int2((int)m, t));
        const Shape& sh = q.shape;
        const int nbx = cdiv(c->nt, (long)kBlock * sh.R);
        for (int by = 0; by < sh.splits; ++by) {
          const int tiles = std::min(sh.tiles_per_split, sh.ntiles - by * sh.tiles_per_split);
          for (int bx = 0; bx < nbx; ++bx) {
            const int tgts = std::min(kBlock * sh.R, c->nt - bx * kBlock * sh.R);
            byR[sh.R].push_back({(long)tiles * tgts, EnsItem{(int)m, bx, by}});
          }
        }
      }
    }
  }
  auto alloc = [&](void** p, size_t n) -> int {
    if (hipMalloc(p, std::max<size_t>(n, 256)) != hipSuccess) return fail(LPM_ERR_NOMEM, "ensemble plan allocation");
    e->plan_allocs.push_back(*p);
    return LPM_OK;
  };
  int rc;
  if (M > e->cap_mem) {
    // grow (doubling); the old buffers stay allocated until destroy
    const size_t cap = std::max<size_t>(M, 2 * e->cap_mem);
    if ((rc = alloc((void**)&e->d_mem_w, sizeof(EnsMember) * cap))) return rc;
    if ((rc = alloc((void**)&e->d_mem_x, sizeof(EnsMember) * cap))) return rc;
    if ((rc = alloc((void**)&e->d_eval_w, sizeof(EnsEval) * cap))) return rc;
    if ((rc = alloc((void**)&e->d_eval_x, sizeof(EnsEval) * cap))) return rc;
    if ((rc = alloc((void**)&e->d_stage, sizeof(StageArgs) * cap))) return rc;
    e->cap_mem = cap;
  }
  if (tblk.size() > e->cap_tblk) {
    const size_t cap = std::max(tblk.size(), 2 * e->cap_tblk);
    if ((rc = alloc((void**)&e->d_tblk, sizeof(int2) * cap))) return rc;
    e->cap_tblk = cap;
  }
  if (sblk.size() > e->cap_sblk) {
    const size_t cap = std::max(sblk.size(), 2 * e->cap_sblk);
    if ((rc = alloc((void**)&e->d_sblk, sizeof(int2) * cap))) return rc;
    e->cap_sblk = cap;
  }
  e->n_tblk = (int)tblk.size();
  e->n_sblk = (int)sblk.size();
  if (!e->d_counters && (rc = alloc((void**)&e->d_counters, sizeof(int) * 4 * 4))) return rc;
  // stream-ordered uploads: the previous step (which reads these buffers) is
  // ahead of them on the ensemble's stream
  LPM_HIP(hipMemcpyAsync(e->d_mem_w, hw.data(), sizeof(EnsMember) * M, hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_mem_x, hx.data(), sizeof(EnsMember) * M, hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_eval_w, vw.data(), sizeof(EnsEval) * M, hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_eval_x, vx.data(), sizeof(EnsEval) * M, hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_stage, hs.data(), sizeof(StageArgs) * M, hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_tblk, tblk.data(), sizeof(int2) * tblk.size(), hipMemcpyHostToDevice, e->stream));
  LPM_HIP(hipMemcpyAsync(e->d_sblk, sblk.data(), sizeof(int2) * sblk.size(), hipMemcpyHostToDevice, e->stream));
  size_t k = 0;
  // one host list per launch, all alive until the uploads below have run (a
  // copy from pageable memory may still be queued behind the running step)
  std::vector<std::vector<EnsItem>> lists(byR.size());
  for (auto& kv : byR) {
    auto& v = kv.second;
    std::stable_sort(v.begin(), v.end(), [](const std::pair<long, EnsItem>& a, const std::pair<long, EnsItem>& b) {
      return a.first > b.first;
    });
    std::vector<EnsItem>& items = lists[k];
    items.reserve(v.size());
    for (auto& pr : v) items.push_back(pr.second);
    if (k >= e->item_bufs.size()) {
      e->item_bufs.push_back(nullptr);
      e->item_caps.push_back(0);
    }
    if (items.size() > e->item_caps[k]) {
      const size_t cap = std::max(items.size(), 2 * e->item_caps[k]);
      if ((rc = alloc((void**)&e->item_bufs[k], sizeof(EnsItem) * cap))) return rc;
      e->item_caps[k] = cap;
    }
    EnsItem* d = e->item_bufs[k];
    LPM_HIP(hipMemcpyAsync(d, items.data(), sizeof(EnsItem) * items.size(), hipMemcpyHostToDevice, e->stream));
    e->plan_R.push_back(kv.first);
    e->plan_nitems.push_back((int)items.size());
    e->plan_grid.push_back(ens_grid(e, kv.first, prec, form, (int)items.size()));
    e->d_items.push_back(d);
    ++k;
  }
  // the host copies above must outlive the uploads
  LPM_HIP(hipStreamSynchronize(e->stream));
  e->plan_generations.clear();
  e->plan_shapes.clear();
  for (auto* c : e->members) {
    e->plan_generations.push_back(c->generation);
    e->plan_shapes.push_back(shape_key(c));
  }
  e->plan_prec = prec;
  e->plan_form = form;
  e->plan_valid = true;
  return LPM_OK;
}

template <int R>
int ens_launch_R(lpm_bve_ensemble* e, int k, const EnsMember* mem, int* counter, int prec, int form, int grid) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, e->stream, counter);
  if (prec == LPM_PREC_FP32C)
    hipLaunchKernelGGL((k_pair_f32c_ens<R>), dim3(grid), dim3(kBlock), 0, e->stream, mem, e->d_items[k],
                       e->plan_nitems[k], counter);
  else if (form == FORM_DOT)
    hipLaunchKernelGGL((k_pair_f64_ens<R, FORM_DOT>), dim3(grid), dim3(kBlock), 0, e->stream, mem, e->d_items[k],
                       e->plan_nitems[k], counter);
  else
    hipLaunchKernelGGL((k_pair_f64_ens<R, FORM_CHORD>), dim3(grid), dim3(kBlock), 0, e->stream, mem,
                       e->d_items[k], e->plan_nitems[k], counter);
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

template <int R>
const void* ens_kernel(int prec, int form) {
  if (prec == LPM_PREC_FP32C) return (const void*)&k_pair_f32c_ens<R>;
  if (form == FORM_DOT) return (const void*)&k_pair_f64_ens<R, FORM_DOT>;
  return (const void*)&k_pair_f64_ens<R, FORM_CHORD>;
}

int64_t ens_total_targets(const lpm_bve_ensemble* e) {
  int64_t n = 0;
  for (const auto* c : e->members) n += c->nt;
  return n;
}

// Workgroups per CU an ensemble runs its persistent kernels with: its own
// setting, else the autotuned rule for its total N, else 0 (occupancy).
int ens_blocks_per_cu(const lpm_bve_ensemble* e) {
  if (e->blocks_per_cu > 0) return e->blocks_per_cu;
  std::lock_guard<std::mutex> g(g_rule_mu);
  auto it = g_ens_rules.find(ens_total_targets(e));
  return it == g_ens_rules.end() ? 0 : it->second;
}

// Occupancy limit (workgroups per CU) of the persistent kernel for R.
int ens_occupancy(int R, int prec, int form) {
  const void* k = R == 1   ? ens_kernel<1>(prec, form)
                  : R == 3 ? ens_kernel<3>(prec, form)
                  : R == 4 ? ens_kernel<4>(prec, form)
                           : ens_kernel<2>(prec, form);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
  return per_cu;
}

int ens_grid(lpm_bve_ensemble* e, int R, int prec, int form, int nitems) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->device) != hipSuccess) return 256;
  int per_cu = ens_blocks_per_cu(e);
  if (per_cu <= 0) per_cu = ens_occupancy(R, prec, form);
  return std::max(1, std::min(nitems, per_cu * prop.multiProcessorCount));
}

// One evaluation of every member: one prep launch over all members' source
// tiles, one persistent pair launch per distinct R over all members' items,
// one finalize launch over all members' targets.  which: 0 = stage
// arguments, 1 = state.
int ens_eval(lpm_bve_ensemble* e, int which, int prec, int form, int pass) {
  int rc;
  const EnsEval* ev = which ? e->d_eval_x : e->d_eval_w;
  const dim3 sg(e->n_sblk), tg(e->n_tblk);
  if (prec == LPM_PREC_FP32C)
    hipLaunchKernelGGL(k_prep_ens<PREP_F32C>, sg, dim3(kTile), 0, e->stream, ev, e->d_sblk);
  else if (form == FORM_DOT)
    hipLaunchKernelGGL(k_prep_ens<PREP_F64_DOT>, sg, dim3(kTile), 0, e->stream, ev, e->d_sblk);
  else
    hipLaunchKernelGGL(k_prep_ens<PREP_F64_TILES>, sg, dim3(kTile), 0, e->stream, ev, e->d_sblk);
  LPM_HIP(hipGetLastError());
  const EnsMember* mem = which ? e->d_mem_x : e->d_mem_w;
  for (size_t k = 0; k < e->plan_R.size(); ++k) {
    int* counter = e->d_counters + 4 * k + pass;
    const int R = e->plan_R[k];
    const int grid = e->plan_grid[k];
    if (R == 1) rc = ens_launch_R<1>(e, (int)k, mem, counter, prec, form, grid);
    else if (R == 3) rc = ens_launch_R<3>(e, (int)k, mem, counter, prec, form, grid);
    else if (R == 4) rc = ens_launch_R<4>(e, (int)k, mem, counter, prec, form, grid);
    else rc = ens_launch_R<2>(e, (int)k, mem, counter, prec, form, grid);
    if (rc) return rc;
  }
  if (prec == LPM_PREC_FP32C)
    hipLaunchKernelGGL(k_finalize_ens<FIN_F32C>, tg, dim3(kBlock), 0, e->stream, ev, e->d_tblk);
  else if (form == FORM_DOT)
    hipLaunchKernelGGL(k_finalize_ens<FIN_DOT>, tg, dim3(kBlock), 0, e->stream, ev, e->d_tblk);
  else
    hipLaunchKernelGGL(k_finalize_ens<FIN_CHORD>, tg, dim3(kBlock), 0, e->stream, ev, e->d_tblk);
  LPM_HIP(hipGetLastError());
  return LPM_OK;
}

int ens_enqueue_step(lpm_bve_ensemble* e, double dt, double omega, int scheme, int prec) {
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  int rc;
  for (int stage = 1; stage <= 4; ++stage) {
    hipLaunchKernelGGL(k_stage_ens, dim3(e->n_tblk), dim3(kBlock), 0, e->stream, e->d_stage, e->d_tblk, stage,
                       scheme == LPM_SCHEME_LAGR ? 1 : 0, dt, omega);
    LPM_HIP(hipGetLastError());
    if ((rc = ens_eval(e, stage == 4 ? 1 : 0, prec, form, stage - 1))) return rc;
  }
  return LPM_OK;
}

}  // namespace
}  // namespace lpm

extern "C" {

int lpm_ens_create(int device, lpm_bve_ensemble** out) {
  if (!out) return fail(LPM_ERR_ARG, "lpm_ens_create: null output");
  int n = 0;
  LPM_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(LPM_ERR_ARG, "lpm_ens_create: no such device");
  LPM_HIP(hipSetDevice(device));
  auto* e = new lpm_bve_ensemble();
  e->device = device;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess) {
    delete e;
    return fail(LPM_ERR_HIP, "lpm_ens_create: stream/event creation failed");
  }
  *out = e;
  return ok();
}

int lpm_ens_stream(lpm_bve_ensemble* e, void** stream) {
  if (!e || !stream) return fail(LPM_ERR_ARG, "lpm_ens_stream: null argument");
  *stream = (void*)e->stream;
  return ok();
}

int lpm_ens_destroy(lpm_bve_ensemble* e) {
  if (!e) return ok();
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->stream);
  for (auto* c : e->members)
    if (c->done_flag_shared == e->done_flag) c->done_flag_shared = nullptr;
  ens_drop_plan(e);
  ens_free_buffers(e);
  (void)hipEventDestroy(e->ev);
  if (e->done_flag) (void)hipHostFree(e->done_flag);
  (void)hipStreamDestroy(e->stream);
  delete e;
  return ok();
}

int lpm_ens_add(lpm_bve_ensemble* e, lpm_bve_ctx* c) {
  if (!e || !c) return fail(LPM_ERR_ARG, "lpm_ens_add: null argument");
  if (!c->x) return fail(LPM_ERR_STATE, "lpm_ens_add: member has no state");
  if (c->device != e->device) return fail(LPM_ERR_ARG, "lpm_ens_add: member is on another device");
  if (std::find(e->members.begin(), e->members.end(), c) != e->members.end())
    return fail(LPM_ERR_ARG, "lpm_ens_add: already a member");
  e->members.push_back(c);
  e->plan_valid = false;
  return ok();
}

int lpm_ens_remove(lpm_bve_ensemble* e, lpm_bve_ctx* c) {
  if (!e || !c) return fail(LPM_ERR_ARG, "lpm_ens_remove: null argument");
  auto it = std::find(e->members.begin(), e->members.end(), c);
  if (it == e->members.end()) return fail(LPM_ERR_ARG, "lpm_ens_remove: not a member");
  LPM_HIP(hipSetDevice(e->device));
  LPM_HIP(hipStreamSynchronize(e->stream));
  if (c->done_flag_shared == e->done_flag) c->done_flag_shared = nullptr;  // its steps have finished
  e->members.erase(it);
  e->plan_valid = false;
  return ok();
}

int lpm_ens_size(lpm_bve_ensemble* e, int* n) {
  if (!e || !n) return fail(LPM_ERR_ARG, "lpm_ens_size: null argument");
  *n = (int)e->members.size();
  return ok();
}

int lpm_ens_set_blocks_per_cu(lpm_bve_ensemble* e, int n) {
  if (!e || n < 0) return fail(LPM_ERR_ARG, "lpm_ens_set_blocks_per_cu: bad argument");
  LPM_HIP(hipSetDevice(e->device));
  LPM_HIP(hipStreamSynchronize(e->stream));
  e->blocks_per_cu = n;
  ens_drop_plan(e);
  return ok();
}

// Time one evaluation of every member at its state per candidate persistent
// grid (1 .. the occupancy limit workgroups per CU) and record the fastest as
// the rule for ensembles of this total target count.
int lpm_ens_autotune(lpm_bve_ensemble* e, int prec, int scheme, int reps, int* best_blocks_per_cu, double* best_ms,
                     double* ms_by_blocks, int nms) {
  if (!e) return fail(LPM_ERR_ARG, "lpm_ens_autotune: null ensemble");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR) return fail(LPM_ERR_ARG, "lpm_ens_autotune: bad scheme");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C) return fail(LPM_ERR_ARG, "lpm_ens_autotune: bad precision");
  if (prec == LPM_PREC_FP32C && scheme != LPM_SCHEME_LAGR)
    return fail(LPM_ERR_ARG, "lpm_ens_autotune: FP32C pair sums require the renormalised (LAGR) scheme");
  if (nms < 0 || (nms > 0 && !ms_by_blocks)) return fail(LPM_ERR_ARG, "lpm_ens_autotune: bad time array");
  if (e->members.empty()) return fail(LPM_ERR_STATE, "lpm_ens_autotune: no members");
  if (reps < 1) reps = 1;
  LPM_HIP(hipSetDevice(e->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  for (int k = 0; k < nms; ++k) ms_by_blocks[k] = -1.0;
  // order after each member's earlier work, as lpm_ens_advance does
  for (auto* c : e->members) {
    if (c->timing) return fail(LPM_ERR_STATE, "lpm_ens_autotune: member has pair timing enabled");
    if (c->stream == e->stream || mark_reached(c->done_flag, c->done_seq)) continue;
    LPM_HIP(hipEventRecord(e->ev, c->stream));
    LPM_HIP(hipStreamWaitEvent(e->stream, e->ev, 0));
  }
  int rc, occ = 1;
  for (auto* c : e->members) {
    if ((rc = ensure_eval_buffers(c, ctx_shape(c)))) return rc;
    occ = std::max(occ, ens_occupancy(ctx_shape(c).R, prec, form));
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) {
    if (t0) (void)hipEventDestroy(t0);
    return fail(LPM_ERR_HIP, "lpm_ens_autotune: event creation failed");
  }
  const int saved = e->blocks_per_cu;
  std::vector<double> times(occ + 1, -1.0);
  rc = LPM_OK;
  for (int n = 1; n <= occ && rc == LPM_OK; ++n) {
    e->blocks_per_cu = n;
    if ((rc = ens_build_plan(e, prec, form))) break;
    if ((rc = ens_eval(e, 1, prec, form, 0))) break;  // warm
    if (hipEventRecord(t0, e->stream) != hipSuccess) rc = fail(LPM_ERR_HIP, "lpm_ens_autotune: event record");
    for (int r = 0; r < reps && rc == LPM_OK; ++r) rc = ens_eval(e, 1, prec, form, 0);
    if (rc) break;
    float ms = 0.0f;
    if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
        hipEventElapsedTime(&ms, t0, t1) != hipSuccess) {
      rc = fail(LPM_ERR_HIP, "lpm_ens_autotune: timing failed");
      break;
    }
    times[n] = ms / reps;
    if (n <= nms) ms_by_blocks[n - 1] = times[n];
  }
  // the fewest workgroups per CU within 0.5 % of the fastest: candidates past
  // the occupancy limit only queue workgroups, and differences that small are
  // run-to-run noise
  double best = 1e300;
  for (int n = 1; n <= occ; ++n)
    if (times[n] >= 0.0) best = std::min(best, times[n]);
  int best_n = occ;
  for (int n = occ; n >= 1; --n)
    if (times[n] >= 0.0 && times[n] <= 1.005 * best) best_n = n;
  (void)hipStreamSynchronize(e->stream);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  e->blocks_per_cu = saved;
  ens_drop_plan(e);
  // the trial evaluations overwrote each member's velocity; its next step
  // recomputes the entry velocity
  for (auto* c : e->members) c->vel_valid = false;
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> g(g_rule_mu);
    g_ens_rules[ens_total_targets(e)] = best_n;
  }
  if (best_blocks_per_cu) *best_blocks_per_cu = best_n;
  if (best_ms) *best_ms = times[best_n];
  return ok();
}

// Launch shapes tuned for the ensemble: for the member sizes N that carry the
// most pair work (until `work_fraction` of it is covered), try launch shapes
// for that N on one evaluation of the whole ensemble and keep the fastest as
// the rule for N.  Members of one N share a rule, and a context run alone
// takes the same rule, so results stay bit-identical alone and in the
// ensemble.
int lpm_ens_autotune_shapes(lpm_bve_ensemble* e, int prec, int scheme, int reps, double work_fraction,
                            int* ntuned) {
  if (!e) return fail(LPM_ERR_ARG, "lpm_ens_autotune_shapes: null ensemble");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR)
    return fail(LPM_ERR_ARG, "lpm_ens_autotune_shapes: bad scheme");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C)
    return fail(LPM_ERR_ARG, "lpm_ens_autotune_shapes: bad precision");
  if (prec == LPM_PREC_FP32C && scheme != LPM_SCHEME_LAGR)
    return fail(LPM_ERR_ARG, "lpm_ens_autotune_shapes: FP32C pair sums require the renormalised (LAGR) scheme");
  if (!(work_fraction > 0.0)) return fail(LPM_ERR_ARG, "lpm_ens_autotune_shapes: work_fraction must be > 0");
  if (e->members.empty()) return fail(LPM_ERR_STATE, "lpm_ens_autotune_shapes: no members");
  if (reps < 1) reps = 1;
  LPM_HIP(hipSetDevice(e->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  for (auto* c : e->members) {
    if (c->timing) return fail(LPM_ERR_STATE, "lpm_ens_autotune_shapes: member has pair timing enabled");
    if (c->R_req || c->splits_req)
      return fail(LPM_ERR_STATE, "lpm_ens_autotune_shapes: member has an explicit launch shape");
    if (c->stream == e->stream || mark_reached(c->done_flag, c->done_seq)) continue;
    LPM_HIP(hipEventRecord(e->ev, c->stream));
    LPM_HIP(hipStreamWaitEvent(e->stream, e->ev, 0));
  }
  // pair work per member size N, largest first
  std::map<int, double> work;
  std::map<int, int> nsrc;
  double total = 0.0;
  for (auto* c : e->members) {
    const double w = (double)c->nt * c->ns;
    work[c->nt] += w;
    nsrc[c->nt] = c->ns;
    total += w;
  }
  std::vector<std::pair<double, int>> order;
  for (auto& kv : work) order.push_back({kv.second, kv.first});
  std::sort(order.rbegin(), order.rend());
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) {
    if (t0) (void)hipEventDestroy(t0);
    return fail(LPM_ERR_HIP, "lpm_ens_autotune_shapes: event creation failed");
  }
  auto time_eval = [&](double* ms_out) -> int {
    int rc;
    for (auto* c : e->members)
      if ((rc = ensure_eval_buffers(c, ctx_shape(c)))) return rc;
    if ((rc = ens_build_plan(e, prec, form))) return rc;
    if ((rc = ens_eval(e, 1, prec, form, 0))) return rc;  // warm
    LPM_HIP(hipEventRecord(t0, e->stream));
    for (int r = 0; r < reps; ++r)
      if ((rc = ens_eval(e, 1, prec, form, 0))) return rc;
    LPM_HIP(hipEventRecord(t1, e->stream));
    LPM_HIP(hipEventSynchronize(t1));
    float ms = 0.0f;
    LPM_HIP(hipEventElapsedTime(&ms, t0, t1));
    *ms_out = ms / reps;
    return LPM_OK;
  };
  int rc = LPM_OK, tuned = 0;
  double covered = 0.0;
  for (auto& ow : order) {
    if (covered >= work_fraction * total) break;
    covered += ow.first;
    const int N = ow.second;
    const int ns = nsrc[N];
    const Shape base = choose_shape(N, std::max(ns, 1), 0, 0);
    std::vector<std::pair<int, int>> cands;
    for (int R : {base.R, 2, 4}) {
      for (int sp : {base.splits, 8, 16, 32, 64}) {
        const Shape sh = choose_shape(N, std::max(ns, 1), R, sp);
        const std::pair<int, int> c2(sh.R, sh.splits);
        if (std::find(cands.begin(), cands.end(), c2) == cands.end()) cands.push_back(c2);
      }
    }
    double best = 1e300;
    std::pair<int, int> bestc = cands.front();
    for (auto& cd : cands) {
      {
        std::lock_guard<std::mutex> g(g_rule_mu);
        g_rules[N] = ShapeRule{cd.first, cd.second};
      }
      double ms = 0.0;
      if ((rc = time_eval(&ms))) break;
      if (ms < best) {
        best = ms;
        bestc = cd;
      }
    }
    {
      std::lock_guard<std::mutex> g(g_rule_mu);
      g_rules[N] = ShapeRule{bestc.first, bestc.second};
    }
    if (rc) break;
    ++tuned;
  }
  (void)hipStreamSynchronize(e->stream);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  ens_drop_plan(e);
  // the trial evaluations overwrote each member's velocity; its next step
  // recomputes the entry velocity
  for (auto* c : e->members) c->vel_valid = false;
  if (rc) return rc;
  if (ntuned) *ntuned = tuned;
  return ok();
}

int lpm_ens_rules_clear(void) {
  std::lock_guard<std::mutex> g(g_rule_mu);
  g_ens_rules.clear();
  return ok();
}

int lpm_ens_advance(lpm_bve_ensemble* e, double dt, double omega, int scheme, int prec, int nsteps, int use_graph) {
  if (!e) return fail(LPM_ERR_ARG, "lpm_ens_advance: null ensemble");
  if (scheme != LPM_SCHEME_REF && scheme != LPM_SCHEME_LAGR) return fail(LPM_ERR_ARG, "lpm_ens_advance: bad scheme");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32C) return fail(LPM_ERR_ARG, "lpm_ens_advance: bad precision");
  if (prec == LPM_PREC_FP32C && scheme != LPM_SCHEME_LAGR)
    return fail(LPM_ERR_ARG, "lpm_ens_advance: FP32C pair sums require the renormalised (LAGR) scheme");
  if (nsteps < 0) return fail(LPM_ERR_ARG, "lpm_ens_advance: nsteps < 0");
  if (e->members.empty() || nsteps == 0) return ok();
  LPM_HIP(hipSetDevice(e->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  // order after each member's earlier asynchronous work (its completion
  // event, if still pending; members on the ensemble's stream are ordered by
  // the stream), then prepare every member
  for (auto* c : e->members) {
    if (c->timing) return fail(LPM_ERR_STATE, "lpm_ens_advance: member has pair timing enabled");
    if (c->stream == e->stream || mark_reached(c->done_flag, c->done_seq)) continue;
    LPM_HIP(hipEventRecord(e->ev, c->stream));
    LPM_HIP(hipStreamWaitEvent(e->stream, e->ev, 0));
  }
  int rc;
  for (auto* c : e->members)
    if ((rc = ctx_prepare_step(c, omega, scheme, prec, e->stream))) return rc;
  if (!ens_plan_current(e, prec, form))
    if ((rc = ens_build_plan(e, prec, form))) return rc;
  if (use_graph) {
    const auto key = std::make_tuple(dt, omega, scheme, prec);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
      hipGraph_t g;
      LPM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      rc = ens_enqueue_step(e, dt, omega, scheme, prec);
      hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return fail(LPM_ERR_HIP, std::string("ensemble capture failed: ") + hipGetErrorString(ce));
      hipGraphExec_t ge;
      hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(LPM_ERR_HIP, std::string("graph instantiate failed: ") + hipGetErrorString(ie));
      it = e->graphs.emplace(key, ge).first;
    }
    for (int s = 0; s < nsteps; ++s) LPM_HIP(hipGraphLaunch(it->second, e->stream));
  } else {
    for (int s = 0; s < nsteps; ++s)
      if ((rc = ens_enqueue_step(e, dt, omega, scheme, prec))) return rc;
  }
  // later work on each member's own stream waits for the ensemble's
  LPM_HIP(hipEventRecord(e->ev, e->stream));
  if ((rc = queue_mark(&e->done_flag, &e->done_seq, e->stream))) return rc;
  for (auto* c : e->members) {
    if (c->stream != e->stream) {
      LPM_HIP(hipStreamWaitEvent(c->stream, e->ev, 0));
      if ((rc = ctx_mark_done(c))) return rc;
    } else {
      // same stream: the ensemble's marker covers the member
      c->done_flag_shared = e->done_flag;
      c->done_seq_shared = e->done_seq;
    }
    c->vel_valid = true;
    c->vel_prec = prec;
    c->vel_form = form;
    c->psi_valid = false;
  }
  return ok();
}

int lpm_ens_synchronize(lpm_bve_ensemble* e) {
  if (!e) return fail(LPM_ERR_ARG, "lpm_ens_synchronize: null ensemble");
  LPM_HIP(hipSetDevice(e->device));
  LPM_HIP(hipStreamSynchronize(e->stream));
  return ok();
}

namespace {
int query_mark(const unsigned long long* flag, unsigned long long seq, int* done, const char* who) {
  if (!done) return fail(LPM_ERR_ARG, std::string(who) + ": null output");
  *done = mark_reached(flag, seq) ? 1 : 0;
  return ok();
}
}  // namespace

int lpm_ens_query(lpm_bve_ensemble* e, int* done) {
  if (!e) return fail(LPM_ERR_ARG, "lpm_ens_query: null ensemble");
  return query_mark(e->done_flag, e->done_seq, done, "lpm_ens_query");
}

int lpm_bve_query(lpm_bve_ctx* c, int* done) {
  if (!c) return fail(LPM_ERR_ARG, "lpm_bve_query: null context");
  const int rc = query_mark(c->done_flag, c->done_seq, done, "lpm_bve_query");
  if (rc == LPM_OK && *done) *done = mark_reached(c->done_flag_shared, c->done_seq_shared) ? 1 : 0;
  return rc;
}

// ---------------------------------------------------------------------------
// launch-shape autotuning, keyed on the number of targets N
// ---------------------------------------------------------------------------
int lpm_shape_rule_set(int64_t ntargets, int targets_per_lane, int source_splits) {
  if (ntargets < 1 || targets_per_lane < 1 || targets_per_lane > 4 ||
      source_splits < 0)
    return fail(LPM_ERR_ARG, "lpm_shape_rule_set: bad arguments");
  std::lock_guard<std::mutex> g(g_rule_mu);
  g_rules[(int)ntargets] = ShapeRule{targets_per_lane, source_splits};
  return ok();
}

int lpm_shape_rule_get(int64_t ntargets, int* targets_per_lane, int* source_splits) {
  std::lock_guard<std::mutex> g(g_rule_mu);
  auto it = g_rules.find((int)ntargets);
  if (it == g_rules.end()) return fail(LPM_ERR_ARG, "lpm_shape_rule_get: no rule for this N");
  if (targets_per_lane) *targets_per_lane = it->second.R;
  if (source_splits) *source_splits = it->second.splits;
  return ok();
}

int lpm_shape_rules_clear(void) {
  std::lock_guard<std::mutex> g(g_rule_mu);
  g_rules.clear();
  return ok();
}

// Time one evaluation per candidate (targets per lane x source splits) on
// this context's problem and record the fastest as the rule for its N.
int lpm_bve_autotune(lpm_bve_ctx* c, int prec, int scheme, int reps, int* best_R, int* best_splits, double* best_ms) {
  if (!c || !c->x) return fail(LPM_ERR_STATE, "lpm_bve_autotune: no state");
  if (reps < 1) reps = 1;
  LPM_HIP(hipSetDevice(c->device));
  const int form = scheme == LPM_SCHEME_LAGR ? FORM_CHORD : FORM_DOT;
  const int saved_R = c->R_req, saved_splits = c->splits_req;
  const bool saved_timing = c->timing;
  const int ntiles = std::max(1, cdiv(c->ns, kTile));
  std::vector<std::pair<int, int>> cands;
  for (int R : {1, 2, 3, 4}) {
    const Shape base = choose_shape(c->nt, std::max(c->ns, 1), R, 0);
    std::vector<int> sp = {base.splits, 1, 2, 4, 8, 16, 32, 64};
    std::sort(sp.begin(), sp.end());
    sp.erase(std::unique(sp.begin(), sp.end()), sp.end());
    for (int s : sp)
      if (s <= ntiles) cands.push_back({R, s});
  }
  double best = 1e300;
  int bR = 2, bS = 0;
  c->timing = true;
  for (auto& cd : cands) {
    c->R_req = cd.first;
    c->splits_req = cd.second;
    for (auto& ev : c->timing_events) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    c->timing_events.clear();
    int rc = ctx_eval(c, c->x, c->zeta, prec, form, false, c->stream);  // warm
    for (int r = 0; r < reps && !rc; ++r) rc = ctx_eval(c, c->x, c->zeta, prec, form, false, c->stream);
    if (rc) {
      c->R_req = saved_R;
      c->splits_req = saved_splits;
      c->timing = saved_timing;
      return rc;
    }
    LPM_HIP(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t k = 1; k < c->timing_events.size(); ++k) {
      float ms = 0.0f;
      LPM_HIP(hipEventElapsedTime(&ms, c->timing_events[k].first, c->timing_events[k].second));
      tot += ms;
    }
    const double mean = tot / reps;
    if (mean < best) {
      best = mean;
      bR = cd.first;
      bS = cd.second;
    }
  }
  for (auto& ev : c->timing_events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  c->timing_events.clear();
  c->timing = saved_timing;
  c->R_req = saved_R;
  c->splits_req = saved_splits;
  {
    std::lock_guard<std::mutex> g(g_rule_mu);
    g_rules[c->nt] = ShapeRule{bR, bS};
  }
  // the trial evaluations overwrote the velocity with the last candidate's
  // rounding; the next step recomputes it with the chosen shape
  c->vel_valid = false;
  if (best_R) *best_R = bR;
  if (best_splits) *best_splits = bS;
  if (best_ms) *best_ms = best;
  return ok();
}

// Device memory one context needs for a problem of this size (bytes).
int lpm_bve_state_bytes(int64_t nv, int64_t nactive, int64_t* bytes) {
  if (nv < 0 || nactive < 1 || !bytes) return fail(LPM_ERR_ARG, "lpm_bve_state_bytes: bad arguments");
  if (nv + nactive > (int64_t)INT32_MAX / 8)
    return fail(LPM_ERR_ARG, "lpm_bve_state_bytes: sizes out of range (the limit of lpm_bve_set_state)");
  const int64_t nt = nv + nactive;
  const Shape sh = choose_shape((int)nt, (int)nactive, 0, 0);
  const BufferPlan bp = plan_buffers((int)nt, sh);
  const int64_t state = 8 * nt * (3 + 3 + 3 + 5 + 16) + 4 * (nt + nactive);
  *bytes = state + (int64_t)(bp.src64 + bp.src32 + bp.ctr + bp.trad + bp.srcq + bp.part) + 8 * 256;
  return ok();
}

}  // extern "C"
