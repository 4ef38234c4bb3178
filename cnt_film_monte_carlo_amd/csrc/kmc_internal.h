// kmc_internal.h -- device data layout and shared host/device arithmetic of the KMC path.
//
// Layout in HBM (all arrays owned by one kmc_ctx):
//   sites      SoA double px,py,pz, ox,oy,oz [N]; uint8 chir[N], flags[N]; int32 left/right[N]
//   cells      int32 cell_of[N]; int64 cell_start[n_cells+1]; int32 cell_sites[N]
//   CSR        site_hdr hdr[N] (16 B: row start, degree, total rate Gamma)
//              double cum[nnz] (FP64 inclusive prefix sums, reference order)
//              edge_rec edge[nnz] (16 B: target | drain bit, float3 displacement)
//   excitons   SoA state [E]; MSD records [n_bins][E]; per-site counts [N]
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace cntkmc {

constexpr double RAND_MAX_D = 2147483647.0;  // glibc RAND_MAX, the divisor of scatterer.cpp:17
// RN(1 / RAND_MAX_D) = 2^-31 (1 + 2^-31) exactly: 1/D = 2^-31 (1 + 2^-31 + 2^-62 + ...).
constexpr double INV_RAND_MAX_D = 1.0 / 2147483647.0;

// x / RAND_MAX_D, correctly rounded (the IEEE quotient, bit for bit), in one multiply and two
// FMAs instead of a divide sequence.  Proof for a normal double x with quotient Q in
// [2^E, 2^(E+1)), D = 2^31 - 1:
//  * x / D is never a rounding midpoint: that would need D k (k odd, 54 bits) to fit in x's
//    53-bit significand.  Both x (ulp >= 2^(E-22), since x < 2^(E+32)) and the midpoints
//    (odd multiples of 2^(E-53)) lie on the 2^(E-53) grid, so x / D is at least 2^(E-53) / D
//    ~ 2^(E-84) from any midpoint.
//  * q0 = RN(x INV) is within one ulp of Q: INV's relative error is ~2^-62.  The remainder
//    x - q0 D is a multiple of 2^(E-52) below 2^(E-21), exact in one FMA.
//  * q0 + rem INV differs from Q = q0 + rem / D by rem |INV - 1/D| < 2^(E-114).  That is far
//    inside the 2^(E-84) distance to a midpoint, so RN of it is RN(Q).
// Checked exhaustively for every r in [1, 2^31 - 1] (dwell draws) and on 5e8 random doubles.
__host__ __device__ __forceinline__ double div_rand_max(double x) {
  const double q0 = x * INV_RAND_MAX_D;
  const double rem = __builtin_fma(-q0, RAND_MAX_D, x);
  return __builtin_fma(rem, INV_RAND_MAX_D, q0);
}
// 1 / d for a positive normal double, within about an ulp: on the device the hardware estimate
// (v_rcp_f64) and two Newton steps r <- r + r (1 - d r), each an FMA pair; on the host the IEEE
// quotient.  Used for 1 / Gamma of the dwell (set_max_rate, scatterer.h:92) and inside log_unit.
__host__ __device__ __forceinline__ double recip(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / d;
#endif
}

// Natural logarithm of a normal double q in (0, 1], the argument of ff_time's log
// (scatterer.h:74-80).  Range reduction of the classic atanh method (fdlibm's e_log.c, Sun
// Microsystems 1993): q = 2^k (1 + f) with sqrt(2)/2 <= 1 + f < sqrt(2), taken from the bits by
// one integer add (a mantissa at or above sqrt(2) carries into the exponent); f is exact.  With
// s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - f^2/2 + s (f^2/2 + R(s^2)) where
// R(z) = sum_j 2 z^j / (2j + 1); ten Taylor terms leave a remainder below 2^-60 of the result for
// |s| <= 3 - 2 sqrt(2).  k ln 2 is added in two parts (k ln2_hi exact for |k| < 2^11).  About
// 40 VALU operations instead of the ~75 of the library log, within two ulps of the exact value
// (checked against glibc on the host); event times need 1e-12 relative.
// fma(a, b, c) as one VOP3 v_fma_f64 on the device.  The compiler otherwise evaluates a Horner
// step with loop-invariant coefficients as a copy of the coefficient plus the two-address
// v_fmac_f64: two instructions where one does.  Same rounding (one FMA).
__host__ __device__ __forceinline__ double fma3(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
#else
  return __builtin_fma(a, b, c);
#endif
}

constexpr double LN2_HI = 0x1.62e42fefa3800p-1;   // ln 2 to 42 bits
constexpr double LN2_LO = 0x1.ef35793c76730p-45;  // ln 2 - LN2_HI
__host__ __device__ __forceinline__ double log_unit(double q) {
#pragma clang fp contract(off)
  const uint64_t b = __builtin_bit_cast(uint64_t, q);
  const uint32_t hi = (uint32_t)(b >> 32) + (0x3FF00000u - 0x3FE6A09Eu);
  const int k = (int)(hi >> 20) - 1023;
  const uint64_t mb = ((uint64_t)((hi & 0x000FFFFFu) + 0x3FE6A09Eu) << 32) | (b & 0xFFFFFFFFull);
  const double f = __builtin_bit_cast(double, mb) - 1.0;
  const double s = f * recip(2.0 + f);
  const double z = s * s;
  double R = 2.0 / 21.0;
  R = fma3(R, z, 2.0 / 19.0);
  R = fma3(R, z, 2.0 / 17.0);
  R = fma3(R, z, 2.0 / 15.0);
  R = fma3(R, z, 2.0 / 13.0);
  R = fma3(R, z, 2.0 / 11.0);
  R = fma3(R, z, 2.0 / 9.0);
  R = fma3(R, z, 2.0 / 7.0);
  R = fma3(R, z, 2.0 / 5.0);
  R = fma3(R, z, 2.0 / 3.0);
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

constexpr uint32_t EDGE_DRAIN_BIT = 0x80000000u;
constexpr uint32_t ROW_UNITS = 8;  // 16-byte units per 128-byte line of the hop kernel's blocks
// Same-tube links of the drift path: no link (tube end) and the bound on site passes per flight.
constexpr uint32_t LINK_NONE = 0xFFFFFFFFu;
constexpr uint32_t FLY_MAX_PASSES = 1u << 22;
constexpr uint32_t EDGE_COL_MASK = 0x7FFFFFFFu;
// Row word of the hop kernel (a target, a source, a link): first line of the site's block |
// TGT_WIDE_BIT when the block continues past its first line | EDGE_DRAIN_BIT for a drain site.
constexpr uint32_t TGT_LINE_MASK = 0x3FFFFFFFu;
constexpr uint32_t TGT_WIDE_BIT = 0x40000000u;
constexpr uint64_t FNV_OFFSET = 0xcbf29ce484222325ull;
constexpr uint64_t FNV_PRIME = 0x100000001b3ull;
constexpr int MAX_CHIR = 16;
#define KMC_COUNT_COPIES 8u  // one departure-count array per XCD
constexpr int64_t COUNT_ALIGN = 1024;  // each copy starts on a 4 KB boundary
// Departure counts are kept per block line (the first line of the departure site's block, in
// the Morton placement of the blocks), so the counters an XCD's excitons touch -- one compact
// region of the film -- form a compact, L2-resident range.  Slot of line L in a copy of
// count_stride (a multiple of 32) counters: the 32 counters of a 128-byte line hold lines 32
// apart, L -> (L % 32) stride / 32 + L / 32, so that excitons hopping among neighbouring blocks
// spread their atomics over many cache lines and L2 channels.  (Slot L for line L, which puts
// neighbouring blocks in one cache line, made a hop launch 1 % faster on a film of 1.6 excitons
// per site and 1.8 times slower on one of 10 per site, whose departures crowd a region's lines.)
__host__ __device__ __forceinline__ uint32_t count_slot(uint32_t L, int64_t count_stride) {
  return (L & 31u) * (uint32_t)(count_stride >> 5) + (L >> 5);
}

struct __align__(16) site_hdr {
  uint32_t rs;   // first CSR entry of the row
  uint32_t n;    // out-degree
  double gamma;  // total out-rate = last cumulative entry (scatterer.h:91)
};

struct __align__(16) edge_rec {
  uint32_t col;  // target site | EDGE_DRAIN_BIT when the target is a drain site
  float dx, dy, dz;
};

// Rate tables in device memory: grid axes (theta, z_shift, axis_shift_1, axis_shift_2)
// concatenated in one small array, rates [n_chir*n_chir][n_th][n_z][n_a1][n_a2].
struct table_desc {
  int32_t n_chir, n_th, n_z, n_a1, n_a2;
  const double* rates;
  const double* th;
  const double* z;
  const double* a1;
  const double* a2;
};

struct cell_desc {
  int32_t nx, ny, nz;
  int32_t periodic_mask;
  double lo[3];       // origin per axis (bounding-box minimum or period_lo)
  double cell[3];     // cell edge per axis (cutoff on open axes, period/n on periodic axes)
  double period[3];   // period length (periodic axes only)
  double cutoff, min_distance;
};

// ---------------------------------------------------------------------------------------------
// FP64 vector arithmetic in Armadillo's order for 3-vectors (op_dot::direct_dot_arma and
// op_norm::vec_norm_2_direct_std accumulate elements 0 and 2 in one accumulator and element 1
// in the other, then add the two).  Contraction is off so no FMA changes the rounding.
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ double dot3(double ax, double ay, double az, double bx,
                                                double by, double bz) {
#pragma clang fp contract(off)
  double v1 = ax * bx;
  const double v2 = ay * by;
  v1 = v1 + az * bz;
  return v1 + v2;
}

__host__ __device__ __forceinline__ double norm3(double x, double y, double z) {
#pragma clang fp contract(off)
  double acc1 = x * x;
  const double acc2 = y * y;
  acc1 = acc1 + z * z;
  return sqrt(acc1 + acc2);
}

// First index of the minimum of |grid - x| (arma::index_min on abs(grid - x),
// scattering_struct.h:42-49); NaN x keeps index 0.
__host__ __device__ __forceinline__ int nearest_index(const double* g, int n, double x) {
#pragma clang fp contract(off)
  double best = INFINITY;
  int idx = 0;
  for (int i = 0; i < n; ++i) {
    const double v = fabs(g[i] - x);
    if (v < best) {
      best = v;
      idx = i;
    }
  }
  return idx;
}

struct pair_geom {
  double theta, z, a1, a2;
};

// Pair geometry of scatterer::find_neighbors (src/monte_carlo/scatterer.cpp:44-61).  p_i is the
// source, p_j the (image of the) target; o_i, o_j unit orientations.
__host__ __device__ __forceinline__ pair_geom pair_geometry(double pix, double piy, double piz,
                                                            double pjx, double pjy, double pjz,
                                                            double oix, double oiy, double oiz,
                                                            double ojx, double ojy, double ojz) {
#pragma clang fp contract(off)
  pair_geom g;
  const double dx = pix - pjx, dy = piy - pjy, dz = piz - pjz;
  const double c = dot3(oix, oiy, oiz, ojx, ojy, ojz);
  if (c == 1.0) {
    g.a1 = 0.0;
    const double d = dot3(dx, dy, dz, oix, oiy, oiz);
    g.a2 = d;
    g.theta = 0.0;
    g.z = norm3(dx - d * oix, dy - d * oiy, dz - d * oiz);
  } else {
    g.theta = acos(c);
    const double y1 = dot3(oix, oiy, oiz, dx, dy, dz);
    const double y2 = dot3(ojx, ojy, ojz, dx, dy, dz);
    const double s = 1.0 - c * c;
    g.a1 = (y1 + y2 * c) / s;
    g.a2 = (y2 + y1 * c) / s;
    g.z = norm3((g.a1 * oix + pix) - (g.a2 * ojx + pjx), (g.a1 * oiy + piy) - (g.a2 * ojy + pjy),
                (g.a1 * oiz + piz) - (g.a2 * ojz + pjz));
  }
  return g;
}

__host__ __device__ __forceinline__ double table_rate(const table_desc& t, int ci, int cj,
                                                      const pair_geom& g) {
  const int ith = nearest_index(t.th, t.n_th, g.theta);
  const int iz = nearest_index(t.z, t.n_z, g.z);
  const int ia1 = nearest_index(t.a1, t.n_a1, g.a1);
  const int ia2 = nearest_index(t.a2, t.n_a2, g.a2);
  const int64_t tab = (int64_t)ci * t.n_chir + cj;
  const int64_t idx = (((tab * t.n_th + ith) * t.n_z + iz) * t.n_a1 + ia1) * t.n_a2 + ia2;
  return t.rates[idx];
}

// Minimum image of p_j seen from p_i on one periodic axis.
__host__ __device__ __forceinline__ double image_of(double pi, double pj, double period) {
#pragma clang fp contract(off)
  const double d = pi - pj;
  const double half = 0.5 * period;
  if (d > half) return pj + period;
  if (d < -half) return pj - period;
  return pj;
}

// Cell coordinate along one axis.  Open axes follow create_scatterer_buckets
// (monte_carlo.h:390-392): int((p - min) / radius).  Periodic axes: floor((p - lo)/cell), clamped.
__host__ __device__ __forceinline__ int cell_coord(const cell_desc& c, int axis, double p) {
#pragma clang fp contract(off)
  const int n = axis == 0 ? c.nx : (axis == 1 ? c.ny : c.nz);
  if (c.periodic_mask & (1 << axis)) {
    int i = (int)floor((p - c.lo[axis]) / c.cell[axis]);
    if (i < 0) i = 0;
    if (i >= n) i = n - 1;
    return i;
  }
  return (int)((p - c.lo[axis]) / c.cell[axis]);
}

// The up to three stencil coordinates along one axis, in the order i-1, i, i+1 with
// out-of-range cells dropped on open axes (monte_carlo.h:402-405) and wrapped (duplicates
// dropped) on periodic axes.  Returns the count.
__host__ __device__ __forceinline__ int stencil_axis(const cell_desc& c, int axis, int i,
                                                     int out[3]) {
  const int n = axis == 0 ? c.nx : (axis == 1 ? c.ny : c.nz);
  int k = 0;
  if (c.periodic_mask & (1 << axis)) {
    for (int d = -1; d <= 1; ++d) {
      const int w = ((i + d) % n + n) % n;
      bool dup = false;
      for (int q = 0; q < k; ++q) dup = dup || (out[q] == w);
      if (!dup) out[k++] = w;
    }
  } else {
    for (int d = -1; d <= 1; ++d) {
      const int w = i + d;
      if (w > -1 && w < n) out[k++] = w;
    }
  }
  return k;
}

}  // namespace cntkmc
