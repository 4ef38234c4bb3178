// kmc_host.hip -- host-only entry points of libcntkmc.so: synthetic CNT networks, the
// Förster/Wong closed-form rate table, arma::linspace grids and the diffusion estimator.
// None of these touch the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cntkmc.h"
#include "kmc_internal.h"
#include "philox.h"

using namespace cntkmc;

namespace {

constexpr uint32_t STREAM_NETGEN = 0x4E455447u;

double uniform53(uint32_t hi, uint32_t lo) {
  const uint64_t b = ((uint64_t)hi << 32 | lo) >> 11;
  return (double)b * (1.0 / 9007199254740992.0);
}

// Unit in-plane orientation with dot(o,o) == 1.0 exactly in the Armadillo dot order.  For
// phi_k = phi + k * 1e-9 rad (k = 0, 1, ...): keep the smaller component of (cos, sin),
// recompute the larger one as sqrt(1 - minor^2) and try it and its neighbours up to 3 ulps
// away.  The first exact candidate wins (k <= 12 in practice); the plain (cos phi, sin phi)
// is returned if none is found.
void exact_unit(double phi, double* ox, double* oy) {
  static const int steps[7] = {0, -1, 1, -2, 2, -3, 3};
  for (int k = 0; k < 256; ++k) {
    const double ph = phi + k * 1e-9;
    const double c0 = std::cos(ph), s0 = std::sin(ph);
    const bool c_major = std::fabs(c0) >= std::fabs(s0);
    const double minor = c_major ? s0 : c0;
    const double major0 = std::copysign(std::sqrt(1.0 - minor * minor), c_major ? c0 : s0);
    for (int q = 0; q < 7; ++q) {
      double m = major0;
      for (int u = 0; u < std::abs(steps[q]); ++u)
        m = std::nextafter(m, steps[q] > 0 ? 2.0 : -2.0);
      const double c = c_major ? m : minor, s = c_major ? minor : m;
      if (dot3(c, s, 0.0, c, s, 0.0) == 1.0) {
        *ox = c;
        *oy = s;
        return;
      }
    }
  }
  *ox = std::cos(phi);
  *oy = std::sin(phi);
}

double wrap_into(double p, double lo, double len) {
  double q = lo + std::fmod(p - lo, len);
  if (q < lo) q += len;
  if (q >= lo + len) q -= len;
  if (q < lo) q = lo;
  return q;
}

// Two-sided 95% Student-t quantiles, df = 1..30.
const double kT975[31] = {0.0,    12.706, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365,
                          2.306,  2.262,  2.228, 2.201, 2.179, 2.160, 2.145, 2.131,
                          2.120,  2.110,  2.101, 2.093, 2.086, 2.080, 2.074, 2.069,
                          2.064,  2.060,  2.056, 2.052, 2.048, 2.045, 2.042};

double t975(int df) {
  if (df <= 0) return INFINITY;
  if (df <= 30) return kT975[df];
  const double z = 1.959964;
  return z + (z * z * z + z) / (4.0 * df);  // Cornish-Fisher first correction
}

}  // namespace

extern "C" {

int kmc_linspace(double start, double end, int32_t n, double* out) {
  // arma::linspace: delta = (end-start)/(n-1); x[i] = start + i*delta; x[n-1] = end.
  if (n < 1 || !out) return KMC_ERR_ARG;
  if (n == 1) {
    out[0] = end;
    return KMC_OK;
  }
  const double delta = (end - start) / double(n - 1);
  for (int32_t i = 0; i < n - 1; ++i) out[i] = start + double(i) * delta;
  out[n - 1] = end;
  return KMC_OK;
}

int kmc_forster_table(double gamma_0, int32_t n_theta, const double* theta, int32_t n_z,
                      const double* z_shift, int32_t n_a1, const double* a1, int32_t n_a2,
                      const double* a2, double* rates) {
  // create_forster_scatt_table (src/monte_carlo/monte_carlo.cpp:174-195)
  if (n_theta < 1 || n_z < 1 || n_a1 < 1 || n_a2 < 1 || !theta || !z_shift || !a1 || !a2 ||
      !rates)
    return KMC_ERR_ARG;
  auto normalise = [](double v[3]) {
    const double nv = norm3(v[0], v[1], v[2]);
    const double d = (nv != 0.0) ? nv : 1.0;
    v[0] = v[0] / d;
    v[1] = v[1] / d;
    v[2] = v[2] / d;
  };
  size_t idx = 0;
  for (int32_t it = 0; it < n_theta; ++it) {
    // g++ -O3 (the reference build, makefile:3-7) merges std::sin/std::cos of one angle into
    // glibc sincos, whose sine differs from sin() by 1 ulp at some angles (135 deg on this
    // grid); call sincos directly so the table matches that build.
    double sin_th, cos_th;
    ::sincos(theta[it], &sin_th, &cos_th);
    for (int32_t iz = 0; iz < n_z; ++iz) {
      for (int32_t i1 = 0; i1 < n_a1; ++i1) {
        for (int32_t i2 = 0; i2 < n_a2; ++i2) {
          double r1[3] = {a1[i1], 0.0, 0.0};
          double r2[3] = {a2[i2] * cos_th, a2[i2] * sin_th, z_shift[iz]};
          double dR[3] = {r1[0] - r2[0], r1[1] - r2[1], r1[2] - r2[2]};
          const double dist = norm3(dR[0], dR[1], dR[2]);
          double n1[3] = {r1[0], r1[1], r1[2]}, n2[3] = {r2[0], r2[1], r2[2]},
                 nd[3] = {dR[0], dR[1], dR[2]};
          normalise(n1);
          normalise(n2);
          normalise(nd);
          const double d1 = dot3(n1[0], n1[1], n1[2], nd[0], nd[1], nd[2]);
          const double d2 = dot3(n2[0], n2[1], n2[2], nd[0], nd[1], nd[2]);
          const double af = cos_th - 3 * d1 * d2;
          rates[idx++] = gamma_0 * std::pow(af, 2) * std::pow(1.e-9 / dist, 6);
        }
      }
    }
  }
  return KMC_OK;
}

int kmc_table_rate(int32_t n_chir, int32_t n_theta, const double* theta, int32_t n_z,
                   const double* z_shift, int32_t n_a1, const double* a1, int32_t n_a2,
                   const double* a2, const double* rates, int64_t n, const uint8_t* ci,
                   const uint8_t* cj, const double* q_theta, const double* q_z, const double* q_a1,
                   const double* q_a2, double* out, int32_t* nodes) {
  // scattering_struct::get_rate (src/monte_carlo/scattering_struct.h:38-52) through the same
  // table_rate the CSR kernels evaluate on the device
  if (n_chir < 1 || n_chir > MAX_CHIR || n_theta < 1 || n_z < 1 || n_a1 < 1 || n_a2 < 1 ||
      n < 0 || !theta || !z_shift || !a1 || !a2 || !rates || (n > 0 && (!q_theta || !q_z ||
      !q_a1 || !q_a2 || !out)))
    return KMC_ERR_ARG;
  table_desc t;
  t.n_chir = n_chir;
  t.n_th = n_theta;
  t.n_z = n_z;
  t.n_a1 = n_a1;
  t.n_a2 = n_a2;
  t.rates = rates;
  t.th = theta;
  t.z = z_shift;
  t.a1 = a1;
  t.a2 = a2;
  for (int64_t k = 0; k < n; ++k) {
    const int c1 = ci ? ci[k] : 0, c2 = cj ? cj[k] : 0;
    if (c1 >= n_chir || c2 >= n_chir) return KMC_ERR_ARG;
    const pair_geom g{q_theta[k], q_z[k], q_a1[k], q_a2[k]};
    out[k] = table_rate(t, c1, c2, g);
    if (nodes) {
      nodes[4 * k] = nearest_index(theta, n_theta, g.theta);
      nodes[4 * k + 1] = nearest_index(z_shift, n_z, g.z);
      nodes[4 * k + 2] = nearest_index(a1, n_a1, g.a1);
      nodes[4 * k + 3] = nearest_index(a2, n_a2, g.a2);
    }
  }
  return KMC_OK;
}

int kmc_ff_time(int64_t n, const double* gamma, const uint32_t* r, double* out) {
#pragma clang fp contract(off)
  // ff_time (src/monte_carlo/scatterer.h:74-80) with the hop kernels' arithmetic
  if (n < 0 || (n > 0 && (!gamma || !r || !out))) return KMC_ERR_ARG;
  for (int64_t k = 0; k < n; ++k) {
    if (!(gamma[k] > 0.0) || r[k] == 0u || r[k] > 0x7FFFFFFFu) return KMC_ERR_ARG;
    out[k] = -recip(gamma[k]) * log_unit(div_rand_max((double)r[k]));
  }
  return KMC_OK;
}

int kmc_generate_network(const kmc_netgen_params* p, int64_t capacity, double* x, double* y,
                         double* z, double* ox, double* oy, double* oz, int32_t* left,
                         int32_t* right, uint8_t* chirality, int32_t* tube_id, int64_t* n_out) {
  if (!p || !n_out) return KMC_ERR_ARG;
  if (p->n_tubes < 0 || p->n_tubes > 0x7FFFFFFF || !(p->site_spacing > 0.0) ||
      !(p->len_min > 0.0) || !(p->len_max >= p->len_min) || p->n_chir < 1 ||
      p->n_chir > MAX_CHIR)
    return KMC_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (!(p->box_hi[a] >= p->box_lo[a])) return KMC_ERR_ARG;
  const bool write = capacity > 0 && x && y && z && ox && oy && oz;
  int64_t count = 0;
  for (int64_t t = 0; t < p->n_tubes; ++t) {
    const uint32_t tt = (uint32_t)t;
    const u32x4 w0 = draw4(p->seed, tt, 0, STREAM_NETGEN);
    const u32x4 w1 = draw4(p->seed, tt, 1, STREAM_NETGEN);
    const u32x4 w2 = draw4(p->seed, tt, 2, STREAM_NETGEN);
    const double u_phi = uniform53(w0.v[0], w0.v[1]);
    const double u_len = uniform53(w0.v[2], w0.v[3]);
    const double u_c[3] = {uniform53(w1.v[0], w1.v[1]), uniform53(w1.v[2], w1.v[3]),
                           uniform53(w2.v[0], w2.v[1])};
    const double u_chir = uniform53(w2.v[2], w2.v[3]);
    const double len = p->len_min + (p->len_max - p->len_min) * u_len;
    int64_t n_seg = (int64_t)std::llround(len / p->site_spacing);
    if (n_seg < 1) n_seg = 1;
    const double phi = 2.0 * M_PI * u_phi;
    double cx[3], o[3];
    for (int a = 0; a < 3; ++a) cx[a] = p->box_lo[a] + (p->box_hi[a] - p->box_lo[a]) * u_c[a];
    exact_unit(phi, &o[0], &o[1]);
    o[2] = 0.0;
    int chir = (int)std::floor(u_chir * p->n_chir);
    if (chir >= p->n_chir) chir = p->n_chir - 1;
    if (write && count + n_seg > capacity) return KMC_ERR_CAPACITY;
    for (int64_t k = 0; k < n_seg; ++k) {
      const int64_t i = count + k;
      if (write) {
        const double off = ((double)k - 0.5 * (double)(n_seg - 1)) * p->site_spacing;
        double pos[3];
        for (int a = 0; a < 3; ++a) {
          pos[a] = cx[a] + off * o[a];
          if (p->wrap_mask & (1 << a))
            pos[a] = wrap_into(pos[a], p->box_lo[a], p->box_hi[a] - p->box_lo[a]);
        }
        x[i] = pos[0];
        y[i] = pos[1];
        z[i] = pos[2];
        ox[i] = o[0];
        oy[i] = o[1];
        oz[i] = o[2];
        if (left) left[i] = k > 0 ? (int32_t)(i - 1) : -1;
        if (right) right[i] = k + 1 < n_seg ? (int32_t)(i + 1) : -1;
        if (chirality) chirality[i] = (uint8_t)chir;
        if (tube_id) tube_id[i] = (int32_t)t;
      }
    }
    count += n_seg;
    if (count >= 0x7FFFFFFF) return KMC_ERR_CAPACITY;
  }
  *n_out = count;
  return KMC_OK;
}

int kmc_estimate_diffusion(const double* red, int32_t n_bins, int32_t n_batches,
                           double t_sample_max, int32_t k_lo, int32_t k_hi, double D[3],
                           double ci95[3]) {
  if (!red || !D || n_bins < 1 || n_batches < 1 || !(t_sample_max > 0.0)) return KMC_ERR_ARG;
  if (k_lo < 0) k_lo = 0;
  if (k_hi > n_bins) k_hi = n_bins;
  if (k_lo >= k_hi) return KMC_ERR_ARG;
  const double dt = t_sample_max / n_bins;
  auto at = [&](int k, int b, int q) { return red[((size_t)k * n_batches + b) * 5 + q]; };
  // batch sizes: the count of bin 0 when complete; a bin is used when every batch is complete
  std::vector<double> full(n_batches, 0.0);
  for (int b = 0; b < n_batches; ++b) {
    double m = 0.0;
    for (int k = 0; k < n_bins; ++k) m = std::max(m, at(k, b, 0));
    full[b] = m;
  }
  std::vector<int> bins;
  for (int k = k_lo; k < k_hi; ++k) {
    bool ok = true;
    for (int b = 0; b < n_batches; ++b) ok = ok && at(k, b, 0) == full[b] && full[b] > 0.0;
    if (ok) bins.push_back(k);
  }
  if (bins.empty()) return KMC_ERR_STATE;
  for (int ax = 0; ax < 3; ++ax) {
    double num = 0.0, den = 0.0;
    for (int k : bins) {
      double s = 0.0, n = 0.0;
      for (int b = 0; b < n_batches; ++b) {
        s += at(k, b, 1 + ax);
        n += at(k, b, 0);
      }
      const double tk = (k + 1) * dt;
      num += tk * (s / n);
      den += tk * tk;
    }
    D[ax] = 0.5 * num / den;
    if (ci95) {
      if (n_batches < 2) {
        ci95[ax] = INFINITY;
        continue;
      }
      double mean = 0.0, m2 = 0.0;
      std::vector<double> db(n_batches);
      for (int b = 0; b < n_batches; ++b) {
        double nb = 0.0;
        for (int k : bins) {
          const double tk = (k + 1) * dt;
          nb += tk * (at(k, b, 1 + ax) / at(k, b, 0));
        }
        db[b] = 0.5 * nb / den;
        mean += db[b];
      }
      mean /= n_batches;
      for (int b = 0; b < n_batches; ++b) m2 += (db[b] - mean) * (db[b] - mean);
      const double sd = std::sqrt(m2 / (n_batches - 1));
      ci95[ax] = t975(n_batches - 1) * sd / std::sqrt((double)n_batches);
    }
  }
  return KMC_OK;
}

}  // extern "C"
