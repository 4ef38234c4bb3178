// cnt_model.hip -- host restatement of the reference's CNT exciton model
// (src/exciton_transfer/cnt.cpp).  FP64 complex throughout, no device work.  Functions cite
// the reference lines they follow; the Hermitian eigensolver stands in for arma::eig_sym.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

#include "cnt_model.h"

namespace cntkmc {

namespace phys {
const double eV = 1.6 * std::pow(10, -19.0);
const double hb = 6.5 * std::pow(10, -16.0) * eV;
const double kb = 1.3865 * std::pow(10, -23.0);
const double eps0 = 8.85 * std::pow(10, -12);
const double q0 = 1.6 * std::pow(10, -19);
}  // namespace phys

namespace {

template <typename F>
void parallel_for(int n, int threads, F&& f) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, std::max(1, n));
  if (threads == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (int i = t; i < n; i += threads) f(i);
    });
  for (auto& th : pool) th.join();
}

// Armadillo-order 2-vector helpers.
inline double dot2(const vec2& a, const vec2& b) {
  double v1 = a.x * b.x;
  double v2 = a.y * b.y;
  return v1 + v2;
}
inline double norm2v(const vec2& a) {
  double acc1 = a.x * a.x;
  double acc2 = a.y * a.y;
  return std::sqrt(acc1 + acc2);
}
inline vec2 add(const vec2& a, const vec2& b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 scale(double s, const vec2& a) { return {s * a.x, s * a.y}; }

// exp(i x) as glibc cexp computes it for a zero real part: (cos x, sin x) from sincos.
inline cplx expi(double x) {
  double s, c;
  ::sincos(x, &s, &c);
  return cplx(c, s);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Hermitian eigensolver: unitary Householder reduction to a complex tridiagonal matrix, a
// diagonal phase transform to a real symmetric tridiagonal one, implicit-shift QL on that, and
// the back transformation.  O(n^3).
// ---------------------------------------------------------------------------------------------
void hermitian_eig(int n, const std::vector<cplx>& a_in, std::vector<double>& evals,
                   std::vector<cplx>& vecs) {
  std::vector<cplx> A(a_in);  // column-major: A[i + n*j]
  auto at = [&](int i, int j) -> cplx& { return A[i + (size_t)n * j]; };
  std::vector<cplx> Qm((size_t)n * n, cplx(0, 0));
  for (int i = 0; i < n; ++i) Qm[i + (size_t)n * i] = 1.0;
  std::vector<cplx> v(n), p(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int r0 = k + 1, r = n - r0;
    double alpha2 = 0.0;
    for (int i = r0; i < n; ++i) alpha2 += std::norm(at(i, k));
    const double alpha = std::sqrt(alpha2);
    if (alpha == 0.0) continue;
    const cplx x0 = at(r0, k);
    const double ax0 = std::abs(x0);
    const cplx ph = ax0 > 0 ? x0 / ax0 : cplx(1, 0);
    for (int i = 0; i < r; ++i) v[i] = at(r0 + i, k);
    v[0] += ph * alpha;
    double vn = 0.0;
    for (int i = 0; i < r; ++i) vn += std::norm(v[i]);
    const double tau = 2.0 / vn;
    // trailing block B = A[r0:, r0:]: p = tau B v, K = tau/2 v^H p, w = p - K v
    for (int i = 0; i < r; ++i) {
      cplx s = 0;
      for (int j = 0; j < r; ++j) s += at(r0 + i, r0 + j) * v[j];
      p[i] = tau * s;
    }
    cplx vhp = 0;
    for (int i = 0; i < r; ++i) vhp += std::conj(v[i]) * p[i];
    const cplx K = 0.5 * tau * vhp;
    for (int i = 0; i < r; ++i) w[i] = p[i] - K * v[i];
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i)
        at(r0 + i, r0 + j) -= v[i] * std::conj(w[j]) + w[i] * std::conj(v[j]);
    // the column becomes -ph * alpha e_1
    at(r0, k) = -ph * alpha;
    at(k, r0) = std::conj(at(r0, k));
    for (int i = r0 + 1; i < n; ++i) {
      at(i, k) = 0;
      at(k, i) = 0;
    }
    // Q <- Q H on columns r0..n-1
    for (int row = 0; row < n; ++row) {
      cplx s = 0;
      for (int j = 0; j < r; ++j) s += Qm[row + (size_t)n * (r0 + j)] * v[j];
      s *= tau;
      for (int j = 0; j < r; ++j) Qm[row + (size_t)n * (r0 + j)] -= s * std::conj(v[j]);
    }
  }
  // real tridiagonal via phases D: d_0 = 1, d_{k+1} = d_k t_{k+1,k} / |t_{k+1,k}|
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cplx> D(n, cplx(1, 0));
  for (int i = 0; i < n; ++i) d[i] = at(i, i).real();
  for (int k = 0; k + 1 < n; ++k) {
    const cplx t = at(k + 1, k);
    const double at_ = std::abs(t);
    e[k] = at_;
    D[k + 1] = at_ > 0 ? D[k] * (t / at_) : D[k];
  }
  // implicit QL on (d, e), e[i] couples d[i] and d[i+1]
  std::vector<double> Z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[i + (size_t)n * i] = 1.0;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 200; ++iter) {
      int mm = l;
      for (; mm < n - 1; ++mm) {
        const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= 2.2e-16 * dd) break;
      }
      if (mm == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double rr = std::hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(rr) : -std::fabs(rr)));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = mm - 1;
      bool early = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        rr = std::hypot(f, g);
        e[i + 1] = rr;
        if (rr == 0.0) {
          d[i + 1] -= pp;
          e[mm] = 0.0;
          early = true;
          break;
        }
        s = f / rr;
        c = g / rr;
        g = d[i + 1] - pp;
        rr = (d[i] - g) * s + 2.0 * c * b;
        pp = s * rr;
        d[i + 1] = g + pp;
        g = c * rr - b;
        double* zi = &Z[(size_t)n * i];
        double* zi1 = &Z[(size_t)n * (i + 1)];
        for (int q = 0; q < n; ++q) {
          const double fz = zi1[q];
          zi1[q] = s * zi[q] + c * fz;
          zi[q] = c * zi[q] - s * fz;
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[mm] = 0.0;
    }
  }
  // ascending order
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  evals.resize(n);
  vecs.assign((size_t)n * n, cplx(0, 0));
  // V = Q D Z
  std::vector<cplx> QD((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) QD[i + (size_t)n * j] = Qm[i + (size_t)n * j] * D[j];
  for (int jj = 0; jj < n; ++jj) {
    const int j = ord[jj];
    evals[jj] = d[j];
    const double* zj = &Z[(size_t)n * j];
    for (int k = 0; k < n; ++k) {
      const double zk = zj[k];
      if (zk == 0.0) continue;
      const cplx* qk = &QD[(size_t)n * k];
      cplx* out = &vecs[(size_t)n * jj];
      for (int i = 0; i < n; ++i) out[i] += qk[i] * zk;
    }
  }
}

// ---------------------------------------------------------------------------------------------
double cnt_model::Au() const { return std::fabs(a1.x * a2.y - a1.y * a2.x); }
double cnt_model::length_in_meter() const { return double(length) * norm2v(t_vec); }

std::string cnt_model::build(int n_, int m_, int length_, int threads, int device) {
  using namespace phys;
  n = n_;
  m = m_;
  length = length_;
  if (n < 1 || m < 0 || m > n || length < 1) return "cnt_model: need n >= 1, 0 <= m <= n, length >= 1";
  a_l = (double)std::sqrt(float(3.0)) * a_cc;  // single-precision sqrt(3) as in cnt.h:26
  e2p = 0.0;
  t0 = 2.7 * eV;
  s0 = 0.0;
  Upp = 11.3 * eV;

  // ---- get_parameters (cnt.cpp:16-117)
  a1 = {a_l * std::sqrt(3.0) / 2.0, +a_l / 2.0};
  a2 = {a_l * std::sqrt(3.0) / 2.0, -a_l / 2.0};
  b1 = {1.0 / std::sqrt(3.0) * 2.0 * pi / a_l, +2.0 * pi / a_l};
  b2 = {1.0 / std::sqrt(3.0) * 2.0 * pi / a_l, -2.0 * pi / a_l};
  aCC = {1.0 / 3.0 * (a1.x + a2.x), 1.0 / 3.0 * (a1.y + a2.y)};
  ch_vec = add(scale(double(n), a1), scale(double(m), a2));
  ch_len = norm2v(ch_vec);
  radius = ch_len / 2.0 / pi;
  const int dR = std::gcd(2 * n + m, n + 2 * m);
  t1 = +(2 * m + n) / dR;
  t2 = -(2 * n + m) / dR;
  t_vec = add(scale(double(t1), a1), scale(double(t2), a2));
  Nu = (int)(2 * (std::pow(n, 2) + std::pow(m, 2) + n * m) / dR);
  const double cos_t = ch_vec.x / norm2v(ch_vec), sin_t = ch_vec.y / norm2v(ch_vec);
  auto rot = [&](const vec2& v) { return vec2{cos_t * v.x + sin_t * v.y, -sin_t * v.x + cos_t * v.y}; };
  ch_vec = rot(ch_vec);
  t_vec = rot(t_vec);
  a1 = rot(a1);
  a2 = rot(a2);
  b1 = rot(b1);
  b2 = rot(b2);
  aCC = rot(aCC);
  t_vec_3d = {0.0, t_vec.y, 0.0};
  K1 = {(-double(t2) * b1.x + double(t1) * b2.x) / double(Nu),
        (-double(t2) * b1.y + double(t1) * b2.y) / double(Nu)};
  K2 = {(double(m) * b1.x - double(n) * b2.x) / double(Nu),
        (double(m) * b1.y - double(n) * b2.y) / double(Nu)};
  nk_K1 = length;
  dk_l = {K2.x / double(nk_K1), K2.y / double(nk_K1)};
  {
    const double p_min = (1. / double(t1) + 1. / double(n)) / (double(m) / double(n) - double(t2) / double(t1));
    const double p_max = (1. / double(t1) + double(Nu) / double(n)) / (double(m) / double(n) - double(t2) / double(t1));
    bool found = false;
    for (int p = (int)std::ceil(p_min); p < (int)std::ceil(p_max); p++) {
      if (((1 + t2 * p) % t1) == 0) {
        const int q = (1 + t2 * p) / t1;
        M = m * p - n * q;
        Q = std::gcd(Nu, M);
        found = true;
        break;
      }
    }
    if (!found) return "cnt_model: K2-extended parameters p, q not found";
  }

  // ---- get_atom_coordinates (cnt.cpp:120-209)
  pos_a.assign(Nu, vec2{});
  pos_b.assign(Nu, vec2{});
  {
    int k = 0;
    for (int i = 0; i <= t1 + n; i++) {
      for (int j = t2; j <= m; j++) {
        const bool f1 = double(t2 * i) / (double)t1 <= double(j);
        const bool f2 = double(m * i) / (double)n >= double(j);
        const bool f3 = double(t2 * (i - n)) / double(t1) > double(j - m);
        const bool f4 = double(m * (i - t1)) / double(n) < double(j - t2);
        if (f1 && f2 && f3 && f4) {
          if (k >= Nu) return "cnt_model: too many atoms in the unit cell";
          vec2 pa = {double(i) * a1.x + double(j) * a2.x, double(i) * a1.y + double(j) * a2.y};
          vec2 pb = {pa.x + aCC.x, pa.y + aCC.y};
          for (vec2* q : {&pa, &pb}) {
            if (q->x > ch_vec.x) q->x -= ch_vec.x;
            if (q->x < 0.0) q->x += ch_vec.x;
            if (q->y > ch_vec.y) q->y -= ch_vec.y;
            if (q->y < 0.0) q->y += ch_vec.y;
          }
          pos_a[k] = pa;
          pos_b[k] = pb;
          k++;
        }
      }
    }
    if (k != Nu) return "cnt_model: wrong number of atoms in the unit cell";
  }
  pos_u_3d.assign(Nu, vec3{});
  for (int i = 0; i < Nu; i++) {
    double s, c;
    ::sincos(pos_a[i].x / radius, &s, &c);
    pos_u_3d[i] = {radius * c, pos_a[i].y, radius * s};
  }

  // ---- electron_energy in the K2-extended zone (cnt.cpp:301-358, 1061-1063)
  nk = Nu / Q * nk_K1;
  n_mu = Q;
  el_energy.assign((size_t)2 * nk * n_mu, 0.0);
  el_wf.assign((size_t)n_mu * nk * 4, cplx(0, 0));
  {
    const vec2 v1 = {(a1.x + a2.x) / 3., (a1.y + a2.y) / 3.};
    const vec2 v2 = {(a1.x - 2. * a2.x) / 3., (a1.y - 2. * a2.y) / 3.};
    const vec2 v3 = {(a2.x - 2. * a1.x) / 3., (a2.y - 2. * a1.y) / 3.};
    const double r2 = 1. / std::sqrt(2.);
    for (int mu = 0; mu < n_mu; mu++) {
      for (int ik = 0; ik < nk; ik++) {
        const vec2 kv = {double(mu) * K1.x + double(ik) * dk_l.x, double(mu) * K1.y + double(ik) * dk_l.y};
        const cplx fk = expi(dot2(kv, v1)) + expi(dot2(kv, v2)) + expi(dot2(kv, v3));
        const double afk = std::hypot(fk.real(), fk.imag());
        el_energy[((size_t)1 * nk + ik) * n_mu + mu] = +t0 * afk;
        el_energy[((size_t)0 * nk + ik) * n_mu + mu] = -t0 * afk;
        cplx* w = &el_wf[(((size_t)mu * nk + ik) * 2) * 2];
        const cplx cf = std::conj(fk);
        w[1 * 2 + 0] = +r2;                         // (iA, ic)
        w[0 * 2 + 0] = +r2;                         // (iA, iv)
        w[1 * 2 + 1] = (-1. / std::sqrt(2.)) * cf / afk;  // (iB, ic)
        w[0 * 2 + 1] = (+1. / std::sqrt(2.)) * cf / afk;  // (iB, iv)
      }
    }
  }

  // ---- find_valleys (cnt.cpp:362-416)
  {
    std::vector<std::array<int, 2>> vl;
    for (int ik = 0; ik < nk; ik++) {
      for (int mu = 0; mu < n_mu; mu++) {
        int ip = ik + 1, im = ik - 1;
        while (ip >= nk) ip -= nk;
        while (im < 0) im += nk;
        if (el(1, ik, mu) < el(1, im, mu) && el(1, ik, mu) < el(1, ip, mu)) vl.push_back({ik, mu});
      }
    }
    std::stable_sort(vl.begin(), vl.end(), [&](const auto& s1, const auto& s2) {
      return el(1, s1[0], s1[1]) < el(1, s2[0], s2[1]);
    });
    valleys.clear();
    for (size_t i = 0; i < vl.size() / 2; i++) valleys.push_back({vl[2 * i], vl[2 * i + 1]});
    if (valleys.empty()) return "cnt_model: no conduction-band valley found";
  }

  // ---- find_relev_ik_range(1 eV) (cnt.cpp:419-517)
  for (int iv = 0; iv < 2; iv++) {
    auto& lst = relev[iv];
    lst.clear();
    const int ik_bottom = valleys[0][iv][0];
    const int mu_bottom = valleys[0][iv][1];
    const double max_energy = el(1, ik_bottom, mu_bottom) + 1. * eV;
    lst.push_back({ik_bottom, mu_bottom});
    bool in_range = true;
    int count = 0;
    while (in_range) {
      in_range = false;
      count++;
      if (count > nk) return "cnt_model: relevant-state search did not end";
      int ik = ik_bottom + count;
      while (ik >= nk) ik -= nk;
      if (el(1, ik, mu_bottom) < max_energy) {
        lst.push_back({ik, mu_bottom});
        in_range = true;
      }
      ik = ik_bottom - count;
      while (ik < 0) ik += nk;
      if (el(1, ik, mu_bottom) < max_energy) {
        lst.insert(lst.begin(), {ik, mu_bottom});
        in_range = true;
      }
    }
  }
  if (relev[0].size() != relev[1].size()) return "cnt_model: valleys differ in relevant states";

  // ---- calculate_vq (cnt.cpp:520-631) over iq in [-(nk-1), nk), mu in [-(Q-1), Q)
  nq = 2 * nk - 1;
  nmu_q = 2 * n_mu - 1;
  {
    int ncells = length;
    if (ncells % 2 == 0) ncells++;
    std::vector<vec2> rel[4];
    const int half = (int)std::floor(double(ncells) / 2.);
    for (int s = 0; s < 4; s++) rel[s].assign((size_t)Nu * ncells, vec2{});
    std::vector<vec2> paa(Nu), pab(Nu), pba(Nu), pbb(Nu);
    for (int i = 0; i < Nu; i++) {
      paa[i] = {pos_a[i].x - pos_a[0].x, pos_a[i].y - pos_a[0].y};
      pab[i] = {pos_a[i].x - pos_b[0].x, pos_a[i].y - pos_b[0].y};
      pba[i] = {pos_b[i].x - pos_a[0].x, pos_b[i].y - pos_a[0].y};
      pbb[i] = {pos_b[i].x - pos_b[0].x, pos_b[i].y - pos_b[0].y};
      for (vec2* q : {&paa[i], &pab[i], &pba[i], &pbb[i]})
        if (q->x > ch_vec.x / 2) q->x -= ch_vec.x;
    }
    for (int i = -half; i <= half; i++) {
      const int idx = (i + half) * Nu;
      for (int j = 0; j < Nu; j++) {
        const vec2 tv = {double(i) * t_vec.x, double(i) * t_vec.y};
        rel[0][idx + j] = {paa[j].x + tv.x, paa[j].y + tv.y};
        rel[1][idx + j] = {pab[j].x + tv.x, pab[j].y + tv.y};
        rel[2][idx + j] = {pba[j].x + tv.x, pba[j].y + tv.y};
        rel[3][idx + j] = {pbb[j].x + tv.x, pbb[j].y + tv.y};
      }
    }
    const double coeff = std::pow(4. * pi * eps0 * Upp / q0 / q0, 2);
    // the Ohno denominators do not depend on q: precompute Upp / sqrt(coeff r^2 + 1)
    std::vector<double> den[4];
    for (int s = 0; s < 4; s++) {
      den[s].resize(rel[s].size());
      for (size_t k = 0; k < rel[s].size(); k++)
        den[s][k] = Upp / std::sqrt(coeff * (std::pow(rel[s][k].x, 2) + std::pow(rel[s][k].y, 2)) + 1);
    }
    const double norm = double(2 * Nu * ncells);
    if (device >= 0) {
      // v(q) and PI(q) on the GPU (cnt_gpu.hip): positions [4][natom][2], denominators [4][natom]
      const int natom = (int)rel[0].size();
      std::vector<double> rel_flat((size_t)4 * natom * 2), den_flat((size_t)4 * natom);
      for (int s = 0; s < 4; s++)
        for (int k = 0; k < natom; k++) {
          rel_flat[((size_t)s * natom + k) * 2] = rel[s][k].x;
          rel_flat[((size_t)s * natom + k) * 2 + 1] = rel[s][k].y;
          den_flat[(size_t)s * natom + k] = den[s][k];
        }
      const std::string e =
          cnt_sums_on_device(device, *this, natom, rel_flat.data(), den_flat.data(), norm, vq, PI);
      if (!e.empty()) return e;
    } else {
    vq.assign((size_t)nq * nmu_q * 4, cplx(0, 0));
    parallel_for(nq, threads, [&](int iqi) {
      const int iq = iqi - (nk - 1);
      for (int mi = 0; mi < nmu_q; mi++) {
        const int mu = mi - (n_mu - 1);
        const vec2 q = {iq * dk_l.x + mu * K1.x, iq * dk_l.y + mu * K1.y};
        for (int s = 0; s < 4; s++) {
          cplx acc(0, 0);
          for (size_t k = 0; k < rel[s].size(); k++)
            acc += expi(dot2(q, rel[s][k])) * den[s][k];
          vq[((size_t)iqi * nmu_q + mi) * 4 + s] = acc / norm;
        }
      }
    });
    }
  }

  // ---- calculate_polarization (cnt.cpp:634-732) and calculate_dielectric (cnt.cpp:735-780)
  eps.assign((size_t)nq * nmu_q, 0.0);
  {
    const int nkM = nk_K1 * M;
    if (device < 0) {
    PI.assign((size_t)nq * nmu_q, 0.0);
    parallel_for(nq, threads, [&](int iqi) {
      const int iq = iqi - (nk - 1);
      for (int mi = 0; mi < nmu_q; mi++) {
        const int mu_q = mi - (n_mu - 1);
        double acc = 0.0;
        for (int ik = 0; ik < nk; ik++) {
          for (int mu_k = 0; mu_k < n_mu; mu_k++) {
            int mu_kq = mu_k + mu_q, ikq = ik + iq;
            while (mu_kq >= n_mu) {
              mu_kq -= n_mu;
              ikq += nkM;
            }
            while (mu_kq < 0) {
              mu_kq += n_mu;
              ikq -= nkM;
            }
            while (ikq >= nk) ikq -= nk;
            while (ikq < 0) ikq += nk;
            // <v(k)|c(k+q)> and <c(k)|v(k+q)> over the two atoms
            const cplx d1 = std::conj(wf(mu_k, ik, 0, 0)) * wf(mu_kq, ikq, 1, 0) +
                            std::conj(wf(mu_k, ik, 0, 1)) * wf(mu_kq, ikq, 1, 1);
            const cplx d2 = std::conj(wf(mu_k, ik, 1, 0)) * wf(mu_kq, ikq, 0, 0) +
                            std::conj(wf(mu_k, ik, 1, 1)) * wf(mu_kq, ikq, 0, 1);
            acc += std::pow(std::abs(d1), 2) / (el(1, ikq, mu_kq) - el(0, ik, mu_k)) +
                   std::pow(std::abs(d2), 2) / (el(1, ik, mu_k) - el(0, ikq, mu_kq));
          }
        }
        PI[(size_t)iqi * nmu_q + mi] = 2 * acc;
      }
    });
    }
    for (size_t i = 0; i < PI.size(); i++) {
      const cplx* v = &vq[i * 4];
      const double mean = (v[0].real() + v[1].real() + v[2].real() + v[3].real()) / 4.0;
      eps[i] = mean * PI[i] + 1.;
    }
  }

  // ---- calculate_A_excitons: A2 singlet only (cnt.cpp:783-1053)
  nk_relev = (int)relev[0].size();
  nk_c = 2 * nk_relev;
  nk_cm = 2 * nk_relev;
  ik_cm0 = -nk_relev;
  ex_energy.assign((size_t)nk_cm * nk_relev, 0.0);
  ex_psi.assign((size_t)nk_cm * nk_relev * nk_c, cplx(0, 0));
  ex_idx.assign((size_t)nk_cm * nk_c * 4, 0);
  std::string err;
  parallel_for(nk_cm, threads, [&](int ic) {
    const std::vector<cplx> H = singlet_kernel(ic);
    std::vector<double> ev;
    std::vector<cplx> vec;
    hermitian_eig(nk_relev, H, ev, vec);
    const double r2 = 1 / std::sqrt(2.);
    for (int nn = 0; nn < nk_relev; nn++) {
      ex_energy[(size_t)ic * nk_relev + nn] = ev[nn];
      cplx* ps = &ex_psi[((size_t)ic * nk_relev + nn) * nk_c];
      for (int r = 0; r < nk_relev; r++) {
        ps[r] = (+r2) * vec[r + (size_t)nk_relev * nn];
        ps[nk_relev + r] = (+r2) * vec[r + (size_t)nk_relev * nn];
      }
    }
    const int ik_cm = ik_cm0 + ic;
    for (int c = 0; c < nk_relev; c++) {
      const int ik_c = relev[0][c][0], mu_c = relev[0][c][1];
      int ik_v = ik_c - ik_cm;
      while (ik_v >= nk) ik_v -= nk;
      while (ik_v < 0) ik_v += nk;
      int32_t* o = &ex_idx[((size_t)ic * nk_c + c) * 4];
      o[0] = ik_c; o[1] = mu_c; o[2] = ik_v; o[3] = mu_c;
    }
    for (int vi = 0; vi < nk_relev; vi++) {
      const int ik_v = relev[1][vi][0], mu_v = relev[1][vi][1];
      int ik_c = ik_v + ik_cm;
      while (ik_c >= nk) ik_c -= nk;
      while (ik_c < 0) ik_c += nk;
      int32_t* o = &ex_idx[((size_t)ic * nk_c + (nk_c - 1 - vi)) * 4];
      o[0] = ik_c; o[1] = mu_v; o[2] = ik_v; o[3] = mu_v;
    }
  });
  return err;
}

std::vector<cplx> cnt_model::singlet_kernel(int ik_cm_idx) const {
  const int ik_cm = ik_cm0 + ik_cm_idx, mu_cm = 0;
  const int nr = nk_relev;
  std::vector<cplx> k11((size_t)nr * nr, 0.0), k12((size_t)nr * nr, 0.0), kx((size_t)nr * nr, 0.0);
  auto idx = [nr](int i, int j) { return (size_t)i + (size_t)nr * j; };
  auto vqa = [&](int iq, int mu, int s) {
    return vq[((size_t)(iq + nk - 1) * nmu_q + (mu + n_mu - 1)) * 4 + s];
  };
  auto epsa = [&](int iq, int mu) { return eps[(size_t)(iq + nk - 1) * nmu_q + (mu + n_mu - 1)]; };
  auto wrap = [&](int k) {
    while (k >= nk) k -= nk;
    while (k < 0) k += nk;
    return k;
  };
  const int iv = 0, icb = 1;
  auto direct = [&](int ik_c, int mu_c, int ik_v, int mu_v, int ik_cp, int mu_cp, int ik_vp,
                    int mu_vp) {
    int ik_d = wrap(ik_c - ik_cp);
    const int mu_d = mu_c - mu_cp;
    cplx acc = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        acc += std::conj(wf(mu_c, ik_c, icb, i)) * wf(mu_v, ik_v, iv, j) * wf(mu_cp, ik_cp, icb, i) *
               std::conj(wf(mu_vp, ik_vp, iv, j)) * vqa(ik_d, mu_d, 2 * i + j) / epsa(ik_d, mu_d);
    return acc;
  };
  auto exchange = [&](int ik_c, int mu_c, int ik_v, int mu_v, int ik_cp, int mu_cp, int ik_vp,
                      int mu_vp) {
    cplx acc = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        acc += std::conj(wf(mu_c, ik_c, icb, i)) * wf(mu_v, ik_v, iv, i) * wf(mu_cp, ik_cp, icb, j) *
               std::conj(wf(mu_vp, ik_vp, iv, j)) * vqa(ik_cm, mu_cm, 2 * i + j);
    return acc;
  };
  for (int c = 0; c < nr; c++) {
    const int ik_c = relev[0][c][0], mu_c = relev[0][c][1];
    const int ik_v = wrap(ik_c - ik_cm), mu_v = mu_c;
    k11[idx(c, c)] += el(icb, ik_c, mu_c) - el(iv, ik_v, mu_c);
    for (int cp = 0; cp <= c; cp++) {
      const int ik_cp = relev[0][cp][0], mu_cp = relev[0][cp][1];
      const int ik_vp = wrap(ik_cp - ik_cm), mu_vp = mu_cp;
      k11[idx(c, cp)] -= direct(ik_c, mu_c, ik_v, mu_v, ik_cp, mu_cp, ik_vp, mu_vp);
      kx[idx(c, cp)] += cplx(2, 0) * exchange(ik_c, mu_c, ik_v, mu_v, ik_cp, mu_cp, ik_vp, mu_vp);
    }
    for (int vp = c; vp < nr; vp++) {
      const int ik_vp = relev[1][vp][0], mu_vp = relev[1][vp][1];
      const int ik_cp = wrap(ik_vp + ik_cm), mu_cp = mu_vp;
      k12[idx(c, nr - 1 - vp)] -= direct(ik_c, mu_c, ik_v, mu_v, ik_cp, mu_cp, ik_vp, mu_vp);
    }
  }
  // K += K^H, diagonal halved (cnt.cpp:900-908)
  auto herm = [&](std::vector<cplx>& K) {
    std::vector<cplx> T(K);
    for (int j = 0; j < nr; j++)
      for (int i = 0; i < nr; i++) K[idx(i, j)] = T[idx(i, j)] + std::conj(T[idx(j, i)]);
    for (int i = 0; i < nr; i++) K[idx(i, i)] /= cplx(2, 0);
  };
  herm(k11);
  herm(k12);
  herm(kx);
  std::vector<cplx> H((size_t)nr * nr);
  for (size_t q = 0; q < H.size(); q++) H[q] = k11[q] + k12[q] + cplx(2, 0) * kx[q];
  return H;
}

}  // namespace cntkmc
