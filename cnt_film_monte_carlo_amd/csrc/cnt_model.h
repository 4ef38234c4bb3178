// cnt_model.h -- carbon-nanotube exciton model (host, FP64): geometry, tight-binding bands in
// the K2-extended zone, Coulomb v(q), polarization, dielectric function and the A2 singlet
// exciton dispersion, restated from src/exciton_transfer/cnt.{h,cpp} of the reference.
#pragma once
#include <array>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace cntkmc {

using cplx = std::complex<double>;

// Truncated constants of the reference (src/exciton_transfer/constants.h:10-18).
namespace phys {
constexpr double pi = 3.141592;
constexpr double inv_pi = 1. / pi;
extern const double eV, hb, kb, eps0, q0;
}  // namespace phys

struct vec2 {
  double x = 0, y = 0;
};
struct vec3 {
  double x = 0, y = 0, z = 0;
};

// Hermitian eigen-decomposition (ascending eigenvalues, eigenvectors in columns of vecs,
// column-major n x n): Householder reduction to real tridiagonal form and implicit QL.
void hermitian_eig(int n, const std::vector<cplx>& a_colmajor, std::vector<double>& evals,
                   std::vector<cplx>& vecs);

struct cnt_model {
  // inputs
  int n = 0, m = 0, length = 0;  // chirality and length in cnt unit cells
  // constants of the cnt class (cnt.h:25-33)
  double a_cc = 1.42e-10, a_l = 0, e2p = 0, t0 = 0, s0 = 0, Upp = 0, kappa = 2.0;
  // geometry (get_parameters, cnt.cpp:16-117)
  vec2 a1, a2, b1, b2, aCC, ch_vec, t_vec, K1, K2, dk_l;
  vec3 t_vec_3d;
  double ch_len = 0, radius = 0;
  int t1 = 0, t2 = 0, Nu = 0, M = 0, Q = 0, nk_K1 = 0;
  // atoms (get_atom_coordinates, cnt.cpp:120-209)
  std::vector<vec2> pos_a, pos_b;  // Nu each
  std::vector<vec3> pos_u_3d;      // Nu
  // K2-extended electrons (electron_energy, cnt.cpp:301-358): ik in [0, nk), mu in [0, n_mu)
  int nk = 0, n_mu = 0;
  std::vector<double> el_energy;   // [band 0/1][ik][mu]
  std::vector<cplx> el_wf;         // [mu][ik][band][atom]
  // valleys and relevant states (cnt.cpp:362-517)
  std::vector<std::array<std::array<int, 2>, 2>> valleys;  // pairs of {ik_idx, mu_idx}
  std::vector<std::array<int, 2>> relev[2];                // per valley: {ik, mu}
  // interaction (cnt.cpp:520-780): iq in [-(nk-1), nk), mu in [-(n_mu-1), n_mu)
  int nq = 0, nmu_q = 0;
  std::vector<cplx> vq;            // [iq][mu][pair 0..3]
  std::vector<double> PI, eps;     // [iq][mu]
  // A2 singlet exciton (calculate_A_excitons, cnt.cpp:783-1053)
  int nk_relev = 0, nk_c = 0, nk_cm = 0;
  int ik_cm0 = 0;                  // ik_cm of index 0 (= -nk_relev)
  std::vector<double> ex_energy;   // [ik_cm_idx][n]   (ascending in n)
  std::vector<cplx> ex_psi;        // [ik_cm_idx][n][ik_c_idx]  (nk_c entries)
  std::vector<int32_t> ex_idx;     // [ik_cm_idx][ik_c_idx][4]  {ik_c, mu_c, ik_v, mu_v} indices

  double el(int band, int ik, int mu) const { return el_energy[((size_t)band * nk + ik) * n_mu + mu]; }
  cplx wf(int mu, int ik, int band, int atom) const {
    return el_wf[(((size_t)mu * nk + ik) * 2 + band) * 2 + atom];
  }
  double Au() const;                // graphene unit-cell area (cnt.h:346-349)
  double length_in_meter() const;   // length * |t_vec| (cnt.h:318-320)

  // Build everything (calculate_exciton_dispersion, cnt.cpp:1056-1081).  threads <= 0: all.
  // device >= 0 computes v(q) and PI(q) on that GPU, the rest on the host.  Returns an empty
  // string on success, else an error message.
  std::string build(int n_, int m_, int length_, int threads, int device = -1);
  // The A2 singlet kernel (kernel_11 + kernel_12 + 2 kernel_exchange) at one ik_cm_idx,
  // column-major nk_relev x nk_relev (for tests of the eigensolver).
  std::vector<cplx> singlet_kernel(int ik_cm_idx) const;
};

// v(q) [nq][nmu_q][4] and PI(q) [nq][nmu_q] of m (its bands, wavefunctions and q grid) on the
// GPU `device`, from the relative atom positions rel [4][natom][2] and the Ohno denominators
// den [4][natom]; norm = 2 Nu ncells.  Returns an empty string on success.
std::string cnt_sums_on_device(int device, const cnt_model& m, int natom, const double* rel,
                               const double* den, double norm, std::vector<cplx>& vq,
                               std::vector<double>& PI);

}  // namespace cntkmc

// opaque handle of the C ABI (include/cntkmc.h)
struct kmc_cnt {
  cntkmc::cnt_model m;
};
