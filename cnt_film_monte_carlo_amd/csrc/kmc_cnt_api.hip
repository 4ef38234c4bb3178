// kmc_cnt_api.hip -- C ABI of the CNT exciton model (host only).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "../../include/cntkmc.h"
#include "cnt_model.h"

using namespace cntkmc;


namespace {
thread_local std::string g_cnt_err;
int cfail(int code, const std::string& msg) {
  g_cnt_err = msg;
  return code;
}
}  // namespace

extern "C" {

const char* kmc_cnt_last_error(void) { return g_cnt_err.c_str(); }

int kmc_cnt_create(int32_t n, int32_t m, int32_t length_cells, int32_t threads, kmc_cnt** out) {
  if (!out) return cfail(KMC_ERR_ARG, "kmc_cnt_create: null output");
  *out = nullptr;
  kmc_cnt* c = new kmc_cnt();
  const std::string err = c->m.build(n, m, length_cells, threads);
  if (!err.empty()) {
    delete c;
    return cfail(KMC_ERR_ARG, err);
  }
  *out = c;
  return KMC_OK;
}

int kmc_cnt_create_on_device(int32_t n, int32_t m, int32_t length_cells, int32_t threads,
                             int32_t device, kmc_cnt** out) {
  if (!out) return cfail(KMC_ERR_ARG, "kmc_cnt_create_on_device: null output");
  *out = nullptr;
  if (device < 0) return cfail(KMC_ERR_ARG, "kmc_cnt_create_on_device: need device >= 0");
  kmc_cnt* c = new kmc_cnt();
  const std::string err = c->m.build(n, m, length_cells, threads, device);
  if (!err.empty()) {
    delete c;
    return cfail(err.rfind("cnt_model on the GPU", 0) == 0 ? KMC_ERR_HIP : KMC_ERR_ARG, err);
  }
  *out = c;
  return KMC_OK;
}

int kmc_cnt_destroy(kmc_cnt* c) {
  delete c;
  return KMC_OK;
}

int kmc_cnt_get_info(const kmc_cnt* c, kmc_cnt_info* info) {
  if (!c || !info) return cfail(KMC_ERR_ARG, "kmc_cnt_get_info: null argument");
  const cnt_model& m = c->m;
  std::memset(info, 0, sizeof(*info));
  info->n = m.n;
  info->m = m.m;
  info->length_cells = m.length;
  info->Nu = m.Nu;
  info->Q = m.Q;
  info->M = m.M;
  info->t1 = m.t1;
  info->t2 = m.t2;
  info->nk = m.nk;
  info->n_mu = m.n_mu;
  info->nk_relev = m.nk_relev;
  info->nk_cm = m.nk_cm;
  info->nk_c = m.nk_c;
  info->ik_cm0 = m.ik_cm0;
  info->radius = m.radius;
  info->length_m = m.length_in_meter();
  info->Au = m.Au();
  info->a_l = m.a_l;
  const vec2* v[6] = {&m.ch_vec, &m.t_vec, &m.aCC, &m.dk_l, &m.K1, &m.K2};
  double* o[6] = {info->ch_vec, info->t_vec, info->aCC, info->dk_l, info->K1, info->K2};
  for (int q = 0; q < 6; ++q) {
    o[q][0] = v[q]->x;
    o[q][1] = v[q]->y;
  }
  double emin = m.ex_energy.empty() ? 0.0 : m.ex_energy[0];
  for (double e : m.ex_energy) emin = e < emin ? e : emin;
  info->e_min = emin;
  return KMC_OK;
}

int kmc_cnt_dispersion(const kmc_cnt* c, double* energy) {
  if (!c || !energy) return cfail(KMC_ERR_ARG, "kmc_cnt_dispersion: null argument");
  std::memcpy(energy, c->m.ex_energy.data(), c->m.ex_energy.size() * sizeof(double));
  return KMC_OK;
}

int kmc_cnt_electrons(const kmc_cnt* c, double* energy, double* wavefunc) {
  if (!c) return cfail(KMC_ERR_ARG, "kmc_cnt_electrons: null argument");
  if (energy) std::memcpy(energy, c->m.el_energy.data(), c->m.el_energy.size() * sizeof(double));
  if (wavefunc) std::memcpy(wavefunc, c->m.el_wf.data(), c->m.el_wf.size() * sizeof(cplx));
  return KMC_OK;
}

int kmc_cnt_interaction(const kmc_cnt* c, double* vq, double* PI, double* eps) {
  if (!c) return cfail(KMC_ERR_ARG, "kmc_cnt_interaction: null argument");
  if (vq) std::memcpy(vq, c->m.vq.data(), c->m.vq.size() * sizeof(cplx));
  if (PI) std::memcpy(PI, c->m.PI.data(), c->m.PI.size() * sizeof(double));
  if (eps) std::memcpy(eps, c->m.eps.data(), c->m.eps.size() * sizeof(double));
  return KMC_OK;
}

int kmc_cnt_exciton_states(const kmc_cnt* c, double* psi, int32_t* idx) {
  if (!c) return cfail(KMC_ERR_ARG, "kmc_cnt_exciton_states: null argument");
  if (psi) std::memcpy(psi, c->m.ex_psi.data(), c->m.ex_psi.size() * sizeof(cplx));
  if (idx) std::memcpy(idx, c->m.ex_idx.data(), c->m.ex_idx.size() * sizeof(int32_t));
  return KMC_OK;
}

int kmc_cnt_singlet_kernel(const kmc_cnt* c, int32_t ik_cm_idx, double* kernel) {
  if (!c || !kernel) return cfail(KMC_ERR_ARG, "kmc_cnt_singlet_kernel: null argument");
  if (ik_cm_idx < 0 || ik_cm_idx >= c->m.nk_cm)
    return cfail(KMC_ERR_ARG, "kmc_cnt_singlet_kernel: ik_cm index out of range");
  const std::vector<cplx> H = c->m.singlet_kernel(ik_cm_idx);
  std::memcpy(kernel, H.data(), H.size() * sizeof(cplx));
  return KMC_OK;
}

int kmc_cnt_atoms(const kmc_cnt* c, double* pos_a, double* pos_b, double* pos_u_3d) {
  if (!c) return cfail(KMC_ERR_ARG, "kmc_cnt_atoms: null argument");
  const cnt_model& m = c->m;
  for (int i = 0; i < m.Nu; ++i) {
    if (pos_a) {
      pos_a[2 * i] = m.pos_a[i].x;
      pos_a[2 * i + 1] = m.pos_a[i].y;
    }
    if (pos_b) {
      pos_b[2 * i] = m.pos_b[i].x;
      pos_b[2 * i + 1] = m.pos_b[i].y;
    }
    if (pos_u_3d) {
      pos_u_3d[3 * i] = m.pos_u_3d[i].x;
      pos_u_3d[3 * i + 1] = m.pos_u_3d[i].y;
      pos_u_3d[3 * i + 2] = m.pos_u_3d[i].z;
    }
  }
  return KMC_OK;
}

int kmc_hermitian_eig(int32_t n, const double* a, double* evals, double* vecs) {
  if (n < 1 || !a || !evals) return cfail(KMC_ERR_ARG, "kmc_hermitian_eig: bad argument");
  std::vector<cplx> A((size_t)n * n);
  std::memcpy(A.data(), a, A.size() * sizeof(cplx));
  std::vector<double> ev;
  std::vector<cplx> V;
  hermitian_eig(n, A, ev, V);
  std::memcpy(evals, ev.data(), n * sizeof(double));
  if (vecs) std::memcpy(vecs, V.data(), V.size() * sizeof(cplx));
  return KMC_OK;
}

}  // extern "C"
