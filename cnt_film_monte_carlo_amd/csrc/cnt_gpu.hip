// cnt_gpu.hip -- the two large sums of the CNT exciton model on the GPU: the Ohno-potential
// Fourier transform v(q) (calculate_vq, src/exciton_transfer/cnt.cpp:520-631) and the RPA
// polarization PI(q) (calculate_polarization, cnt.cpp:634-732).  Together they are ~95 % of the
// host model's time (v(q): (2 nk - 1)(2 n_mu - 1) x 4 sums of Nu ncells complex exponentials;
// PI(q): (2 nk - 1)(2 n_mu - 1) sums of nk n_mu band-pair terms).  One workgroup per q; the
// arguments are formed with the host model's expressions (contraction off), so only the
// order of the sums and the device sincos / hypot differ from the host's.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cnt_model.h"

namespace cntkmc {
namespace {

constexpr int CNT_WG = 256;

// Tree sum of the NV values per thread over the workgroup; thread 0 receives the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*lds)[CNT_WG]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NV; ++i) lds[i][t] = v[i];
  __syncthreads();
  for (int w = CNT_WG / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int i = 0; i < NV; ++i) lds[i][t] += lds[i][t + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = lds[i][0];
}

struct vq_args {
  int nk, n_mu, nmu_q, natom;
  double dkx, dky, k1x, k1y, norm;
  const double* rel;  // [4][natom][2] relative atom positions (cnt.cpp:555-601)
  const double* den;  // [4][natom]    Upp / sqrt(coeff r^2 + 1)
  double* vq;         // [nq][nmu_q][4] complex, interleaved
};

// v(q) for one (iq, mu): sum_k exp(i q . r_k) Upp / sqrt(coeff |r_k|^2 + 1) / (2 Nu ncells) for
// each of the four sublattice pairs (cnt.cpp:603-628).
__global__ __launch_bounds__(CNT_WG) void k_cnt_vq(vq_args a) {
#pragma clang fp contract(off)
  __shared__ double lds[8][CNT_WG];
  const int iqi = blockIdx.x / a.nmu_q, mi = blockIdx.x % a.nmu_q;
  const int iq = iqi - (a.nk - 1), mu = mi - (a.n_mu - 1);
  const double qx = double(iq) * a.dkx + double(mu) * a.k1x;
  const double qy = double(iq) * a.dky + double(mu) * a.k1y;
  double acc[8];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double re = 0.0, im = 0.0;
    const double* r = a.rel + (size_t)s * a.natom * 2;
    const double* d = a.den + (size_t)s * a.natom;
    for (int k = threadIdx.x; k < a.natom; k += CNT_WG) {
      const double arg = qx * r[2 * k] + qy * r[2 * k + 1];
      double sn, cs;
      sincos(arg, &sn, &cs);
      re += cs * d[k];
      im += sn * d[k];
    }
    acc[2 * s] = re;
    acc[2 * s + 1] = im;
  }
  block_sum<8>(acc, lds);
  if (threadIdx.x < 8) a.vq[(size_t)blockIdx.x * 8 + threadIdx.x] = acc[threadIdx.x] / a.norm;
}

struct pi_args {
  int nk, n_mu, nmu_q, nkM;
  const double* el;  // [band][ik][mu]
  const double* wf;  // [mu][ik][band][atom] complex, interleaved
  double* pi;        // [nq][nmu_q]
};

// PI(q) = 2 sum_{k, mu_k} |<v(k)|c(k+q)>|^2 / (E_c(k+q) - E_v(k)) + |<c(k)|v(k+q)>|^2 /
// (E_c(k) - E_v(k+q)) with k + q folded back into the K2-extended zone (cnt.cpp:660-728).
__global__ __launch_bounds__(CNT_WG) void k_cnt_pi(pi_args a) {
#pragma clang fp contract(off)
  __shared__ double lds[1][CNT_WG];
  const int iqi = blockIdx.x / a.nmu_q, mi = blockIdx.x % a.nmu_q;
  const int iq = iqi - (a.nk - 1), mu_q = mi - (a.n_mu - 1);
  const int n_states = a.nk * a.n_mu;
  double acc[1] = {0.0};
  for (int j = threadIdx.x; j < n_states; j += CNT_WG) {
    const int ik = j / a.n_mu, mu_k = j % a.n_mu;
    int mu_kq = mu_k + mu_q, ikq = ik + iq;
    while (mu_kq >= a.n_mu) {
      mu_kq -= a.n_mu;
      ikq += a.nkM;
    }
    while (mu_kq < 0) {
      mu_kq += a.n_mu;
      ikq -= a.nkM;
    }
    while (ikq >= a.nk) ikq -= a.nk;
    while (ikq < 0) ikq += a.nk;
    const double* w0 = a.wf + (((size_t)mu_k * a.nk + ik) * 2) * 4;     // bands 0,1 at k
    const double* w1 = a.wf + (((size_t)mu_kq * a.nk + ikq) * 2) * 4;   // bands 0,1 at k+q
    // conj(v(k)) . c(k+q) and conj(c(k)) . v(k+q), summed over the two atoms
    const double d1r = (w0[0] * w1[4] + w0[1] * w1[5]) + (w0[2] * w1[6] + w0[3] * w1[7]);
    const double d1i = (w0[0] * w1[5] - w0[1] * w1[4]) + (w0[2] * w1[7] - w0[3] * w1[6]);
    const double d2r = (w0[4] * w1[0] + w0[5] * w1[1]) + (w0[6] * w1[2] + w0[7] * w1[3]);
    const double d2i = (w0[4] * w1[1] - w0[5] * w1[0]) + (w0[6] * w1[3] - w0[7] * w1[2]);
    const double h1 = hypot(d1r, d1i), h2 = hypot(d2r, d2i);
    const size_t band1 = (size_t)a.nk * a.n_mu;
    const double e_c_kq = a.el[band1 + (size_t)ikq * a.n_mu + mu_kq];
    const double e_v_k = a.el[(size_t)ik * a.n_mu + mu_k];
    const double e_c_k = a.el[band1 + (size_t)ik * a.n_mu + mu_k];
    const double e_v_kq = a.el[(size_t)ikq * a.n_mu + mu_kq];
    acc[0] += (h1 * h1) / (e_c_kq - e_v_k) + (h2 * h2) / (e_c_k - e_v_kq);
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) a.pi[blockIdx.x] = 2 * acc[0];
}

// Device buffers of one call, freed on every return path.
struct dev_bufs {
  std::vector<void*> p;
  ~dev_bufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t get(T** out, size_t n) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, n * sizeof(T) + 1);
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T*>(q);
    return e;
  }
};

std::string hip_msg(const char* what, hipError_t e) {
  return std::string("cnt_model on the GPU: ") + what + ": " + hipGetErrorString(e);
}

#define CNT_HIP(what, call)                      \
  do {                                           \
    const hipError_t e_ = (call);                \
    if (e_ != hipSuccess) return hip_msg(what, e_); \
  } while (0)

}  // namespace

std::string cnt_sums_on_device(int device, const cnt_model& m, int natom, const double* rel,
                               const double* den, double norm, std::vector<cplx>& vq,
                               std::vector<double>& PI) {
  int n_dev = 0;
  CNT_HIP("hipGetDeviceCount", hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev) return "cnt_model on the GPU: no device " + std::to_string(device);
  CNT_HIP("hipSetDevice", hipSetDevice(device));
  const size_t n_q = (size_t)m.nq * m.nmu_q;
  if (n_q == 0 || n_q > 0x7FFFFFFF) return "cnt_model on the GPU: q grid out of range";
  dev_bufs b;
  double *d_rel, *d_den, *d_vq, *d_el, *d_wf, *d_pi;
  CNT_HIP("hipMalloc", b.get(&d_rel, (size_t)4 * natom * 2));
  CNT_HIP("hipMalloc", b.get(&d_den, (size_t)4 * natom));
  CNT_HIP("hipMalloc", b.get(&d_vq, n_q * 8));
  CNT_HIP("hipMalloc", b.get(&d_el, m.el_energy.size()));
  CNT_HIP("hipMalloc", b.get(&d_wf, m.el_wf.size() * 2));
  CNT_HIP("hipMalloc", b.get(&d_pi, n_q));
  hipStream_t st = nullptr;
  CNT_HIP("hipStreamCreate", hipStreamCreate(&st));
  struct stream_guard {
    hipStream_t s;
    ~stream_guard() { (void)hipStreamDestroy(s); }
  } guard{st};
  CNT_HIP("copy", hipMemcpyAsync(d_rel, rel, sizeof(double) * 4 * natom * 2, hipMemcpyHostToDevice, st));
  CNT_HIP("copy", hipMemcpyAsync(d_den, den, sizeof(double) * 4 * natom, hipMemcpyHostToDevice, st));
  CNT_HIP("copy", hipMemcpyAsync(d_el, m.el_energy.data(), sizeof(double) * m.el_energy.size(),
                                 hipMemcpyHostToDevice, st));
  CNT_HIP("copy", hipMemcpyAsync(d_wf, m.el_wf.data(), sizeof(cplx) * m.el_wf.size(),
                                 hipMemcpyHostToDevice, st));
  vq_args va{m.nk, m.n_mu, m.nmu_q, natom, m.dk_l.x, m.dk_l.y, m.K1.x, m.K1.y, norm,
             d_rel, d_den, d_vq};
  hipLaunchKernelGGL(k_cnt_vq, dim3((unsigned)n_q), dim3(CNT_WG), 0, st, va);
  CNT_HIP("k_cnt_vq", hipGetLastError());
  pi_args pa{m.nk, m.n_mu, m.nmu_q, m.nk_K1 * m.M, d_el, d_wf, d_pi};
  hipLaunchKernelGGL(k_cnt_pi, dim3((unsigned)n_q), dim3(CNT_WG), 0, st, pa);
  CNT_HIP("k_cnt_pi", hipGetLastError());
  vq.assign(n_q * 4, cplx(0, 0));
  PI.assign(n_q, 0.0);
  CNT_HIP("copy", hipMemcpyAsync(vq.data(), d_vq, sizeof(cplx) * n_q * 4, hipMemcpyDeviceToHost, st));
  CNT_HIP("copy", hipMemcpyAsync(PI.data(), d_pi, sizeof(double) * n_q, hipMemcpyDeviceToHost, st));
  CNT_HIP("hipStreamSynchronize", hipStreamSynchronize(st));
  return std::string();
}

}  // namespace cntkmc
