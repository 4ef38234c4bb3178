// kmc_transfer.hip -- exact first-order exciton-transfer rate (exciton_transfer::first_order,
// src/exciton_transfer/exciton_transfer.cpp:380-441) for batches of pair geometries.
//
// Host: relevant states, partition function, energy matching and the Q factors of every
// matched pair (exciton_transfer.cpp:226-304, exciton_transfer.h:167-196) collapse into one
// coefficient per pair of centre-of-mass momenta (a, b):
//     rate = sum_ab C_ab |J_ab|^2,
//     C_ab = sum_{matched (i,f) with momenta (a,b)} (2 pi / hb) (exp(-E_i/kT) / Z)
//                 |Q_if|^2 / (L_i L_f) * Lorentzian(E_i - E_f),
// since M = |Q J| / sqrt(L_i L_f) (exciton_transfer.cpp:415) and J depends on the two states only
// through their momenta.  Device: J_ab = sum_i sum_j u_a(i) v_b(j) / |R_i - R_j| over the A-site
// unit cells of both tubes (exciton_transfer.cpp:307-377), one workgroup per geometry, each donor
// atom moved into the acceptor's frame (the inverse of make_Ru_3d's shift / rotation) so that the
// acceptor atoms and their phases are read unchanged for every geometry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cntkmc.h"
#include "cnt_model.h"

using namespace cntkmc;


namespace {

constexpr int CHUNK = 16;     // acceptor momenta per kernel pass (W_b held in registers)
constexpr int MAXTERMS = 256; // nonzero C_ab terms per pass (one owning thread each)
constexpr int GEOM_BATCH = 65536;  // geometries per launch (one workgroup each)
// Launches of fewer geometries than this (the surrogate's exact fallback, single-geometry calls)
// give each (geometry, 256-atom donor chunk) its own workgroup, so that the GPU holds enough waves
// to cover the scalar-load latency; the chunks' shares of J go through PART_CAP entries of device
// memory and a finishing kernel.
constexpr int64_t SPLIT_BELOW = 4096;
constexpr size_t PART_CAP = size_t(1) << 21;

struct ex_state {
  int ik_cm_idx, n;
  double energy;
  int ik_cm;
};

// get_relevant_states (exciton_transfer.cpp:226-270)
std::vector<ex_state> relevant_states(const cnt_model& c, double min_energy, double T) {
  const double thr = min_energy + std::abs(std::log(1.e-3) * phys::kb * T);
  std::vector<ex_state> out;
  for (int n = 0; n < c.nk_relev; n++)
    for (int ic = 0; ic < c.nk_cm; ic++) {
      const double e = c.ex_energy[(size_t)ic * c.nk_relev + n];
      if (e <= thr) out.push_back({ic, n, e, c.ik_cm0 + ic});
    }
  std::stable_sort(out.begin(), out.end(),
                   [](const ex_state& a, const ex_state& b) { return a.energy < b.energy; });
  return out;
}

// Q_partial of calculate_Q (exciton_transfer.cpp:276-296)
cplx q_partial(const cnt_model& c, const ex_state& s) {
  const double kx = s.ik_cm * c.dk_l.x, ky = s.ik_cm * c.dk_l.y;
  const cplx ef0 = std::exp(cplx(0., +1.) * (kx * 0.0 + ky * 0.0));
  const cplx ef1 = std::exp(cplx(0., +1.) * (kx * c.aCC.x + ky * c.aCC.y));
  cplx qp = 0;
  const cplx* psi = &c.ex_psi[((size_t)s.ik_cm_idx * c.nk_relev + s.n) * c.nk_c];
  for (int k = 0; k < c.nk_c; k++) {
    const int32_t* id = &c.ex_idx[((size_t)s.ik_cm_idx * c.nk_c + k) * 4];
    const cplx ccA = c.wf(id[1], id[0], 1, 0), ccB = c.wf(id[1], id[0], 1, 1);
    const cplx cvA = c.wf(id[3], id[2], 0, 0), cvB = c.wf(id[3], id[2], 0, 1);
    qp += psi[k] * (ccA * std::conj(cvA) * ef0 + ccB * std::conj(cvB) * ef1);
  }
  return qp;
}

// make_Ru_3d before shift and rotation (exciton_transfer.cpp:310-322): all unit cells, centred
// in y.  make_Ru_2d (exciton_transfer.cpp:341-351): the unrolled positions.
void tube_atoms(const cnt_model& c, std::vector<double>& r3, std::vector<vec2>& r2) {
  const int n = c.Nu * c.length;
  r3.assign((size_t)n * 4, 0.0);
  r2.assign(n, vec2{});
  for (int i = 0; i < c.length; i++)
    for (int a = 0; a < c.Nu; a++) {
      const int k = i * c.Nu + a;
      r3[4 * k + 0] = c.pos_u_3d[a].x + i * c.t_vec_3d.x;
      r3[4 * k + 1] = c.pos_u_3d[a].y + i * c.t_vec_3d.y;
      r3[4 * k + 2] = c.pos_u_3d[a].z + i * c.t_vec_3d.z;
      r2[k] = {c.pos_a[a].x + i * c.t_vec.x, c.pos_a[a].y + i * c.t_vec.y};
    }
  double ymax = r3[1], ymin = r3[1];
  for (int k = 0; k < n; k++) {
    ymax = std::max(ymax, r3[4 * k + 1]);
    ymin = std::min(ymin, r3[4 * k + 1]);
  }
  const double mid = (ymax + ymin) / 2.;
  for (int k = 0; k < n; k++) r3[4 * k + 1] -= mid;
}

// ---------------------------------------------------------------------------------------------
// device: one workgroup per geometry
// ---------------------------------------------------------------------------------------------
struct exact_args {
  int32_t na, nb;                 // atoms per tube
  int32_t nmb;                    // acceptor momenta of this pass (the kernel's NMB)
  int32_t n_terms;                // nonzero (a, b) terms of this pass
  int32_t accumulate;             // 0: out = pass sum; 1: out += pass sum
  const double4* donor;           // [na] centred positions (w unused)
  const double2* u;               // [nma][na]  exp(-i k_a . r)
  const double* rec;              // acceptor records of this pass: 4 + nb (2 nmb + 4) doubles
  const int32_t* term_a;          // [n_terms] donor momentum slot
  const int32_t* term_b;          // [n_terms] acceptor momentum slot within the pass
  const double* term_c;           // [n_terms] C_ab
  int64_t n_geom;
  const double *theta, *z, *a1, *a2;
  const int64_t* sel;             // optional: geometry g reads entry sel[g] and writes out[sel[g]]
  const double2* trig;            // [n_geom] (cos, sin) of geometry g's theta
  int32_t split;                  // 1: workgroup (g, blockIdx.y) does donor chunk blockIdx.y only
  double2* part;                  // split: [n_geom][chunks][n_terms] shares of J_ab
  double* out;
  double clamp_lo[3], clamp_hi[3];  // z, a1, a2 box
};

// 1/sqrt(x): the hardware estimate y (relative error about 2^-23) and one third-order step of the
// binomial series, 1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3 e^2/8 + ...) with e = 1 - x y^2;
// the dropped 5 e^3/16 term is below 2^-60, so the result is within rounding of FP64 (5 VALU
// operations where two Newton steps take 7).
__device__ __forceinline__ double rsqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(y * e, fma(0.375, e, 0.5), y);
}

// Wave-wide sum; every lane ends with the same value.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// rate = sum over terms of C_ab |J_ab|^2 (thread t < n_terms holds term t's J), written or added
// to out[e] by thread 0.  Products are written as explicit FMAs here and in the J sums, so that
// every kernel that shares this arithmetic rounds it the same way.
__device__ __forceinline__ void finish_rate(const exact_args& a, int64_t e, bool owner, double jr,
                                            double ji, double* s_red, int tid) {
  double part = owner ? a.term_c[tid] * fma(ji, ji, jr * jr) : 0.0;
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if ((tid & 63) == 0) s_red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    const double sum = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
    a.out[e] = a.accumulate ? a.out[e] + sum : sum;
  }
  __syncthreads();
}

// One workgroup of 256 threads per geometry.  Donor atoms go in chunks of 256 (one per thread).
// make_Ru_3d places the acceptor at R_j = Rz(theta) (b_j + s2 y) + z_s z and the donor at
// r_i + s1 y; since Rz leaves z and lengths alone, |R_i - R_j| = |q_i - b_j| with the donor atom
// taken into the acceptor's frame, q_i = Rz(-theta) (r_i + s1 y - z_s z) - s2 y.  The acceptor
// records (b_j and v_b(j) of the pass's momenta) are then the same for every geometry and are
// read with wave-uniform addresses (scalar loads): nothing is staged in LDS and the per-pair cost
// is the distance, one reciprocal square root and 2 NMB FMAs on the VALU.  For its atom i a
// thread accumulates W_b(i) = sum_j v_b(j) / |q_i - b_j|; the chunk's share of
// J_ab = sum_i u_a(i) W_b(i) is then summed per term across each wave and the four wave sums
// are added by the term's owning thread (t < n_terms).  Split launches run one chunk per workgroup
// and leave each chunk's four-wave sum in part[] for k_exact_finish, which adds the chunks in the
// same order, so both launch shapes give the same bits.
template <int NMB>
__global__ __launch_bounds__(256) void k_exact_rate(exact_args a, const double* __restrict__ rec,
                                                    const int32_t* __restrict__ term_a,
                                                    const int32_t* __restrict__ term_b) {
  constexpr int S = 4 + 2 * NMB;   // doubles per acceptor record: (Re, Im) v_b, next x, y, z, 0
  __shared__ double2 s_part[4][MAXTERMS];
  __shared__ double s_red[4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int64_t g = blockIdx.x; g < a.n_geom; g += gridDim.x) {
    const int64_t e = a.sel ? a.sel[g] : g;
    const double zs = fmin(fmax(a.z[e], a.clamp_lo[0]), a.clamp_hi[0]);
    const double s1 = fmin(fmax(a.a1[e], a.clamp_lo[1]), a.clamp_hi[1]);
    const double s2 = fmin(fmax(a.a2[e], a.clamp_lo[2]), a.clamp_hi[2]);
    const double cth = a.trig[g].x, sth = a.trig[g].y;
    const bool owner = tid < a.n_terms;
    double jr = 0.0, ji = 0.0;
    const int i_lo = a.split ? 256 * (int)blockIdx.y : 0;
    const int i_hi = a.split ? min(a.na, i_lo + 256) : a.na;
    for (int i0 = i_lo; i0 < i_hi; i0 += 256) {
      const int i = i0 + tid;
      const bool live = i < a.na;
      const double4 ri = live ? a.donor[i] : make_double4(0, 0, 0, 0);
      const double ry = ri.y + s1;
      const double qx = cth * ri.x + sth * ry;
      const double qy = cth * ry - sth * ri.x - s2;
      const double qz = ri.z - zs;
      double wr[NMB], wi[NMB];
#pragma unroll
      for (int q = 0; q < NMB; q++) wr[q] = wi[q] = 0.0;
      // record j carries v_b(j) and the position of atom j + 1 (the first position leads the
      // array), so each pair's scalar loads are issued together and waited on once
      double bx = rec[0], by = rec[1], bz = rec[2];
      asm volatile("" : "+s"(bx), "+s"(by), "+s"(bz));  // loaded before the loop, not pending in it
      const double* b = rec + 4;
      for (int j = 0; j < a.nb; j++) {
        // the record's loads go out first and the distance and square root cover them (a
        // scheduling fence keeps them there, a second one keeps the position hand-over after
        // the FMAs so that the loads are waited on only there)
        double v[2 * NMB];
#pragma unroll
        for (int k = 0; k < 2 * NMB; k++) v[k] = b[k];
        const double nx = b[2 * NMB], ny = b[2 * NMB + 1], nz = b[2 * NMB + 2];
        __builtin_amdgcn_sched_barrier(0);
        const double dx = qx - bx, dy = qy - by, dz = qz - bz;
        const double r = rsqrt_f64(fma(dz, dz, fma(dy, dy, dx * dx)));
#pragma unroll
        for (int q = 0; q < NMB; q++) {
          wr[q] += v[2 * q] * r;
          wi[q] += v[2 * q + 1] * r;
        }
        __builtin_amdgcn_sched_barrier(0);
        bx = nx;
        by = ny;
        bz = nz;
        // carry the next position in SGPRs (the compiler would otherwise reload it at the top
        // of the next pair)
        asm volatile("" : "+s"(bx), "+s"(by), "+s"(bz));
        b += S;
      }
      // this chunk's u_a(i) W_b(i), summed over the wave for every term (a, b)
#pragma unroll
      for (int q = 0; q < NMB; q++) {
        const double w_r = live ? wr[q] : 0.0, w_i = live ? wi[q] : 0.0;
        for (int t = 0; t < a.n_terms; t++) {
          if (term_b[t] != q) continue;
          const double2 u = live ? a.u[(size_t)term_a[t] * a.na + i] : make_double2(0, 0);
          const double pr = wave_sum(fma(u.x, w_r, -(u.y * w_i)));
          const double pi = wave_sum(fma(u.x, w_i, u.y * w_r));
          if (lane == 0) s_part[wave][t] = make_double2(pr, pi);
        }
      }
      __syncthreads();
      if (owner) {
        const double pr = (s_part[0][tid].x + s_part[1][tid].x) + (s_part[2][tid].x + s_part[3][tid].x);
        const double pi = (s_part[0][tid].y + s_part[1][tid].y) + (s_part[2][tid].y + s_part[3][tid].y);
        if (a.split) {
          a.part[((size_t)g * gridDim.y + blockIdx.y) * a.n_terms + tid] = make_double2(pr, pi);
        } else {
          jr += pr;
          ji += pi;
        }
      }
      __syncthreads();
    }
    if (!a.split) finish_rate(a, e, owner, jr, ji, s_red, tid);
  }
}

// Split launches: J_ab = sum over the chunks' shares in chunk order, then the rate.
__global__ __launch_bounds__(256) void k_exact_finish(exact_args a, int n_chunks) {
  __shared__ double s_red[4];
  const int tid = threadIdx.x;
  for (int64_t g = blockIdx.x; g < a.n_geom; g += gridDim.x) {
    const int64_t e = a.sel ? a.sel[g] : g;
    const bool owner = tid < a.n_terms;
    double jr = 0.0, ji = 0.0;
    if (owner)
      for (int c = 0; c < n_chunks; c++) {
        const double2 p = a.part[((size_t)g * n_chunks + c) * a.n_terms + tid];
        jr += p.x;
        ji += p.y;
      }
    finish_rate(a, e, owner, jr, ji, s_red, tid);
  }
}

int pass_width(int nmb) { return std::max(1, std::min(nmb, CHUNK)); }

// cos and sin of each geometry's theta, ahead of the rate passes (sincos' argument reduction would
// otherwise set the register budget, and with it the occupancy, of every rate kernel).
__global__ void k_trig(int64_t n, const int64_t* sel, const double* theta, double2* trig) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  double s, c;
  sincos(theta[sel ? sel[g] : g], &s, &c);
  trig[g] = make_double2(c, s);
}

// Geometries whose transfer has no matched state pair: the rate is exactly zero.
__global__ void k_zero_rates(int64_t n, const int64_t* sel, double* out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) out[sel ? sel[g] : g] = 0.0;
}

// One pass over a chunk of at most 16 acceptor momenta.
hipError_t launch_pass(const exact_args& a, hipStream_t st) {
  const unsigned grid = (unsigned)std::min<int64_t>(a.n_geom, GEOM_BATCH);
  const int chunks = (a.na + 255) / 256;
  const dim3 g(grid, a.split ? chunks : 1), b(256);
  switch (a.nmb) {
#define EXACT_CASE(n) \
  case n: hipLaunchKernelGGL(k_exact_rate<n>, g, b, 0, st, a, a.rec, a.term_a, a.term_b); break;
    EXACT_CASE(1) EXACT_CASE(2) EXACT_CASE(3) EXACT_CASE(4) EXACT_CASE(5) EXACT_CASE(6)
    EXACT_CASE(7) EXACT_CASE(8) EXACT_CASE(9) EXACT_CASE(10) EXACT_CASE(11) EXACT_CASE(12)
    EXACT_CASE(13) EXACT_CASE(14) EXACT_CASE(15) EXACT_CASE(16)
#undef EXACT_CASE
    default: return hipErrorInvalidValue;
  }
  if (a.split) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_exact_finish, dim3(grid), b, 0, st, a, chunks);
  }
  return hipGetLastError();
}

thread_local std::string g_tr_err;
int tfail(int code, const std::string& m) {
  g_tr_err = m;
  return code;
}

}  // namespace

struct kmc_transfer {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  kmc_transfer_info info{};
  std::vector<int> mom_a, mom_b;            // ik_cm of each momentum slot
  std::vector<int32_t> term_a, term_b;      // nonzero terms
  std::vector<double> term_c;
  // acceptor slots [b0, b0 + 16), terms [t0, t0 + nt), records at h_rec[rec0]
  struct pass { int b0, width, t0, nt; size_t rec0; };
  std::vector<pass> passes;
  double clamp_lo[3] = {-1e300, -1e300, -1e300}, clamp_hi[3] = {1e300, 1e300, 1e300};
  std::vector<double> h_donor, h_acc;
  std::vector<vec2> r2d, r2a;
  std::vector<cplx> h_u;
  std::vector<double> h_rec;                // per pass: 4 + nb (2 width + 4) doubles of acceptor records
  bool on_device = false;
  double4* d_donor = nullptr;
  double2* d_u = nullptr;
  double* d_rec = nullptr;
  double2* d_trig = nullptr;                // [GEOM_BATCH] (cos, sin) of the batch's geometries
  double2* d_part = nullptr;                // [PART_CAP] chunk shares of J for split launches
  int32_t *d_ta = nullptr, *d_tb = nullptr;
  double* d_tc = nullptr;
  double* d_geo = nullptr;  // scratch for host-array calls: 4 n doubles
  double* d_out = nullptr;
  size_t geo_cap = 0;
};

namespace {

template <typename T>
hipError_t put(T** d, const std::vector<T>& h) {
  hipError_t e = hipMalloc((void**)d, std::max<size_t>(1, h.size()) * sizeof(T));
  if (e == hipSuccess && !h.empty()) e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

int upload(kmc_transfer* t) {
  if (t->on_device) return KMC_OK;
  hipError_t e = hipSetDevice(t->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&t->ev0);
  if (e == hipSuccess) e = hipEventCreate(&t->ev1);
  if (e == hipSuccess) e = put((double**)&t->d_donor, t->h_donor);
  if (e == hipSuccess) e = put((cplx**)&t->d_u, t->h_u);
  if (e == hipSuccess) e = put(&t->d_rec, t->h_rec);
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_trig, GEOM_BATCH * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_part, PART_CAP * sizeof(double2));
  if (e == hipSuccess) e = put(&t->d_ta, t->term_a);
  if (e == hipSuccess) e = put(&t->d_tb, t->term_b);
  if (e == hipSuccess) e = put(&t->d_tc, t->term_c);
  if (e != hipSuccess)
    return tfail(e == hipErrorNoDevice ? KMC_ERR_NODEVICE : KMC_ERR_HIP,
                 std::string("kmc_transfer: device setup failed: ") + hipGetErrorString(e));
  t->on_device = true;
  return KMC_OK;
}

exact_args args_of(const kmc_transfer* t) {
  exact_args a{};
  a.na = t->info.n_atoms_donor;
  a.nb = t->info.n_atoms_acceptor;
  a.donor = t->d_donor;
  a.u = t->d_u;
  a.trig = t->d_trig;
  a.part = t->d_part;
  for (int k = 0; k < 3; k++) {
    a.clamp_lo[k] = t->clamp_lo[k];
    a.clamp_hi[k] = t->clamp_hi[k];
  }
  return a;
}

// rate = sum over passes of sum_{terms of the pass} C_ab |J_ab|^2; the passes run in order on
// one stream, the first writes and the rest add.  Geometries go in batches (GEOM_BATCH, or what
// the chunk shares of a split launch fit in PART_CAP), each with its (cos, sin) table first.
hipError_t launch_exact(const kmc_transfer* t, exact_args a, hipStream_t st) {
  if (a.n_geom == 0) return hipSuccess;
  if (t->passes.empty()) {
    const unsigned grid = (unsigned)((a.n_geom + 255) / 256);
    hipLaunchKernelGGL(k_zero_rates, dim3(grid), dim3(256), 0, st, a.n_geom, a.sel, a.out);
    return hipGetLastError();
  }
  int max_terms = 1;
  for (const auto& p : t->passes) max_terms = std::max(max_terms, p.nt);
  const size_t per_geom = (size_t)((a.na + 255) / 256) * max_terms;  // chunk shares per geometry
  a.split = a.n_geom < SPLIT_BELOW && per_geom <= PART_CAP;
  const int64_t batch = a.split ? std::min<int64_t>(GEOM_BATCH, (int64_t)(PART_CAP / per_geom))
                                : GEOM_BATCH;
  hipError_t e = hipSuccess;
  for (int64_t g0 = 0; g0 < a.n_geom && e == hipSuccess; g0 += batch) {
    exact_args b = a;
    b.n_geom = std::min<int64_t>(batch, a.n_geom - g0);
    if (a.sel) {
      b.sel = a.sel + g0;
    } else {
      b.theta += g0;
      b.z += g0;
      b.a1 += g0;
      b.a2 += g0;
      b.out += g0;
    }
    hipLaunchKernelGGL(k_trig, dim3((unsigned)((b.n_geom + 255) / 256)), dim3(256), 0, st, b.n_geom,
                       b.sel, b.theta, t->d_trig);
    e = hipGetLastError();
    for (size_t k = 0; k < t->passes.size() && e == hipSuccess; k++) {
      const auto& p = t->passes[k];
      b.nmb = p.width;
      b.n_terms = p.nt;
      b.accumulate = k > 0;
      b.rec = t->d_rec + p.rec0;
      b.term_a = t->d_ta + p.t0;
      b.term_b = t->d_tb + p.t0;
      b.term_c = t->d_tc + p.t0;
      e = launch_pass(b, st);
    }
  }
  return e;
}

}  // namespace

extern "C" {

const char* kmc_transfer_last_error(void) { return g_tr_err.c_str(); }

int kmc_transfer_create(int32_t device, const kmc_cnt* donor, const kmc_cnt* acceptor,
                        double temperature, double broadening, kmc_transfer** out) {
  if (!donor || !acceptor || !out || !(temperature > 0) || !(broadening > 0))
    return tfail(KMC_ERR_ARG, "kmc_transfer_create: bad argument");
  *out = nullptr;
  const cnt_model& D = donor->m;
  const cnt_model& A = acceptor->m;
  using namespace phys;
  // first_order (exciton_transfer.cpp:385-400)
  double min_energy = D.ex_energy[0];
  for (double e : D.ex_energy) min_energy = std::min(min_energy, e);
  const std::vector<ex_state> ds = relevant_states(D, min_energy, temperature);
  const std::vector<ex_state> as = relevant_states(A, min_energy, temperature);
  double Z = 0;
  for (const auto& s : ds) Z += std::exp(-s.energy / (kb * temperature));
  auto lorentzian = [&](double e) { return inv_pi * broadening / (e * e + broadening * broadening); };
  kmc_transfer* t = new kmc_transfer();
  t->device = device;
  // calculate_Q coefficient (exciton_transfer.cpp:298-300)
  const double coeff = (std::pow(q0, 2) * D.Au() * A.Au()) /
                       (16 * std::pow(pi, 3) * eps0 * D.radius * A.radius *
                        std::sqrt(D.length_in_meter() * A.length_in_meter()));
  const double LiLf = D.length_in_meter() * A.length_in_meter();
  std::vector<cplx> qd(ds.size()), qa(as.size());
  for (size_t i = 0; i < ds.size(); i++) qd[i] = q_partial(D, ds[i]);
  for (size_t f = 0; f < as.size(); f++) qa[f] = q_partial(A, as[f]);
  auto slot = [](std::vector<int>& v, int k) {
    for (size_t i = 0; i < v.size(); i++)
      if (v[i] == k) return (int)i;
    v.push_back(k);
    return (int)v.size() - 1;
  };
  struct term { int p, q; double c; };
  std::vector<term> acc;  // coefficients per momentum pair, in first-match order
  int n_pairs = 0;
  // match_states (exciton_transfer.h:167-196) and the sum of first_order (:411-427)
  for (size_t i = 0; i < ds.size(); i++)
    for (size_t f = 0; f < as.size(); f++) {
      const double de = ds[i].energy - as[f].energy;
      if (!(lorentzian(de) > 1.e-2 * lorentzian(0))) continue;
      ++n_pairs;
      const cplx Q = cplx(coeff) * std::conj(qd[i]) * qa[f];
      const double w = (2 * pi / hb) * (std::exp(-ds[i].energy / (kb * temperature)) / Z) *
                       std::norm(Q) / LiLf * lorentzian(de);
      const int p = slot(t->mom_a, ds[i].ik_cm), q = slot(t->mom_b, as[f].ik_cm);
      auto it = std::find_if(acc.begin(), acc.end(), [&](const term& x) { return x.p == p && x.q == q; });
      if (it == acc.end()) acc.push_back({p, q, w});
      else it->c += w;
    }
  if (t->mom_a.empty()) t->mom_a.push_back(0);
  if (t->mom_b.empty()) t->mom_b.push_back(0);
  const int nma = (int)t->mom_a.size(), nmb = (int)t->mom_b.size();
  // terms ordered by (donor slot, acceptor slot) and grouped into passes of 16 acceptor slots
  std::sort(acc.begin(), acc.end(), [](const term& x, const term& y) {
    return x.q / CHUNK != y.q / CHUNK ? x.q / CHUNK < y.q / CHUNK : x.p != y.p ? x.p < y.p : x.q < y.q;
  });
  for (const term& x : acc) {
    if (x.c == 0.0) continue;
    const int b0 = x.q / CHUNK * CHUNK;
    if (t->passes.empty() || t->passes.back().b0 != b0)
      t->passes.push_back({b0, pass_width(std::min(CHUNK, nmb - b0)), (int)t->term_c.size(), 0, 0});
    if (++t->passes.back().nt > MAXTERMS) {
      delete t;
      return tfail(KMC_ERR_CAPACITY, "kmc_transfer_create: more than 256 momentum terms "
                                     "within 16 acceptor momenta");
    }
    t->term_a.push_back(x.p);
    t->term_b.push_back(x.q - b0);
    t->term_c.push_back(x.c);
  }
  const int v_rows = t->passes.empty() ? 0 : t->passes.back().b0 + t->passes.back().width;
  kmc_transfer_info& in = t->info;
  in.n_donor_states = (int)ds.size();
  in.n_acceptor_states = (int)as.size();
  in.n_pairs = n_pairs;
  in.n_mom_donor = nma;
  in.n_mom_acceptor = nmb;
  in.n_terms = (int)t->term_c.size();
  in.n_atoms_donor = D.Nu * D.length;
  in.n_atoms_acceptor = A.Nu * A.length;
  in.partition_Z = Z;
  in.min_energy = min_energy;
  // host copies of the device data; uploaded on the first rate call
  tube_atoms(D, t->h_donor, t->r2d);
  tube_atoms(A, t->h_acc, t->r2a);
  t->h_u.assign((size_t)nma * t->r2d.size(), cplx(0, 0));
  std::vector<cplx> h_v((size_t)std::max(v_rows, nmb) * t->r2a.size(), cplx(0, 0));
  for (int p = 0; p < nma; p++) {
    const double kx = t->mom_a[p] * D.dk_l.x, ky = t->mom_a[p] * D.dk_l.y;
    for (size_t i = 0; i < t->r2d.size(); i++)
      t->h_u[p * t->r2d.size() + i] = std::exp(-cplx(0, 1) * (kx * t->r2d[i].x + ky * t->r2d[i].y));
  }
  for (int q = 0; q < nmb; q++) {
    const double kx = t->mom_b[q] * A.dk_l.x, ky = t->mom_b[q] * A.dk_l.y;
    for (size_t j = 0; j < t->r2a.size(); j++)
      h_v[q * t->r2a.size() + j] = std::exp(+cplx(0, 1) * (kx * t->r2a[j].x + ky * t->r2a[j].y));
  }
  // acceptor records of each pass: the first atom's position (x, y, z, 0), then for every atom
  // j its v_b(j) over the pass's momenta followed by the position of atom j + 1 (zeros after the
  // last atom)
  const size_t nb = t->r2a.size();
  auto put_pos = [&](size_t j) {
    for (int k = 0; k < 3; k++) t->h_rec.push_back(j < nb ? t->h_acc[4 * j + k] : 0.0);
    t->h_rec.push_back(0.0);
  };
  for (auto& p : t->passes) {
    p.rec0 = t->h_rec.size();
    put_pos(0);
    for (size_t j = 0; j < nb; j++) {
      for (int q = 0; q < p.width; q++) {
        t->h_rec.push_back(h_v[(size_t)(p.b0 + q) * nb + j].real());
        t->h_rec.push_back(h_v[(size_t)(p.b0 + q) * nb + j].imag());
      }
      put_pos(j + 1);
    }
  }
  *out = t;
  return KMC_OK;
}

int kmc_transfer_destroy(kmc_transfer* t) {
  if (!t) return KMC_OK;
  if (t->on_device) {
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    void* bufs[] = {t->d_donor, t->d_u, t->d_rec, t->d_trig, t->d_part, t->d_ta, t->d_tb, t->d_tc, t->d_geo, t->d_out};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    if (t->stream) (void)hipStreamDestroy(t->stream);
  }
  delete t;
  return KMC_OK;
}

int kmc_transfer_get_info(const kmc_transfer* t, kmc_transfer_info* info) {
  if (!t || !info) return tfail(KMC_ERR_ARG, "kmc_transfer_get_info: null argument");
  *info = t->info;
  return KMC_OK;
}

int kmc_transfer_terms(const kmc_transfer* t, int32_t* ik_cm_donor, int32_t* ik_cm_acceptor,
                       double* coef) {
  if (!t) return tfail(KMC_ERR_ARG, "kmc_transfer_terms: null argument");
  for (int k = 0; k < t->info.n_terms; k++) {
    if (ik_cm_donor) ik_cm_donor[k] = t->mom_a[t->term_a[k]];
    if (ik_cm_acceptor) ik_cm_acceptor[k] = t->mom_b[t->term_b[k]];
    if (coef) coef[k] = t->term_c[k];
  }
  return KMC_OK;
}

int kmc_transfer_set_clamp(kmc_transfer* t, const double lo[3], const double hi[3]) {
  if (!t || !lo || !hi) return tfail(KMC_ERR_ARG, "kmc_transfer_set_clamp: null argument");
  for (int k = 0; k < 3; k++) {
    if (!(lo[k] <= hi[k])) return tfail(KMC_ERR_ARG, "kmc_transfer_set_clamp: lo > hi");
    t->clamp_lo[k] = lo[k];
    t->clamp_hi[k] = hi[k];
  }
  return KMC_OK;
}

int kmc_transfer_rates_device(kmc_transfer* t, int64_t n, const double* d_theta, const double* d_z,
                              const double* d_a1, const double* d_a2, const int64_t* d_sel,
                              double* d_rate) {
  if (!t || n < 0 || (n > 0 && (!d_theta || !d_z || !d_a1 || !d_a2 || !d_rate)))
    return tfail(KMC_ERR_ARG, "kmc_transfer_rates_device: bad argument");
  if (n == 0) return KMC_OK;
  {
    const int s = upload(t);
    if (s != KMC_OK) return s;
  }
  exact_args a = args_of(t);
  a.n_geom = n;
  a.theta = d_theta;
  a.z = d_z;
  a.a1 = d_a1;
  a.a2 = d_a2;
  a.sel = d_sel;
  a.out = d_rate;
  hipError_t e = hipSetDevice(t->device);
  if (e == hipSuccess) e = launch_exact(t, a, t->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  if (e != hipSuccess) return tfail(KMC_ERR_HIP, std::string("kmc_transfer_rates: ") + hipGetErrorString(e));
  return KMC_OK;
}

int kmc_transfer_rates(kmc_transfer* t, int64_t n, const double* theta, const double* z,
                       const double* a1, const double* a2, double* rate, double* kernel_ms) {
  if (!t || n < 0 || (n > 0 && (!theta || !z || !a1 || !a2 || !rate)))
    return tfail(KMC_ERR_ARG, "kmc_transfer_rates: bad argument");
  if (n == 0) return KMC_OK;
  {
    const int s = upload(t);
    if (s != KMC_OK) return s;
  }
  hipError_t e = hipSetDevice(t->device);
  if (e == hipSuccess && (size_t)n > t->geo_cap) {
    if (t->d_geo) (void)hipFree(t->d_geo);
    if (t->d_out) (void)hipFree(t->d_out);
    t->d_geo = nullptr;
    t->d_out = nullptr;
    e = hipMalloc((void**)&t->d_geo, 4 * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&t->d_out, n * sizeof(double));
    t->geo_cap = e == hipSuccess ? (size_t)n : 0;
  }
  const double* src[4] = {theta, z, a1, a2};
  for (int q = 0; q < 4 && e == hipSuccess; q++)
    e = hipMemcpyAsync(t->d_geo + q * n, src[q], n * sizeof(double), hipMemcpyHostToDevice, t->stream);
  exact_args a = args_of(t);
  a.n_geom = n;
  a.theta = t->d_geo;
  a.z = t->d_geo + n;
  a.a1 = t->d_geo + 2 * n;
  a.a2 = t->d_geo + 3 * n;
  a.sel = nullptr;
  a.out = t->d_out;
  if (e == hipSuccess) e = hipEventRecord(t->ev0, t->stream);
  if (e == hipSuccess) e = launch_exact(t, a, t->stream);
  if (e == hipSuccess) e = hipEventRecord(t->ev1, t->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(rate, t->d_out, n * sizeof(double), hipMemcpyDeviceToHost, t->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  if (e == hipSuccess && kernel_ms) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, t->ev0, t->ev1);
    *kernel_ms = ms;
  }
  if (e != hipSuccess) return tfail(KMC_ERR_HIP, std::string("kmc_transfer_rates: ") + hipGetErrorString(e));
  return KMC_OK;
}

int kmc_transfer_table(kmc_transfer* t, int32_t nt, const double* th, int32_t nz, const double* z,
                       int32_t n1, const double* a1, int32_t n2, const double* a2, double* rates,
                       double* elapsed_ms) {
  if (!t || nt < 1 || nz < 1 || n1 < 1 || n2 < 1 || !th || !z || !a1 || !a2 || !rates)
    return tfail(KMC_ERR_ARG, "kmc_transfer_table: bad argument");
  const auto t0 = std::chrono::steady_clock::now();
  // nodes in the order of create_davoody_scatt_table's loops (monte_carlo.cpp:110-122)
  const int64_t n = (int64_t)nt * nz * n1 * n2;
  std::vector<double> g(4 * n);
  int64_t e = 0;
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nz; ++j)
      for (int k = 0; k < n1; ++k)
        for (int q = 0; q < n2; ++q, ++e) {
          g[e] = th[i];
          g[n + e] = z[j];
          g[2 * n + e] = a1[k];
          g[3 * n + e] = a2[q];
        }
  const int s = kmc_transfer_rates(t, n, g.data(), g.data() + n, g.data() + 2 * n, g.data() + 3 * n,
                                   rates, nullptr);
  if (s != KMC_OK) return s;
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KMC_OK;
}

}  // extern "C"

