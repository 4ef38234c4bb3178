// kmc_mlp.hip -- learned surrogate of the first-order transfer rate.  No reference counterpart:
// it stands beside exciton_transfer::first_order (src/exciton_transfer/exciton_transfer.cpp:
// 380-441), which stays the fallback for flagged pairs.
//
// Network: 16 feature slots -> 4 hidden layers of 128 (SiLU) -> 1, predicting a normalised
// log10 rate.  All products run on the CDNA4 matrix cores as v_mfma_f32_32x32x16_bf16 with BF16
// operands and FP32 accumulation; master weights, biases and the output layer are FP32.
//
// Training (layer by layer, activations in HBM, L2-resident at these batch sizes):
//   forward   Z_l = H_{l-1} W_l + b_l, H_l = bf16(silu(Z_l))         k_gemm<.., EPI_FWD>
//   head      y = H_3 . w4 + b4, dy = 2 (y - t) / n,
//             dZ_3 = bf16(dy w4 silu'(Z_3))                           k_head
//   backward  dW_l = H_{l-1}^T dZ_l (split over the batch, summed in a fixed order),
//             db_l = colsum(dZ_l), dZ_{l-1} = bf16((dZ_l W_l^T) silu'(Z_{l-1}))
//   Adam on the FP32 masters, which refreshes both BF16 copies W_l and W_l^T.
// Inference: one fused kernel, one 32-pair tile per wave.  The BF16 weights (100 KB) sit in LDS,
// pre-arranged as MFMA A fragments; activations never leave registers because the hidden units
// of layers 1..3 are taken in the order in which the accumulator of the previous layer holds
// them (see k_pack), so a C tile becomes the next layer's B operand without any data movement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cntkmc.h"
#include "kmc_mlp.h"

namespace {

using bf16 = __bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int IN = KMC_MLP_IN, WD = KMC_MLP_WIDTH, NH = KMC_MLP_HIDDEN, NCLS = KMC_MLP_CLASSES;
constexpr int64_t NP = KMC_MLP_PARAMS;
constexpr int NBIN_PAIR = NCLS * NCLS;

__host__ __device__ constexpr int fan_in(int l) { return l == 0 ? IN : WD; }
// offsets in the FP32 parameter vector
__host__ __device__ constexpr int64_t w_off(int l) {
  return l == 0 ? 0 : w_off(l - 1) + (int64_t)fan_in(l - 1) * WD + WD;
}
__host__ __device__ constexpr int64_t b_off(int l) { return w_off(l) + (int64_t)fan_in(l) * WD; }
constexpr int64_t W4_OFF = w_off(NH), B4_OFF = W4_OFF + WD;
static_assert(B4_OFF + 1 == NP, "parameter layout");
// offsets of layer l in the BF16 weight copies (W_l, W_l^T and the packed inference layout)
__host__ __device__ constexpr int64_t q_off(int l) { return l == 0 ? 0 : q_off(l - 1) + (int64_t)fan_in(l - 1) * WD; }
constexpr int64_t NQ = q_off(NH);  // 51200 BF16 weights

// logistic function from the hardware exp2 and reciprocal (about 1 ulp each; the results feed
// BF16 operands), shared by training and inference so both evaluate the same network
__device__ __forceinline__ float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ f32x16 mfma32(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------------------------
// features and residual-map bins (shared by host and device)
// ---------------------------------------------------------------------------------------------
struct domain_consts {
  double lo[3], hi[3], t_lo, t_hi, y_shift, y_scale;
};

__host__ __device__ inline double box_coord(double v, const domain_consts& d, int k) {
  const double c = fmin(fmax(v, d.lo[k]), d.hi[k]);
  return (2.0 * c - (d.lo[k] + d.hi[k])) / (d.hi[k] - d.lo[k]);
}
// Features 2..4: the box coordinates of z and of (a1, a2) taken with the sign that makes a2 > 0
// (a1 > 0 when a2 == 0); the a1/a2 box is symmetric, so clamping commutes with the flip.
__host__ __device__ inline void shift_features(double z, double a1, double a2, const domain_consts& d,
                                               float* f) {
  const double c1 = fmin(fmax(a1, d.lo[1]), d.hi[1]), c2 = fmin(fmax(a2, d.lo[2]), d.hi[2]);
  const bool flip = c2 < 0.0 || (c2 == 0.0 && c1 < 0.0);
  f[0] = (float)box_coord(z, d, 0);
  f[1] = (float)box_coord(flip ? -c1 : c1, d, 1);
  f[2] = (float)box_coord(flip ? -c2 : c2, d, 2);
}
__host__ __device__ inline double temp_feature(double T, const domain_consts& d) {
  return log(T / sqrt(d.t_lo * d.t_hi)) / log(sqrt(d.t_hi / d.t_lo));
}
__host__ __device__ inline double inv_temp_feature(double T, const domain_consts& d) {
  return (2.0 / T - (1.0 / d.t_lo + 1.0 / d.t_hi)) / (1.0 / d.t_lo - 1.0 / d.t_hi);
}
__host__ __device__ inline int bin_of(double u, int n) {  // u in [0, 1]
  const int b = (int)(u * n);
  return b < 0 ? 0 : b >= n ? n - 1 : b;
}
// Flattened (pair class, T, z, theta, max |a|) bin of one geometry.
__host__ __device__ inline int64_t bin_index(const int32_t* dims, int cd, int ca, double T, double th,
                                             double z, double a1, double a2, const domain_consts& d) {
  const double zc = fmin(fmax(z, d.lo[0]), d.hi[0]);
  const double c1 = fabs(fmin(fmax(a1, d.lo[1]), d.hi[1])), c2 = fabs(fmin(fmax(a2, d.lo[2]), d.hi[2]));
  const double amax = fmax(fmax(fabs(d.lo[1]), fabs(d.hi[1])), fmax(fabs(d.lo[2]), fabs(d.hi[2])));
  const int iT = bin_of((T - d.t_lo) / (d.t_hi - d.t_lo), dims[1]);
  const int iz = bin_of((zc - d.lo[0]) / (d.hi[0] - d.lo[0]), dims[2]);
  const int ith = bin_of(th / M_PI, dims[3]);
  const double ua = fmax(c1, c2) / amax;
  const int ia = bin_of(ua * ua, dims[4]);  // equal-area rings of the (a1, a2) square
  return ((((int64_t)(cd * NCLS + ca) * dims[1] + iT) * dims[2] + iz) * dims[3] + ith) * dims[4] + ia;
}

// ---------------------------------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------------------------------
enum { EPI_FWD = 0, EPI_DH = 1, EPI_PART = 2 };

struct gemm_args {
  int M, N, K, k_split;
  const bf16* A;  // element (m, k): AT ? A[k * lda + m] : A[m * lda + k]
  int lda;
  const bf16* B;  // element (n, k) of B^T: BT ? B[k * ldb + n] : B[n * ldb + k]
  int ldb;
  const float* bias;   // EPI_FWD: Z = C + bias[n], H = bf16(silu(Z))
  float* Z;
  bf16* H;
  const float* Zprev;  // EPI_DH: dZ = bf16(C * silu'(Zprev))
  bf16* dZ;
  float* part;         // EPI_PART: part[blockIdx.z][M][N] = C over this block's K range
};

// C[M][N] = A B^T with K a multiple of 16.  Four waves per workgroup, each a 32 x 32 tile of
// one MFMA accumulator; rows/columns beyond M/N read zeros and are not stored.
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(gemm_args g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int h = lane >> 5, c = lane & 31;
  const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
  if (m0 >= g.M || n0 >= g.N) return;
  const int k_lo = blockIdx.z * g.k_split, k_hi = min(g.K, k_lo + g.k_split);
  const int am = m0 + c, bn = n0 + c;
  const bool a_ok = am < g.M, b_ok = bn < g.N;
  f32x16 acc = {};
  for (int k0 = k_lo; k0 < k_hi; k0 += 16) {
    const int kk = k0 + 8 * h;
    bf16x8 a = {}, b = {};
    if (a_ok) {
      if (AT) {
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = g.A[(int64_t)(kk + e) * g.lda + am];
      } else {
        a = *reinterpret_cast<const bf16x8*>(g.A + (int64_t)am * g.lda + kk);
      }
    }
    if (b_ok) {
      if (BT) {
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = g.B[(int64_t)(kk + e) * g.ldb + bn];
      } else {
        b = *reinterpret_cast<const bf16x8*>(g.B + (int64_t)bn * g.ldb + kk);
      }
    }
    acc = mfma32(a, b, acc);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + 8 * (r >> 2) + 4 * h + (r & 3), col = n0 + c;
    if (row >= g.M || col >= g.N) continue;
    const int64_t o = (int64_t)row * g.N + col;
    if (EPI == EPI_FWD) {
      const float v = acc[r] + g.bias[col];
      g.Z[o] = v;
      g.H[o] = (bf16)silu_f(v);
    } else if (EPI == EPI_DH) {
      g.dZ[o] = (bf16)(acc[r] * dsilu_f(g.Zprev[o]));
    } else {
      g.part[(int64_t)blockIdx.z * g.M * g.N + o] = acc[r];
    }
  }
}

// y = h3 . w4 + b4 per row (one wave per row, the 128 products reduced by a butterfly).  With
// targets: dy = 2 (y - t) / n_valid for rows < n_valid (0 beyond), dZ3 = bf16(dy w4 silu'(Z3)),
// and loss_part[wave] = sum of (y - t)^2 over the wave's rows.
__global__ __launch_bounds__(256) void k_head(int n, int n_valid, const bf16* __restrict__ H3,
                                              const float* __restrict__ Z3, const float* __restrict__ prm,
                                              const float* __restrict__ tgt, float inv_n, float* y,
                                              float* dy, bf16* dZ3, double* loss_part) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  const float wa = prm[W4_OFF + 2 * lane], wb = prm[W4_OFF + 2 * lane + 1], b4 = prm[B4_OFF];
  double lsum = 0.0;
  for (int s = wave; s < n; s += nw) {
    const bf16* hr = H3 + (int64_t)s * WD;
    float p = (float)hr[2 * lane] * wa + (float)hr[2 * lane + 1] * wb;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) p += __shfl_xor(p, off, 64);
    const float yy = p + b4;
    if (lane == 0) y[s] = yy;
    if (tgt) {
      const float e = s < n_valid ? yy - tgt[s] : 0.0f;
      const float d = 2.0f * e * inv_n;
      if (lane == 0) {
        dy[s] = d;
        lsum += (double)e * (double)e;
      }
      const float* zr = Z3 + (int64_t)s * WD;
      dZ3[(int64_t)s * WD + 2 * lane] = (bf16)(d * wa * dsilu_f(zr[2 * lane]));
      dZ3[(int64_t)s * WD + 2 * lane + 1] = (bf16)(d * wb * dsilu_f(zr[2 * lane + 1]));
    }
  }
  if (loss_part && lane == 0) loss_part[wave] = lsum;
}

// part[block][j] = sum over the block's rows r of (X ? X[r][j] : 1) * (w ? w[r] : 1), cols <= 128.
__global__ __launch_bounds__(256) void k_colsum(int n, int cols, const bf16* __restrict__ X,
                                                const float* __restrict__ w, int chunk, float* part) {
  __shared__ float s[256];
  const int t = threadIdx.x, j = t & 127, half = t >> 7;
  const int r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
  float acc = 0.0f;
  if (j < cols)
    for (int r = r0 + half; r < r1; r += 2) {
      const float x = X ? (float)X[(int64_t)r * cols + j] : 1.0f;
      acc += w ? x * w[r] : x;
    }
  s[t] = acc;
  __syncthreads();
  if (t < cols) part[(int64_t)blockIdx.x * cols + t] = s[t] + s[t + 128];
}

// out[i] = sum_{z < nz} part[z][i], in order.
__global__ void k_sum_parts(int64_t len, int nz, const float* __restrict__ part, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float a = 0.0f;
  for (int z = 0; z < nz; ++z) a += part[(int64_t)z * len + i];
  out[i] = a;
}

__device__ __forceinline__ void write_copies(int64_t p, float w, bf16* Wb, bf16* Wtb) {
#pragma unroll
  for (int l = 0; l < NH; ++l)
    if (p >= w_off(l) && p < b_off(l)) {
      const int64_t q = p - w_off(l);
      const int i = (int)(q / WD), o = (int)(q % WD);
      Wb[q_off(l) + q] = (bf16)w;
      Wtb[q_off(l) + (int64_t)o * fan_in(l) + i] = (bf16)w;
    }
}

__global__ void k_refresh(const float* __restrict__ prm, bf16* Wb, bf16* Wtb) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NP) write_copies(p, prm[p], Wb, Wtb);
}

// Adam with bias corrections c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t).
__global__ void k_adam(float* prm, const float* __restrict__ grad, float* m, float* v, float lr,
                       float b1, float b2, float eps, float c1, float c2, bf16* Wb, bf16* Wtb) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.0f - b1) * g;
  const float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float w = prm[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
  prm[p] = w;
  write_copies(p, w, Wb, Wtb);
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ x, bf16* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

// rows idx[r] of the data set into the batch
__global__ void k_gather(int nb, const int32_t* __restrict__ idx, const bf16* __restrict__ dx,
                         const float* __restrict__ dt, bf16* X, float* tgt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  const int64_t s = idx[r];
  const bf16x8* src = reinterpret_cast<const bf16x8*>(dx + s * IN);
  bf16x8* dst = reinterpret_cast<bf16x8*>(X + (int64_t)r * IN);
  dst[0] = src[0];
  dst[1] = src[1];
  tgt[r] = dt[s];
}

// ---------------------------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------------------------
// Packed A fragments: element (l, ob, kb, lane, e) = W_l^T[ob*32 + lane%32][kb*16 + u(e)] with
// u = 8 h + e for layer 0 (features in natural order, h = lane / 32) and
// u = 8 (e / 4) + 4 h + e % 4 for layers 1..3.  The accumulator of a 32 x 32 MFMA holds, in lane
// (h, c) and register r, row 8 (r/4) + 4 h + r%4; registers 0..7 and 8..15 of output block ob
// are handed on unchanged as the B operands of k-blocks 2 ob and 2 ob + 1, whose slot 8 h + e
// therefore carries hidden unit 16 kb + 8 (e/4) + 4 h + e%4 -- the u above.
__global__ void k_pack(const bf16* __restrict__ Wtb, bf16* pk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NQ) return;
  int l = 0;
  while (l + 1 < NH && t >= q_off(l + 1)) ++l;
  const int64_t q = t - q_off(l);
  const int e = (int)(q & 7), lane = (int)((q >> 3) & 63);
  const int nkb = fan_in(l) / 16;
  const int rest = (int)(q >> 9), kb = rest % nkb, ob = rest / nkb;
  const int h = lane >> 5, r = lane & 31;
  const int u = l == 0 ? 8 * h + e : 8 * (e >> 2) + 4 * h + (e & 3);
  pk[t] = Wtb[q_off(l) + (int64_t)(ob * 32 + r) * fan_in(l) + kb * 16 + u];
}

struct infer_args {
  int64_t n;
  const double *theta, *z, *a1, *a2;
  const uint8_t* pair_cls;  // optional: class pair of entry p, ci * n_chir + cj
  int32_t n_chir;
  int32_t cls_d, cls_a;     // used without pair_cls
  int8_t cmap[16];          // context chirality id -> surrogate class
  double T;
  float tfeat, tinv;
  bool t_ok;
  const bf16* pk;
  const float* prm;
  domain_consts dom;
  const uint8_t* rmap;
  int32_t rdims[5];
  double* rate;
  uint8_t* flag;
};

constexpr int PK_VEC = (int)(NQ / 8);      // 6400 A fragments of 16 bytes
constexpr int SB_N = NH * WD + WD + 1;     // hidden biases, w4, b4

// Output block ob of a layer with its biases as the accumulator's start: element r of lane
// (h, c) of a 32 x 32 accumulator is output unit ob 32 + 8 (r / 4) + 4 h + r % 4 of entry c.
// The biases are read from LDS where each accumulator starts: an opaque copy of h keeps the
// compiler from hoisting the loads (256 values per lane) out of the tile loop.
__device__ __forceinline__ f32x16 bias_acc(const float* bias, int ob, int h) {
  asm volatile("" : "+v"(h));
  f32x16 a;
#pragma unroll
  for (int r = 0; r < 16; ++r) a[r] = bias[ob * 32 + 8 * (r >> 2) + 4 * h + (r & 3)];
  return a;
}

// SiLU of one 32 x 32 accumulator (biases included) into the B operands of k-blocks 2 ob, 2 ob + 1
__device__ __forceinline__ void to_operands(const f32x16& acc, bf16x8& lo, bf16x8& hi) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const bf16 s = (bf16)silu_f(acc[r]);
    if (r < 8) lo[r] = s;
    else hi[r - 8] = s;
  }
}

template <int L>
__device__ __forceinline__ void hidden_layer(const bf16x8* s_w, const float* s_b, int lane, int h,
                                             bf16x8 (&x)[8]) {
  // the fragment offsets are the same for every tile: an opaque copy per layer keeps the
  // compiler from hoisting all 128 fragment loads (512 registers) out of the tile loop
  asm volatile("" : "+v"(lane));
  bf16x8 y[8];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x16 acc = bias_acc(s_b + L * WD, ob, h);
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) acc = mfma32(s_w[q_off(L) / 8 + (ob * 8 + kb) * 64 + lane], x[kb], acc);
    to_operands(acc, y[2 * ob], y[2 * ob + 1]);
  }
#pragma unroll
  for (int kb = 0; kb < 8; ++kb) x[kb] = y[kb];
}

// Waves per workgroup sharing one LDS copy of the weights.  The 103 KB copy allows one workgroup
// per CU, and the kernel's 168 VGPRs allow three waves per SIMD: twelve waves fill both.
constexpr int INFER_WAVES = 12;

// One wave per tile of 32 entries (the 32 columns of the MFMA tiles); the workgroup's waves
// share the LDS copy of the weights and stride over the tiles.
__global__ __launch_bounds__(64 * INFER_WAVES) void k_mlp_infer(infer_args a) {
  __shared__ bf16x8 s_w[PK_VEC];
  __shared__ float s_b[SB_N];
  {
    const bf16x8* src = reinterpret_cast<const bf16x8*>(a.pk);
    for (int i = threadIdx.x; i < PK_VEC; i += blockDim.x) s_w[i] = src[i];
    for (int i = threadIdx.x; i < NH * WD; i += blockDim.x) s_b[i] = a.prm[b_off(i / WD) + i % WD];
    for (int i = threadIdx.x; i < WD; i += blockDim.x) s_b[NH * WD + i] = a.prm[W4_OFF + i];
    if (threadIdx.x == 0) s_b[NH * WD + WD] = a.prm[B4_OFF];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  const int64_t n_tiles = (a.n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * INFER_WAVES + (threadIdx.x >> 6); tile < n_tiles;
       tile += (int64_t)gridDim.x * INFER_WAVES) {
    const int64_t e = tile * 32 + c;
    const bool live = e < a.n;
    int cd = -1, ca = -1;
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (live) {
      if (a.pair_cls) {
        const int pc = a.pair_cls[e];
        cd = a.cmap[pc / a.n_chir];
        ca = a.cmap[pc % a.n_chir];
      } else {
        cd = a.cls_d;
        ca = a.cls_a;
      }
      if (h == 0) {
        double st, ct;
        sincos(a.theta[e], &st, &ct);
        f[0] = (float)ct;
        f[1] = (float)st;
        shift_features(a.z[e], a.a1[e], a.a2[e], a.dom, f + 2);
#pragma unroll
        for (int j = 0; j < NCLS; ++j) f[5 + j] = cd == j ? 1.0f : 0.0f;
      } else {
#pragma unroll
        for (int j = 0; j < NCLS; ++j) f[j] = ca == j ? 1.0f : 0.0f;
        f[3] = a.tfeat;
        f[4] = a.tinv;
      }
    }
    bf16x8 x[8];
    // layer 0: one k-block of the 16 feature slots
    {
      int wl0 = lane;
      asm volatile("" : "+v"(wl0));
      bf16x8 x0;
#pragma unroll
      for (int k = 0; k < 8; ++k) x0[k] = (bf16)f[k];
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) {
        const f32x16 acc = mfma32(s_w[ob * 64 + wl0], x0, bias_acc(s_b, ob, h));
        to_operands(acc, x[2 * ob], x[2 * ob + 1]);
      }
    }
    hidden_layer<1>(s_w, s_b, lane, h, x);
    hidden_layer<2>(s_w, s_b, lane, h, x);
    // layer 3 and the output unit
    float ys = 0.0f;
    int wl = lane;
    asm volatile("" : "+v"(wl));
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      f32x16 acc = bias_acc(s_b + 3 * WD, ob, h);
#pragma unroll
      for (int kb = 0; kb < 8; ++kb) acc = mfma32(s_w[q_off(3) / 8 + (ob * 8 + kb) * 64 + wl], x[kb], acc);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int u = ob * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
        ys += (float)(bf16)silu_f(acc[r]) * s_b[NH * WD + u];
      }
    }
    const float yy = ys + __shfl_xor(ys, 32, 64) + s_b[NH * WD + WD];
    if (h != 0 || !live) continue;
    const double l10 = a.dom.y_shift + a.dom.y_scale * (double)yy;
    a.rate[e] = exp(l10 * 2.302585092994045684);
    if (a.flag) {
      uint8_t fl = 0;
      if (cd < 0 || ca < 0 || !a.t_ok) {
        fl = KMC_MLP_FLAG_DOMAIN;
      } else if (a.rmap) {
        const int64_t b = bin_index(a.rdims, cd, ca, a.T, a.theta[e], a.z[e], a.a1[e], a.a2[e], a.dom);
        if (a.rmap[b]) fl = KMC_MLP_FLAG_RESIDUAL;
      }
      a.flag[e] = fl;
    }
  }
}

thread_local std::string g_mlp_err;
int mfail(int code, const std::string& m) {
  g_mlp_err = m;
  return code;
}

}  // namespace

struct kmc_mlp {
  int device = 0;
  int n_cu = 1;  // compute units of the device (grid of the inference kernel)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  domain_consts dom{};
  float *d_prm = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
  bf16 *d_wb = nullptr, *d_wtb = nullptr, *d_pk = nullptr;
  bool pk_dirty = true;
  int64_t adam_t = 0;
  uint8_t* d_rmap = nullptr;
  int32_t rdims[5] = {0, 0, 0, 0, 0};
  // training workspace for batches of up to `cap` rows
  int64_t cap = 0;
  bf16* X = nullptr;
  float* tgt = nullptr;
  float* Z[NH] = {};
  bf16* H[NH] = {};
  float *y = nullptr, *dy = nullptr;
  bf16 *dZa = nullptr, *dZb = nullptr;
  float *part = nullptr, *cpart = nullptr;
  double* lpart = nullptr;
  float* fstage = nullptr;  // FP32 staging of host feature rows
  // host-array inference staging
  int64_t geo_cap = 0;
  double* d_geo = nullptr;
  double* d_rate = nullptr;
  uint8_t* d_flag = nullptr;
};

namespace {

constexpr int K_SPLIT = 512;    // batch rows per split of the weight-gradient products
constexpr int HEAD_WAVES = 256; // waves of k_head (its loss_part length)
constexpr int COL_BLOCKS = 64;  // row blocks of k_colsum

#define MLP_HIP(call)                                                                      \
  do {                                                                                     \
    const hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                                  \
      return mfail(e_ == hipErrorNoDevice ? KMC_ERR_NODEVICE : KMC_ERR_HIP,                \
                   std::string(__func__) + ": " + hipGetErrorString(e_));                  \
  } while (0)

template <typename T>
hipError_t grow(T** p, size_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

int64_t round_up(int64_t n, int64_t m) { return (n + m - 1) / m * m; }

int ensure_cap(kmc_mlp* m, int64_t rows) {
  rows = round_up(std::max<int64_t>(rows, 128), 128);
  if (rows <= m->cap) return KMC_OK;
  MLP_HIP(grow(&m->X, rows * IN));
  MLP_HIP(grow(&m->tgt, rows));
  for (int l = 0; l < NH; ++l) {
    MLP_HIP(grow(&m->Z[l], rows * WD));
    MLP_HIP(grow(&m->H[l], rows * WD));
  }
  MLP_HIP(grow(&m->y, rows));
  MLP_HIP(grow(&m->dy, rows));
  MLP_HIP(grow(&m->dZa, rows * WD));
  MLP_HIP(grow(&m->dZb, rows * WD));
  MLP_HIP(grow(&m->part, round_up(rows, K_SPLIT) / K_SPLIT * WD * WD));
  MLP_HIP(grow(&m->fstage, rows * IN));
  m->cap = rows;
  return KMC_OK;
}

hipError_t launch_gemm(const gemm_args& g, bool at, bool bt, int epi, hipStream_t st) {
  const int nz = (g.K + g.k_split - 1) / g.k_split;
  const dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, nz);
  if (!at && !bt && epi == EPI_FWD) hipLaunchKernelGGL((k_gemm<false, false, EPI_FWD>), grid, dim3(256), 0, st, g);
  else if (!at && !bt && epi == EPI_DH) hipLaunchKernelGGL((k_gemm<false, false, EPI_DH>), grid, dim3(256), 0, st, g);
  else if (at && bt && epi == EPI_PART) hipLaunchKernelGGL((k_gemm<true, true, EPI_PART>), grid, dim3(256), 0, st, g);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t colsum(kmc_mlp* m, int n, int cols, const bf16* X, const float* w, float* out) {
  const int chunk = (n + COL_BLOCKS - 1) / COL_BLOCKS;
  hipLaunchKernelGGL(k_colsum, dim3(COL_BLOCKS), dim3(256), 0, m->stream, n, cols, X, w, chunk, m->cpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(128), 0, m->stream, (int64_t)cols, COL_BLOCKS, m->cpart, out);
  return hipGetLastError();
}

// Forward over rows [0, nb) of m->X; with targets also the head gradient, backward and the
// gradient in m->d_grad.  nb is a multiple of 128.
hipError_t step(kmc_mlp* m, int nb, int n_valid, bool backward) {
  hipStream_t st = m->stream;
  for (int l = 0; l < NH; ++l) {
    gemm_args g{};
    g.M = nb;
    g.N = WD;
    g.K = fan_in(l);
    g.k_split = g.K;
    g.A = l == 0 ? m->X : m->H[l - 1];
    g.lda = fan_in(l);
    g.B = m->d_wtb + q_off(l);
    g.ldb = fan_in(l);
    g.bias = m->d_prm + b_off(l);
    g.Z = m->Z[l];
    g.H = m->H[l];
    hipError_t e = launch_gemm(g, false, false, EPI_FWD, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_head, dim3(HEAD_WAVES / 4), dim3(256), 0, st, nb, n_valid, m->H[NH - 1],
                     m->Z[NH - 1], m->d_prm, backward ? m->tgt : nullptr,
                     1.0f / (float)std::max(n_valid, 1), m->y, m->dy, m->dZa,
                     backward ? m->lpart : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !backward) return e;
  e = colsum(m, nb, WD, m->H[NH - 1], m->dy, m->d_grad + W4_OFF);
  if (e == hipSuccess) e = colsum(m, nb, 1, nullptr, m->dy, m->d_grad + B4_OFF);
  bf16 *dz = m->dZa, *dz_next = m->dZb;
  for (int l = NH - 1; l >= 0 && e == hipSuccess; --l) {
    gemm_args g{};
    g.M = fan_in(l);
    g.N = WD;
    g.K = nb;
    g.k_split = K_SPLIT;
    g.A = l == 0 ? m->X : m->H[l - 1];
    g.lda = fan_in(l);
    g.B = dz;
    g.ldb = WD;
    g.part = m->part;
    e = launch_gemm(g, true, true, EPI_PART, st);
    if (e != hipSuccess) break;
    const int64_t len = (int64_t)fan_in(l) * WD;
    const int nz = (nb + K_SPLIT - 1) / K_SPLIT;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, nz,
                       m->part, m->d_grad + w_off(l));
    e = colsum(m, nb, WD, dz, nullptr, m->d_grad + b_off(l));
    if (e != hipSuccess || l == 0) break;
    gemm_args d{};
    d.M = nb;
    d.N = fan_in(l);
    d.K = WD;
    d.k_split = WD;
    d.A = dz;
    d.lda = WD;
    d.B = m->d_wb + q_off(l);
    d.ldb = WD;
    d.Zprev = m->Z[l - 1];
    d.dZ = dz_next;
    e = launch_gemm(d, false, false, EPI_DH, st);
    std::swap(dz, dz_next);
  }
  return e;
}

hipError_t adam(kmc_mlp* m, double lr, double b1, double b2, double eps) {
  ++m->adam_t;
  const double c1 = 1.0 / (1.0 - std::pow(b1, (double)m->adam_t));
  const double c2 = 1.0 / (1.0 - std::pow(b2, (double)m->adam_t));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, m->stream, m->d_prm,
                     m->d_grad, m->d_m, m->d_v, (float)lr, (float)b1, (float)b2, (float)eps,
                     (float)c1, (float)c2, m->d_wb, m->d_wtb);
  m->pk_dirty = true;
  return hipGetLastError();
}

// feature rows (host FP32) -> m->X (BF16) with rows [n, nb) zero
int load_rows(kmc_mlp* m, int64_t n, int64_t nb, const float* feat, const float* targets) {
  MLP_HIP(hipMemsetAsync(m->fstage, 0, nb * IN * sizeof(float), m->stream));
  MLP_HIP(hipMemcpyAsync(m->fstage, feat, n * IN * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nb * IN + 255) / 256)), dim3(256), 0, m->stream,
                     nb * IN, m->fstage, m->X);
  MLP_HIP(hipGetLastError());
  if (targets) {
    MLP_HIP(hipMemsetAsync(m->tgt, 0, nb * sizeof(float), m->stream));
    MLP_HIP(hipMemcpyAsync(m->tgt, targets, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
  }
  return KMC_OK;
}

infer_args base_infer(const kmc_mlp* m, double T) {
  infer_args a{};
  a.T = T;
  a.t_ok = T >= m->dom.t_lo && T <= m->dom.t_hi;
  a.tfeat = (float)temp_feature(T, m->dom);
  a.tinv = (float)inv_temp_feature(T, m->dom);
  a.pk = m->d_pk;
  a.prm = m->d_prm;
  a.dom = m->dom;
  a.rmap = m->d_rmap;
  for (int k = 0; k < 5; ++k) a.rdims[k] = m->rdims[k];
  return a;
}

// At most one workgroup per CU, each striding over the tiles, so that every CU copies the
// weights into LDS once per launch.
hipError_t launch_infer(const infer_args& a, int n_cu, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
  const int64_t tiles = (a.n + 31) / 32;
  const unsigned grid = (unsigned)std::min<int64_t>((tiles + INFER_WAVES - 1) / INFER_WAVES, n_cu);
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64 * INFER_WAVES), 0, st, a);
  return hipGetLastError();
}

int stage_geometry(kmc_mlp* m, int64_t n, const double* th, const double* z, const double* a1,
                   const double* a2) {
  if (n > m->geo_cap) {
    MLP_HIP(grow(&m->d_geo, 4 * n));
    MLP_HIP(grow(&m->d_rate, n));
    MLP_HIP(grow(&m->d_flag, n));
    m->geo_cap = n;
  }
  const double* src[4] = {th, z, a1, a2};
  for (int k = 0; k < 4; ++k)
    MLP_HIP(hipMemcpyAsync(m->d_geo + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  return KMC_OK;
}

void grid_geometry(int32_t nt, const double* th, int32_t nz, const double* z, int32_t n1,
                   const double* a1, int32_t n2, const double* a2, std::vector<double>& g) {
  const int64_t n = (int64_t)nt * nz * n1 * n2;
  g.resize(4 * n);
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
}

}  // namespace

namespace cntkmc {

int mlp_device(const kmc_mlp* m) { return m->device; }

int mlp_prepare(kmc_mlp* m) {
  if (!m->pk_dirty) return KMC_OK;
  MLP_HIP(hipSetDevice(m->device));
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((NQ + 255) / 256)), dim3(256), 0, m->stream, m->d_wtb, m->d_pk);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->pk_dirty = false;
  return KMC_OK;
}

hipError_t mlp_launch_pairs(const kmc_mlp* m, int64_t n, const double* theta, const double* z,
                            const double* a1, const double* a2, const uint8_t* pair_cls,
                            int32_t n_chir, const int32_t* chir_class, double temperature,
                            double* rate, uint8_t* flag, hipStream_t st) {
  if (n_chir < 1 || n_chir > 16) return hipErrorInvalidValue;
  infer_args a = base_infer(m, temperature);
  a.n = n;
  a.theta = theta;
  a.z = z;
  a.a1 = a1;
  a.a2 = a2;
  a.pair_cls = pair_cls;
  a.n_chir = n_chir;
  for (int k = 0; k < n_chir; ++k)
    a.cmap[k] = (int8_t)(chir_class[k] >= 0 && chir_class[k] < NCLS ? chir_class[k] : -1);
  a.rate = rate;
  a.flag = flag;
  return launch_infer(a, m->n_cu, st);
}

}  // namespace cntkmc

extern "C" {

const char* kmc_mlp_last_error(void) { return g_mlp_err.c_str(); }

int kmc_mlp_create(int32_t device, const kmc_mlp_domain* dom, kmc_mlp** out) {
  if (!dom || !out) return mfail(KMC_ERR_ARG, "kmc_mlp_create: null argument");
  *out = nullptr;
  for (int k = 0; k < 3; ++k)
    if (!(dom->box_lo[k] < dom->box_hi[k])) return mfail(KMC_ERR_ARG, "kmc_mlp_create: empty box");
  if (dom->box_lo[1] != -dom->box_hi[1] || dom->box_lo[2] != -dom->box_hi[2])
    return mfail(KMC_ERR_ARG, "kmc_mlp_create: the axis-shift box must be symmetric about 0");
  if (!(dom->t_lo > 0 && dom->t_lo < dom->t_hi) || !(dom->y_scale != 0))
    return mfail(KMC_ERR_ARG, "kmc_mlp_create: bad temperature range or output scale");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return mfail(KMC_ERR_NODEVICE, "kmc_mlp_create: no HIP device");
  if (device < 0 || device >= ndev) return mfail(KMC_ERR_ARG, "kmc_mlp_create: bad device index");
  kmc_mlp* m = new kmc_mlp();
  m->device = device;
  for (int k = 0; k < 3; ++k) {
    m->dom.lo[k] = dom->box_lo[k];
    m->dom.hi[k] = dom->box_hi[k];
  }
  m->dom.t_lo = dom->t_lo;
  m->dom.t_hi = dom->t_hi;
  m->dom.y_shift = dom->y_shift;
  m->dom.y_scale = dom->y_scale;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&m->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&m->ev0);
  if (e == hipSuccess) e = hipEventCreate(&m->ev1);
  if (e == hipSuccess) e = grow(&m->d_prm, NP);
  if (e == hipSuccess) e = grow(&m->d_grad, NP);
  if (e == hipSuccess) e = grow(&m->d_m, NP);
  if (e == hipSuccess) e = grow(&m->d_v, NP);
  if (e == hipSuccess) e = grow(&m->d_wb, NQ);
  if (e == hipSuccess) e = grow(&m->d_wtb, NQ);
  if (e == hipSuccess) e = grow(&m->d_pk, NQ);
  if (e == hipSuccess) e = grow(&m->cpart, COL_BLOCKS * WD);
  if (e == hipSuccess) e = grow(&m->lpart, HEAD_WAVES);
  for (float* b : {m->d_prm, m->d_grad, m->d_m, m->d_v})
    if (e == hipSuccess) e = hipMemsetAsync(b, 0, NP * sizeof(float), m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_refresh, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, m->stream,
                       m->d_prm, m->d_wb, m->d_wtb);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    kmc_mlp_destroy(m);
    return mfail(KMC_ERR_HIP, std::string("kmc_mlp_create: ") + hipGetErrorString(e));
  }
  *out = m;
  return KMC_OK;
}

int kmc_mlp_destroy(kmc_mlp* m) {
  if (!m) return KMC_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* bufs[] = {m->d_prm, m->d_grad, m->d_m,  m->d_v,  m->d_wb,    m->d_wtb, m->d_pk,
                  m->d_rmap, m->X,     m->tgt,  m->y,    m->dy,      m->dZa,   m->dZb,
                  m->part,  m->cpart,  m->lpart, m->fstage, m->d_geo, m->d_rate, m->d_flag};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int l = 0; l < NH; ++l) {
    if (m->Z[l]) (void)hipFree(m->Z[l]);
    if (m->H[l]) (void)hipFree(m->H[l]);
  }
  if (m->ev0) (void)hipEventDestroy(m->ev0);
  if (m->ev1) (void)hipEventDestroy(m->ev1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return KMC_OK;
}

int kmc_mlp_set_params(kmc_mlp* m, const float* params) {
  if (!m || !params) return mfail(KMC_ERR_ARG, "kmc_mlp_set_params: null argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(m->d_prm, params, NP * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemsetAsync(m->d_m, 0, NP * sizeof(float), m->stream));
  MLP_HIP(hipMemsetAsync(m->d_v, 0, NP * sizeof(float), m->stream));
  hipLaunchKernelGGL(k_refresh, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, m->stream,
                     m->d_prm, m->d_wb, m->d_wtb);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  m->pk_dirty = true;
  return KMC_OK;
}

int kmc_mlp_get_params(kmc_mlp* m, float* params) {
  if (!m || !params) return mfail(KMC_ERR_ARG, "kmc_mlp_get_params: null argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(params, m->d_prm, NP * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return KMC_OK;
}

int kmc_mlp_set_residual_map(kmc_mlp* m, const int32_t dims[5], const uint8_t* flags) {
  if (!m) return mfail(KMC_ERR_ARG, "kmc_mlp_set_residual_map: null mlp");
  MLP_HIP(hipSetDevice(m->device));
  if (m->d_rmap) (void)hipFree(m->d_rmap);
  m->d_rmap = nullptr;
  if (!flags) return KMC_OK;
  if (!dims || dims[0] != NBIN_PAIR) return mfail(KMC_ERR_ARG, "kmc_mlp_set_residual_map: dims[0] must be 9");
  int64_t len = 1;
  for (int k = 0; k < 5; ++k) {
    if (dims[k] < 1 || dims[k] > 4096) return mfail(KMC_ERR_ARG, "kmc_mlp_set_residual_map: bad dims");
    len *= dims[k];
  }
  if (len > (int64_t)1 << 26) return mfail(KMC_ERR_ARG, "kmc_mlp_set_residual_map: map too large");
  MLP_HIP(hipMalloc((void**)&m->d_rmap, len));
  MLP_HIP(hipMemcpyAsync(m->d_rmap, flags, len, hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  for (int k = 0; k < 5; ++k) m->rdims[k] = dims[k];
  return KMC_OK;
}

int kmc_mlp_features(const kmc_mlp* m, int64_t n, const double* th, const double* z, const double* a1,
                     const double* a2, const int32_t* cd, const int32_t* ca, const double* T,
                     float* f) {
  if (!m || n < 0 || (n > 0 && (!th || !z || !a1 || !a2 || !cd || !ca || !T || !f)))
    return mfail(KMC_ERR_ARG, "kmc_mlp_features: bad argument");
  for (int64_t i = 0; i < n; ++i) {
    float* r = f + i * IN;
    std::fill(r, r + IN, 0.0f);
    r[0] = (float)std::cos(th[i]);
    r[1] = (float)std::sin(th[i]);
    shift_features(z[i], a1[i], a2[i], m->dom, r + 2);
    for (int j = 0; j < NCLS; ++j) {
      r[5 + j] = cd[i] == j ? 1.0f : 0.0f;
      r[8 + j] = ca[i] == j ? 1.0f : 0.0f;
    }
    r[11] = (float)temp_feature(T[i], m->dom);
    r[12] = (float)inv_temp_feature(T[i], m->dom);
  }
  return KMC_OK;
}

int kmc_mlp_forward(kmc_mlp* m, int64_t n, const float* feat, float* y) {
  if (!m || n < 0 || (n > 0 && (!feat || !y))) return mfail(KMC_ERR_ARG, "kmc_mlp_forward: bad argument");
  MLP_HIP(hipSetDevice(m->device));
  constexpr int64_t CH = 65536;
  for (int64_t s = 0; s < n; s += CH) {
    const int64_t k = std::min(CH, n - s), nb = round_up(k, 128);
    int st = ensure_cap(m, nb);
    if (st == KMC_OK) st = load_rows(m, k, nb, feat + s * IN, nullptr);
    if (st != KMC_OK) return st;
    MLP_HIP(step(m, (int)nb, (int)k, false));
    MLP_HIP(hipMemcpyAsync(y + s, m->y, k * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
  }
  return KMC_OK;
}

int kmc_mlp_loss_grad(kmc_mlp* m, int64_t n, const float* feat, const float* targets, double* loss,
                      float* grad) {
  if (!m || n < 1 || n > (1 << 24) || !feat || !targets)
    return mfail(KMC_ERR_ARG, "kmc_mlp_loss_grad: bad argument");
  MLP_HIP(hipSetDevice(m->device));
  const int64_t nb = round_up(n, 128);
  int st = ensure_cap(m, nb);
  if (st == KMC_OK) st = load_rows(m, n, nb, feat, targets);
  if (st != KMC_OK) return st;
  MLP_HIP(step(m, (int)nb, (int)n, true));
  std::vector<double> lp(HEAD_WAVES);
  MLP_HIP(hipMemcpyAsync(lp.data(), m->lpart, HEAD_WAVES * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grad) MLP_HIP(hipMemcpyAsync(grad, m->d_grad, NP * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (loss) {
    double s = 0;
    for (double v : lp) s += v;
    *loss = s / (double)n;
  }
  return KMC_OK;
}

int kmc_mlp_adam_step(kmc_mlp* m, const float* grad, double lr, double b1, double b2, double eps,
                      int64_t step_no) {
  if (!m || !grad || step_no < 1) return mfail(KMC_ERR_ARG, "kmc_mlp_adam_step: bad argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(m->d_grad, grad, NP * sizeof(float), hipMemcpyHostToDevice, m->stream));
  m->adam_t = step_no - 1;
  MLP_HIP(adam(m, lr, b1, b2, eps));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return KMC_OK;
}

int kmc_mlp_train(kmc_mlp* m, int64_t n, const float* feat, const float* targets, int32_t epochs,
                  int32_t batch, double lr, double lr_min, uint64_t seed, double* epoch_loss,
                  double* elapsed_ms) {
  if (!m || !feat || !targets || epochs < 1 || batch < 128 || batch % 128 || n < batch ||
      n > ((int64_t)1 << 30) || !(lr > 0) || lr_min < 0)
    return mfail(KMC_ERR_ARG, "kmc_mlp_train: bad argument (batch a multiple of 128, n >= batch)");
  const auto t0 = std::chrono::steady_clock::now();
  MLP_HIP(hipSetDevice(m->device));
  int st = ensure_cap(m, batch);
  if (st != KMC_OK) return st;
  // the data set on the device: BF16 rows, FP32 targets, per-epoch permutation
  float* d_f = nullptr;
  bf16* d_x = nullptr;
  float* d_t = nullptr;
  int32_t* d_idx = nullptr;
  auto release = [&]() {
    for (void* b : {(void*)d_f, (void*)d_x, (void*)d_t, (void*)d_idx})
      if (b) (void)hipFree(b);
  };
  hipError_t e = hipMalloc((void**)&d_f, n * IN * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&d_x, n * IN * sizeof(bf16));
  if (e == hipSuccess) e = hipMalloc((void**)&d_t, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&d_idx, n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemcpyAsync(d_f, feat, n * IN * sizeof(float), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_t, targets, n * sizeof(float), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n * IN + 255) / 256)), dim3(256), 0, m->stream,
                       n * IN, d_f, d_x);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (d_f) (void)hipFree(d_f);
  d_f = nullptr;
  const int64_t per_epoch = n / batch;
  const double total = (double)epochs * per_epoch;
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  uint64_t rs = seed ^ 0x9E3779B97F4A7C15ull;
  auto next = [&]() {  // splitmix64
    uint64_t z = (rs += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  std::vector<double> lp(HEAD_WAVES * per_epoch);
  double* d_lp = nullptr;
  if (e == hipSuccess) e = hipMalloc((void**)&d_lp, lp.size() * sizeof(double));
  int64_t it = 0;
  for (int32_t ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
    e = hipMemcpyAsync(d_idx, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, m->stream);
    for (int64_t b = 0; b < per_epoch && e == hipSuccess; ++b, ++it) {
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, m->stream,
                         batch, d_idx + b * batch, d_x, d_t, m->X, m->tgt);
      double* keep = m->lpart;
      m->lpart = d_lp + b * HEAD_WAVES;
      e = step(m, batch, batch, true);
      m->lpart = keep;
      const double rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + std::cos(M_PI * (double)it / total));
      if (e == hipSuccess) e = adam(m, rate, 0.9, 0.999, 1e-8);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(lp.data(), d_lp, lp.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess && epoch_loss) {
      double s = 0;
      for (double v : lp) s += v;
      epoch_loss[ep] = s / ((double)per_epoch * batch);
    }
  }
  if (d_lp) (void)hipFree(d_lp);
  release();
  if (e != hipSuccess) return mfail(KMC_ERR_HIP, std::string("kmc_mlp_train: ") + hipGetErrorString(e));
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KMC_OK;
}

int kmc_mlp_rates_device(kmc_mlp* m, int64_t n, const double* d_theta, const double* d_z,
                         const double* d_a1, const double* d_a2, int32_t cls_d, int32_t cls_a,
                         double T, double* d_rate, uint8_t* d_flag) {
  if (!m || n < 0 || (n > 0 && (!d_theta || !d_z || !d_a1 || !d_a2 || !d_rate)) || !(T > 0))
    return mfail(KMC_ERR_ARG, "kmc_mlp_rates_device: bad argument");
  const int s = cntkmc::mlp_prepare(m);
  if (s != KMC_OK) return s;
  infer_args a = base_infer(m, T);
  a.n = n;
  a.theta = d_theta;
  a.z = d_z;
  a.a1 = d_a1;
  a.a2 = d_a2;
  a.cls_d = cls_d >= 0 && cls_d < NCLS ? cls_d : -1;
  a.cls_a = cls_a >= 0 && cls_a < NCLS ? cls_a : -1;
  a.rate = d_rate;
  a.flag = d_flag;
  MLP_HIP(launch_infer(a, m->n_cu, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return KMC_OK;
}

int kmc_mlp_rates(kmc_mlp* m, int64_t n, const double* th, const double* z, const double* a1,
                  const double* a2, int32_t cls_d, int32_t cls_a, double T, double* rate,
                  uint8_t* flag, double* kernel_ms) {
  if (!m || n < 0 || (n > 0 && (!th || !z || !a1 || !a2 || !rate)) || !(T > 0))
    return mfail(KMC_ERR_ARG, "kmc_mlp_rates: bad argument");
  if (n == 0) return KMC_OK;
  int s = cntkmc::mlp_prepare(m);
  if (s == KMC_OK) s = stage_geometry(m, n, th, z, a1, a2);
  if (s != KMC_OK) return s;
  infer_args a = base_infer(m, T);
  a.n = n;
  a.theta = m->d_geo;
  a.z = m->d_geo + n;
  a.a1 = m->d_geo + 2 * n;
  a.a2 = m->d_geo + 3 * n;
  a.cls_d = cls_d >= 0 && cls_d < NCLS ? cls_d : -1;
  a.cls_a = cls_a >= 0 && cls_a < NCLS ? cls_a : -1;
  a.rate = m->d_rate;
  a.flag = m->d_flag;
  MLP_HIP(hipEventRecord(m->ev0, m->stream));
  MLP_HIP(launch_infer(a, m->n_cu, m->stream));
  MLP_HIP(hipEventRecord(m->ev1, m->stream));
  MLP_HIP(hipMemcpyAsync(rate, m->d_rate, n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (flag) MLP_HIP(hipMemcpyAsync(flag, m->d_flag, n, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (kernel_ms) {
    float ms = 0.f;
    MLP_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    *kernel_ms = ms;
  }
  return KMC_OK;
}

int kmc_mlp_table(kmc_mlp* m, int32_t cls_d, int32_t cls_a, double T, kmc_transfer* fb, int32_t nt,
                  const double* th, int32_t nz, const double* z, int32_t n1, const double* a1,
                  int32_t n2, const double* a2, double* rates, int64_t* n_fallback,
                  double* elapsed_ms) {
  if (!m || nt < 1 || nz < 1 || n1 < 1 || n2 < 1 || !th || !z || !a1 || !a2 || !rates)
    return mfail(KMC_ERR_ARG, "kmc_mlp_table: bad argument");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<double> g;
  grid_geometry(nt, th, nz, z, n1, a1, n2, a2, g);
  const int64_t n = (int64_t)nt * nz * n1 * n2;
  std::vector<uint8_t> fl(n);
  int s = kmc_mlp_rates(m, n, g.data(), g.data() + n, g.data() + 2 * n, g.data() + 3 * n, cls_d,
                        cls_a, T, rates, fl.data(), nullptr);
  if (s != KMC_OK) return s;
  std::vector<double> sub;
  std::vector<int64_t> idx;
  // coaxial nodes (theta = 0, a1 = 0) serve every same-tube hop: always exact
  for (int64_t e = 0; e < n; ++e)
    if (fl[e] || (g[e] == 0.0 && g[2 * n + e] == 0.0)) idx.push_back(e);
  if (fb && !idx.empty()) {
    const int64_t k = (int64_t)idx.size();
    sub.resize(4 * k);
    for (int64_t i = 0; i < k; ++i)
      for (int q = 0; q < 4; ++q) sub[q * k + i] = g[q * n + idx[i]];
    std::vector<double> r(k);
    s = kmc_transfer_rates(fb, k, sub.data(), sub.data() + k, sub.data() + 2 * k, sub.data() + 3 * k,
                           r.data(), nullptr);
    if (s != KMC_OK) return mfail(s, std::string("kmc_mlp_table: ") + kmc_transfer_last_error());
    for (int64_t i = 0; i < k; ++i) rates[idx[i]] = r[i];
  }
  if (n_fallback) *n_fallback = fb ? (int64_t)idx.size() : 0;
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KMC_OK;
}

}  // extern "C"
