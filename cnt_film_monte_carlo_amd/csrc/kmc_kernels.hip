// kmc_kernels.hip -- CDNA4 (gfx950) kernels of the KMC exciton-hopping path.
//
//   (a) cell list          k_cell_assign, k_cell_scatter, k_cell_sort
//   (b) CSR out-rate table  k_count_pairs, k_fill_pairs, k_finish_pair_rates
//   (c) hop kernel          k_inject, k_hop
//   (d) MSD/current reduce  k_reduce_samples
// plus a three-phase exclusive scan used by (a) and (b).
//
// Exactness: every FP64 expression that decides a neighbour, a rate, a selection or a dwell
// time is evaluated with contraction off (no FMA), in the operation order of the reference
// (src/monte_carlo/scatterer.cpp, scatterer.h, particle.cpp), so the CPU oracle reproduces
// it bit for bit.
#include <hip/hip_runtime.h>

#include "kmc_internal.h"
#include "kmc_kernels.h"
#include "philox.h"

namespace cntkmc {

// =============================================================================================
// Exclusive scan: uint32 counts -> int64 offsets, out[n] = total.
// =============================================================================================
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
  // Wave-level inclusive scan with shuffles (64-wide wavefronts), then across the 4 waves.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  int64_t wave_off = 0, tot = 0;
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    if (w < wave) wave_off += lds[w];
    tot += lds[w];
  }
  __syncthreads();
  *total = tot;
  return wave_off + x - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                             int64_t* __restrict__ out,
                                                             int64_t* __restrict__ tile_sums,
                                                             int64_t n) {
  __shared__ int64_t lds[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  int64_t sum = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    sum += v[i];
  }
  int64_t total;
  int64_t run = block_exclusive_scan(sum, lds, &total);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(int64_t* __restrict__ tile_sums,
                                                                 int64_t n_tiles,
                                                                 int64_t* __restrict__ grand_total) {
  __shared__ int64_t lds[SCAN_THREADS / 64];
  int64_t carry = 0;
  for (int64_t base = 0; base < n_tiles; base += SCAN_THREADS) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = (i < n_tiles) ? tile_sums[i] : 0;
    int64_t total;
    const int64_t ex = block_exclusive_scan(v, lds, &total);
    if (i < n_tiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *grand_total = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(int64_t* __restrict__ out,
                                                           const int64_t* __restrict__ tile_sums,
                                                           int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  const int64_t off = tile_sums[blockIdx.x];
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
    if (base + i < n) out[base + i] += off;
  }
}

hipError_t exclusive_scan(const uint32_t* d_in, int64_t* d_out, int64_t n, int64_t* d_tmp,
                          hipStream_t st) {
  // d_out has n+1 entries; d_tmp holds n_tiles entries.
  const int64_t n_tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (n_tiles > 0) {
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, st, d_in,
                       d_out, d_tmp, n);
  }
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(SCAN_THREADS), 0, st, d_tmp, n_tiles,
                     d_out + n);
  if (n_tiles > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, st, d_out,
                       d_tmp, n);
  }
  return hipGetLastError();
}

int64_t scan_tmp_elems(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// =============================================================================================
// (a) Cell list (create_scatterer_buckets, src/monte_carlo/monte_carlo.h:368-416)
// =============================================================================================
__global__ __launch_bounds__(256) void k_cell_assign(site_arrays s, cell_desc c,
                                                     int32_t* __restrict__ cell_of,
                                                     uint32_t* __restrict__ cell_count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  const int ix = cell_coord(c, 0, s.px[i]);
  const int iy = cell_coord(c, 1, s.py[i]);
  const int iz = cell_coord(c, 2, s.pz[i]);
  const int64_t idx = (int64_t)ix + (int64_t)iy * c.nx + (int64_t)iz * c.nx * c.ny;
  cell_of[i] = (int32_t)idx;
  atomicAdd(&cell_count[idx], 1u);
}

__global__ __launch_bounds__(256) void k_cell_scatter(int64_t n, const int32_t* __restrict__ cell_of,
                                                      const int64_t* __restrict__ cell_start,
                                                      uint32_t* __restrict__ cell_fill,
                                                      int32_t* __restrict__ cell_sites) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = cell_of[i];
  const uint32_t slot = atomicAdd(&cell_fill[c], 1u);
  cell_sites[cell_start[c] + slot] = (int32_t)i;
}

// Buckets hold their sites in ascending index order (the reference pushes sites in list
// order, monte_carlo.h:389-395).  Buckets are short: insertion sort per cell.
__global__ __launch_bounds__(256) void k_cell_sort(int64_t n_cells,
                                                   const int64_t* __restrict__ cell_start,
                                                   int32_t* __restrict__ cell_sites) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const int64_t b = cell_start[c], e = cell_start[c + 1];
  for (int64_t i = b + 1; i < e; ++i) {
    const int32_t v = cell_sites[i];
    int64_t j = i - 1;
    while (j >= b && cell_sites[j] > v) {
      cell_sites[j + 1] = cell_sites[j];
      --j;
    }
    cell_sites[j + 1] = v;
  }
}

// =============================================================================================
// (b) CSR build (find_neighbors + set_max_rate, src/monte_carlo/scatterer.cpp:34-83,
//     src/monte_carlo/scatterer.h:89-93)
// =============================================================================================
// Visit every candidate of site i in the reference order and call f(j, image, distance)
// for those inside (min_distance, cutoff).
template <typename F>
__device__ __forceinline__ void for_each_neighbor(const site_arrays& s, const cell_desc& c,
                                                  const int64_t* __restrict__ cell_start,
                                                  const int32_t* __restrict__ cell_sites,
                                                  int64_t i, F&& f) {
#pragma clang fp contract(off)
  const double pix = s.px[i], piy = s.py[i], piz = s.pz[i];
  int sx[3], sy[3], sz[3];
  const int nsx = stencil_axis(c, 0, cell_coord(c, 0, pix), sx);
  const int nsy = stencil_axis(c, 1, cell_coord(c, 1, piy), sy);
  const int nsz = stencil_axis(c, 2, cell_coord(c, 2, piz), sz);
  for (int a = 0; a < nsx; ++a) {
    for (int b = 0; b < nsy; ++b) {
      for (int d = 0; d < nsz; ++d) {
        const int64_t cell = (int64_t)sx[a] + (int64_t)sy[b] * c.nx + (int64_t)sz[d] * c.nx * c.ny;
        const int64_t beg = cell_start[cell], end = cell_start[cell + 1];
        for (int64_t q = beg; q < end; ++q) {
          const int32_t j = cell_sites[q];
          double pjx = s.px[j], pjy = s.py[j], pjz = s.pz[j];
          if (c.periodic_mask & 1) pjx = image_of(pix, pjx, c.period[0]);
          if (c.periodic_mask & 2) pjy = image_of(piy, pjy, c.period[1]);
          if (c.periodic_mask & 4) pjz = image_of(piz, pjz, c.period[2]);
          const double dist = norm3(pix - pjx, piy - pjy, piz - pjz);
          if ((dist < c.cutoff) && (dist > c.min_distance)) f(j, pjx, pjy, pjz);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_count_pairs(site_arrays s, cell_desc c,
                                                     const int64_t* __restrict__ cell_start,
                                                     const int32_t* __restrict__ cell_sites,
                                                     uint32_t* __restrict__ deg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  uint32_t count = 0;
  for_each_neighbor(s, c, cell_start, cell_sites, i,
                    [&](int32_t, double, double, double) { ++count; });
  deg[i] = count;
}

// mode 0 (table): rate from the nearest-node table and the running FP64 prefix sum
// (scatterer.cpp:63, 78-80).  mode 1 (pair): store the geometry for an external rate.
template <int MODE>
__global__ __launch_bounds__(256) void k_fill_pairs(site_arrays s, cell_desc c,
                                                    const int64_t* __restrict__ cell_start,
                                                    const int32_t* __restrict__ cell_sites,
                                                    const int64_t* __restrict__ row_ptr,
                                                    table_desc t, csr_arrays out,
                                                    pair_geom_arrays geo,
                                                    int32_t* __restrict__ err) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  const double pix = s.px[i], piy = s.py[i], piz = s.pz[i];
  const double oix = s.ox[i], oiy = s.oy[i], oiz = s.oz[i];
  const int ci = s.chir[i];
  int64_t pos = row_ptr[i];
  double cum = 0.0;
  bool first = true;
  for_each_neighbor(s, c, cell_start, cell_sites, i,
                    [&](int32_t j, double pjx, double pjy, double pjz) {
                      const pair_geom g = pair_geometry(pix, piy, piz, pjx, pjy, pjz, oix, oiy,
                                                        oiz, s.ox[j], s.oy[j], s.oz[j]);
                      edge_rec e;
                      e.col = (uint32_t)j | ((s.flags[j] & 2u) ? EDGE_DRAIN_BIT : 0u);
                      e.dx = (float)(pjx - pix);
                      e.dy = (float)(pjy - piy);
                      e.dz = (float)(pjz - piz);
                      out.edge[pos] = e;
                      if (MODE == 0) {
                        const double r = table_rate(t, ci, s.chir[j], g);
                        if (!(r >= 0.0) || isinf(r)) atomicOr(err, 1);
                        cum = first ? r : cum + r;
                        first = false;
                        out.cum[pos] = cum;
                      } else {
                        geo.theta[pos] = g.theta;
                        geo.z[pos] = g.z;
                        geo.a1[pos] = g.a1;
                        geo.a2[pos] = g.a2;
                      }
                      ++pos;
                    });
  site_hdr h;
  h.rs = (uint32_t)row_ptr[i];
  h.n = (uint32_t)(row_ptr[i + 1] - row_ptr[i]);
  h.gamma = (MODE == 0 && h.n > 0) ? cum : 0.0;
  out.hdr[i] = h;
}

template __global__ void k_fill_pairs<0>(site_arrays, cell_desc, const int64_t*, const int32_t*,
                                         const int64_t*, table_desc, csr_arrays, pair_geom_arrays,
                                         int32_t*);
template __global__ void k_fill_pairs<1>(site_arrays, cell_desc, const int64_t*, const int32_t*,
                                         const int64_t*, table_desc, csr_arrays, pair_geom_arrays,
                                         int32_t*);

// Sequential FP64 inclusive prefix sum per row from per-pair rates (scatterer.cpp:78-80).
__global__ __launch_bounds__(256) void k_finish_pair_rates(int64_t n, const double* __restrict__ rates,
                                                           csr_arrays out,
                                                           int32_t* __restrict__ err) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  site_hdr h = out.hdr[i];
  double cum = 0.0;
  for (uint32_t k = 0; k < h.n; ++k) {
    const double r = rates[h.rs + k];
    if (!(r >= 0.0) || isinf(r)) atomicOr(err, 1);
    cum = (k == 0) ? r : cum + r;
    out.cum[h.rs + k] = cum;
  }
  h.gamma = h.n > 0 ? cum : 0.0;
  out.hdr[i] = h;
}

// =============================================================================================
// (c) Injection and the persistent hop kernel
// =============================================================================================
// Dwell time of ff_time (scatterer.h:74-80): -(1/Gamma) * log(r / RAND_MAX), r in [1, RAND_MAX].
// A site without neighbours never hops again (the reference's set_max_rate is undefined there).
__device__ __forceinline__ double dwell_time(uint32_t n, double gamma, uint32_t r) {
#pragma clang fp contract(off)
  if (n == 0u) return INFINITY;
  const double inv = 1.0 / gamma;
  return -inv * log((double)r / RAND_MAX_D);
}

// Site header as three scalars from one 16-byte load.
struct hdr_regs {
  uint32_t rs, n;
  double gamma;
};

__device__ __forceinline__ hdr_regs load_hdr(const site_hdr* __restrict__ hdr, int32_t s) {
  const uint4 v = *(reinterpret_cast<const uint4*>(hdr) + s);
  hdr_regs h;
  h.rs = v.x;
  h.n = v.y;
  h.gamma = __hiloint2double((int)v.w, (int)v.z);
  return h;
}

__device__ __forceinline__ void load_edge(const edge_rec* __restrict__ edge, uint32_t p,
                                          uint32_t& col, float& ex, float& ey, float& ez) {
  const uint4 v = *(reinterpret_cast<const uint4*>(edge) + p);
  col = v.x;
  ex = __uint_as_float(v.y);
  ey = __uint_as_float(v.z);
  ez = __uint_as_float(v.w);
}

__global__ __launch_bounds__(256) void k_inject(exciton_arrays x, const site_hdr* __restrict__ hdr,
                                                const int32_t* __restrict__ src, uint32_t n_src,
                                                uint64_t seed) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= x.n) return;
  const u32x4 w = draw4(seed, e, 0, STREAM_INIT);
  const int32_t s = src[r31(w.v[0]) % n_src];
  const hdr_regs h = load_hdr(hdr, s);
  const uint32_t r = nonzero_draw(w.v[1], w.v[3], seed, e, 0, STREAM_INIT_REDRAW);
  x.site[e] = s;
  x.t_next[e] = 0.0 + dwell_time(h.n, h.gamma, r);
  x.hops[e] = 0;
  x.dx[e] = 0.0;
  x.dy[e] = 0.0;
  x.dz[e] = 0.0;
  x.drains[e] = 0;
  x.hash[e] = (FNV_OFFSET ^ (uint64_t)(uint32_t)s) * FNV_PRIME;
  x.k_rec[e] = 0;
  x.heading[e] = (uint8_t)(r31(w.v[2]) & 1u);
}

// One thread per exciton; each thread runs its exciton through its whole hop budget
// (event-driven form of particle::step, src/monte_carlo/particle.cpp:57-80, with v = 0).
template <bool COUNT, bool SAMPLE, bool TRACE>
__global__ __launch_bounds__(256) void k_hop(hop_args a) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.x.n) return;
  int32_t s = a.x.site[e];
  double t = a.x.t_next[e];
  uint64_t h = a.x.hops[e];
  double dx = a.x.dx[e], dy = a.x.dy[e], dz = a.x.dz[e];
  uint32_t nd = a.x.drains[e];
  uint64_t hs = a.x.hash[e];
  uint32_t k = SAMPLE ? a.x.k_rec[e] : 0u;
  int64_t tr = 0;
  if (TRACE && e < a.n_trace) tr = a.tr_n[e];
  hdr_regs H = load_hdr(a.hdr, s);
  const uint64_t h_end = h + a.hop_budget;
  const size_t E = a.x.n;

  while (true) {
    if (SAMPLE) {
      while (k < (uint32_t)a.n_bins) {
        const double tk = (double)(k + 1u) * a.dt_sample;
        if (!(tk < t)) break;
        const size_t r = (size_t)k * E + e;
        a.rec_dx[r] = dx;
        a.rec_dy[r] = dy;
        a.rec_dz[r] = dz;
        a.rec_drain[r] = nd;
        ++k;
      }
    }
    if (h >= h_end || H.n == 0u || !(t <= a.t_max)) break;

    // event selection (scatterer::update_state, scatterer.cpp:17-30)
    const u32x4 w = draw4(a.seed, e, h, STREAM_HOP);
    const double dice = (H.gamma * (double)r31(w.v[0])) / RAND_MAX_D;
    const double* row = a.cum + H.rs;
    uint32_t left = 0xFFFFFFFFu, right = H.n - 1u;
    while (left + 1u < right) {
      const uint32_t mid = (left + right) >> 1;
      if (row[mid] <= dice)
        left = mid;
      else
        right = mid;
    }
    uint32_t col;
    float ex, ey, ez;
    load_edge(a.edge, H.rs + right, col, ex, ey, ez);
    if (COUNT) atomicAdd(&a.site_count[s], 1u);
    const int32_t j = (int32_t)(col & EDGE_COL_MASK);
    dx = dx + (double)ex;
    dy = dy + (double)ey;
    dz = dz + (double)ez;
    hs = (hs ^ (uint64_t)(uint32_t)j) * FNV_PRIME;
    const double t_hop = t;

    // dwell at the new site (particle.cpp:183 -> ff_time)
    H = load_hdr(a.hdr, j);
    const uint32_t r = nonzero_draw(w.v[1], w.v[3], a.seed, e, h, STREAM_HOP_REDRAW);
    t = t + dwell_time(H.n, H.gamma, r);
    s = j;
    // drain: count current and re-inject, keeping displacement and dwell (kubo_step,
    // src/monte_carlo/monte_carlo.cpp:330-336)
    if (col & EDGE_DRAIN_BIT) {
      ++nd;
      s = a.src[r31(w.v[2]) % a.n_src];
      hs = (hs ^ ((uint64_t)(uint32_t)s | (1ull << 32))) * FNV_PRIME;
      H = load_hdr(a.hdr, s);
    }
    if (TRACE && e < a.n_trace && tr < a.tr_max) {
      a.tr_site[(size_t)e * a.tr_max + tr] = s;
      a.tr_time[(size_t)e * a.tr_max + tr] = t_hop;
      ++tr;
    }
    ++h;
  }

  a.x.site[e] = s;
  a.x.t_next[e] = t;
  a.x.hops[e] = h;
  a.x.dx[e] = dx;
  a.x.dy[e] = dy;
  a.x.dz[e] = dz;
  a.x.drains[e] = nd;
  a.x.hash[e] = hs;
  if (SAMPLE) a.x.k_rec[e] = k;
  if (TRACE && e < a.n_trace) a.tr_n[e] = tr;
}

hipError_t launch_hop(const hop_args& a, bool count, bool sample, bool trace, hipStream_t st) {
  const dim3 grid((unsigned)((a.x.n + 255) / 256)), block(256);
#define KMC_HOP_CASE(C, S, T)                                        \
  if (count == C && sample == S && trace == T) {                     \
    hipLaunchKernelGGL((k_hop<C, S, T>), grid, block, 0, st, a);     \
    return hipGetLastError();                                        \
  }
  KMC_HOP_CASE(false, false, false)
  KMC_HOP_CASE(true, false, false)
  KMC_HOP_CASE(false, true, false)
  KMC_HOP_CASE(true, true, false)
  KMC_HOP_CASE(false, false, true)
  KMC_HOP_CASE(true, false, true)
  KMC_HOP_CASE(false, true, true)
  KMC_HOP_CASE(true, true, true)
#undef KMC_HOP_CASE
  return hipErrorInvalidValue;
}

// =============================================================================================
// (d) MSD / current reduction: block (k, b) reduces bin k over exciton batch b in a fixed
//     order (per-thread strided sums, wavefront shuffle tree, LDS across waves).
// =============================================================================================
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_reduce_samples(reduce_args a) {
  __shared__ double lds[5][4];
  const int k = blockIdx.x, b = blockIdx.y;
  const int64_t E = a.n_exc;
  const int64_t e0 = (int64_t)b * E / a.n_batches, e1 = (int64_t)(b + 1) * E / a.n_batches;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    if ((uint32_t)k < a.k_rec[e]) {
      const size_t r = (size_t)k * E + e;
      const double x = a.rec_dx[r], y = a.rec_dy[r], z = a.rec_dz[r];
      acc[0] += 1.0;
      acc[1] += x * x;
      acc[2] += y * y;
      acc[3] += z * z;
      acc[4] += (double)a.rec_drain[r];
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) lds[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int q = threadIdx.x;
    const double v = (lds[q][0] + lds[q][1]) + (lds[q][2] + lds[q][3]);
    a.out[((size_t)k * a.n_batches + b) * 5 + q] = v;
  }
}

hipError_t launch_reduce(const reduce_args& a, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_samples, dim3(a.n_bins, a.n_batches), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// launch wrappers for (a) and (b)
// ---------------------------------------------------------------------------------------------
static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_cell_assign(const site_arrays& s, const cell_desc& c, int32_t* cell_of,
                              uint32_t* cell_count, hipStream_t st) {
  if (s.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_assign, grid_for(s.n), dim3(256), 0, st, s, c, cell_of, cell_count);
  return hipGetLastError();
}

hipError_t launch_cell_scatter(int64_t n, const int32_t* cell_of, const int64_t* cell_start,
                               uint32_t* cell_fill, int32_t* cell_sites, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_scatter, grid_for(n), dim3(256), 0, st, n, cell_of, cell_start,
                     cell_fill, cell_sites);
  return hipGetLastError();
}

hipError_t launch_cell_sort(int64_t n_cells, const int64_t* cell_start, int32_t* cell_sites,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_cell_sort, grid_for(n_cells), dim3(256), 0, st, n_cells, cell_start,
                     cell_sites);
  return hipGetLastError();
}

hipError_t launch_count_pairs(const site_arrays& s, const cell_desc& c, const int64_t* cell_start,
                              const int32_t* cell_sites, uint32_t* deg, hipStream_t st) {
  if (s.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_pairs, grid_for(s.n), dim3(256), 0, st, s, c, cell_start,
                     cell_sites, deg);
  return hipGetLastError();
}

hipError_t launch_fill_pairs(int mode, const site_arrays& s, const cell_desc& c,
                             const int64_t* cell_start, const int32_t* cell_sites,
                             const int64_t* row_ptr, const table_desc& t, const csr_arrays& out,
                             const pair_geom_arrays& geo, int32_t* err, hipStream_t st) {
  if (s.n == 0) return hipSuccess;
  if (mode == 0) {
    hipLaunchKernelGGL(k_fill_pairs<0>, grid_for(s.n), dim3(256), 0, st, s, c, cell_start,
                       cell_sites, row_ptr, t, out, geo, err);
  } else {
    hipLaunchKernelGGL(k_fill_pairs<1>, grid_for(s.n), dim3(256), 0, st, s, c, cell_start,
                       cell_sites, row_ptr, t, out, geo, err);
  }
  return hipGetLastError();
}

hipError_t launch_finish_pair_rates(int64_t n, const double* rates, const csr_arrays& out,
                                    int32_t* err, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finish_pair_rates, grid_for(n), dim3(256), 0, st, n, rates, out, err);
  return hipGetLastError();
}

hipError_t launch_inject(const exciton_arrays& x, const site_hdr* hdr, const int32_t* src,
                         uint32_t n_src, uint64_t seed, hipStream_t st) {
  if (x.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_inject, grid_for(x.n), dim3(256), 0, st, x, hdr, src, n_src, seed);
  return hipGetLastError();
}

}  // namespace cntkmc
