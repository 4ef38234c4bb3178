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
// A site without neighbours (the reference's set_max_rate is undefined there) or with total
// rate 0 (1/Gamma = inf in the reference, so ff_time never falls below dt) never hops again.
__device__ __forceinline__ bool trapped(uint32_t n, double gamma) { return n == 0u || !(gamma > 0.0); }

__device__ __forceinline__ double dwell_time(uint32_t n, double gamma, uint32_t r) {
#pragma clang fp contract(off)
  if (trapped(n, gamma)) return INFINITY;
  return -recip(gamma) * log_unit(div_rand_max((double)r));
}

// ---------------------------------------------------------------------------------------------
// Row blocks: the hop kernel's copy of the CSR.  Each site owns whole 128-byte lines (the L2
// line), so a hop with at most ROW_INLINE neighbours touches exactly one line.  First line, in
// 16-byte units:
//   u0      n | site | Gamma (FP64)
//   u1..u4  cum[0..8)   FP64 prefix sums; slots n-1..7 hold +inf (the last entry is Gamma,
//                       which the header carries: a count of the slots <= dice is then the
//                       selected index without a comparison against n)
//   u5, u6  tgt[0..8)   row word of the target: first line of its block | TGT_WIDE_BIT when
//                       that block has continuation lines | EDGE_DRAIN_BIT for a drain target
//   u7      x, y, z     FP32 position of the site | 0
// A site with n > ROW_INLINE continues on the next lines: cum[8..n) then tgt[8..n).  The
// displacement of a hop is the difference of the two sites' FP32 positions, so the block needs
// no per-pair payload.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t ROW_INLINE = 8;

__host__ __device__ __forceinline__ uint32_t block_lines(uint32_t n) {
  const uint32_t m = n > ROW_INLINE ? n - ROW_INLINE : 0u;
  return 1u + (12u * m + 127u) / 128u;
}

// Row word of a site whose block starts at line R.
__device__ __forceinline__ uint32_t row_word(uint32_t R, uint32_t n) {
  return R | (n > ROW_INLINE ? TGT_WIDE_BIT : 0u);
}

// Placement order of the blocks: sites sorted by the Morton (Z-order) code of their cell, so that
// any contiguous range of blocks -- one XCD's share of the sorted excitons, a workgroup's rows --
// covers a compact region of the film rather than a slab.  Only the memory layout follows this
// order; the neighbour order of every row stays the reference's.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ __launch_bounds__(256) void k_morton_keys(int64_t n, const int32_t* __restrict__ cell_of,
                                                     int32_t nx, int32_t ny,
                                                     uint64_t* __restrict__ keys,
                                                     int32_t* __restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c = cell_of[i];
  const uint32_t ix = (uint32_t)(c % nx), iy = (uint32_t)((c / nx) % ny), iz = (uint32_t)(c / ((int64_t)nx * ny));
  keys[i] = spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
  ids[i] = (int32_t)i;
}

// Lines of the q-th block in placement order (order[q] = its site; identity when NULL).
__global__ __launch_bounds__(256) void k_block_sizes(int64_t n, const site_hdr* __restrict__ hdr,
                                                     const int32_t* __restrict__ order,
                                                     uint32_t* __restrict__ lines) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  lines[q] = block_lines(hdr[order ? order[q] : q].n);
}

// roff[site] = first line of its block, from the scanned line counts in placement order.
__global__ __launch_bounds__(256) void k_place_blocks(int64_t n, const int32_t* __restrict__ order,
                                                      const int64_t* __restrict__ first,
                                                      int64_t* __restrict__ roff) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > n) return;
  if (q == n) roff[n] = first[n];
  else roff[order[q]] = first[q];
}

// Sort keys of the excitons: the first line of their current site's block.
// An exciton is idle in a launch if its next hop falls after t_max and no MSD record falls due
// before that hop (the hop kernel would only reload and store its state).  Idle excitons key
// IDLE_KEY and sort after every block line; `alive` counts the others (the hop kernel deals only
// those positions to the XCDs) and `alive_next`, the next launch's count, starts at zero.
constexpr uint32_t IDLE_KEY = 0xFFFFFFFFu;
__global__ __launch_bounds__(256) void k_exciton_keys(int64_t n, const int32_t* __restrict__ site,
                                                      const int64_t* __restrict__ roff,
                                                      idle_test it, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ live) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e == 0) {
    if (live) *live = 0u;
    if (it.alive_next) *it.alive_next = 0u;
  }
  bool alive = false;
  if (e < n) {
    uint32_t key = (uint32_t)roff[site[e]];
    if (it.alive) {
      const double t = it.t_next[e];
      bool idle = !(t <= it.t_max);
      if (idle && it.k_rec) {  // the hop kernel's record test: t_rec = (k + 1) dt_sample < t
        const uint32_t k = it.k_rec[e];
        idle = !(k < (uint32_t)it.n_bins && (double)(k + 1u) * it.dt_sample < t);
      }
      if (idle) key = IDLE_KEY;
      alive = !idle;
    }
    keys[e] = key;
    ids[e] = (uint32_t)e;
  }
  if (it.alive) {
    const uint64_t b = __ballot(alive);
    if ((threadIdx.x & 63u) == 0u && b != 0ull) atomicAdd(it.alive, (uint32_t)__popcll(b));
  }
}

__global__ __launch_bounds__(256) void k_pack_rows(int64_t n, const site_hdr* __restrict__ hdr,
                                                   const double* __restrict__ cum,
                                                   const edge_rec* __restrict__ edge,
                                                   const int64_t* __restrict__ roff,
                                                   const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz,
                                                   uint4* __restrict__ blocks) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const site_hdr h = hdr[i];
  uint4* b = blocks + (size_t)ROW_UNITS * (size_t)roff[i];
  uint4 head;
  head.x = h.n;
  head.y = (uint32_t)i;
  head.z = (uint32_t)__double2loint(h.gamma);
  head.w = (uint32_t)__double2hiint(h.gamma);
  b[0] = head;
  double* c = reinterpret_cast<double*>(b + 1);
  uint32_t* t = reinterpret_cast<uint32_t*>(b + 5);
  const uint32_t m = h.n > ROW_INLINE ? h.n - ROW_INLINE : 0u;
  double* ce = reinterpret_cast<double*>(b + ROW_UNITS);
  uint32_t* te = reinterpret_cast<uint32_t*>(ce + m);
  for (uint32_t k = (h.n > 0u ? h.n - 1u : 0u); k < ROW_INLINE; ++k) c[k] = INFINITY;
  for (uint32_t k = 0; k < h.n; ++k) {
    const edge_rec e = edge[h.rs + k];
    const uint32_t j = e.col & EDGE_COL_MASK;
    const uint32_t tg = row_word((uint32_t)roff[j], hdr[j].n) | (e.col & EDGE_DRAIN_BIT);
    if (k < ROW_INLINE) {
      if (k + 1u < h.n) c[k] = cum[h.rs + k];
      t[k] = tg;
    } else {
      ce[k - ROW_INLINE] = cum[h.rs + k];
      te[k - ROW_INLINE] = tg;
    }
  }
  b[7] = make_uint4(__float_as_uint((float)px[i]), __float_as_uint((float)py[i]),
                    __float_as_uint((float)pz[i]), 0u);
}

__global__ __launch_bounds__(256) void k_gather_offsets(int64_t n, const int32_t* __restrict__ idx,
                                                        const int64_t* __restrict__ roff,
                                                        const site_hdr* __restrict__ hdr,
                                                        uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = row_word((uint32_t)roff[idx[i]], hdr[idx[i]].n);
}

struct row_regs {
  uint32_t n, site;
  double gamma;
  double c[8];
  uint4 t0, t1;   // tgt[0..4), tgt[4..8)
  float x, y, z;
};

// Registers of a row from the eight 16-byte units of its first line.
__device__ __forceinline__ row_regs row_from_units(const uint4& h, const uint4& q0, const uint4& q1,
                                                   const uint4& q2, const uint4& q3, const uint4& t0,
                                                   const uint4& t1, const uint4& p) {
  row_regs r;
  r.t0 = t0;
  r.t1 = t1;
  r.n = h.x;
  r.site = h.y;
  r.gamma = __hiloint2double((int)h.w, (int)h.z);
  r.c[0] = __hiloint2double((int)q0.y, (int)q0.x);
  r.c[1] = __hiloint2double((int)q0.w, (int)q0.z);
  r.c[2] = __hiloint2double((int)q1.y, (int)q1.x);
  r.c[3] = __hiloint2double((int)q1.w, (int)q1.z);
  r.c[4] = __hiloint2double((int)q2.y, (int)q2.x);
  r.c[5] = __hiloint2double((int)q2.w, (int)q2.z);
  r.c[6] = __hiloint2double((int)q3.y, (int)q3.x);
  r.c[7] = __hiloint2double((int)q3.w, (int)q3.z);
  r.x = __uint_as_float(p.x);
  r.y = __uint_as_float(p.y);
  r.z = __uint_as_float(p.z);
  return r;
}

// The eight 16-byte loads of one block's first line (row word W), issued together.
__device__ __forceinline__ row_regs load_row(const uint4* __restrict__ blocks, uint32_t W) {
  const uint4* b = blocks + (size_t)ROW_UNITS * (W & TGT_LINE_MASK);
  return row_from_units(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]);
}

// Wave-cooperative form of load_row: every lane of the wave (all 64 active) passes the row word
// of the row it needs; the wave brings the 64 first lines into its 8 KB of LDS with eight
// direct-to-LDS loads, each fetching eight whole lines (lane j of load k takes one 16-byte unit
// of row 8k + j/8, so eight lanes cover one line), and each lane then reads its own row.  Load
// k writes lane j's unit to LDS unit 64k + j = 8 (8k + j/8) + j%8, i.e. row r owns units
// [8r, 8r + 8); the unit a lane fetches is rotated by r (unit u of row r sits in slot (u + r) % 8)
// so that the per-lane reads of one unit fall in different banks.  Replaces eight scattered
// 16-byte gathers per lane (64 lines per instruction) by eight coalesced loads (8 lines each).
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// The row as the wave loop holds it: header, the eight inline cumulative rates and the site's
// position.  Its targets stay in the wave's LDS copy of the line and are read by index
// (target_lds), so they take no registers.
struct coop_row {
  uint32_t n, site;
  double gamma;
  double c[8];
  float x, y, z;
};

// Byte offsets into the row blocks are 32-bit (pack_rows refuses blocks of 4 GiB or more), so
// each load takes the blocks' base from scalar registers and one 32-bit offset per lane.
// coop_issue starts the eight direct-to-LDS loads of the wave's 64 first lines; coop_finish waits
// for them (and for any other vector memory operation of the wave) and reads the lane's row.
__device__ __forceinline__ void coop_issue(const uint4* __restrict__ blocks, uint32_t W,
                                           uint4* lds_wave, uint32_t lane) {
  const uint32_t line = W & TGT_LINE_MASK;
  const uint32_t unit = ((lane & 7u) - (lane >> 3)) & 7u;
  // the eight source lines first (eight lane exchanges in flight together), then the loads
  uint32_t L[8];
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) L[k] = (uint32_t)__shfl((int)line, (int)(8u * k + (lane >> 3)), 64);
  // earlier reads of this LDS region are complete before it is overwritten
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  const char* base = reinterpret_cast<const char*>(blocks);
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t off = L[k] * (uint32_t)(16u * ROW_UNITS) + unit * 16u;
    __builtin_amdgcn_global_load_lds((global_void_t*)(base + off), (lds_void_t*)(lds_wave + 64u * k),
                                     16, 0, 0);
  }
}

__device__ __forceinline__ coop_row coop_finish(const uint4* lds_wave, uint32_t lane) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  const uint4* my = lds_wave + 8u * lane;
  const uint4 h = my[lane & 7u], q0 = my[(1u + lane) & 7u], q1 = my[(2u + lane) & 7u],
              q2 = my[(3u + lane) & 7u], q3 = my[(4u + lane) & 7u], p = my[(7u + lane) & 7u];
  coop_row r;
  r.n = h.x;
  r.site = h.y;
  r.gamma = __hiloint2double((int)h.w, (int)h.z);
  r.c[0] = __hiloint2double((int)q0.y, (int)q0.x);
  r.c[1] = __hiloint2double((int)q0.w, (int)q0.z);
  r.c[2] = __hiloint2double((int)q1.y, (int)q1.x);
  r.c[3] = __hiloint2double((int)q1.w, (int)q1.z);
  r.c[4] = __hiloint2double((int)q2.y, (int)q2.x);
  r.c[5] = __hiloint2double((int)q2.w, (int)q2.z);
  r.c[6] = __hiloint2double((int)q3.y, (int)q3.x);
  r.c[7] = __hiloint2double((int)q3.w, (int)q3.z);
  r.x = __uint_as_float(p.x);
  r.y = __uint_as_float(p.y);
  r.z = __uint_as_float(p.z);
  return r;
}

// Target word idx < ROW_INLINE of the lane's row from the wave's LDS copy (units 5 and 6,
// rotated by the lane as coop_issue stores them).  Valid until the next coop_issue.
__device__ __forceinline__ uint32_t target_lds(const uint4* lds_wave, uint32_t lane, uint32_t idx) {
  const uint32_t slot = (5u + (idx >> 2) + lane) & 7u;
  return reinterpret_cast<const uint32_t*>(lds_wave + 8u * lane + slot)[idx & 3u];
}

// The number of entries <= d of a non-decreasing c[0..8), in 0..8: a three-step binary search
// over registers (selects, no indexing) and one more comparison for 8.  Step one tells count >= 4,
// step two count >= base + 2 (base = 0 or 4), step three count >= base + 2 b1 + 1.
__device__ __forceinline__ uint32_t count_le8(const double* c, double d) {
  const bool b2 = c[3] <= d;
  const bool b1 = (b2 ? c[5] : c[1]) <= d;
  const bool b0 = (b2 ? (b1 ? c[6] : c[4]) : (b1 ? c[2] : c[0])) <= d;
  const uint32_t idx = (b2 ? 4u : 0u) | (b1 ? 2u : 0u) | (b0 ? 1u : 0u);
  return (idx == 7u && c[7] <= d) ? 8u : idx;
}

// Index picked by update_state's binary search (scatterer.cpp:17-30): the first entry whose
// cumulative rate exceeds dice, or n-1.  Rows are non-decreasing (rates >= 0 are enforced when
// the CSR is built), so that index equals the number of leading entries k < n-1 with
// cum[k] <= dice.  The inline slots n-1..7 hold +inf, so the count over the eight inline slots
// is that number for n <= 9 without any comparison against n; a longer row continues on its
// next line.
__device__ __forceinline__ uint32_t select_entry(const uint4* __restrict__ blocks, uint32_t W,
                                                 const double* c8, uint32_t n, double dice) {
  uint32_t idx = count_le8(c8, dice);
  if (idx == 8u && n > 9u) {
    // cum[8..16): the first 64 bytes of the continuation line, which a row with n > 9 always
    // owns, read in one batch of four 16-byte loads before any comparison; then further entries
    // one at a time
    const uint32_t lim = n - 1u;
    const uint4* cu = blocks + (size_t)ROW_UNITS * (W & TGT_LINE_MASK) + ROW_UNITS;
    const uint4 e0 = cu[0], e1 = cu[1], e2 = cu[2], e3 = cu[3];
    const double c[8] = {__hiloint2double((int)e0.y, (int)e0.x), __hiloint2double((int)e0.w, (int)e0.z),
                         __hiloint2double((int)e1.y, (int)e1.x), __hiloint2double((int)e1.w, (int)e1.z),
                         __hiloint2double((int)e2.y, (int)e2.x), __hiloint2double((int)e2.w, (int)e2.z),
                         __hiloint2double((int)e3.y, (int)e3.x), __hiloint2double((int)e3.w, (int)e3.z)};
    uint32_t k = 8;
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) k += (8u + q < lim && c[q] <= dice) ? 1u : 0u;
    if (k == 16u) {
      const double* ce = reinterpret_cast<const double*>(cu);
      while (k < lim && ce[k - 8u] <= dice) ++k;
    }
    idx = k;
  }
  return idx;
}

// Target word of a continuation entry idx >= ROW_INLINE of a row with n entries.
__device__ __forceinline__ uint32_t target_wide(const uint4* __restrict__ blocks, uint32_t W,
                                                uint32_t n, uint32_t idx) {
  const uint32_t m = n - ROW_INLINE;
  const double* ce =
      reinterpret_cast<const double*>(blocks + (size_t)ROW_UNITS * (W & TGT_LINE_MASK) + ROW_UNITS);
  return reinterpret_cast<const uint32_t*>(ce + m)[idx - ROW_INLINE];
}

// Target word of entry idx: from the registers for idx < 8, else from the continuation lines.
__device__ __forceinline__ uint32_t target_of(const uint4* __restrict__ blocks, uint32_t W,
                                              const row_regs& r, uint32_t idx) {
  if (idx >= ROW_INLINE) return target_wide(blocks, W, r.n, idx);
  const uint4 q = idx < 4u ? r.t0 : r.t1;
  const uint32_t k = idx & 3u;
  return k == 0u ? q.x : k == 1u ? q.y : k == 2u ? q.z : q.w;
}

// One coordinate of a hop's displacement: FP64 difference of the two FP32 positions (exact for
// sites this close), shifted by one period on a periodic axis when it exceeds half of it.
// `half` is half the period on a periodic axis and +inf on an open one (a kernel argument, so
// it stays in scalar registers); the period is half + half, exact.  One comparison of |d| on
// the common path; the shift itself is rare.
__device__ __forceinline__ double min_image(double d, double half) {
#pragma clang fp contract(off)
  if (__builtin_expect(fabs(d) > half, 0)) d = d > half ? d - (half + half) : d + (half + half);
  return d;
}

// Per-thread write-back cache of departure counts: four (site, count) entries in LRU order,
// flushed with one atomic add into this XCD's private copy of the count array when evicted.
struct count_cache {
  uint32_t s0 = 0xFFFFFFFFu, s1 = 0xFFFFFFFFu, s2 = 0xFFFFFFFFu, s3 = 0xFFFFFFFFu;
  uint32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;

  // Selects rather than a branch per entry: the lanes of a wave hit different entries, and a
  // branch chain would run every arm under partial masks and rotate the eight registers at
  // its merge.  The entries above the hit (all four on a miss) move down one place.
  __device__ __forceinline__ void add(uint32_t s, uint32_t* __restrict__ base) {
    const bool h0 = s == s0, h1 = s == s1, h2 = s == s2, h3 = s == s3;
    if (!(h0 || h1 || h2 || h3) && n3 != 0u) atomicAdd(base + s3, n3);  // the evicted entry
    // at most one entry holds s (entries are distinct lines; empty ones hold no line), so
    // the hit's count is a sum of independent selects
    const uint32_t c = (h0 ? n0 : 0u) + (h1 ? n1 : 0u) + (h2 ? n2 : 0u) + (h3 ? n3 : 0u) + 1u;
    const bool m1 = !h0, m2 = !(h0 || h1), m3 = !(h0 || h1 || h2);
    s3 = m3 ? s2 : s3;
    n3 = m3 ? n2 : n3;
    s2 = m2 ? s1 : s2;
    n2 = m2 ? n1 : n2;
    s1 = m1 ? s0 : s1;
    n1 = m1 ? n0 : n1;
    s0 = s;
    n0 = c;
  }
  __device__ __forceinline__ void flush(uint32_t* __restrict__ base) {
    if (n0) atomicAdd(base + s0, n0);
    if (n1) atomicAdd(base + s1, n1);
    if (n2) atomicAdd(base + s2, n2);
    if (n3) atomicAdd(base + s3, n3);
  }
};

// Hardware id of the XCD running this wave (gfx950: 8 XCDs).  Only picks which private copy
// of the count array a wave updates; correctness does not depend on it (device-scope atomics).
__device__ __forceinline__ uint32_t xcd_id() {
  return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (KMC_COUNT_COPIES - 1u);
}

__global__ __launch_bounds__(256) void k_inject(exciton_arrays x, const uint4* __restrict__ blocks,
                                                const uint32_t* __restrict__ src_roff,
                                                uint32_t n_src, uint64_t seed) {
#pragma clang fp contract(off)
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= x.n) return;
  const u32x4 w = draw4(seed, e, 0, STREAM_INIT);
  const uint4 h = blocks[(size_t)ROW_UNITS * (src_roff[r31(w.v[0]) % n_src] & TGT_LINE_MASK)];
  const uint32_t r = nonzero_draw(w.v[1], w.v[3], seed, e, 0, STREAM_INIT_REDRAW);
  x.site[e] = (int32_t)h.y;
  x.t_next[e] = 0.0 + dwell_time(h.x, __hiloint2double((int)h.w, (int)h.z), r);
  x.hops[e] = 0;
  x.dx[e] = 0.0;
  x.dy[e] = 0.0;
  x.dz[e] = 0.0;
  x.drains[e] = 0;
  x.hash[e] = (FNV_OFFSET ^ (uint64_t)h.y) * FNV_PRIME;
  x.k_rec[e] = 0;
  x.heading[e] = (uint8_t)(r31(w.v[2]) & 1u);
  x.s_off[e] = 0.0;
  x.t_last[e] = 0.0;
}

// Same-tube links as block lines with the target's drain bit (the flight's view of
// scatterer::left/right, scatterer.h:29-30).
__global__ __launch_bounds__(256) void k_links(int64_t n, const int32_t* __restrict__ left,
                                               const int32_t* __restrict__ right,
                                               const uint8_t* __restrict__ flags,
                                               const int64_t* __restrict__ roff,
                                               const site_hdr* __restrict__ hdr,
                                               uint2* __restrict__ links) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t l = left[i], r = right[i];
  uint2 o;
  o.x = l < 0 ? LINK_NONE
              : (row_word((uint32_t)roff[l], hdr[l].n) | ((flags[l] & 2u) ? EDGE_DRAIN_BIT : 0u));
  o.y = r < 0 ? LINK_NONE
              : (row_word((uint32_t)roff[r], hdr[r].n) | ((flags[r] & 2u) ? EDGE_DRAIN_BIT : 0u));
  links[i] = o;
}

// ---------------------------------------------------------------------------------------------
// Drift along the tube (particle::fly, src/monte_carlo/particle.cpp:9-54) for v > 0.  The
// exciton sits s_off beyond its current site on the segment toward the next site along its
// heading; its displacement is kept as the displacement of the current site's position (hops
// and passed segments) plus that offset.  One flight covers the whole dwell before a hop (the
// reference cuts it into 1 fs steps; the path is the same).
// ---------------------------------------------------------------------------------------------
struct flight_state {
  uint32_t R;         // first line of the current site's block
  uint32_t site;
  float x, y, z;      // the current site's FP32 position
  double s;           // distance travelled from the current site toward the next one
  uint32_t heading;   // 1 = toward `right`
  double dx, dy, dz;  // displacement of the current site's position
  uint32_t nd;        // drains
  uint64_t hs;        // visited-site hash
};

// Next site along the heading; a tube end reverses the heading (particle.cpp:18-37).
__device__ __forceinline__ uint32_t fly_next(uint2 lk, uint32_t& heading) {
  const uint32_t nx = heading ? (lk.y != LINK_NONE ? lk.y : lk.x) : (lk.x != LINK_NONE ? lk.x : lk.y);
  heading = (nx == lk.y) ? 1u : 0u;
  return nx;
}

__device__ __forceinline__ void fly_segment(const net_view& v, const flight_state& f, uint32_t Rn,
                                            float& px, float& py, float& pz, double& sx,
                                            double& sy, double& sz) {
#pragma clang fp contract(off)
  const uint4 p = v.blocks[(size_t)ROW_UNITS * (Rn & TGT_LINE_MASK) + 7];
  px = __uint_as_float(p.x);
  py = __uint_as_float(p.y);
  pz = __uint_as_float(p.z);
  sx = min_image((double)px - (double)f.x, v.half_per[0]);
  sy = min_image((double)py - (double)f.y, v.half_per[1]);
  sz = min_image((double)pz - (double)f.z, v.half_per[2]);
}

// Flight of duration tau.  A site reached on the way becomes the current site; a drain site
// counts a drain and re-injects on a source site (draw of stream STREAM_FLY_DRAIN + j for the
// j-th drain of this flight, hop counter h), keeping the displacement, and the flight goes on
// from there, as the reference's next 1 fs step would.  False when the flight would pass more
// than FLY_MAX_PASSES sites (the state is then left where the bound stopped it).
__device__ bool fly(const net_view& v, flight_state& f, double tau, uint32_t e, uint64_t h) {
#pragma clang fp contract(off)
  uint2 lk = v.links[f.site];
  if (lk.x == LINK_NONE && lk.y == LINK_NONE) return true;  // particle.cpp:11-12
  uint32_t j = 0;
  for (uint32_t pass = 0; pass < FLY_MAX_PASSES; ++pass) {
    const uint32_t nx = fly_next(lk, f.heading);
    const uint32_t Rn = nx & ~EDGE_DRAIN_BIT;
    float px, py, pz;
    double sx, sy, sz;
    fly_segment(v, f, Rn, px, py, pz, sx, sy, sz);
    const double len = norm3(sx, sy, sz);
    const double dist = len - f.s;
    if (!(dist / v.velocity < tau)) {  // stops on this segment (particle.cpp:46-49)
      f.s = f.s + v.velocity * tau;
      return true;
    }
    tau = tau - dist / v.velocity;
    f.dx = f.dx + sx;
    f.dy = f.dy + sy;
    f.dz = f.dz + sz;
    f.s = 0.0;
    f.R = Rn;
    f.site = v.blocks[(size_t)ROW_UNITS * (Rn & TGT_LINE_MASK)].y;
    f.x = px;
    f.y = py;
    f.z = pz;
    f.hs = (f.hs ^ ((uint64_t)f.site | (2ull << 32))) * FNV_PRIME;
    if (nx & EDGE_DRAIN_BIT) {
      ++f.nd;
      const u32x4 w = draw4(v.seed, e, h, STREAM_FLY_DRAIN + j);
      ++j;
      f.R = v.src_roff[r31(w.v[0]) % v.n_src];
      const uint4 hd = v.blocks[(size_t)ROW_UNITS * (f.R & TGT_LINE_MASK)];
      const uint4 p = v.blocks[(size_t)ROW_UNITS * (f.R & TGT_LINE_MASK) + 7];
      f.site = hd.y;
      f.x = __uint_as_float(p.x);
      f.y = __uint_as_float(p.y);
      f.z = __uint_as_float(p.z);
      f.hs = (f.hs ^ ((uint64_t)f.site | (1ull << 32))) * FNV_PRIME;
    }
    lk = v.links[f.site];
    if (lk.x == LINK_NONE && lk.y == LINK_NONE) return true;
  }
  return false;
}

// Position of the exciton relative to its current site: s along the segment toward the next one.
__device__ __forceinline__ void fly_offset(const net_view& v, const flight_state& f, double& ox,
                                           double& oy, double& oz) {
#pragma clang fp contract(off)
  ox = oy = oz = 0.0;
  if (!(f.s != 0.0)) return;
  uint32_t hd = f.heading;
  const uint32_t Rn = fly_next(v.links[f.site], hd) & ~EDGE_DRAIN_BIT;
  float px, py, pz;
  double sx, sy, sz;
  fly_segment(v, f, Rn, px, py, pz, sx, sy, sz);
  const double q = f.s / norm3(sx, sy, sz);
  ox = sx * q;
  oy = sy * q;
  oz = sz * q;
}

// One exciton through its hop budget (event-driven form of particle::step,
// src/monte_carlo/particle.cpp:57-80; FLY adds the drift of particle::fly for v > 0).  v holds everything of the exciton's network;
// it is uniform across the wave, so its fields live in scalar registers.
// Returns whether a further launch could move the exciton (t <= t_max, a dwell pending).
template <bool COUNT, bool SAMPLE, bool TRACE, bool FLY = false>
__device__ __forceinline__ bool hop_exciton(const net_view& v, uint32_t e, uint64_t hop_budget,
                                            double t_max, const trace_view& tv) {
#pragma clang fp contract(off)
  double t = v.x.t_next[e];
  double tl = FLY ? v.x.t_last[e] : 0.0;
  double so = FLY ? v.x.s_off[e] : 0.0;
  uint32_t hd = FLY ? (uint32_t)v.x.heading[e] : 0u;
  uint64_t h = v.x.hops[e];
  double dx = v.x.dx[e], dy = v.x.dy[e], dz = v.x.dz[e];
  uint32_t nd = v.x.drains[e];
  uint64_t hs = v.x.hash[e];
  uint32_t k = SAMPLE ? v.x.k_rec[e] : 0u;
  int64_t tr = 0;
  if (TRACE && e < tv.n_trace) tr = tv.tr_n[e];
  const uint32_t R0 = (uint32_t)v.roff[v.x.site[e]];
  uint32_t R = row_word(R0, v.blocks[(size_t)ROW_UNITS * R0].x);
  row_regs row = load_row(v.blocks, R);
  const uint64_t h_end = h + hop_budget;
  const size_t E = v.x.n;
  count_cache cc;
  hop_words hw;
  uint32_t* cbase = COUNT ? v.site_count + (size_t)xcd_id() * v.count_stride : nullptr;

  while (true) {
    if (SAMPLE) {
      while (k < (uint32_t)v.n_bins) {
        const double tk = (double)(k + 1u) * v.dt_sample;
        if (!(tk < t)) break;
        const size_t r = (size_t)(k - (uint32_t)v.bin0) * E + e;
        if (FLY) {  // the flight from the last event up to tk
          flight_state g{R, row.site, row.x, row.y, row.z, so, hd, dx, dy, dz, nd, hs};
          if (!fly(v, g, tk - tl, e, h)) {
            atomicOr(v.fly_err, 1);
            break;
          }
          double ox, oy, oz;
          fly_offset(v, g, ox, oy, oz);
          v.rec_dx[r] = g.dx + ox;
          v.rec_dy[r] = g.dy + oy;
          v.rec_dz[r] = g.dz + oz;
          v.rec_drain[r] = g.nd;
        } else {
          v.rec_dx[r] = dx;
          v.rec_dy[r] = dy;
          v.rec_dz[r] = dz;
          v.rec_drain[r] = nd;
        }
        ++k;
      }
    }
    if (FLY) {
      if (h >= h_end || !(t <= t_max) || !(t < INFINITY)) break;
      // fly through the dwell, then hop from the site reached (particle.cpp:62-75)
      flight_state g{R, row.site, row.x, row.y, row.z, so, hd, dx, dy, dz, nd, hs};
      if (!fly(v, g, t - tl, e, h)) {
        atomicOr(v.fly_err, 1);
        break;
      }
      if (g.R != R) {
        R = g.R;
        row = load_row(v.blocks, R);
      }
      so = g.s;
      hd = g.heading;
      dx = g.dx;
      dy = g.dy;
      dz = g.dz;
      nd = g.nd;
      hs = g.hs;
      if (trapped(row.n, row.gamma)) {
        // no rates here: the event leaves the exciton in place (scatterer.cpp:14-15) and no
        // dwell ends again; it counts as a hop so that later flights draw on fresh counters
        tl = t;
        t = INFINITY;
        ++h;
        hw.have = false;
        continue;
      }
    } else if (h >= h_end || trapped(row.n, row.gamma) || !(t <= t_max)) {
      break;
    }

    // event selection (scatterer::update_state, scatterer.cpp:17-30)
    uint32_t w_sel, w_dw;
    hw.get(v.seed, e, h, w_sel, w_dw);
    const double dice = div_rand_max(row.gamma * (double)r31(w_sel));
    const uint32_t idx = select_entry(v.blocks, R, row.c, row.n, dice);
    const uint32_t tg = target_of(v.blocks, R, row, idx);
    if (COUNT) cc.add(count_slot(R & TGT_LINE_MASK, v.count_stride), cbase);
    const float sx = row.x, sy = row.y, sz = row.z;
    const double t_hop = t;

    // dwell at the new site (particle.cpp:183 -> ff_time)
    R = tg & ~EDGE_DRAIN_BIT;
    row = load_row(v.blocks, R);
    dx = dx + min_image((double)row.x - (double)sx, v.half_per[0]);
    dy = dy + min_image((double)row.y - (double)sy, v.half_per[1]);
    dz = dz + min_image((double)row.z - (double)sz, v.half_per[2]);
    hs = (hs ^ (uint64_t)row.site) * FNV_PRIME;
    const uint32_t r = nonzero_draw1(w_dw, v.seed, e, h, STREAM_HOP_REDRAW);
    t = t + dwell_time(row.n, row.gamma, r);
    // drain: count current and re-inject, keeping displacement and dwell (kubo_step,
    // src/monte_carlo/monte_carlo.cpp:330-336)
    if (tg & EDGE_DRAIN_BIT) {
      ++nd;
      R = v.src_roff[r31(draw4(v.seed, e, h, STREAM_REINJECT).v[0]) % v.n_src];
      row = load_row(v.blocks, R);
      hs = (hs ^ ((uint64_t)row.site | (1ull << 32))) * FNV_PRIME;
    }
    if (FLY) {
      so = 0.0;
      tl = t_hop;
    }
    if (TRACE && e < tv.n_trace && tr < tv.tr_max) {
      tv.tr_site[(size_t)e * tv.tr_max + tr] = (int32_t)row.site;
      tv.tr_time[(size_t)e * tv.tr_max + tr] = t_hop;
      ++tr;
    }
    ++h;
  }
  if (COUNT) cc.flush(cbase);

  v.x.site[e] = (int32_t)row.site;
  v.x.t_next[e] = t;
  v.x.hops[e] = h;
  v.x.dx[e] = dx;
  v.x.dy[e] = dy;
  v.x.dz[e] = dz;
  v.x.drains[e] = nd;
  v.x.hash[e] = hs;
  if (SAMPLE) v.x.k_rec[e] = k;
  if (TRACE && e < tv.n_trace) tv.tr_n[e] = tr;
  if (FLY) {
    v.x.s_off[e] = so;
    v.x.t_last[e] = tl;
    v.x.heading[e] = (uint8_t)hd;
    return t <= t_max && t < INFINITY;
  }
  return !trapped(row.n, row.gamma) && t <= t_max;
}

// The hop loop of one wave (v = 0), built around the cooperative row load: every lane takes part
// in every row load, so the loop runs until no lane of the wave can hop; a lane without an
// exciton (valid false) or whose exciton has stopped keeps loading some row.  Each iteration
// takes the row of the exciton's current site and
//   (A) selects the next hop from it (draws, dice, target) and issues the load of the next row
//       -- the target's, or the source site's after a drain -- before anything else, so the load
//       is in flight while
//   (B) the hop into this row finishes (displacement, visit hash, dwell at this site) and the
//       MSD records fall due, and
//   (C) the hop selected in (A) happens if the exciton's clock allows it (t <= t_max), else the
//       exciton stops at this site and the issued load is simply not used.
// Every exciton's arithmetic and its order are those of hop_exciton (and of the oracle); only
// the selection is computed before the clock test that decides whether it is used.  A drain
// target's header and position are gathered by the lanes concerned.
// COUNT: departures from the block lines [win_base, win_base + COUNT_WINDOW_LINES) are counted
// in the workgroup's LDS window `win` (zeroed and flushed by the kernel), the rest through a
// four-entry register LRU into the XCD's copy of the counts.  The window holds 16-bit counters,
// two lines per word, so that 8 KB of LDS cover 4,096 lines.  A counter never wraps: the
// departure whose add returns 0x7FFF (taking the counter to 0x8000) subtracts 0x8000 from it
// again and adds 0x8000 to the line's device count.  Exactly one lane sees each such return,
// and its subtraction follows its add within a few iterations of the other waves, far fewer
// than the 0x7FFF departures that would carry into the neighbouring counter.
constexpr uint32_t COUNT_WINDOW_WORDS = 2048;
constexpr uint32_t COUNT_WINDOW_LINES = 2 * COUNT_WINDOW_WORDS;
constexpr uint32_t COUNT_NONE = 0xFFFFFFFFu;

// One departure from window line `off` (block line `line`).
__device__ __forceinline__ void window_count(uint32_t* win, uint32_t off, uint32_t line,
                                             uint32_t* __restrict__ cbase, int64_t stride) {
  const uint32_t sh = (off & 1u) << 4;
  const uint32_t old = atomicAdd(win + (off >> 1), 1u << sh);
  if (__builtin_expect(((old >> sh) & 0xFFFFu) == 0x7FFFu, 0)) {
    atomicSub(win + (off >> 1), 0x8000u << sh);
    atomicAdd(cbase + count_slot(line, stride), 0x8000u);
  }
}

// The workgroup's window into the XCD's copy of the counts, one atomic per nonzero line (after
// a barrier that follows every window_count of the workgroup).
__device__ __forceinline__ void window_flush(const uint32_t* win, uint32_t base,
                                             uint32_t* __restrict__ cbase, int64_t stride) {
  for (uint32_t i = threadIdx.x; i < COUNT_WINDOW_WORDS; i += blockDim.x) {
    const uint32_t c = win[i];
    if (c & 0xFFFFu) atomicAdd(cbase + count_slot(base + 2u * i, stride), c & 0xFFFFu);
    if (c >> 16) atomicAdd(cbase + count_slot(base + 2u * i + 1u, stride), c >> 16);
  }
}

template <bool COUNT, bool SAMPLE, bool TRACE>
__device__ __forceinline__ void hop_wave(const net_view& v, uint32_t e, bool valid,
                                         uint64_t hop_budget, double t_max, const trace_view& tv,
                                         uint4* lds_wave, uint32_t lane, uint32_t* win,
                                         uint32_t win_base) {
#pragma clang fp contract(off)
  double t = INFINITY, dx = 0.0, dy = 0.0, dz = 0.0;
  uint64_t h = 0, hs = 0;
  uint32_t nd = 0, k = 0, W = 0;
  int64_t tr = 0;
  if (valid) {
    t = v.x.t_next[e];
    h = v.x.hops[e];
    dx = v.x.dx[e];
    dy = v.x.dy[e];
    dz = v.x.dz[e];
    nd = v.x.drains[e];
    hs = v.x.hash[e];
    if (SAMPLE) k = v.x.k_rec[e];
    if (TRACE && e < tv.n_trace) tr = tv.tr_n[e];
    const uint32_t R0 = (uint32_t)v.roff[v.x.site[e]];
    W = row_word(R0, v.blocks[(size_t)ROW_UNITS * R0].x);
  }
  const uint64_t h_end = h + hop_budget;
  const size_t E = v.x.n;
  // the next MSD record falls due once the clock passes t_rec = (k + 1) dt_sample
  double t_rec = INFINITY;
  if (SAMPLE && k < (uint32_t)v.n_bins) t_rec = (double)(k + 1u) * v.dt_sample;
  bool active = valid;
  bool live = false;           // when stopped: able to hop in a later launch
  uint32_t site_now = 0;       // when stopped: the site it stopped at
  bool pend = false;           // a hop waits for its target row
  uint32_t r_dw = 0;           // its dwell draw
  float px = 0.f, py = 0.f, pz = 0.f;  // position of the site it left
  double t_hop = 0.0;
  hop_words hw;
  count_cache cc;
  uint32_t* cbase = COUNT ? v.site_count + (size_t)xcd_id() * v.count_stride : nullptr;
  // the departure of the last hop is counted at the top of the next iteration, so that an
  // eviction's device atomic has a whole iteration to complete before the wave waits on its
  // next row load (one counter tracks both)
  uint32_t count_line = COUNT_NONE;
  auto count_departure = [&]() {
    const uint32_t off = count_line - win_base;
    if (off < COUNT_WINDOW_LINES) window_count(win, off, count_line, cbase, v.count_stride);
    else cc.add(count_slot(count_line, v.count_stride), cbase);
    count_line = COUNT_NONE;
  };

  coop_issue(v.blocks, W, lds_wave, lane);
  while (true) {
    // the wave issues at raised priority from its row's arrival until the next row load is out
    // (the stretch its exciton's hops wait on); the dwell and records of (B) and (C) run at the
    // normal one, so the SIMD's other waves reach their own loads first
    __builtin_amdgcn_s_setprio(2);
    const coop_row row = coop_finish(lds_wave, lane);
    if (COUNT && count_line != COUNT_NONE) count_departure();
    // (A) the hop from this site, if the exciton may still hop (scatterer::update_state,
    // scatterer.cpp:17-30), and the load of the row it leads to
    const bool can = active && h < h_end && !trapped(row.n, row.gamma);
    uint32_t tg = 0, r_next = 0, W_next = W;
    if (can) {
      uint32_t w_sel, w_dw;
      hw.get(v.seed, e, h, w_sel, w_dw);
      const double dice = div_rand_max(row.gamma * (double)r31(w_sel));
      const uint32_t idx = select_entry(v.blocks, W, row.c, row.n, dice);
      tg = idx < ROW_INLINE ? target_lds(lds_wave, lane, idx) : target_wide(v.blocks, W, row.n, idx);
      r_next = nonzero_draw1(w_dw, v.seed, e, h, STREAM_HOP_REDRAW);
      W_next = (tg & EDGE_DRAIN_BIT) ? v.src_roff[r31(draw4(v.seed, e, h, STREAM_REINJECT).v[0]) % v.n_src]
                                     : (tg & ~EDGE_DRAIN_BIT);
    }
    // a lane whose exciton has stopped never reads its row again: it loads block line 0, which
    // every such lane of the wave shares, instead of its own row
    coop_issue(v.blocks, active ? W_next : 0u, lds_wave, lane);
    __builtin_amdgcn_s_setprio(0);
    // (B) the hop into this site finishes: dwell at the new site (particle.cpp:183 -> ff_time)
    if (pend) {
      dx = dx + min_image((double)row.x - (double)px, v.half_per[0]);
      dy = dy + min_image((double)row.y - (double)py, v.half_per[1]);
      dz = dz + min_image((double)row.z - (double)pz, v.half_per[2]);
      hs = (hs ^ (uint64_t)row.site) * FNV_PRIME;
      t = t + dwell_time(row.n, row.gamma, r_dw);
      if (TRACE && e < tv.n_trace && tr < tv.tr_max) {
        tv.tr_site[(size_t)e * tv.tr_max + tr] = (int32_t)row.site;
        tv.tr_time[(size_t)e * tv.tr_max + tr] = t_hop;
        ++tr;
      }
      pend = false;
    }
    if (SAMPLE && active) {
      while (t_rec < t) {
        const size_t r = (size_t)(k - (uint32_t)v.bin0) * E + e;
        v.rec_dx[r] = dx;
        v.rec_dy[r] = dy;
        v.rec_dz[r] = dz;
        v.rec_drain[r] = nd;
        ++k;
        t_rec = k < (uint32_t)v.n_bins ? (double)(k + 1u) * v.dt_sample : INFINITY;
      }
    }
    // (C) the hop of (A) happens at t if t <= t_max; otherwise the exciton stops here
    const bool hop = can && t <= t_max;
    if (active && !hop) {
      active = false;
      site_now = row.site;
      live = !trapped(row.n, row.gamma) && t <= t_max;
    }
    if (hop) {
      if (COUNT) count_line = W & TGT_LINE_MASK;
      t_hop = t;
      if (tg & EDGE_DRAIN_BIT) {
        // drain: the dwell drawn at the drain site, then re-injection keeping displacement and
        // dwell (kubo_step, src/monte_carlo/monte_carlo.cpp:330-336)
        const uint4* bd = v.blocks + (size_t)ROW_UNITS * (tg & TGT_LINE_MASK);
        const uint4 hd = bd[0], pd = bd[7];
        dx = dx + min_image((double)__uint_as_float(pd.x) - (double)row.x, v.half_per[0]);
        dy = dy + min_image((double)__uint_as_float(pd.y) - (double)row.y, v.half_per[1]);
        dz = dz + min_image((double)__uint_as_float(pd.z) - (double)row.z, v.half_per[2]);
        hs = (hs ^ (uint64_t)hd.y) * FNV_PRIME;
        t = t + dwell_time(hd.x, __hiloint2double((int)hd.w, (int)hd.z), r_next);
        ++nd;
        const uint32_t src_site = v.blocks[(size_t)ROW_UNITS * (W_next & TGT_LINE_MASK)].y;
        hs = (hs ^ ((uint64_t)src_site | (1ull << 32))) * FNV_PRIME;
        if (TRACE && e < tv.n_trace && tr < tv.tr_max) {
          tv.tr_site[(size_t)e * tv.tr_max + tr] = (int32_t)src_site;
          tv.tr_time[(size_t)e * tv.tr_max + tr] = t_hop;
          ++tr;
        }
      } else {
        pend = true;
        r_dw = r_next;
        px = row.x;
        py = row.y;
        pz = row.z;
      }
      W = W_next;
      ++h;
    }
    if (__ballot(active) == 0) break;
  }
  // the last issued load lands in this wave's LDS rows: wait for it before the wave moves on
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  if (COUNT && count_line != COUNT_NONE) count_departure();
  if (COUNT) cc.flush(cbase);
  if (v.live) {
    // excitons that a further launch could move (the budget aside)
    const uint64_t b = __ballot(valid && live);
    if (lane == 0u && b != 0ull) atomicAdd(v.live, (uint32_t)__popcll(b));
  }
  if (!valid) return;
  v.x.site[e] = (int32_t)site_now;
  v.x.t_next[e] = t;
  v.x.hops[e] = h;
  v.x.dx[e] = dx;
  v.x.dy[e] = dy;
  v.x.dz[e] = dz;
  v.x.drains[e] = nd;
  v.x.hash[e] = hs;
  if (SAMPLE) v.x.k_rec[e] = k;
  if (TRACE && e < tv.n_trace) tv.tr_n[e] = tr;
}

// Position of this thread in the sorted exciton order.  Workgroups are dealt to the XCDs in turn,
// so workgroup b takes the b%8-th eighth of the order: each XCD's L2 serves excitons that sit on
// neighbouring blocks.  gridDim.x is a multiple of KMC_COUNT_COPIES.
__device__ __forceinline__ uint32_t xcd_position() {
  const uint32_t per_xcd = gridDim.x / KMC_COUNT_COPIES;
  const uint32_t chunk = (blockIdx.x % KMC_COUNT_COPIES) * per_xcd + blockIdx.x / KMC_COUNT_COPIES;
  return chunk * blockDim.x + threadIdx.x;
}

// One network: thread -> position -> exciton perm[position] (its id keys every draw, so the order
// changes nothing but locality).
template <bool COUNT, bool SAMPLE, bool TRACE>
__global__ __launch_bounds__(256, 4) void k_hop(hop_args a) {
  __shared__ uint4 rows[4 * 64 * ROW_UNITS];  // 8 KB of staged rows per wave
  __shared__ uint32_t win[COUNT_WINDOW_WORDS]; // 8 KB: the workgroup's departure-count window
  // the positions taken: all excitons, or the ones not idle in a clock-bounded launch (sorted
  // first), dealt in eighths to the XCDs like all of them
  uint32_t pos, n;
  if (a.n_alive) {
    n = *a.n_alive;
    const uint32_t per_xcd = ((n + 255u) / 256u + KMC_COUNT_COPIES - 1) / KMC_COUNT_COPIES;
    const uint32_t rank = blockIdx.x / KMC_COUNT_COPIES;
    pos = rank < per_xcd
              ? ((blockIdx.x % KMC_COUNT_COPIES) * per_xcd + rank) * blockDim.x + threadIdx.x
              : 0xFFFFFFFFu;
  } else {
    n = (uint32_t)a.v.x.n;
    pos = xcd_position();
  }
  const uint32_t lane = threadIdx.x & 63u;
  const bool wave_live = __builtin_amdgcn_readfirstlane(pos - lane) < n;
  const bool valid = pos < n;
  const uint32_t e = valid ? (a.perm ? a.perm[pos] : pos) : 0u;
  uint4* lds_wave = rows + 64u * ROW_UNITS * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if constexpr (!COUNT) {
    if (wave_live)
      hop_wave<COUNT, SAMPLE, TRACE>(a.v, e, valid, a.hop_budget, a.t_max, a.tv, lds_wave, lane,
                                     nullptr, 0u);
  } else {
    // The workgroup's excitons are consecutive in the order sorted by block line, and their
    // hops in one sub-launch stay mostly among nearby blocks: departures from the window's
    // lines around the first one's line are summed in LDS and flushed once, per line, at the end.
    for (uint32_t i = threadIdx.x; i < COUNT_WINDOW_WORDS; i += blockDim.x) win[i] = 0u;
    const uint32_t p0 = pos - threadIdx.x;
    uint32_t base = 0u;
    if (p0 < n) {
      const uint32_t e0 = a.perm ? a.perm[p0] : p0;
      const uint32_t L0 = (uint32_t)a.v.roff[a.v.x.site[e0]];
      base = L0 > COUNT_WINDOW_LINES / 2 ? L0 - COUNT_WINDOW_LINES / 2 : 0u;
    }
    __syncthreads();
    if (wave_live)
      hop_wave<COUNT, SAMPLE, TRACE>(a.v, e, valid, a.hop_budget, a.t_max, a.tv, lds_wave, lane,
                                     win, base);
    __syncthreads();
    window_flush(win, base, a.v.site_count + (size_t)xcd_id() * a.v.count_stride,
                 a.v.count_stride);
  }
}

// The same with drift along the tubes (v > 0).
template <bool COUNT, bool SAMPLE, bool TRACE>
__global__ __launch_bounds__(256) void k_hop_fly(hop_args a) {
  const uint32_t pos = xcd_position();
  bool live = false;
  if (pos < a.v.x.n) {
    const uint32_t e = a.perm ? a.perm[pos] : pos;
    live = hop_exciton<COUNT, SAMPLE, TRACE, true>(a.v, e, a.hop_budget, a.t_max, a.tv);
  }
  if (a.v.live) {
    const uint64_t b = __ballot(live);
    if ((threadIdx.x & 63u) == 0u && b != 0ull) atomicAdd(a.v.live, (uint32_t)__popcll(b));
  }
}

// Several networks in one launch.  Network k's sorted excitons occupy positions
// [pbase[k], pbase[k] + n_k) of a position space whose network ranges start on workgroup
// boundaries, so every workgroup belongs to one network and reads that network's view with
// scalar loads; sorted entries of network k sit at [ebase[k], ebase[k] + n_k) of perm, which
// holds ensemble-wide exciton numbers ebase[k] + local id.  Departures are counted as in k_hop:
// an LDS window over the block lines around the workgroup's first exciton, flushed into that
// network's counts.
template <bool COUNT, bool SAMPLE>
__global__ __launch_bounds__(256, 4) void k_hop_ens(ens_args a) {
  __shared__ uint4 rows[4 * 64 * ROW_UNITS];  // 8 KB of staged rows per wave
  __shared__ uint32_t win[COUNT_WINDOW_WORDS]; // 8 KB: the workgroup's departure-count window
  const uint32_t pos = xcd_position();
  const uint32_t p0 = pos - threadIdx.x;       // the workgroup's first position
  if (p0 >= a.n_positions) return;             // the whole workgroup is past the end
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t net = __builtin_amdgcn_readfirstlane(a.wave_net[p0 >> 6]);
  const net_view v = a.nets[net];
  const int64_t q = (int64_t)pos - a.pbase[net];
  const bool valid = q < v.x.n;
  const bool wave_live = __builtin_amdgcn_readfirstlane((uint32_t)(q - lane < v.x.n));
  const int64_t eb = a.ebase[net];
  const uint32_t e = valid ? (uint32_t)((int64_t)a.perm[eb + q] - eb) : 0u;
  uint4* lds_wave = rows + 64u * ROW_UNITS * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if constexpr (!COUNT) {
    if (wave_live)
      hop_wave<COUNT, SAMPLE, false>(v, e, valid, a.hop_budget, a.t_max, trace_view{}, lds_wave,
                                     lane, nullptr, 0u);
  } else {
    for (uint32_t i = threadIdx.x; i < COUNT_WINDOW_WORDS; i += blockDim.x) win[i] = 0u;
    const int64_t q0 = (int64_t)p0 - a.pbase[net];
    uint32_t base = 0u;
    if (q0 < v.x.n) {
      const uint32_t e0 = (uint32_t)((int64_t)a.perm[eb + q0] - eb);
      const uint32_t L0 = (uint32_t)v.roff[v.x.site[e0]];
      base = L0 > COUNT_WINDOW_LINES / 2 ? L0 - COUNT_WINDOW_LINES / 2 : 0u;
    }
    __syncthreads();
    if (wave_live)
      hop_wave<COUNT, SAMPLE, false>(v, e, valid, a.hop_budget, a.t_max, trace_view{}, lds_wave,
                                     lane, win, base);
    __syncthreads();
    window_flush(win, base, v.site_count + (size_t)xcd_id() * v.count_stride, v.count_stride);
  }
}

hipError_t launch_hop_ens(const ens_args& a, bool count, bool sample, hipStream_t st) {
  const uint64_t groups = (a.n_positions + 255) / 256;
  const dim3 grid((unsigned)((groups + KMC_COUNT_COPIES - 1) / KMC_COUNT_COPIES * KMC_COUNT_COPIES)),
      block(256);
  if (count && sample) hipLaunchKernelGGL((k_hop_ens<true, true>), grid, block, 0, st, a);
  else if (count) hipLaunchKernelGGL((k_hop_ens<true, false>), grid, block, 0, st, a);
  else if (sample) hipLaunchKernelGGL((k_hop_ens<false, true>), grid, block, 0, st, a);
  else hipLaunchKernelGGL((k_hop_ens<false, false>), grid, block, 0, st, a);
  return hipGetLastError();
}

// Sort keys of an ensemble's excitons: (network << 32) | first line of the current site's block.
__global__ __launch_bounds__(256) void k_ens_keys(int64_t n, const net_view* __restrict__ nets,
                                                  const uint16_t* __restrict__ exc_net,
                                                  const int64_t* __restrict__ ebase,
                                                  uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ ids,
                                                  uint32_t* __restrict__ live) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0 && live) *live = 0u;
  if (g >= n) return;
  const uint32_t k = exc_net[g];
  const net_view& v = nets[k];
  const int64_t e = g - ebase[k];
  keys[g] = ((uint64_t)k << 32) | (uint64_t)(uint32_t)v.roff[v.x.site[e]];
  ids[g] = (uint32_t)g;
}

hipError_t launch_ens_keys(int64_t n, const net_view* nets, const uint16_t* exc_net,
                           const int64_t* ebase, uint64_t* keys, uint32_t* ids, uint32_t* live,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ens_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, nets,
                     exc_net, ebase, keys, ids, live);
  return hipGetLastError();
}

hipError_t launch_hop(const hop_args& a, bool count, bool sample, bool trace, bool fly,
                      hipStream_t st) {
  // a multiple of KMC_COUNT_COPIES workgroups, so the per-XCD chunks tile the excitons exactly
  const uint64_t groups = (a.v.x.n + 255) / 256;
  const dim3 grid((unsigned)((groups + KMC_COUNT_COPIES - 1) / KMC_COUNT_COPIES * KMC_COUNT_COPIES)),
      block(256);
#define KMC_HOP_CASE(C, S, T)                                            \
  if (count == C && sample == S && trace == T) {                         \
    if (fly) hipLaunchKernelGGL((k_hop_fly<C, S, T>), grid, block, 0, st, a); \
    else hipLaunchKernelGGL((k_hop<C, S, T>), grid, block, 0, st, a);    \
    return hipGetLastError();                                            \
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

// The per-XCD copies of the departure counts folded into the 64-bit per-site totals and zeroed.
// Thread q owns block line q (slots and lines correspond one to one) in all copies; only a
// block's first line is ever counted, and its header gives the site.
__global__ __launch_bounds__(256) void k_fold_counts(int64_t n_lines, int64_t stride,
                                                     const uint4* __restrict__ blocks,
                                                     uint32_t* __restrict__ copies,
                                                     uint64_t* __restrict__ totals) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_lines) return;
  uint64_t s = 0;
  const uint32_t slot = count_slot((uint32_t)q, stride);
#pragma unroll
  for (int x = 0; x < KMC_COUNT_COPIES; ++x) {
    uint32_t* c = copies + (size_t)x * stride + slot;
    const uint32_t v = *c;
    if (v) {
      s += v;
      *c = 0u;
    }
  }
  if (s) totals[blocks[(size_t)ROW_UNITS * q].y] += s;
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
    if ((uint32_t)(a.bin0 + k) < a.k_rec[e]) {
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
// =============================================================================================
// Contact (steady-state) mode: monte_carlo::step, save_population_profile,
// save_currents and repopulate_contacts (src/monte_carlo/monte_carlo.h:343-355, 443-516, 525-630).
// A particle's draws are keyed by its id (unique over the run) and its own hop count, so slots
// can be compacted and reused without any draw repeating.
// =============================================================================================
__device__ __forceinline__ int contact_section(double y, double ymin, double dy, int n_seg) {
#pragma clang fp contract(off)
  // save_population_profile (monte_carlo.h:567-568): truncation toward zero, then clamped
  int i = (int)((y - ymin) / dy);
  return i < 0 ? 0 : (i < n_seg ? i : n_seg - 1);
}

// The population and crossing counts of one particle's step (save_population_profile,
// save_currents: monte_carlo.h:565-571, 598-626) into the workgroup's LDS counters.  Interface k
// lies at ymin + dy (k + 1); +1 for a crossing upward (old < y_k <= new), -1 downward.
// Interfaces more than one section away from both positions cannot lie between them, so only
// those near the two sections are compared.
__device__ __forceinline__ void contact_tally(const contact_args& a, double y_old, double y_new,
                                              int32_t* s_pop, int32_t* s_cur) {
#pragma clang fp contract(off)
  const int i_new = contact_section(y_new, a.ymin, a.dy, a.n_seg);
  atomicAdd(&s_pop[i_new], 1);
  if (y_old != y_new) {
    const int i_old = contact_section(y_old, a.ymin, a.dy, a.n_seg);
    const int k_lo = max(0, min(i_old, i_new) - 1);
    const int k_hi = min(a.n_seg - 2, max(i_old, i_new) + 1);
    for (int k = k_lo; k <= k_hi; ++k) {
      const double yk = a.ymin + a.dy * (double)(k + 1);
      if (y_old < yk && y_new >= yk) atomicAdd(&s_cur[k], 1);
      else if (y_old >= yk && y_new < yk) atomicAdd(&s_cur[k], -1);
    }
  }
}

// particle::step (src/monte_carlo/particle.cpp:57-80) over one step with drift (v > 0), one
// thread per particle: the flight up to each due hop, the hop from the site reached, then the
// flight to the step's end, which is taken (particle::fly(dt), particle.cpp:77) and leaves the
// particle between sites.
__global__ __launch_bounds__(256) void k_contact_fly(contact_args a) {
#pragma clang fp contract(off)
  __shared__ int32_t s_pop[CONTACT_MAX_SEG], s_cur[CONTACT_MAX_SEG];
  __shared__ unsigned long long s_made;
  for (int i = threadIdx.x; i < a.n_seg; i += blockDim.x) s_pop[i] = s_cur[i] = 0;
  if (threadIdx.x == 0) s_made = 0ull;
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < *a.n_live) {
    // slots in the order of their sites' rows: neighbouring lanes read neighbouring rows
    const int64_t e = a.perm ? (int64_t)a.perm[j] : j;
    const uint32_t id = a.s.id[e];
    double t = a.s.t_next[e];
    uint64_t h = a.s.hops[e];
    const uint32_t L0 = (uint32_t)a.roff[a.s.site[e]];
    uint32_t R = row_word(L0, a.blocks[(size_t)ROW_UNITS * L0].x);
    row_regs row = load_row(a.blocks, R);
    const double y_old = a.s.y[e];
    double tl = a.s.t_last[e];
    flight_state g{R, row.site, row.x, row.y, row.z, a.s.s_off[e], (uint32_t)a.s.heading[e],
                   0.0, 0.0, 0.0, 0u, 0ull};
    uint64_t made = 0;
    hop_words hw;
    while (t <= a.t_end) {
      if (!fly(a.fv, g, t - tl, id, h)) {
        atomicOr(a.err, CONTACT_ERR_FLY);
        break;
      }
      if (g.R != R) {
        R = g.R;
        row = load_row(a.blocks, R);
      }
      if (trapped(row.n, row.gamma)) {
        // no rates: the particle stays on its tube and flies on (scatterer.cpp:14-15)
        tl = t;
        t = INFINITY;
        ++h;
        break;
      }
      if (made == a.max_hops) {
        atomicOr(a.err, CONTACT_ERR_HOPS);
        break;
      }
      uint32_t w_sel, w_dw;
      hw.get(a.seed, id, h, w_sel, w_dw);
      const double dice = div_rand_max(row.gamma * (double)r31(w_sel));
      const uint32_t idx = select_entry(a.blocks, R, row.c, row.n, dice);
      R = target_of(a.blocks, R, row, idx) & ~EDGE_DRAIN_BIT;
      row = load_row(a.blocks, R);
      const uint32_t r = nonzero_draw1(w_dw, a.seed, id, h, STREAM_HOP_REDRAW);
      g = flight_state{R, row.site, row.x, row.y, row.z, 0.0, g.heading, 0.0, 0.0, 0.0, 0u, 0ull};
      tl = t;
      t = t + dwell_time(row.n, row.gamma, r);
      ++h;
      ++made;
    }
    if (!fly(a.fv, g, a.t_end - tl, id, h)) atomicOr(a.err, CONTACT_ERR_FLY);
    tl = a.t_end;
    double ox, oy, oz;
    fly_offset(a.fv, g, ox, oy, oz);
    const double y_new = a.py[g.site] + oy;
    a.s.site[e] = (int32_t)g.site;
    a.s.heading[e] = (uint8_t)g.heading;
    a.s.s_off[e] = g.s;
    a.s.t_last[e] = tl;
    a.s.t_next[e] = t;
    a.s.hops[e] = h;
    a.s.y[e] = y_new;
    if (made) atomicAdd(&s_made, (unsigned long long)made);
    contact_tally(a, y_old, y_new, s_pop, s_cur);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.n_seg; i += blockDim.x) {
    if (s_pop[i]) atomicAdd(a.pop + i, s_pop[i]);
    if (i + 1 < a.n_seg && s_cur[i]) atomicAdd(a.cur + i, s_cur[i]);
  }
  if (threadIdx.x == 0 && s_made) atomicAdd(a.hops_made, s_made);
}

__global__ __launch_bounds__(256) void k_contact_keys(int64_t capacity, const int64_t* n_live,
                                                      const int32_t* __restrict__ site,
                                                      const int64_t* __restrict__ roff,
                                                      const uint8_t* __restrict__ chir,
                                                      contact_key_layout kl,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ ids) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= capacity) return;
  uint32_t k = kl.dead;
  if (e < *n_live) {
    const int32_t s = site[e];
    const uint32_t row = (uint32_t)roff[s];
    k = row;
    if (kl.regions) {
      const uint32_t region = (uint32_t)((uint64_t)row * KMC_COUNT_COPIES / kl.lines);
      k |= (region << (kl.chir_bits + kl.row_bits)) | ((uint32_t)chir[s] << kl.row_bits);
    }
  }
  keys[e] = k;
  ids[e] = (uint32_t)e;
}

// particle::step of every live particle for v = 0, built like the Kubo kernel's wave loop around
// the cooperative row load: every lane of a wave takes part in every row load, and the wave
// loops until no lane has a hop due by the step's end (a lane whose particle has stopped keeps
// taking part in the loads).  The next hop is selected as soon as a row arrives and its row load issued
// before the dwell at the current site is drawn and the clock tested, so the load is in flight
// meanwhile.  Each particle's arithmetic and its order are those of the oracle's step.  The live
// slots, sorted by region, chirality and row, are dealt to the XCDs in eighths (from the
// device-side live count), so each XCD's L2 serves one compact region of the film.  The grid is
// a multiple of 8 workgroups covering the capacity.
__global__ __launch_bounds__(256, 4) void k_contact_wave(contact_args a) {
#pragma clang fp contract(off)
  __shared__ uint4 rows[4 * 64 * ROW_UNITS];  // 8 KB of staged rows per wave
  __shared__ unsigned long long s_made;
  // the section counters are sized by the launch (2 n_seg words), so that four workgroups fit
  // a CU's LDS beside the staged rows
  extern __shared__ int32_t s_counts[];
  int32_t* s_pop = s_counts;
  int32_t* s_cur = s_counts + a.n_seg;
  for (int i = threadIdx.x; i < a.n_seg; i += blockDim.x) s_pop[i] = s_cur[i] = 0;
  if (threadIdx.x == 0) s_made = 0ull;
  __syncthreads();
  const int64_t n_live = *a.n_live;
  const uint32_t groups = (uint32_t)((n_live + 255) / 256);
  const uint32_t per_xcd = (groups + KMC_COUNT_COPIES - 1) / KMC_COUNT_COPIES;
  const uint32_t b = blockIdx.x;
  const int64_t j = (b / KMC_COUNT_COPIES < per_xcd)
                        ? ((int64_t)((b % KMC_COUNT_COPIES) * per_xcd + b / KMC_COUNT_COPIES) * 256 +
                           threadIdx.x)
                        : INT64_MAX;
  const bool valid = j < n_live;
  const uint32_t lane = threadIdx.x & 63u;
  if (__ballot(valid) != 0) {
    uint4* lds_wave = rows + 64u * ROW_UNITS * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t e = valid ? (a.perm ? (int64_t)a.perm[j] : j) : 0;
    uint32_t id = 0, W = 0, r_dw = 0;
    double t = INFINITY, y_old = 0.0;
    uint64_t h = 0, made = 0;
    if (valid) {
      id = a.s.id[e];
      t = a.s.t_next[e];
      h = a.s.hops[e];
      y_old = a.s.y[e];
      const uint32_t L0 = (uint32_t)a.roff[a.s.site[e]];
      W = row_word(L0, a.blocks[(size_t)ROW_UNITS * L0].x);
    }
    bool active = valid, pend = false;
    uint32_t site_now = 0;  // when stopped: the site it stopped at (the row loaded after that
                            // may be the one of the hop selected but not made)
    hop_words hw;
    coop_issue(a.blocks, W, lds_wave, lane);
    while (true) {
      __builtin_amdgcn_s_setprio(2);  // raised until the next row load is out, as in hop_wave
      const coop_row row = coop_finish(lds_wave, lane);
      // (A) the hop from this site, selected before the clock test that decides whether it is
      // made, and the load of the row it leads to, in flight while (B) and (C) run
      const bool can = active && !trapped(row.n, row.gamma) && made < a.max_hops;
      uint32_t W_next = W, r_next = 0;
      if (can) {
        uint32_t w_sel, w_dw;
        hw.get(a.seed, id, h, w_sel, w_dw);
        const double dice = div_rand_max(row.gamma * (double)r31(w_sel));
        const uint32_t idx = select_entry(a.blocks, W, row.c, row.n, dice);
        const uint32_t tg = idx < ROW_INLINE ? target_lds(lds_wave, lane, idx)
                                             : target_wide(a.blocks, W, row.n, idx);
        W_next = tg & ~EDGE_DRAIN_BIT;
        r_next = nonzero_draw1(w_dw, a.seed, id, h, STREAM_HOP_REDRAW);
      }
      // a lane whose particle has stopped never reads its row again: it loads block line 0,
      // which every such lane of the wave shares, instead of its own row
      coop_issue(a.blocks, active ? W_next : 0u, lds_wave, lane);
      __builtin_amdgcn_s_setprio(0);
      // (B) the dwell at this site of the hop that led here
      if (pend) {
        t = t + dwell_time(row.n, row.gamma, r_dw);
        pend = false;
      }
      // (C) the hop of (A) is made if it falls due by the step's end
      if (active) {
        if (!(t <= a.t_end) || trapped(row.n, row.gamma)) {
          active = false;
        } else if (made == a.max_hops) {
          atomicOr(a.err, CONTACT_ERR_HOPS);
          active = false;
        }
        if (!active) site_now = row.site;
      }
      if (active) {
        W = W_next;
        r_dw = r_next;
        pend = true;
        ++h;
        ++made;
      }
      if (__ballot(active) == 0) break;
    }
    // the last issued load lands in this wave's LDS rows: wait for it before the wave moves on
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (valid) {
      const double y_new = a.py[site_now];
      a.s.site[e] = (int32_t)site_now;
      a.s.t_next[e] = t;
      a.s.hops[e] = h;
      a.s.y[e] = y_new;
      if (made) atomicAdd(&s_made, (unsigned long long)made);
      contact_tally(a, y_old, y_new, s_pop, s_cur);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.n_seg; i += blockDim.x) {
    if (s_pop[i]) atomicAdd(a.pop + i, s_pop[i]);
    if (i + 1 < a.n_seg && s_cur[i]) atomicAdd(a.cur + i, s_cur[i]);
  }
  if (threadIdx.x == 0 && s_made) atomicAdd(a.hops_made, s_made);
}

// repopulate (monte_carlo.h:461-474): a particle whose y lies in [y_lo, y_hi] leaves
__global__ __launch_bounds__(256) void k_contact_keep(contact_args a, int64_t n, double y_lo,
                                                      double y_hi, uint32_t* keep) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint32_t k = 0;
  if (e < *a.n_live) {
    const double y = a.s.y[e];
    k = (y >= y_lo && y <= y_hi) ? 0u : 1u;
  }
  keep[e] = k;
}

// The kept particles, in slot order, to the front of the other buffer.
__global__ __launch_bounds__(256) void k_contact_compact(int64_t capacity, contact_slots src,
                                                         contact_slots dst,
                                                         const uint32_t* __restrict__ keep,
                                                         const int64_t* __restrict__ off) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= capacity || !keep[e]) return;
  const int64_t o = off[e];
  dst.site[o] = src.site[e];
  dst.t_next[o] = src.t_next[e];
  dst.hops[o] = src.hops[e];
  dst.id[o] = src.id[e];
  dst.y[o] = src.y[e];
  dst.heading[o] = src.heading[e];
  dst.s_off[o] = src.s_off[e];
  dst.t_last[o] = src.t_last[e];
}

// `count` new particles in slots *base + k on sites of `sites`, as the particle constructor
// makes them (particle.h:49-52): a random site of the list (rand() % size), the first dwell
// from the site (ff_time), then the heading (rand() % 2).  Ids continue from *next_id.
__global__ __launch_bounds__(256) void k_contact_inject(contact_args a, const int64_t* base,
                                                        int64_t count,
                                                        const int32_t* __restrict__ sites,
                                                        int64_t n_sites,
                                                        const uint64_t* __restrict__ next_id,
                                                        double t_now) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int64_t slot = *base + k;
  const uint64_t id64 = *next_id + (uint64_t)k;
  if (slot >= a.capacity) {
    atomicOr(a.err, CONTACT_ERR_CAPACITY);
    return;
  }
  if (id64 > 0xFFFFFFFFull) {
    atomicOr(a.err, CONTACT_ERR_IDS);
    return;
  }
  const uint32_t id = (uint32_t)id64;
  const u32x4 w = draw4(a.seed, id, 0, STREAM_CONTACT);
  const int32_t site = sites[r31(w.v[0]) % (uint32_t)n_sites];
  const uint4 hh = a.blocks[(size_t)ROW_UNITS * a.roff[site]];
  const uint32_t r = nonzero_draw(w.v[1], w.v[3], a.seed, id, 0, STREAM_CONTACT_REDRAW);
  a.s.site[slot] = (int32_t)hh.y;
  a.s.t_next[slot] = t_now + dwell_time(hh.x, __hiloint2double((int)hh.w, (int)hh.z), r);
  a.s.hops[slot] = 0;
  a.s.id[slot] = id;
  a.s.y[slot] = a.py[site];
  a.s.heading[slot] = (uint8_t)(r31(w.v[2]) & 1u);
  a.s.s_off[slot] = 0.0;
  a.s.t_last[slot] = t_now;
}

__global__ void k_contact_finish(const int64_t* base, int64_t count, int64_t capacity,
                                 uint64_t* next_id, int64_t* n_live) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t n = *base + count;
  *n_live = n < capacity ? n : capacity;
  *next_id += (uint64_t)count;
}

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

hipError_t launch_inject(const exciton_arrays& x, const uint4* blocks, const uint32_t* src_roff,
                         uint32_t n_src, uint64_t seed, hipStream_t st) {
  if (x.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_inject, grid_for(x.n), dim3(256), 0, st, x, blocks, src_roff, n_src, seed);
  return hipGetLastError();
}

hipError_t launch_block_sizes(int64_t n, const site_hdr* hdr, const int32_t* order, uint32_t* lines,
                              hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_block_sizes, grid_for(n), dim3(256), 0, st, n, hdr, order, lines);
  return hipGetLastError();
}

hipError_t launch_place_blocks(int64_t n, const int32_t* order, const int64_t* first, int64_t* roff,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_place_blocks, grid_for(n + 1), dim3(256), 0, st, n, order, first, roff);
  return hipGetLastError();
}

hipError_t launch_exciton_keys(int64_t n, const int32_t* site, const int64_t* roff,
                               const idle_test& it, uint32_t* keys, uint32_t* ids, uint32_t* live,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_exciton_keys, grid_for(n), dim3(256), 0, st, n, site, roff, it, keys, ids,
                     live);
  return hipGetLastError();
}

hipError_t launch_fill_blocks(int64_t n, const site_hdr* hdr, const double* cum,
                              const edge_rec* edge, const int64_t* roff, const double* px,
                              const double* py, const double* pz, uint4* blocks, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_rows, grid_for(n), dim3(256), 0, st, n, hdr, cum, edge, roff, px, py,
                     pz, blocks);
  return hipGetLastError();
}

// Chirality class (donor * n_chir + acceptor) of every CSR entry, and per-class index lists.
// Class counts are summed per workgroup in LDS first (a film of one chirality puts every pair
// in one class), then one device atomic per class and workgroup.
constexpr int MAX_PAIR_CLASSES = 256;

__device__ __forceinline__ void clear_class_counts(uint32_t* lds) {
  for (int c = threadIdx.x; c < MAX_PAIR_CLASSES; c += blockDim.x) lds[c] = 0u;
  __syncthreads();
}

__device__ __forceinline__ void flush_class_counts(const uint32_t* lds, uint32_t* __restrict__ counts,
                                                   int n_classes) {
  __syncthreads();
  for (int c = threadIdx.x; c < n_classes; c += blockDim.x)
    if (lds[c]) atomicAdd(&counts[c], lds[c]);
}

__global__ __launch_bounds__(256) void k_pair_class(int64_t n, const site_hdr* __restrict__ hdr,
                                                    const edge_rec* __restrict__ edge,
                                                    const uint8_t* __restrict__ chir, int32_t n_chir,
                                                    uint8_t* __restrict__ cls,
                                                    uint32_t* __restrict__ counts) {
  __shared__ uint32_t lds[MAX_PAIR_CLASSES];
  clear_class_counts(lds);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const site_hdr h = hdr[i];
    const int ci = chir[i];
    for (uint32_t k = 0; k < h.n; ++k) {
      const int cj = chir[edge[h.rs + k].col & EDGE_COL_MASK];
      const int c = ci * n_chir + cj;
      cls[h.rs + k] = (uint8_t)c;
      atomicAdd(&lds[c], 1u);
    }
  }
  flush_class_counts(lds, counts, n_chir * n_chir);
}

// entries with mask[p] == 0 (mask optional) are left out
// Slots within a class: a workgroup first numbers its own entries per class in LDS, then
// reserves one range per class with a single device atomic.
__global__ __launch_bounds__(256) void k_pair_scatter(int64_t nnz, const uint8_t* __restrict__ cls,
                                                      const uint8_t* __restrict__ mask,
                                                      const int64_t* __restrict__ starts,
                                                      uint32_t* __restrict__ fill,
                                                      int64_t* __restrict__ sel) {
  __shared__ uint32_t local[MAX_PAIR_CLASSES];
  __shared__ uint32_t base[MAX_PAIR_CLASSES];
  clear_class_counts(local);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool take = p < nnz && (!mask || mask[p]);
  const int c = take ? (int)cls[p] : 0;
  const uint32_t slot = take ? atomicAdd(&local[c], 1u) : 0u;
  __syncthreads();
  for (int q = threadIdx.x; q < MAX_PAIR_CLASSES; q += blockDim.x)
    base[q] = local[q] ? atomicAdd(&fill[q], local[q]) : 0u;
  __syncthreads();
  if (take) sel[starts[c] + base[c] + slot] = p;
}

hipError_t launch_pair_classes(int64_t n, int64_t nnz, const site_hdr* hdr, const edge_rec* edge,
                               const uint8_t* chir, int32_t n_chir, uint8_t* cls, uint32_t* counts,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_class, grid_for(n), dim3(256), 0, st, n, hdr, edge, chir, n_chir, cls,
                     counts);
  (void)nnz;
  return hipGetLastError();
}

hipError_t launch_pair_scatter(int64_t nnz, const uint8_t* cls, const uint8_t* mask,
                               const int64_t* starts, uint32_t* fill, int64_t* sel, hipStream_t st) {
  if (nnz == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_scatter, grid_for(nnz), dim3(256), 0, st, nnz, cls, mask, starts, fill, sel);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_mask_count(int64_t nnz, const uint8_t* __restrict__ cls,
                                                    const uint8_t* __restrict__ mask,
                                                    uint32_t* __restrict__ counts) {
  __shared__ uint32_t lds[MAX_PAIR_CLASSES];
  clear_class_counts(lds);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz && mask[p]) atomicAdd(&lds[cls[p]], 1u);
  flush_class_counts(lds, counts, MAX_PAIR_CLASSES);
}

hipError_t launch_mask_count(int64_t nnz, const uint8_t* cls, const uint8_t* mask, uint32_t* counts,
                             hipStream_t st) {
  if (nnz == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mask_count, grid_for(nnz), dim3(256), 0, st, nnz, cls, mask, counts);
  return hipGetLastError();
}

hipError_t launch_links(int64_t n, const int32_t* left, const int32_t* right, const uint8_t* flags,
                        const int64_t* roff, const site_hdr* hdr, uint2* links, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_links, grid_for(n), dim3(256), 0, st, n, left, right, flags, roff, hdr,
                     links);
  return hipGetLastError();
}

hipError_t launch_morton_keys(int64_t n, const int32_t* cell_of, int32_t nx, int32_t ny,
                              uint64_t* keys, int32_t* ids, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_keys, grid_for(n), dim3(256), 0, st, n, cell_of, nx, ny, keys, ids);
  return hipGetLastError();
}

hipError_t launch_gather_offsets(int64_t n, const int32_t* idx, const int64_t* roff,
                                 const site_hdr* hdr, uint32_t* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_offsets, grid_for(n), dim3(256), 0, st, n, idx, roff, hdr, out);
  return hipGetLastError();
}

hipError_t launch_fold_counts(int64_t n_lines, int64_t stride, const uint4* blocks,
                              uint32_t* copies, uint64_t* totals, hipStream_t st) {
  if (n_lines == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fold_counts, grid_for(n_lines), dim3(256), 0, st, n_lines, stride, blocks,
                     copies, totals);
  return hipGetLastError();
}

hipError_t launch_contact_step(const contact_args& a, bool fly, int64_t n_max, hipStream_t st) {
  if (n_max == 0) return hipSuccess;
  if (a.n_seg < 1 || a.n_seg > CONTACT_MAX_SEG || n_max < 0 || n_max > a.capacity)
    return hipErrorInvalidValue;
  if (fly)
    hipLaunchKernelGGL(k_contact_fly, grid_for(n_max), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_contact_wave,
                       dim3((unsigned)((n_max + 256 * KMC_COUNT_COPIES - 1) /
                                       (256 * KMC_COUNT_COPIES) * KMC_COUNT_COPIES)),
                       dim3(256), (size_t)2 * a.n_seg * sizeof(int32_t), st, a);
  return hipGetLastError();
}

hipError_t launch_contact_keys(int64_t capacity, const int64_t* n_live, const int32_t* site,
                               const int64_t* roff, const uint8_t* chir,
                               const contact_key_layout& kl, uint32_t* keys, uint32_t* ids,
                               hipStream_t st) {
  if (capacity == 0) return hipSuccess;
  hipLaunchKernelGGL(k_contact_keys, grid_for(capacity), dim3(256), 0, st, capacity, n_live, site,
                     roff, chir, kl, keys, ids);
  return hipGetLastError();
}

hipError_t launch_contact_keep(const contact_args& a, int64_t n, double y_lo, double y_hi,
                               uint32_t* keep, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (n < 0 || n > a.capacity) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_contact_keep, grid_for(n), dim3(256), 0, st, a, n, y_lo, y_hi, keep);
  return hipGetLastError();
}

hipError_t launch_contact_compact(int64_t capacity, const contact_slots& src,
                                  const contact_slots& dst, const uint32_t* keep,
                                  const int64_t* off, hipStream_t st) {
  if (capacity == 0) return hipSuccess;
  hipLaunchKernelGGL(k_contact_compact, grid_for(capacity), dim3(256), 0, st, capacity, src, dst,
                     keep, off);
  return hipGetLastError();
}

hipError_t launch_contact_inject(const contact_args& a, const int64_t* base, int64_t count,
                                 const int32_t* sites, int64_t n_sites, uint64_t* next_id,
                                 double t_now, int64_t* n_live, hipStream_t st) {
  if (count > 0) {
    if (n_sites < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_contact_inject, grid_for(count), dim3(256), 0, st, a, base, count, sites,
                       n_sites, (const uint64_t*)next_id, t_now);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_contact_finish, dim3(1), dim3(64), 0, st, base, count, a.capacity, next_id,
                     n_live);
  return hipGetLastError();
}

}  // namespace cntkmc
