// kmc_api.hip -- extern "C" entry points of libcntkmc.so (declared in include/cntkmc.h).
//
// One kmc_ctx owns one HIP stream, two timing events and every device buffer of one KMC
// problem: sites, cell list, rate tables, CSR, excitons, MSD records and trace.  Inputs are
// caller-owned host arrays copied in; outputs go to caller-provided host arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cntkmc.h"
#include "kmc_internal.h"
#include "kmc_kernels.h"
#include "kmc_mlp.h"

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_reduce.hpp>
#include <mutex>
#include "philox.h"

using namespace cntkmc;

// Scratch of the allocation-free MSD reduction (kmc_reduce_samples_noalloc): the partial sums
// of one call, n_bins * n_batches * 5 doubles at most.  Every device holds its own copy until
// the process ends; host code uses only the address hipGetSymbolAddress returns for the
// device (the array's name is a host-side stand-in, not a device address).
constexpr size_t REDUCE_SCRATCH_DOUBLES = 32768;  // 256 KB per device
static __device__ __align__(256) double scratch[REDUCE_SCRATCH_DOUBLES];

namespace {
thread_local std::string g_last_error;

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Once per device, with that device current (kmc_create): the device address of `scratch`.
hipError_t scratch_setup(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return hipSuccess;  // the no-alloc entry refuses
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[dev] = static_cast<double*>(p);
  return e;
}

double* scratch_of(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}
}  // namespace

namespace cntkmc {
// The message kmc_last_error(NULL) returns, for entry points without a context.
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace cntkmc

namespace {

template <typename T>
struct dbuf {
  T* p = nullptr;
  size_t n = 0;
};
}  // namespace

struct kmc_ctx {
  kmc_config cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  // sites
  int64_t n_sites = 0;
  std::vector<double> h_pos[3], h_ori[3];
  std::vector<uint8_t> h_chir, h_flags;
  std::vector<int32_t> h_left, h_right, h_src;
  double bb_lo[3] = {0, 0, 0}, bb_hi[3] = {0, 0, 0};
  dbuf<double> d_pos[3], d_ori[3];
  dbuf<uint8_t> d_chir, d_flags;
  dbuf<int32_t> d_left, d_right, d_src;

  // cells
  bool cells_ok = false;
  cell_desc cells{};
  int64_t n_cells = 0;
  dbuf<int32_t> d_cell_of, d_cell_sites;
  dbuf<uint32_t> d_cell_count;
  dbuf<int64_t> d_cell_start, d_scan_tmp;

  // rate tables
  bool table_ok = false;
  int32_t tab_n[5] = {0, 0, 0, 0, 0};  // n_chir, n_th, n_z, n_a1, n_a2
  dbuf<double> d_tab_rates, d_tab_grid;

  // CSR
  int csr_state = 0;  // 0 none, 1 waiting for pair rates, 2 ready
  int64_t nnz = 0;
  dbuf<uint32_t> d_deg;
  dbuf<int64_t> d_row_ptr;
  dbuf<site_hdr> d_hdr;
  dbuf<double> d_cum, d_pair_rates;
  dbuf<edge_rec> d_edge;
  dbuf<double> d_geo[4];
  dbuf<int32_t> d_err;
  // row blocks read by the hop kernel
  int64_t n_units = 0;
  dbuf<uint32_t> d_units, d_src_roff;
  dbuf<int64_t> d_roff;
  dbuf<uint4> d_blocks;
  dbuf<int64_t> d_first;         // block offsets in placement order (scan output)
  dbuf<uint64_t> d_pkeys;        // block placement: Morton keys (2 x n_sites, in and out)
  dbuf<int32_t> d_place;         // block placement: site ids (2 x n_sites, in and out)
  dbuf<uint32_t> d_keys, d_ids;  // exciton sort: 2 x n_exc each (input and output halves)
  dbuf<uint8_t> d_sort_tmp;
  size_t sort_tmp_bytes = 0;
  int64_t resort_hops = 625;     // hops per sub-launch between exciton sorts (0: one launch)
  // per-chirality-class pair lists for externally evaluated rates
  dbuf<uint8_t> d_cls;
  dbuf<uint32_t> d_cls_count;
  dbuf<int64_t> d_cls_start, d_sel;
  dbuf<uint8_t> d_mlp_flag;  // surrogate flags of the last surrogate build
  // drift along the tubes (particle::fly): velocity [m/s], links as block lines, error word
  double velocity = 0.0;      // set by kmc_set_velocity
  double exc_velocity = 0.0;  // the velocity of the injected excitons (fixed by kmc_inject)
  dbuf<uint2> d_links;
  dbuf<int32_t> d_fly_err;

  // excitons
  int64_t n_exc = 0;
  dbuf<int32_t> x_site;
  dbuf<double> x_t, x_dx, x_dy, x_dz;
  dbuf<uint64_t> x_hops, x_hash;
  dbuf<uint32_t> x_drains, x_krec;
  dbuf<uint8_t> x_heading;
  dbuf<double> x_soff, x_tlast;
  dbuf<uint32_t> d_site_count;  // KMC_COUNT_COPIES copies, count_stride apart, per block line
  dbuf<uint64_t> d_counts64;    // running 64-bit totals the copies are folded into
  int64_t count_stride = 0;
  uint64_t count_pending = 0;   // bound on any copy's count since the last fold
  dbuf<uint32_t> d_live;        // [0] excitons able to hop after the last launch; [1], [2] the
                                // excitons not idle at a launch's start, used alternately
  int64_t run_launches = 0;     // hop launches of the last kmc_run
  // hop-kernel timing of a timed kmc_run: one (start, stop) event pair per launch
  std::vector<hipEvent_t> hop_ev;
  bool time_hops = false;
  double hop_ms = -1.0;         // summed hop-kernel time of the last timed kmc_run

  // MSD records: the grid of n_bins sample times, recorded win_bins at a time from bin0
  int32_t n_bins = 0;
  int32_t bin0 = 0, win_bins = 0;
  double t_sample_max = 0.0;
  dbuf<uint32_t> d_min;         // smallest next-sample index, when a window is moved on
  dbuf<double> r_dx, r_dy, r_dz, d_reduced;
  dbuf<uint32_t> r_drain;

  // trace
  int64_t n_trace = 0, tr_max = 0;
  dbuf<int32_t> tr_site;
  dbuf<double> tr_time;
  dbuf<int64_t> tr_n;

  // contact (steady-state) mode
  bool c_on = false;
  int32_t c_nseg = 0;
  int64_t c_pop[2] = {0, 0}, c_cap = 0;
  double c_ymin = 0.0, c_dy = 0.0, c_time = 0.0, c_vel = 0.0;
  double c_rm_lo[2] = {0, 0}, c_rm_hi[2] = {0, 0};  // repopulate ranges of the two contacts
  std::vector<double> c_area;
  int64_t c_list_off[2] = {0, 0}, c_list_n[2] = {0, 0};
  dbuf<int32_t> c_sites;                  // the two contacts' site lists, then the sections'
  dbuf<int32_t> c_site[2];
  dbuf<double> c_t[2];
  dbuf<uint64_t> c_hops[2];
  dbuf<uint32_t> c_id[2];
  dbuf<double> c_y[2], c_soff[2], c_tlast[2];
  dbuf<uint8_t> c_head[2];
  int c_buf = 0;                          // the buffer holding the live particles
  dbuf<uint32_t> c_keep;
  dbuf<int64_t> c_off, c_scan_tmp, c_live, c_live_out;
  dbuf<uint64_t> c_next;
  dbuf<int32_t> c_err, c_pop_out, c_cur_out;
  dbuf<uint64_t> c_made;                  // hops made since kmc_contact_init
  dbuf<uint32_t> c_skeys, c_sids;         // [2 capacity] slot sort (contact_key_layout_of): keys, ids
  dbuf<uint8_t> c_sort_tmp;
  double c_step_ms = 0.0;                 // device time of the step kernels since then
  std::vector<hipEvent_t> c_ev;           // CONTACT_EV_PAIRS (start, stop) pairs, reused
  std::vector<char> c_ev_busy;            // pair recorded and not yet read
  // live counts after recent steps, copied back to pinned memory (CONTACT_LIVE_LAG slots, each
  // with the event recorded after its copy)
  int64_t* c_live_host = nullptr;
  std::vector<hipEvent_t> c_live_ev;
};

// step-kernel timing: event pairs reused round robin; a pair is read before it is recorded
// again, so the host waits only on a step CONTACT_EV_PAIRS steps back
constexpr int CONTACT_EV_PAIRS = 64;
// the live count after step k - CONTACT_LIVE_LAG bounds the slots step k's sort, step and
// compaction cover; the host waits for it only that many steps back
constexpr int CONTACT_LIVE_LAG = 2;

namespace {

int fail(kmc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

int hip_fail(kmc_ctx* c, hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return fail(c, KMC_ERR_HIP, m);
}

#define KMC_HIP(c, call)                                    \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail((c), e_, #call);  \
  } while (0)

template <typename T>
void dfree(dbuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
}

template <typename T>
int dalloc(kmc_ctx* c, dbuf<T>& b, size_t n) {
  if (b.p && b.n == n) return KMC_OK;
  dfree(b);
  if (n == 0) return KMC_OK;
  hipError_t e = hipMalloc((void**)&b.p, n * sizeof(T));
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, KMC_ERR_ALLOC, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
  }
  b.n = n;
  return KMC_OK;
}

#define KMC_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ != KMC_OK) return s_; \
  } while (0)

site_arrays sites_of(kmc_ctx* c) {
  site_arrays s;
  s.n = c->n_sites;
  s.px = c->d_pos[0].p;
  s.py = c->d_pos[1].p;
  s.pz = c->d_pos[2].p;
  s.ox = c->d_ori[0].p;
  s.oy = c->d_ori[1].p;
  s.oz = c->d_ori[2].p;
  s.chir = c->d_chir.p;
  s.flags = c->d_flags.p;
  return s;
}

csr_arrays csr_of(kmc_ctx* c) {
  csr_arrays a;
  a.hdr = c->d_hdr.p;
  a.cum = c->d_cum.p;
  a.edge = c->d_edge.p;
  return a;
}

exciton_arrays excitons_of(kmc_ctx* c) {
  exciton_arrays x;
  x.n = c->n_exc;
  x.site = c->x_site.p;
  x.t_next = c->x_t.p;
  x.hops = c->x_hops.p;
  x.dx = c->x_dx.p;
  x.dy = c->x_dy.p;
  x.dz = c->x_dz.p;
  x.drains = c->x_drains.p;
  x.hash = c->x_hash.p;
  x.k_rec = c->x_krec.p;
  x.heading = c->x_heading.p;
  x.s_off = c->x_soff.p;
  x.t_last = c->x_tlast.p;
  return x;
}

// The hop loop's view of one context (its excitons, blocks, samples and counts).
// Bins of the open sample window (the last window of the grid may be shorter).
static inline int32_t window_bins(const kmc_ctx* c) {
  return std::min(c->win_bins, c->n_bins - c->bin0);
}

// Zeroes the first nrec records (a window's worth).
static int clear_records(kmc_ctx* c, size_t nrec) {
  if (nrec == 0) return KMC_OK;
  KMC_HIP(c, hipMemsetAsync(c->r_dx.p, 0, nrec * sizeof(double), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->r_dy.p, 0, nrec * sizeof(double), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->r_dz.p, 0, nrec * sizeof(double), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->r_drain.p, 0, nrec * sizeof(uint32_t), c->stream));
  return KMC_OK;
}

net_view view_of(kmc_ctx* c) {
  net_view v{};
  v.x = excitons_of(c);
  v.blocks = c->d_blocks.p;
  v.roff = c->d_roff.p;
  v.src_roff = c->d_src_roff.p;
  v.n_src = (uint32_t)c->h_src.size();
  v.seed = c->cfg.seed;
  for (int k = 0; k < 3; ++k) {
    v.half_per[k] = (c->cfg.periodic_mask & (1 << k)) ? 0.5 * c->cfg.period_len[k] : INFINITY;
  }
  v.bin0 = c->bin0;
  v.n_bins = c->n_bins > 0 ? c->bin0 + window_bins(c) : 0;  // the window's end on the grid
  v.dt_sample = c->n_bins > 0 ? c->t_sample_max / c->n_bins : 0.0;
  v.rec_dx = c->r_dx.p;
  v.rec_dy = c->r_dy.p;
  v.rec_dz = c->r_dz.p;
  v.rec_drain = c->r_drain.p;
  v.site_count = c->d_site_count.p;
  v.count_stride = c->count_stride;
  v.velocity = c->exc_velocity;
  v.links = c->d_links.p;
  v.fly_err = c->d_fly_err.p;
  v.live = c->d_live.p;
  return v;
}

void reset_downstream(kmc_ctx* c, int level) {
  // level 0: sites changed; 1: cells/tables changed; 2: CSR changed
  if (level <= 0) c->cells_ok = false;
  if (level <= 1) c->csr_state = 0;
  c->n_exc = 0;
  c->n_bins = c->bin0 = c->win_bins = 0;
  c->n_trace = 0;
}

int copy_in(kmc_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return KMC_OK;
  KMC_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return KMC_OK;
}

int copy_out(kmc_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return KMC_OK;
  KMC_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  return KMC_OK;
}

// Row blocks of the hop kernel from the finished CSR (hdr, cum, edge), plus the block offsets of
// the source sites.  Blocks are laid out in cell-list order when the cell list is current, so
// sites that are close in space are close in memory.
int pack_rows(kmc_ctx* c) {
  const int64_t n = c->n_sites;
  KMC_TRY(dalloc(c, c->d_units, n));
  KMC_TRY(dalloc(c, c->d_roff, n + 1));
  KMC_TRY(dalloc(c, c->d_first, n + 1));
  KMC_TRY(dalloc(c, c->d_scan_tmp, scan_tmp_elems(n)));
  const int32_t* order = nullptr;
  if (c->cells_ok) {
    // sites in Morton order of their cells (ties: ascending site index, the sort is stable)
    KMC_TRY(dalloc(c, c->d_pkeys, 2 * n));
    KMC_TRY(dalloc(c, c->d_place, 2 * n));
    KMC_HIP(c, launch_morton_keys(n, c->d_cell_of.p, c->cells.nx, c->cells.ny, c->d_pkeys.p,
                                  c->d_place.p, c->stream));
    size_t need = 0;
    KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, need, c->d_pkeys.p, c->d_pkeys.p + n,
                                                  c->d_place.p, c->d_place.p + n, (int)n, 0, 64,
                                                  c->stream));
    if (need > c->sort_tmp_bytes) {
      KMC_TRY(dalloc(c, c->d_sort_tmp, need));
      c->sort_tmp_bytes = need;
    }
    size_t have = c->sort_tmp_bytes;
    KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->d_sort_tmp.p, have, c->d_pkeys.p,
                                                  c->d_pkeys.p + n, c->d_place.p,
                                                  c->d_place.p + n, (int)n, 0, 64, c->stream));
    order = c->d_place.p + n;
  }
  if (order) {
    KMC_HIP(c, launch_block_sizes(n, c->d_hdr.p, order, c->d_units.p, c->stream));
    KMC_HIP(c, exclusive_scan(c->d_units.p, c->d_first.p, n, c->d_scan_tmp.p, c->stream));
    KMC_HIP(c, launch_place_blocks(n, order, c->d_first.p, c->d_roff.p, c->stream));
  } else {
    KMC_HIP(c, launch_block_sizes(n, c->d_hdr.p, nullptr, c->d_units.p, c->stream));
    KMC_HIP(c, exclusive_scan(c->d_units.p, c->d_roff.p, n, c->d_scan_tmp.p, c->stream));
  }
  int64_t lines = 0;
  KMC_TRY(copy_out(c, &lines, c->d_roff.p + n, sizeof(int64_t)));
  // the wave loop addresses the blocks with 32-bit byte offsets: at most 2^25 lines (4 GiB)
  if (lines > ((int64_t)1 << 25))
    return fail(c, KMC_ERR_CAPACITY, "pack_rows: row blocks beyond 4 GiB (2^25 lines of 128 B)");
  c->n_units = lines * ROW_UNITS;
  KMC_TRY(dalloc(c, c->d_blocks, c->n_units));
  KMC_HIP(c, hipMemsetAsync(c->d_blocks.p, 0, c->n_units * sizeof(uint4), c->stream));
  KMC_HIP(c, launch_fill_blocks(n, c->d_hdr.p, c->d_cum.p, c->d_edge.p, c->d_roff.p,
                                c->d_pos[0].p, c->d_pos[1].p, c->d_pos[2].p, c->d_blocks.p,
                                c->stream));
  KMC_TRY(dalloc(c, c->d_src_roff, c->h_src.size()));
  KMC_HIP(c, launch_gather_offsets((int64_t)c->h_src.size(), c->d_src.p, c->d_roff.p, c->d_hdr.p,
                                   c->d_src_roff.p, c->stream));
  KMC_TRY(dalloc(c, c->d_links, n));
  KMC_HIP(c, launch_links(n, c->d_left.p, c->d_right.p, c->d_flags.p, c->d_roff.p, c->d_hdr.p,
                          c->d_links.p, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  return KMC_OK;
}

int check_err_flag(kmc_ctx* c, const char* what) {
  int32_t h = 0;
  KMC_TRY(copy_out(c, &h, c->d_err.p, sizeof(int32_t)));
  if (h != 0) return fail(c, KMC_ERR_RATE, std::string(what) + ": negative, NaN or infinite rate");
  return KMC_OK;
}

// Every same-tube link must have a nonzero length in the FP32 positions a flight uses (a flight
// over a zero-length segment would pass sites without using time).
int check_links(kmc_ctx* c, const char* what) {
  double per[3];
  for (int k = 0; k < 3; ++k)
    per[k] = (c->cfg.periodic_mask & (1 << k)) ? c->cfg.period_len[k] : 0.0;
  for (int64_t i = 0; i < c->n_sites; ++i) {
    const int32_t nb[2] = {c->h_left[i], c->h_right[i]};
    for (int32_t j : nb) {
      if (j < 0) continue;
      double d[3];
      for (int k = 0; k < 3; ++k) {
        d[k] = (double)(float)c->h_pos[k][j] - (double)(float)c->h_pos[k][i];
        if (per[k] > 0.0) {
          if (d[k] > 0.5 * per[k]) d[k] -= per[k];
          else if (d[k] < -0.5 * per[k]) d[k] += per[k];
        }
      }
      if (!(norm3(d[0], d[1], d[2]) > 0.0))
        return fail(c, KMC_ERR_GEOMETRY, std::string(what) + ": linked sites " +
                                             std::to_string(i) + " and " + std::to_string(j) +
                                             " coincide");
    }
  }
  return KMC_OK;
}

}  // namespace

// =============================================================================================
extern "C" {

const char* kmc_version(void) { return "cntkmc 0.1 (gfx950)"; }

const char* kmc_last_error(const kmc_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

int kmc_create(const kmc_config* cfg, kmc_ctx** out) {
  if (!cfg || !out) return fail(nullptr, KMC_ERR_ARG, "kmc_create: null argument");
  *out = nullptr;
  if (!(cfg->cutoff > 0.0) || !std::isfinite(cfg->cutoff) || !(cfg->min_distance >= 0.0) ||
      !(cfg->min_distance < cfg->cutoff))
    return fail(nullptr, KMC_ERR_ARG, "kmc_create: need 0 <= min_distance < cutoff < inf");
  for (int a = 0; a < 3; ++a) {
    if (cfg->periodic_mask & (1 << a)) {
      if (!(cfg->period_len[a] > 2.0 * cfg->cutoff) || !std::isfinite(cfg->period_len[a]) ||
          !std::isfinite(cfg->period_lo[a]))
        return fail(nullptr, KMC_ERR_GEOMETRY,
                    "kmc_create: a periodic axis needs period_len > 2 * cutoff");
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, KMC_ERR_NODEVICE, "kmc_create: no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(nullptr, KMC_ERR_ARG, "kmc_create: device ordinal out of range");
  kmc_ctx* c = new kmc_ctx();
  c->cfg = *cfg;
  hipError_t e = hipSetDevice(cfg->device);
  if (e == hipSuccess) e = scratch_setup(cfg->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_err.p, sizeof(int32_t));
  if (e != hipSuccess) {
    int code = hip_fail(nullptr, e, "kmc_create");
    kmc_destroy(c);
    return code;
  }
  c->d_err.n = 1;
  *out = c;
  return KMC_OK;
}

int kmc_destroy(kmc_ctx* c) {
  if (!c) return KMC_OK;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int a = 0; a < 3; ++a) {
    dfree(c->d_pos[a]);
    dfree(c->d_ori[a]);
  }
  dfree(c->d_chir); dfree(c->d_flags); dfree(c->d_left); dfree(c->d_right); dfree(c->d_src);
  dfree(c->d_cell_of); dfree(c->d_cell_sites); dfree(c->d_cell_count); dfree(c->d_cell_start);
  dfree(c->d_scan_tmp); dfree(c->d_tab_rates); dfree(c->d_tab_grid); dfree(c->d_deg);
  dfree(c->d_row_ptr); dfree(c->d_hdr); dfree(c->d_cum); dfree(c->d_pair_rates);
  dfree(c->d_edge);
  for (int q = 0; q < 4; ++q) dfree(c->d_geo[q]);
  dfree(c->d_pkeys); dfree(c->d_place);
  dfree(c->d_units); dfree(c->d_src_roff); dfree(c->d_roff); dfree(c->d_blocks);
  dfree(c->d_first); dfree(c->d_keys); dfree(c->d_ids); dfree(c->d_sort_tmp);
  dfree(c->d_counts64); dfree(c->d_cls); dfree(c->d_cls_count); dfree(c->d_cls_start);
  dfree(c->d_sel); dfree(c->d_mlp_flag);
  dfree(c->d_err); dfree(c->x_site); dfree(c->x_t); dfree(c->x_dx); dfree(c->x_dy);
  dfree(c->x_dz); dfree(c->x_hops); dfree(c->x_hash); dfree(c->x_drains); dfree(c->x_krec);
  dfree(c->x_heading); dfree(c->x_soff); dfree(c->x_tlast); dfree(c->d_links); dfree(c->d_fly_err);
  dfree(c->d_site_count); dfree(c->r_dx); dfree(c->r_dy); dfree(c->r_dz);
  dfree(c->d_reduced); dfree(c->r_drain); dfree(c->tr_site); dfree(c->tr_time); dfree(c->tr_n);
  dfree(c->d_min);
  dfree(c->c_sites); dfree(c->c_keep); dfree(c->c_off); dfree(c->c_scan_tmp); dfree(c->c_live);
  dfree(c->c_live_out); dfree(c->c_next); dfree(c->c_err); dfree(c->c_pop_out); dfree(c->c_cur_out);
  dfree(c->c_made); dfree(c->c_skeys); dfree(c->c_sids); dfree(c->c_sort_tmp);
  for (hipEvent_t e : c->c_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->c_live_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->c_live_host) (void)hipHostFree(c->c_live_host);
  for (hipEvent_t e : c->hop_ev)
    if (e) (void)hipEventDestroy(e);
  for (int b = 0; b < 2; ++b) {
    dfree(c->c_site[b]); dfree(c->c_t[b]); dfree(c->c_hops[b]); dfree(c->c_id[b]);
    dfree(c->c_y[b]); dfree(c->c_soff[b]); dfree(c->c_tlast[b]); dfree(c->c_head[b]);
  }
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KMC_OK;
}

int kmc_sync(kmc_ctx* c) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_sync: null context");
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  return KMC_OK;
}

// ---------------------------------------------------------------------------------------------
int kmc_set_sites(kmc_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                  const double* ox, const double* oy, const double* oz, const int32_t* left,
                  const int32_t* right, const uint8_t* chirality, const uint8_t* flags) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_set_sites: null context");
  if (n <= 0 || n >= (int64_t)EDGE_COL_MASK || !x || !y || !z || !ox || !oy || !oz)
    return fail(c, KMC_ERR_ARG, "kmc_set_sites: need 0 < n < 2^31 and non-null x,y,z,ox,oy,oz");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  c->c_on = false;  // a contact run's sections and lists belong to the old sites
  const double* pin[3] = {x, y, z};
  const double* oin[3] = {ox, oy, oz};
  for (int a = 0; a < 3; ++a) {
    c->h_pos[a].assign(pin[a], pin[a] + n);
    c->h_ori[a].assign(oin[a], oin[a] + n);
  }
  c->h_left.assign(n, -1);
  c->h_right.assign(n, -1);
  c->h_chir.assign(n, 0);
  c->h_flags.assign(n, 0);
  if (left) c->h_left.assign(left, left + n);
  if (right) c->h_right.assign(right, right + n);
  if (chirality) c->h_chir.assign(chirality, chirality + n);
  if (flags) c->h_flags.assign(flags, flags + n);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const double p = c->h_pos[a][i];
      if (!std::isfinite(p) || !std::isfinite(c->h_ori[a][i]))
        return fail(c, KMC_ERR_ARG, "kmc_set_sites: non-finite position or orientation");
      if (c->cfg.periodic_mask & (1 << a)) {
        if (!(p >= c->cfg.period_lo[a] && p < c->cfg.period_lo[a] + c->cfg.period_len[a]))
          return fail(c, KMC_ERR_GEOMETRY,
                      "kmc_set_sites: periodic coordinate outside [period_lo, period_lo+len)");
      }
    }
    if (c->h_left[i] < -1 || c->h_left[i] >= n || c->h_right[i] < -1 || c->h_right[i] >= n)
      return fail(c, KMC_ERR_ARG, "kmc_set_sites: link index out of range");
    if (c->h_chir[i] >= MAX_CHIR)
      return fail(c, KMC_ERR_ARG, "kmc_set_sites: chirality index too large");
  }
  // bounding box (find_simulation_domain, monte_carlo.h:328-340)
  for (int a = 0; a < 3; ++a) {
    c->bb_lo[a] = c->bb_hi[a] = c->h_pos[a][0];
    for (int64_t i = 0; i < n; ++i) {
      const double p = c->h_pos[a][i];
      c->bb_lo[a] = c->bb_lo[a] > p ? p : c->bb_lo[a];
      c->bb_hi[a] = c->bb_hi[a] < p ? p : c->bb_hi[a];
    }
  }
  c->h_src.clear();
  for (int64_t i = 0; i < n; ++i)
    if (c->h_flags[i] & KMC_SITE_SOURCE) c->h_src.push_back((int32_t)i);
  c->n_sites = n;
  for (int a = 0; a < 3; ++a) {
    KMC_TRY(dalloc(c, c->d_pos[a], n));
    KMC_TRY(dalloc(c, c->d_ori[a], n));
    KMC_TRY(copy_in(c, c->d_pos[a].p, c->h_pos[a].data(), n * sizeof(double)));
    KMC_TRY(copy_in(c, c->d_ori[a].p, c->h_ori[a].data(), n * sizeof(double)));
  }
  KMC_TRY(dalloc(c, c->d_chir, n));
  KMC_TRY(dalloc(c, c->d_flags, n));
  KMC_TRY(dalloc(c, c->d_left, n));
  KMC_TRY(dalloc(c, c->d_right, n));
  KMC_TRY(copy_in(c, c->d_chir.p, c->h_chir.data(), n));
  KMC_TRY(copy_in(c, c->d_flags.p, c->h_flags.data(), n));
  KMC_TRY(copy_in(c, c->d_left.p, c->h_left.data(), n * sizeof(int32_t)));
  KMC_TRY(copy_in(c, c->d_right.p, c->h_right.data(), n * sizeof(int32_t)));
  KMC_TRY(dalloc(c, c->d_src, c->h_src.size()));
  KMC_TRY(copy_in(c, c->d_src.p, c->h_src.data(), c->h_src.size() * sizeof(int32_t)));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  reset_downstream(c, 0);
  return KMC_OK;
}

// ---------------------------------------------------------------------------------------------
int kmc_build_cells(kmc_ctx* c) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_build_cells: null context");
  if (c->n_sites == 0) return fail(c, KMC_ERR_STATE, "kmc_build_cells: call kmc_set_sites first");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  cell_desc& d = c->cells;
  d = cell_desc{};
  d.periodic_mask = c->cfg.periodic_mask;
  d.cutoff = c->cfg.cutoff;
  d.min_distance = c->cfg.min_distance;
  int32_t dims[3];
  for (int a = 0; a < 3; ++a) {
    if (d.periodic_mask & (1 << a)) {
      const int n = (int)std::floor(c->cfg.period_len[a] / c->cfg.cutoff);
      dims[a] = n < 1 ? 1 : n;
      d.lo[a] = c->cfg.period_lo[a];
      d.period[a] = c->cfg.period_len[a];
      d.cell[a] = c->cfg.period_len[a] / dims[a];
    } else {
      // n = ceil(extent / radius) + 1 (monte_carlo.h:377-385)
      const double ext = (c->bb_hi[a] - c->bb_lo[a]) / c->cfg.cutoff;
      if (!(ext < 1.0e9)) return fail(c, KMC_ERR_GEOMETRY, "kmc_build_cells: domain too large");
      dims[a] = (int)std::ceil(ext) + 1;
      d.lo[a] = c->bb_lo[a];
      d.cell[a] = c->cfg.cutoff;
      d.period[a] = 0.0;
    }
  }
  d.nx = dims[0];
  d.ny = dims[1];
  d.nz = dims[2];
  const int64_t n_cells = (int64_t)d.nx * d.ny * d.nz;
  if (n_cells >= (int64_t)1 << 31)
    return fail(c, KMC_ERR_GEOMETRY, "kmc_build_cells: more than 2^31 cells");
  c->n_cells = n_cells;
  const int64_t n = c->n_sites;
  KMC_TRY(dalloc(c, c->d_cell_of, n));
  KMC_TRY(dalloc(c, c->d_cell_sites, n));
  KMC_TRY(dalloc(c, c->d_cell_count, n_cells));
  KMC_TRY(dalloc(c, c->d_cell_start, n_cells + 1));
  KMC_TRY(dalloc(c, c->d_scan_tmp, std::max(scan_tmp_elems(n_cells), scan_tmp_elems(n))));
  KMC_HIP(c, hipMemsetAsync(c->d_cell_count.p, 0, n_cells * sizeof(uint32_t), c->stream));
  site_arrays s = sites_of(c);
  KMC_HIP(c, launch_cell_assign(s, d, c->d_cell_of.p, c->d_cell_count.p, c->stream));
  KMC_HIP(c, exclusive_scan(c->d_cell_count.p, c->d_cell_start.p, n_cells, c->d_scan_tmp.p,
                            c->stream));
  KMC_HIP(c, hipMemsetAsync(c->d_cell_count.p, 0, n_cells * sizeof(uint32_t), c->stream));
  KMC_HIP(c, launch_cell_scatter(n, c->d_cell_of.p, c->d_cell_start.p, c->d_cell_count.p,
                                 c->d_cell_sites.p, c->stream));
  KMC_HIP(c, launch_cell_sort(n_cells, c->d_cell_start.p, c->d_cell_sites.p, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  c->cells_ok = true;
  reset_downstream(c, 1);
  return KMC_OK;
}

int kmc_get_cells(kmc_ctx* c, int32_t dims[3], int64_t* n_cells, int64_t* cell_start,
                  int32_t* cell_sites) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_cells: null context");
  if (!c->cells_ok) return fail(c, KMC_ERR_STATE, "kmc_get_cells: cells not built");
  if (dims) {
    dims[0] = c->cells.nx;
    dims[1] = c->cells.ny;
    dims[2] = c->cells.nz;
  }
  if (n_cells) *n_cells = c->n_cells;
  if (cell_start)
    KMC_TRY(copy_out(c, cell_start, c->d_cell_start.p, (c->n_cells + 1) * sizeof(int64_t)));
  if (cell_sites)
    KMC_TRY(copy_out(c, cell_sites, c->d_cell_sites.p, c->n_sites * sizeof(int32_t)));
  return KMC_OK;
}

// ---------------------------------------------------------------------------------------------
int kmc_set_rate_table(kmc_ctx* c, int32_t n_chir, int32_t n_theta, const double* theta,
                       int32_t n_z, const double* z_shift, int32_t n_a1, const double* a1,
                       int32_t n_a2, const double* a2, const double* rates) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_set_rate_table: null context");
  if (n_chir < 1 || n_chir > MAX_CHIR || n_theta < 1 || n_z < 1 || n_a1 < 1 || n_a2 < 1 ||
      !theta || !z_shift || !a1 || !a2 || !rates)
    return fail(c, KMC_ERR_ARG, "kmc_set_rate_table: bad sizes or null arrays");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t nr = (int64_t)n_chir * n_chir * n_theta * n_z * n_a1 * n_a2;
  for (int64_t i = 0; i < nr; ++i)
    if (!(rates[i] >= 0.0) || !std::isfinite(rates[i]))
      return fail(c, KMC_ERR_RATE, "kmc_set_rate_table: negative, NaN or infinite rate");
  std::vector<double> grid;
  grid.insert(grid.end(), theta, theta + n_theta);
  grid.insert(grid.end(), z_shift, z_shift + n_z);
  grid.insert(grid.end(), a1, a1 + n_a1);
  grid.insert(grid.end(), a2, a2 + n_a2);
  KMC_TRY(dalloc(c, c->d_tab_grid, grid.size()));
  KMC_TRY(dalloc(c, c->d_tab_rates, nr));
  KMC_TRY(copy_in(c, c->d_tab_grid.p, grid.data(), grid.size() * sizeof(double)));
  KMC_TRY(copy_in(c, c->d_tab_rates.p, rates, nr * sizeof(double)));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  c->tab_n[0] = n_chir;
  c->tab_n[1] = n_theta;
  c->tab_n[2] = n_z;
  c->tab_n[3] = n_a1;
  c->tab_n[4] = n_a2;
  c->table_ok = true;
  return KMC_OK;
}

int kmc_build_rates(kmc_ctx* c, int32_t mode) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_build_rates: null context");
  if (mode != KMC_RATE_TABLE && mode != KMC_RATE_PAIR)
    return fail(c, KMC_ERR_ARG, "kmc_build_rates: unknown mode");
  if (!c->cells_ok) return fail(c, KMC_ERR_STATE, "kmc_build_rates: call kmc_build_cells first");
  if (mode == KMC_RATE_TABLE) {
    if (!c->table_ok) return fail(c, KMC_ERR_STATE, "kmc_build_rates: no rate table set");
    for (int64_t i = 0; i < c->n_sites; ++i)
      if (c->h_chir[i] >= c->tab_n[0])
        return fail(c, KMC_ERR_ARG, "kmc_build_rates: site chirality outside the table set");
  }
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t n = c->n_sites;
  site_arrays s = sites_of(c);
  KMC_TRY(dalloc(c, c->d_deg, n));
  KMC_TRY(dalloc(c, c->d_row_ptr, n + 1));
  KMC_HIP(c, launch_count_pairs(s, c->cells, c->d_cell_start.p, c->d_cell_sites.p, c->d_deg.p,
                                c->stream));
  KMC_HIP(c, exclusive_scan(c->d_deg.p, c->d_row_ptr.p, n, c->d_scan_tmp.p, c->stream));
  int64_t nnz = 0;
  KMC_TRY(copy_out(c, &nnz, c->d_row_ptr.p + n, sizeof(int64_t)));
  if (nnz >= (int64_t)0xFFFFFFFFll)
    return fail(c, KMC_ERR_CAPACITY, "kmc_build_rates: more than 2^32 pairs");
  c->nnz = nnz;
  KMC_TRY(dalloc(c, c->d_hdr, n));
  KMC_TRY(dalloc(c, c->d_cum, std::max<int64_t>(nnz, 1)));
  KMC_TRY(dalloc(c, c->d_edge, std::max<int64_t>(nnz, 1)));
  pair_geom_arrays geo{nullptr, nullptr, nullptr, nullptr};
  if (mode == KMC_RATE_PAIR) {
    for (int q = 0; q < 4; ++q) KMC_TRY(dalloc(c, c->d_geo[q], std::max<int64_t>(nnz, 1)));
    geo.theta = c->d_geo[0].p;
    geo.z = c->d_geo[1].p;
    geo.a1 = c->d_geo[2].p;
    geo.a2 = c->d_geo[3].p;
  }
  table_desc t{};
  if (c->table_ok) {
    t.n_chir = c->tab_n[0];
    t.n_th = c->tab_n[1];
    t.n_z = c->tab_n[2];
    t.n_a1 = c->tab_n[3];
    t.n_a2 = c->tab_n[4];
    t.rates = c->d_tab_rates.p;
    t.th = c->d_tab_grid.p;
    t.z = t.th + t.n_th;
    t.a1 = t.z + t.n_z;
    t.a2 = t.a1 + t.n_a1;
  }
  KMC_HIP(c, hipMemsetAsync(c->d_err.p, 0, sizeof(int32_t), c->stream));
  KMC_HIP(c, launch_fill_pairs(mode, s, c->cells, c->d_cell_start.p, c->d_cell_sites.p,
                               c->d_row_ptr.p, t, csr_of(c), geo, c->d_err.p, c->stream));
  KMC_TRY(check_err_flag(c, "kmc_build_rates"));
  if (mode == KMC_RATE_TABLE) KMC_TRY(pack_rows(c));
  c->csr_state = (mode == KMC_RATE_TABLE) ? 2 : 1;
  reset_downstream(c, 2);
  return KMC_OK;
}

int kmc_get_nnz(kmc_ctx* c, int64_t* nnz) {
  if (!c || !nnz) return fail(c, KMC_ERR_ARG, "kmc_get_nnz: null argument");
  if (c->csr_state == 0) return fail(c, KMC_ERR_STATE, "kmc_get_nnz: no CSR built");
  *nnz = c->nnz;
  return KMC_OK;
}

int kmc_get_pair_geometry(kmc_ctx* c, double* theta, double* z_shift, double* a1, double* a2,
                          uint8_t* chir_src, uint8_t* chir_dst) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_pair_geometry: null context");
  if (c->csr_state == 0 || !c->d_geo[0].p)
    return fail(c, KMC_ERR_STATE, "kmc_get_pair_geometry: build with KMC_RATE_PAIR first");
  double* outs[4] = {theta, z_shift, a1, a2};
  for (int q = 0; q < 4; ++q)
    if (outs[q]) KMC_TRY(copy_out(c, outs[q], c->d_geo[q].p, c->nnz * sizeof(double)));
  if (chir_src || chir_dst) {
    std::vector<site_hdr> hdr(c->n_sites);
    std::vector<edge_rec> edge(c->nnz);
    KMC_TRY(copy_out(c, hdr.data(), c->d_hdr.p, c->n_sites * sizeof(site_hdr)));
    KMC_TRY(copy_out(c, edge.data(), c->d_edge.p, c->nnz * sizeof(edge_rec)));
    for (int64_t i = 0; i < c->n_sites; ++i) {
      for (uint32_t k = 0; k < hdr[i].n; ++k) {
        const int64_t p = hdr[i].rs + k;
        if (chir_src) chir_src[p] = c->h_chir[i];
        if (chir_dst) chir_dst[p] = c->h_chir[edge[p].col & EDGE_COL_MASK];
      }
    }
  }
  return KMC_OK;
}

int kmc_set_pair_rates_device(kmc_ctx* c, const double* d_rates) {
  if (!c || !d_rates) return fail(c, KMC_ERR_ARG, "kmc_set_pair_rates: null argument");
  if (c->csr_state == 0)
    return fail(c, KMC_ERR_STATE, "kmc_set_pair_rates: build with KMC_RATE_PAIR first");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  KMC_HIP(c, hipMemsetAsync(c->d_err.p, 0, sizeof(int32_t), c->stream));
  KMC_HIP(c, launch_finish_pair_rates(c->n_sites, d_rates, csr_of(c), c->d_err.p, c->stream));
  KMC_TRY(check_err_flag(c, "kmc_set_pair_rates"));
  KMC_TRY(pack_rows(c));
  c->csr_state = 2;
  reset_downstream(c, 2);
  return KMC_OK;
}

int kmc_pair_classes(kmc_ctx* c, int32_t n_chir, int64_t* class_start, const int64_t** d_sel) {
  if (!c || n_chir < 1 || n_chir > MAX_CHIR || !class_start)
    return fail(c, KMC_ERR_ARG, "kmc_pair_classes: bad argument");
  if (c->csr_state == 0) return fail(c, KMC_ERR_STATE, "kmc_pair_classes: no pair geometry");
  for (int64_t i = 0; i < c->n_sites; ++i)
    if (c->h_chir[i] >= n_chir) return fail(c, KMC_ERR_ARG, "kmc_pair_classes: chirality >= n_chir");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int nc = n_chir * n_chir;
  const int64_t nnz = std::max<int64_t>(c->nnz, 1);
  KMC_TRY(dalloc(c, c->d_cls, nnz));
  KMC_TRY(dalloc(c, c->d_cls_count, nc));
  KMC_TRY(dalloc(c, c->d_cls_start, nc + 1));
  KMC_TRY(dalloc(c, c->d_sel, nnz));
  KMC_HIP(c, hipMemsetAsync(c->d_cls_count.p, 0, nc * sizeof(uint32_t), c->stream));
  KMC_HIP(c, launch_pair_classes(c->n_sites, c->nnz, c->d_hdr.p, c->d_edge.p, c->d_chir.p, n_chir,
                                 c->d_cls.p, c->d_cls_count.p, c->stream));
  std::vector<uint32_t> cnt(nc);
  KMC_TRY(copy_out(c, cnt.data(), c->d_cls_count.p, nc * sizeof(uint32_t)));
  class_start[0] = 0;
  for (int k = 0; k < nc; ++k) class_start[k + 1] = class_start[k] + cnt[k];
  KMC_TRY(copy_in(c, c->d_cls_start.p, class_start, (nc + 1) * sizeof(int64_t)));
  KMC_HIP(c, hipMemsetAsync(c->d_cls_count.p, 0, nc * sizeof(uint32_t), c->stream));
  KMC_HIP(c, launch_pair_scatter(c->nnz, c->d_cls.p, nullptr, c->d_cls_start.p, c->d_cls_count.p,
                                 c->d_sel.p, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  if (d_sel) *d_sel = c->d_sel.p;
  return KMC_OK;
}

int kmc_build_rates_exact(kmc_ctx* c, int32_t n_chir, kmc_transfer* const* transfers,
                          double* elapsed_ms) {
  if (!c || !transfers) return fail(c, KMC_ERR_ARG, "kmc_build_rates_exact: null argument");
  if (c->csr_state != 1 || !c->d_geo[0].p)
    return fail(c, KMC_ERR_STATE, "kmc_build_rates_exact: build with KMC_RATE_PAIR first");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int64_t> start(n_chir * n_chir + 1);
  const int64_t* sel = nullptr;
  KMC_TRY(kmc_pair_classes(c, n_chir, start.data(), &sel));
  KMC_TRY(dalloc(c, c->d_pair_rates, std::max<int64_t>(c->nnz, 1)));
  KMC_HIP(c, hipMemsetAsync(c->d_pair_rates.p, 0, c->nnz * sizeof(double), c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < n_chir * n_chir; ++k) {
    const int64_t n = start[k + 1] - start[k];
    if (n == 0 || !transfers[k]) continue;
    const int s = kmc_transfer_rates_device(transfers[k], n, c->d_geo[0].p, c->d_geo[1].p,
                                            c->d_geo[2].p, c->d_geo[3].p, sel + start[k],
                                            c->d_pair_rates.p);
    if (s != KMC_OK) return fail(c, s, std::string("kmc_build_rates_exact: ") + kmc_transfer_last_error());
  }
  KMC_TRY(kmc_set_pair_rates_device(c, c->d_pair_rates.p));
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KMC_OK;
}

int kmc_build_rates_mlp(kmc_ctx* c, kmc_mlp* mlp, int32_t n_chir, const int32_t* chir_class,
                        double temperature, kmc_transfer* const* transfers, int64_t* n_fallback,
                        double* elapsed_ms) {
  if (!c || !mlp || !chir_class || !transfers || !(temperature > 0))
    return fail(c, KMC_ERR_ARG, "kmc_build_rates_mlp: bad argument");
  if (c->csr_state != 1 || !c->d_geo[0].p)
    return fail(c, KMC_ERR_STATE, "kmc_build_rates_mlp: build with KMC_RATE_PAIR first");
  if (cntkmc::mlp_device(mlp) != c->cfg.device)
    return fail(c, KMC_ERR_ARG, "kmc_build_rates_mlp: surrogate lives on another device");
  const auto t0 = std::chrono::steady_clock::now();
  const int nc = n_chir * n_chir;
  std::vector<int64_t> start(nc + 1);
  KMC_TRY(kmc_pair_classes(c, n_chir, start.data(), nullptr));
  {
    const int s = cntkmc::mlp_prepare(mlp);
    if (s != KMC_OK) return fail(c, s, std::string("kmc_build_rates_mlp: ") + kmc_mlp_last_error());
  }
  const int64_t nnz = std::max<int64_t>(c->nnz, 1);
  KMC_TRY(dalloc(c, c->d_pair_rates, nnz));
  KMC_TRY(dalloc(c, c->d_mlp_flag, nnz));
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  KMC_HIP(c, cntkmc::mlp_launch_pairs(mlp, c->nnz, c->d_geo[0].p, c->d_geo[1].p, c->d_geo[2].p,
                                      c->d_geo[3].p, c->d_cls.p, n_chir, chir_class, temperature,
                                      c->d_pair_rates.p, c->d_mlp_flag.p, c->stream));
  // flagged entries grouped by chirality class, then the exact routine on each group
  KMC_HIP(c, hipMemsetAsync(c->d_cls_count.p, 0, nc * sizeof(uint32_t), c->stream));
  KMC_HIP(c, launch_mask_count(c->nnz, c->d_cls.p, c->d_mlp_flag.p, c->d_cls_count.p, c->stream));
  std::vector<uint32_t> cnt(nc);
  KMC_TRY(copy_out(c, cnt.data(), c->d_cls_count.p, nc * sizeof(uint32_t)));
  start[0] = 0;
  for (int k = 0; k < nc; ++k) start[k + 1] = start[k] + cnt[k];
  for (int k = 0; k < nc; ++k)
    if (cnt[k] && !transfers[k])
      return fail(c, KMC_ERR_ARG, "kmc_build_rates_mlp: flagged pairs of a class without a transfer");
  if (start[nc] > 0) {
    KMC_TRY(copy_in(c, c->d_cls_start.p, start.data(), (nc + 1) * sizeof(int64_t)));
    KMC_HIP(c, hipMemsetAsync(c->d_cls_count.p, 0, nc * sizeof(uint32_t), c->stream));
    KMC_HIP(c, launch_pair_scatter(c->nnz, c->d_cls.p, c->d_mlp_flag.p, c->d_cls_start.p,
                                   c->d_cls_count.p, c->d_sel.p, c->stream));
    KMC_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < nc; ++k) {
      if (!cnt[k]) continue;
      const int s = kmc_transfer_rates_device(transfers[k], cnt[k], c->d_geo[0].p, c->d_geo[1].p,
                                              c->d_geo[2].p, c->d_geo[3].p, c->d_sel.p + start[k],
                                              c->d_pair_rates.p);
      if (s != KMC_OK) return fail(c, s, std::string("kmc_build_rates_mlp: ") + kmc_transfer_last_error());
    }
  }
  KMC_TRY(kmc_set_pair_rates_device(c, c->d_pair_rates.p));
  if (n_fallback) *n_fallback = start[nc];
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KMC_OK;
}

int kmc_get_pair_rates(kmc_ctx* c, double* rates) {
  if (!c || !rates) return fail(c, KMC_ERR_ARG, "kmc_get_pair_rates: null argument");
  if (!c->d_pair_rates.p) return fail(c, KMC_ERR_STATE, "kmc_get_pair_rates: no pair rates");
  return copy_out(c, rates, c->d_pair_rates.p, c->nnz * sizeof(double));
}

int kmc_set_pair_rates(kmc_ctx* c, const double* rates) {
  if (!c || !rates) return fail(c, KMC_ERR_ARG, "kmc_set_pair_rates: null argument");
  if (c->csr_state == 0)
    return fail(c, KMC_ERR_STATE, "kmc_set_pair_rates: build with KMC_RATE_PAIR first");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  KMC_TRY(dalloc(c, c->d_pair_rates, std::max<int64_t>(c->nnz, 1)));
  KMC_TRY(copy_in(c, c->d_pair_rates.p, rates, c->nnz * sizeof(double)));
  return kmc_set_pair_rates_device(c, c->d_pair_rates.p);
}

int kmc_get_csr(kmc_ctx* c, int64_t* row_ptr, int32_t* col, double* cum, float* disp,
                double* gamma) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_csr: null context");
  if (c->csr_state != 2) return fail(c, KMC_ERR_STATE, "kmc_get_csr: CSR not complete");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t n = c->n_sites, nnz = c->nnz;
  std::vector<site_hdr> hdr(n);
  KMC_TRY(copy_out(c, hdr.data(), c->d_hdr.p, n * sizeof(site_hdr)));
  if (row_ptr) {
    for (int64_t i = 0; i < n; ++i) row_ptr[i] = hdr[i].rs;
    row_ptr[n] = nnz;
  }
  if (gamma)
    for (int64_t i = 0; i < n; ++i) gamma[i] = hdr[i].gamma;
  if (cum) KMC_TRY(copy_out(c, cum, c->d_cum.p, nnz * sizeof(double)));
  if (col || disp) {
    std::vector<edge_rec> edge(nnz);
    KMC_TRY(copy_out(c, edge.data(), c->d_edge.p, nnz * sizeof(edge_rec)));
    for (int64_t p = 0; p < nnz; ++p) {
      if (col) col[p] = (int32_t)(edge[p].col & EDGE_COL_MASK);
      if (disp) {
        disp[3 * p + 0] = edge[p].dx;
        disp[3 * p + 1] = edge[p].dy;
        disp[3 * p + 2] = edge[p].dz;
      }
    }
  }
  return KMC_OK;
}

int kmc_set_csr(kmc_ctx* c, int64_t n_sites, const int64_t* row_ptr, const int32_t* col,
                const double* cum, const float* disp) {
  if (!c || !row_ptr || !col || !cum || !disp)
    return fail(c, KMC_ERR_ARG, "kmc_set_csr: null argument");
  if (n_sites != c->n_sites || n_sites == 0)
    return fail(c, KMC_ERR_STATE, "kmc_set_csr: site count differs from kmc_set_sites");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t nnz = row_ptr[n_sites];
  if (row_ptr[0] != 0 || nnz < 0 || nnz >= (int64_t)0xFFFFFFFFll)
    return fail(c, KMC_ERR_ARG, "kmc_set_csr: bad row_ptr");
  std::vector<site_hdr> hdr(n_sites);
  std::vector<edge_rec> edge(std::max<int64_t>(nnz, 1));
  for (int64_t i = 0; i < n_sites; ++i) {
    if (row_ptr[i + 1] < row_ptr[i]) return fail(c, KMC_ERR_ARG, "kmc_set_csr: row_ptr decreases");
    hdr[i].rs = (uint32_t)row_ptr[i];
    hdr[i].n = (uint32_t)(row_ptr[i + 1] - row_ptr[i]);
    hdr[i].gamma = hdr[i].n ? cum[row_ptr[i + 1] - 1] : 0.0;
    double prev = 0.0;
    for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
      if (col[p] < 0 || col[p] >= n_sites) return fail(c, KMC_ERR_ARG, "kmc_set_csr: bad column");
      if (!(cum[p] >= prev) || !std::isfinite(cum[p]))
        return fail(c, KMC_ERR_RATE, "kmc_set_csr: cumulative rates must be finite and rise");
      prev = cum[p];
      edge[p].col = (uint32_t)col[p] | ((c->h_flags[col[p]] & KMC_SITE_DRAIN) ? EDGE_DRAIN_BIT : 0u);
      edge[p].dx = disp[3 * p + 0];
      edge[p].dy = disp[3 * p + 1];
      edge[p].dz = disp[3 * p + 2];
    }
  }
  c->nnz = nnz;
  KMC_TRY(dalloc(c, c->d_hdr, n_sites));
  KMC_TRY(dalloc(c, c->d_cum, std::max<int64_t>(nnz, 1)));
  KMC_TRY(dalloc(c, c->d_edge, std::max<int64_t>(nnz, 1)));
  KMC_TRY(copy_in(c, c->d_hdr.p, hdr.data(), n_sites * sizeof(site_hdr)));
  KMC_TRY(copy_in(c, c->d_cum.p, cum, nnz * sizeof(double)));
  KMC_TRY(copy_in(c, c->d_edge.p, edge.data(), nnz * sizeof(edge_rec)));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  if (!c->d_scan_tmp.p || (int64_t)c->d_scan_tmp.n < scan_tmp_elems(n_sites))
    KMC_TRY(dalloc(c, c->d_scan_tmp, scan_tmp_elems(n_sites)));
  KMC_TRY(pack_rows(c));
  c->csr_state = 2;
  reset_downstream(c, 2);
  return KMC_OK;
}

// ---------------------------------------------------------------------------------------------
int kmc_inject(kmc_ctx* c, int64_t n_excitons) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_inject: null context");
  if (c->csr_state != 2) return fail(c, KMC_ERR_STATE, "kmc_inject: CSR not complete");
  if (n_excitons <= 0 || n_excitons > (int64_t)0x7FFFFFFF)
    return fail(c, KMC_ERR_ARG, "kmc_inject: need 0 < n_excitons < 2^31");
  if (c->h_src.empty()) return fail(c, KMC_ERR_STATE, "kmc_inject: no source (injection) sites");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t E = n_excitons;
  KMC_TRY(dalloc(c, c->x_site, E));
  KMC_TRY(dalloc(c, c->x_t, E));
  KMC_TRY(dalloc(c, c->x_dx, E));
  KMC_TRY(dalloc(c, c->x_dy, E));
  KMC_TRY(dalloc(c, c->x_dz, E));
  KMC_TRY(dalloc(c, c->x_hops, E));
  KMC_TRY(dalloc(c, c->x_hash, E));
  KMC_TRY(dalloc(c, c->x_drains, E));
  KMC_TRY(dalloc(c, c->x_krec, E));
  KMC_TRY(dalloc(c, c->x_heading, E));
  KMC_TRY(dalloc(c, c->x_soff, E));
  KMC_TRY(dalloc(c, c->x_tlast, E));
  const int64_t n_lines = c->n_units / ROW_UNITS;  // counters per block line
  c->count_stride = (n_lines + COUNT_ALIGN - 1) / COUNT_ALIGN * COUNT_ALIGN;
  KMC_TRY(dalloc(c, c->d_site_count, (size_t)KMC_COUNT_COPIES * c->count_stride));
  KMC_HIP(c, hipMemsetAsync(c->d_site_count.p, 0, c->d_site_count.n * sizeof(uint32_t),
                            c->stream));
  KMC_TRY(dalloc(c, c->d_counts64, c->n_sites));
  KMC_HIP(c, hipMemsetAsync(c->d_counts64.p, 0, c->n_sites * sizeof(uint64_t), c->stream));
  c->count_pending = 0;
  KMC_TRY(dalloc(c, c->d_live, 3));
  KMC_HIP(c, hipMemsetAsync(c->d_live.p, 0, 3 * sizeof(uint32_t), c->stream));
  c->run_launches = 0;
  c->n_exc = E;
  c->n_bins = c->bin0 = c->win_bins = 0;
  c->n_trace = 0;
  if (c->velocity > 0.0) KMC_TRY(check_links(c, "kmc_inject"));
  c->exc_velocity = c->velocity;
  KMC_HIP(c, launch_inject(excitons_of(c), c->d_blocks.p, c->d_src_roff.p,
                           (uint32_t)c->h_src.size(), c->cfg.seed, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  return KMC_OK;
}

// Allocates and clears the records of one window of `window` bins (the whole grid when window
// equals n_bins) and starts it at bin 0; every exciton's next sample is bin 0.
static int set_samples(kmc_ctx* c, int32_t n_bins, double t_sample_max, int32_t window,
                       const char* who) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, std::string(who) + ": null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, std::string(who) + ": call kmc_inject first");
  if (n_bins < 0 || (n_bins > 0 && !(t_sample_max > 0.0 && std::isfinite(t_sample_max))))
    return fail(c, KMC_ERR_ARG,
                std::string(who) + ": need n_bins >= 0 and 0 < t_sample_max < inf");
  if (n_bins > 0 && (window < 1 || window > n_bins))
    return fail(c, KMC_ERR_ARG, std::string(who) + ": need 1 <= window_bins <= n_bins");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const size_t nrec = (size_t)(n_bins > 0 ? window : 0) * c->n_exc;
  KMC_TRY(dalloc(c, c->r_dx, nrec));
  KMC_TRY(dalloc(c, c->r_dy, nrec));
  KMC_TRY(dalloc(c, c->r_dz, nrec));
  KMC_TRY(dalloc(c, c->r_drain, nrec));
  KMC_HIP(c, hipMemsetAsync(c->x_krec.p, 0, c->n_exc * sizeof(uint32_t), c->stream));
  KMC_TRY(clear_records(c, nrec));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  c->n_bins = n_bins;
  c->bin0 = 0;
  c->win_bins = n_bins > 0 ? window : 0;
  c->t_sample_max = t_sample_max;
  return KMC_OK;
}

int kmc_set_samples(kmc_ctx* c, int32_t n_bins, double t_sample_max) {
  return set_samples(c, n_bins, t_sample_max, n_bins, "kmc_set_samples");
}

int kmc_set_samples_window(kmc_ctx* c, int32_t n_bins, double t_sample_max,
                           int32_t window_bins) {
  return set_samples(c, n_bins, t_sample_max, window_bins, "kmc_set_samples_window");
}

int kmc_get_sample_window(kmc_ctx* c, int32_t* first_bin, int32_t* n_window) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_sample_window: null context");
  if (c->n_bins == 0) return fail(c, KMC_ERR_STATE, "kmc_get_sample_window: no sampling grid");
  if (first_bin) *first_bin = c->bin0;
  if (n_window) *n_window = window_bins(c);
  return KMC_OK;
}

int kmc_next_sample_window(kmc_ctx* c) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_next_sample_window: null context");
  if (c->n_bins == 0) return fail(c, KMC_ERR_STATE, "kmc_next_sample_window: no sampling grid");
  const int32_t end = c->bin0 + window_bins(c);
  if (end >= c->n_bins)
    return fail(c, KMC_ERR_STATE, "kmc_next_sample_window: the last window is already open");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  // every exciton must have recorded the whole window (its clock past the window's last sample
  // time), or its records would be lost: the smallest next-sample index decides
  KMC_TRY(dalloc(c, c->d_min, 1));
  size_t need = 0;
  KMC_HIP(c, hipcub::DeviceReduce::Min(nullptr, need, c->x_krec.p, c->d_min.p, (int)c->n_exc,
                                       c->stream));
  if (need > c->sort_tmp_bytes) {
    KMC_TRY(dalloc(c, c->d_sort_tmp, need));
    c->sort_tmp_bytes = need;
  }
  KMC_HIP(c, hipcub::DeviceReduce::Min(c->d_sort_tmp.p, need, c->x_krec.p, c->d_min.p,
                                       (int)c->n_exc, c->stream));
  uint32_t lo = 0;
  KMC_TRY(copy_out(c, &lo, c->d_min.p, sizeof(uint32_t)));
  if (lo < (uint32_t)end)
    return fail(c, KMC_ERR_STATE,
                "kmc_next_sample_window: some exciton has recorded only " + std::to_string(lo) +
                    " of the " + std::to_string(end) +
                    " samples up to the window's end (run to t >= " +
                    std::to_string(c->t_sample_max * end / c->n_bins) + " s first)");
  c->bin0 = end;
  KMC_TRY(clear_records(c, (size_t)c->win_bins * c->n_exc));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  return KMC_OK;
}

int kmc_set_trace(kmc_ctx* c, int64_t n_trace, int64_t max_hops) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_set_trace: null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_set_trace: call kmc_inject first");
  if (n_trace < 0 || max_hops < 0 || n_trace > c->n_exc)
    return fail(c, KMC_ERR_ARG, "kmc_set_trace: need 0 <= n_trace <= n_excitons, max_hops >= 0");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  KMC_TRY(dalloc(c, c->tr_site, (size_t)n_trace * max_hops));
  KMC_TRY(dalloc(c, c->tr_time, (size_t)n_trace * max_hops));
  KMC_TRY(dalloc(c, c->tr_n, n_trace));
  if (n_trace) KMC_HIP(c, hipMemsetAsync(c->tr_n.p, 0, n_trace * sizeof(int64_t), c->stream));
  if (n_trace && max_hops) {
    const size_t nt = (size_t)n_trace * max_hops;
    KMC_HIP(c, hipMemsetAsync(c->tr_site.p, 0, nt * sizeof(int32_t), c->stream));
    KMC_HIP(c, hipMemsetAsync(c->tr_time.p, 0, nt * sizeof(double), c->stream));
  }
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  c->n_trace = n_trace;
  c->tr_max = max_hops;
  return KMC_OK;
}

// Departure counts: each sub-launch adds at most L x (its hops) to any slot of a 32-bit copy,
// L the excitons a launch could still move (E, or fewer once the live count has been read), so
// launches are capped at UINT32_MAX / L hops while counting, and the copies are folded into the
// 64-bit totals before a launch that could take `pending` past UINT32_MAX.
constexpr uint64_t COUNT_COPY_MAX = 0xFFFFFFFFull;

// Once at most E / TAIL_FRACTION excitons can still move (a clock-bounded run whose remaining
// excitons have long hop chains), sub-launches grow TAIL_STRETCH-fold: the few chains left
// gain nothing from the sort, and each launch's sort and start-up would otherwise cost about as
// much as its hops.
constexpr int64_t TAIL_FRACTION = 64, TAIL_STRETCH = 16;

// Reads of the device's live-exciton count (each a host synchronisation) after launches 1, 2,
// 4 and every 8th: a run stops early once no exciton can hop (all past t_max or trapped).
constexpr int64_t LIVE_CHECK_EVERY = 8;
static inline bool live_check_due(int64_t launches) {
  return (launches & (launches - 1)) == 0 || launches % LIVE_CHECK_EVERY == 0;
}

static int fold_counts(kmc_ctx* c, hipStream_t st) {
  KMC_HIP(c, launch_fold_counts(c->n_units / ROW_UNITS, c->count_stride, c->d_blocks.p,
                                c->d_site_count.p, c->d_counts64.p, st));
  c->count_pending = 0;
  return KMC_OK;
}

// The hop budget runs as sub-launches of at most `chunk` hops (0: one launch); before each, the
// excitons are sorted by the block of their current site (radix sort of 32-bit keys) so that
// neighbouring threads and each XCD's L2 share rows.  Draws are keyed by exciton and hop, so the
// split and the order leave every trajectory unchanged.  With `fold` the counts go to the
// context's own copies (folded into its totals as needed); without, to a scratch array.
static int run_hops(kmc_ctx* c, hop_args a, int64_t hops, int64_t chunk, bool count, bool sample,
                    bool trace, bool fold) {
  const int64_t E = c->n_exc;
  const bool fly = c->exc_velocity > 0.0;
  if (fly) KMC_HIP(c, hipMemsetAsync(c->d_fly_err.p, 0, sizeof(int32_t), c->stream));
  KMC_TRY(dalloc(c, c->d_keys, 2 * E));
  KMC_TRY(dalloc(c, c->d_ids, 2 * E));
  size_t need = 0;
  KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, need, c->d_keys.p, c->d_keys.p + E,
                                                c->d_ids.p, c->d_ids.p + E, (int)E, 0, 32,
                                                c->stream));
  if (need > c->sort_tmp_bytes) {
    KMC_TRY(dalloc(c, c->d_sort_tmp, need));
    c->sort_tmp_bytes = need;
  }
  a.perm = c->d_ids.p + E;
  // A clock-bounded run leaves out of each launch the excitons whose next hop falls after t_max
  // (with no MSD record due before it): they sort after the others, which the hop kernel alone
  // deals to the XCDs, so the waves hold excitons that still hop.
  const bool leave_idle = !fly && std::isfinite(a.t_max);
  idle_test it{};
  if (leave_idle) {
    it.t_next = c->x_t.p;
    it.t_max = a.t_max;
    it.k_rec = sample ? c->x_krec.p : nullptr;
    it.n_bins = a.v.n_bins;
    it.dt_sample = a.v.dt_sample;
    KMC_HIP(c, hipMemsetAsync(c->d_live.p + 1, 0, 2 * sizeof(uint32_t), c->stream));
  }
  const int64_t step = chunk > 0 ? chunk : std::max<int64_t>(hops, 1);
  int64_t done = 0, launches = 0;
  int64_t movable = E;  // an upper bound on the excitons a launch can move (the last live count)
  do {
    int64_t len = step;
    if (chunk > 0 && movable * TAIL_FRACTION <= E && step <= INT64_MAX / TAIL_STRETCH)
      len = step * TAIL_STRETCH;
    if (count)
      len = std::min<int64_t>(len, (int64_t)(COUNT_COPY_MAX / (uint64_t)std::max<int64_t>(movable, 1)));
    const int64_t sub = std::min(len, hops - done);
    if (count && fold) {
      const uint64_t add = (uint64_t)movable * (uint64_t)sub;
      if (c->count_pending + add > COUNT_COPY_MAX) KMC_TRY(fold_counts(c, c->stream));
      c->count_pending += add;
    }
    if (leave_idle) {
      it.alive = c->d_live.p + 1 + (launches & 1);
      it.alive_next = c->d_live.p + 1 + ((launches + 1) & 1);
      a.n_alive = it.alive;
    }
    KMC_HIP(c, launch_exciton_keys(E, c->x_site.p, c->d_roff.p, it, c->d_keys.p, c->d_ids.p,
                                   c->d_live.p, c->stream));
    size_t have = c->sort_tmp_bytes;
    KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->d_sort_tmp.p, have, c->d_keys.p,
                                                  c->d_keys.p + E, c->d_ids.p, c->d_ids.p + E,
                                                  (int)E, 0, 32, c->stream));
    a.hop_budget = (uint64_t)sub;
    if (c->time_hops) {
      while ((int64_t)c->hop_ev.size() < 2 * (launches + 1)) {
        hipEvent_t ev = nullptr;
        KMC_HIP(c, hipEventCreate(&ev));
        c->hop_ev.push_back(ev);
      }
      KMC_HIP(c, hipEventRecord(c->hop_ev[2 * launches], c->stream));
    }
    KMC_HIP(c, launch_hop(a, count, sample, trace, fly, c->stream));
    if (c->time_hops) KMC_HIP(c, hipEventRecord(c->hop_ev[2 * launches + 1], c->stream));
    done += sub;
    ++launches;
    if (done < hops && live_check_due(launches)) {
      uint32_t live = 0;
      KMC_TRY(copy_out(c, &live, c->d_live.p, sizeof(uint32_t)));
      if (live == 0u) break;
      movable = std::min<int64_t>(movable, (int64_t)live);
    }
  } while (done < hops);
  c->run_launches = launches;
  if (c->time_hops) {
    double sum = 0.0;
    for (int64_t i = 0; i < launches; ++i) {
      float ms = 0.f;
      KMC_HIP(c, hipEventSynchronize(c->hop_ev[2 * i + 1]));
      KMC_HIP(c, hipEventElapsedTime(&ms, c->hop_ev[2 * i], c->hop_ev[2 * i + 1]));
      sum += ms;
    }
    c->hop_ms = sum;
  }
  if (fly) {
    int32_t flag = 0;
    KMC_TRY(copy_out(c, &flag, c->d_fly_err.p, sizeof(int32_t)));
    if (flag)
      return fail(c, KMC_ERR_STATE, "kmc_run: a flight passed more than 2^22 sites "
                                    "(velocity x dwell far beyond the tube lengths)");
  }
  return KMC_OK;
}

int kmc_set_hop_schedule(kmc_ctx* c, int64_t resort_hops) {
  if (!c || resort_hops < 0) return fail(c, KMC_ERR_ARG, "kmc_set_hop_schedule: bad argument");
  c->resort_hops = resort_hops;
  return KMC_OK;
}

int kmc_set_velocity(kmc_ctx* c, double velocity) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_set_velocity: null context");
  if (!(velocity >= 0.0) || !std::isfinite(velocity))
    return fail(c, KMC_ERR_ARG, "kmc_set_velocity: need 0 <= velocity < inf");
  if (velocity > 0.0) {
    if (c->n_sites == 0) return fail(c, KMC_ERR_STATE, "kmc_set_velocity: call kmc_set_sites first");
    KMC_TRY(check_links(c, "kmc_set_velocity"));
    KMC_HIP(c, hipSetDevice(c->cfg.device));
    if (!c->d_fly_err.p) KMC_TRY(dalloc(c, c->d_fly_err, 1));
  }
  c->velocity = velocity;
  return KMC_OK;
}

int kmc_get_flight(kmc_ctx* c, double* offset, uint8_t* heading, double* t_last) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_flight: null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_get_flight: no excitons");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t E = c->n_exc;
  if (offset) KMC_TRY(copy_out(c, offset, c->x_soff.p, E * sizeof(double)));
  if (heading) KMC_TRY(copy_out(c, heading, c->x_heading.p, E));
  if (t_last) KMC_TRY(copy_out(c, t_last, c->x_tlast.p, E * sizeof(double)));
  return KMC_OK;
}

int kmc_run(kmc_ctx* c, int64_t hops_per_exciton, double t_max, int32_t count_sites,
            double* kernel_ms) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_run: null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_run: call kmc_inject first");
  if (hops_per_exciton < 0 || std::isnan(t_max))
    return fail(c, KMC_ERR_ARG, "kmc_run: need hops_per_exciton >= 0 and t_max not NaN");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  hop_args a{};
  a.v = view_of(c);
  a.t_max = t_max;
  a.tv.n_trace = c->n_trace;
  a.tv.tr_max = c->tr_max;
  a.tv.tr_site = c->tr_site.p;
  a.tv.tr_time = c->tr_time.p;
  a.tv.tr_n = c->tr_n.p;
  KMC_HIP(c, hipEventRecord(c->ev0, c->stream));
  c->time_hops = kernel_ms != nullptr;
  c->hop_ms = -1.0;
  const int st = run_hops(c, a, hops_per_exciton, c->resort_hops, count_sites != 0,
                          c->n_bins > 0, c->n_trace > 0 && c->tr_max > 0, true);
  c->time_hops = false;
  if (st != KMC_OK) return st;
  KMC_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (kernel_ms) {
    KMC_HIP(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    KMC_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *kernel_ms = ms;
  }
  return KMC_OK;
}

int kmc_get_excitons(kmc_ctx* c, int32_t* site, double* t_next, uint64_t* hops, double* disp,
                     uint32_t* drains, uint64_t* seq_hash) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_excitons: null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_get_excitons: no excitons");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t E = c->n_exc;
  if (site) KMC_TRY(copy_out(c, site, c->x_site.p, E * sizeof(int32_t)));
  if (t_next) KMC_TRY(copy_out(c, t_next, c->x_t.p, E * sizeof(double)));
  if (hops) KMC_TRY(copy_out(c, hops, c->x_hops.p, E * sizeof(uint64_t)));
  if (drains) KMC_TRY(copy_out(c, drains, c->x_drains.p, E * sizeof(uint32_t)));
  if (seq_hash) KMC_TRY(copy_out(c, seq_hash, c->x_hash.p, E * sizeof(uint64_t)));
  if (disp) {
    std::vector<double> v(E);
    const dbuf<double>* src[3] = {&c->x_dx, &c->x_dy, &c->x_dz};
    for (int a = 0; a < 3; ++a) {
      KMC_TRY(copy_out(c, v.data(), src[a]->p, E * sizeof(double)));
      for (int64_t e = 0; e < E; ++e) disp[3 * e + a] = v[e];
    }
  }
  return KMC_OK;
}

int kmc_get_hop_counts(kmc_ctx* c, uint64_t* counts) {
  if (!c || !counts) return fail(c, KMC_ERR_ARG, "kmc_get_hop_counts: null argument");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_get_hop_counts: no excitons");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  if (c->count_pending) KMC_TRY(fold_counts(c, c->stream));
  KMC_TRY(copy_out(c, counts, c->d_counts64.p, c->n_sites * sizeof(uint64_t)));
  return KMC_OK;
}

int kmc_get_hop_time(kmc_ctx* c, double* hop_kernel_ms) {
  if (!c || !hop_kernel_ms) return fail(c, KMC_ERR_ARG, "kmc_get_hop_time: bad argument");
  if (c->hop_ms < 0.0)
    return fail(c, KMC_ERR_STATE, "kmc_get_hop_time: the last kmc_run was not timed");
  *hop_kernel_ms = c->hop_ms;
  return KMC_OK;
}

int kmc_get_run_stats(kmc_ctx* c, int64_t* launches, int64_t* live_excitons) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_run_stats: null context");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_get_run_stats: no excitons");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  if (launches) *launches = c->run_launches;
  if (live_excitons) {
    uint32_t live = 0;
    KMC_TRY(copy_out(c, &live, c->d_live.p, sizeof(uint32_t)));
    *live_excitons = c->run_launches > 0 ? (int64_t)live : -1;
  }
  return KMC_OK;
}

int kmc_get_samples(kmc_ctx* c, double* disp, uint32_t* drains, uint32_t* n_recorded) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_samples: null context");
  if (c->n_bins == 0) return fail(c, KMC_ERR_STATE, "kmc_get_samples: no sampling grid");
  const size_t nrec = (size_t)window_bins(c) * c->n_exc;
  if (disp) {
    std::vector<double> v(nrec);
    const dbuf<double>* src[3] = {&c->r_dx, &c->r_dy, &c->r_dz};
    for (int a = 0; a < 3; ++a) {
      KMC_TRY(copy_out(c, v.data(), src[a]->p, nrec * sizeof(double)));
      for (size_t r = 0; r < nrec; ++r) disp[3 * r + a] = v[r];
    }
  }
  if (drains) KMC_TRY(copy_out(c, drains, c->r_drain.p, nrec * sizeof(uint32_t)));
  if (n_recorded) KMC_TRY(copy_out(c, n_recorded, c->x_krec.p, c->n_exc * sizeof(uint32_t)));
  return KMC_OK;
}

int kmc_get_trace(kmc_ctx* c, int32_t* sites, double* times, int64_t* n_hops) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_get_trace: null context");
  if (c->n_trace == 0) return fail(c, KMC_ERR_STATE, "kmc_get_trace: no trace buffer");
  const size_t nt = (size_t)c->n_trace * c->tr_max;
  if (sites) KMC_TRY(copy_out(c, sites, c->tr_site.p, nt * sizeof(int32_t)));
  if (times) KMC_TRY(copy_out(c, times, c->tr_time.p, nt * sizeof(double)));
  if (n_hops) KMC_TRY(copy_out(c, n_hops, c->tr_n.p, c->n_trace * sizeof(int64_t)));
  return KMC_OK;
}

int kmc_reduce_samples(kmc_ctx* c, int32_t n_batches, double* out, double* kernel_ms) {
  if (!c || !out) return fail(c, KMC_ERR_ARG, "kmc_reduce_samples: null argument");
  if (c->n_bins == 0) return fail(c, KMC_ERR_STATE, "kmc_reduce_samples: no sampling grid");
  if (n_batches < 1 || n_batches > c->n_exc || n_batches > 65535)
    return fail(c, KMC_ERR_ARG, "kmc_reduce_samples: need 1 <= n_batches <= min(E, 65535)");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const size_t nout = (size_t)window_bins(c) * n_batches * 5;
  KMC_TRY(dalloc(c, c->d_reduced, nout));
  reduce_args a{};
  a.n_exc = c->n_exc;
  a.bin0 = c->bin0;
  a.n_bins = window_bins(c);
  a.n_batches = n_batches;
  a.k_rec = c->x_krec.p;
  a.rec_dx = c->r_dx.p;
  a.rec_dy = c->r_dy.p;
  a.rec_dz = c->r_dz.p;
  a.rec_drain = c->r_drain.p;
  a.out = c->d_reduced.p;
  KMC_HIP(c, hipEventRecord(c->ev0, c->stream));
  KMC_HIP(c, launch_reduce(a, c->stream));
  KMC_HIP(c, hipEventRecord(c->ev1, c->stream));
  KMC_TRY(copy_out(c, out, c->d_reduced.p, nout * sizeof(double)));
  if (kernel_ms) {
    float ms = 0.f;
    KMC_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *kernel_ms = ms;
  }
  return KMC_OK;
}

int kmc_reduce_samples_noalloc(kmc_ctx* c, int32_t n_batches, double* out, double* kernel_ms) {
  // kmc_reduce_samples with the partial sums in the device's static scratch: no allocation
  if (!c || !out) return fail(c, KMC_ERR_ARG, "kmc_reduce_samples_noalloc: null argument");
  if (c->n_bins == 0) return fail(c, KMC_ERR_STATE, "kmc_reduce_samples_noalloc: no sampling grid");
  if (n_batches < 1 || n_batches > c->n_exc || n_batches > 65535)
    return fail(c, KMC_ERR_ARG, "kmc_reduce_samples_noalloc: need 1 <= n_batches <= min(E, 65535)");
  const size_t nout = (size_t)window_bins(c) * n_batches * 5;
  if (nout > REDUCE_SCRATCH_DOUBLES)
    return fail(c, KMC_ERR_CAPACITY,
                "kmc_reduce_samples_noalloc: window bins * n_batches * 5 = " + std::to_string(nout) +
                    " exceeds the static scratch of " + std::to_string(REDUCE_SCRATCH_DOUBLES) +
                    " doubles (use kmc_reduce_samples)");
  double* ptr = scratch_of(c->cfg.device);
  if (!ptr)
    return fail(c, KMC_ERR_STATE, "kmc_reduce_samples_noalloc: no static scratch on this device");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  reduce_args a{};
  a.n_exc = c->n_exc;
  a.bin0 = c->bin0;
  a.n_bins = window_bins(c);
  a.n_batches = n_batches;
  a.k_rec = c->x_krec.p;
  a.rec_dx = c->r_dx.p;
  a.rec_dy = c->r_dy.p;
  a.rec_dz = c->r_dz.p;
  a.rec_drain = c->r_drain.p;
  a.out = ptr;  // every one of the nout partials is written by the kernel: nothing to clear
  KMC_HIP(c, hipEventRecord(c->ev0, c->stream));
  KMC_HIP(c, launch_reduce(a, c->stream));
  KMC_HIP(c, hipEventRecord(c->ev1, c->stream));
  KMC_TRY(copy_out(c, out, ptr, nout * sizeof(double)));
  if (kernel_ms) {
    float ms = 0.f;
    KMC_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *kernel_ms = ms;
  }
  return KMC_OK;
}


}  // extern "C"

// =============================================================================================
// Launch-schedule autotuning and ensembles (no reference counterpart)
// =============================================================================================
namespace {

// Copy of one context's exciton state, so that trial runs can be undone.
struct state_copy {
  uint8_t* buf = nullptr;
  int64_t n = 0;
  static constexpr size_t per = 4 + 8 + 8 + 24 + 4 + 8 + 4 + 1 + 16;  // bytes of state per exciton
};

int save_state(kmc_ctx* c, state_copy& s, bool restore) {
  const int64_t E = c->n_exc;
  if (!restore) {
    KMC_HIP(c, hipMalloc((void**)&s.buf, (size_t)E * state_copy::per));
    s.n = E;
  }
  struct part { void* p; size_t b; };
  const part parts[] = {{c->x_site.p, 4}, {c->x_t.p, 8},      {c->x_hops.p, 8}, {c->x_dx.p, 8},
                        {c->x_dy.p, 8},   {c->x_dz.p, 8},     {c->x_drains.p, 4}, {c->x_hash.p, 8},
                        {c->x_krec.p, 4}, {c->x_heading.p, 1}, {c->x_soff.p, 8},
                        {c->x_tlast.p, 8}};
  size_t off = 0;
  for (const part& q : parts) {
    void* dev = q.p;
    void* bak = s.buf + off;
    KMC_HIP(c, hipMemcpyAsync(restore ? dev : bak, restore ? bak : dev, (size_t)E * q.b,
                              hipMemcpyDeviceToDevice, c->stream));
    off += (size_t)E * q.b;
  }
  return KMC_OK;
}

}  // namespace

struct kmc_ensemble {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<kmc_ctx*> ctxs;
  std::vector<int64_t> n_exc;    // per network, fixed at creation
  std::vector<net_view> views;   // uploaded before every run
  int64_t n_total = 0, n_positions = 0;
  net_view* d_nets = nullptr;
  uint16_t *d_wave_net = nullptr, *d_exc_net = nullptr;
  int64_t *d_pbase = nullptr, *d_ebase = nullptr;
  uint64_t* d_keys = nullptr;    // 2 n_total: sort input and output
  uint32_t* d_ids = nullptr;     // 2 n_total
  void* d_sort_tmp = nullptr;
  size_t sort_bytes = 0;
  int end_bit = 33;
  int64_t resort_hops = 625;
  uint32_t* d_live = nullptr;    // excitons of all members able to hop after the last launch
  int64_t max_exc = 0;           // the largest member (bounds a launch's count per slot)
};

namespace {

thread_local std::string g_ens_err;
int efail(int code, const std::string& m) {
  g_ens_err = m;
  return code;
}
#define ENS_HIP(call)                                                                            \
  do {                                                                                           \
    const hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return efail(KMC_ERR_HIP, std::string(__func__) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Members unchanged since creation, their pending work finished, their current views uploaded
// (modify may adjust each view, e.g. for trial runs).
template <typename F>
int ens_upload(kmc_ensemble* m, F modify) {
  for (size_t k = 0; k < m->ctxs.size(); ++k) {
    kmc_ctx* c = m->ctxs[k];
    if (c->n_exc != m->n_exc[k] || c->csr_state != 2)
      return efail(KMC_ERR_STATE, "kmc_ensemble: a member was re-injected or rebuilt since "
                                  "kmc_ensemble_create");
    if (c->n_trace > 0) return efail(KMC_ERR_STATE, "kmc_ensemble: members may not record traces");
    if (c->exc_velocity > 0.0)
      return efail(KMC_ERR_STATE, "kmc_ensemble: members must have velocity 0 (run drifting "
                                  "networks with kmc_run)");
    ENS_HIP(hipStreamSynchronize(c->stream));
    m->views[k] = view_of(c);
    m->views[k].live = m->d_live;
    modify(k, m->views[k]);
  }
  ENS_HIP(hipMemcpyAsync(m->d_nets, m->views.data(), m->views.size() * sizeof(net_view),
                         hipMemcpyHostToDevice, m->stream));
  return KMC_OK;
}

// As run_hops, for all members in one launch per sub-launch.  With `fold` the counts go to the
// members' own copies, folded into their totals as needed; without, to the views' scratch.
int ens_run_hops(kmc_ensemble* m, int64_t hops, int64_t chunk, bool count, bool sample,
                 double t_max, bool fold) {
  ens_args a{};
  a.nets = m->d_nets;
  a.wave_net = m->d_wave_net;
  a.pbase = m->d_pbase;
  a.ebase = m->d_ebase;
  a.perm = m->d_ids + m->n_total;
  a.n_positions = (uint64_t)m->n_positions;
  a.t_max = t_max;
  int64_t step = chunk > 0 ? chunk : std::max<int64_t>(hops, 1);
  if (count) step = std::min<int64_t>(step, (int64_t)(COUNT_COPY_MAX / (uint64_t)m->max_exc));
  int64_t done = 0, launches = 0;
  do {
    const int64_t sub = std::min(step, hops - done);
    if (count && fold) {
      for (kmc_ctx* c : m->ctxs) {
        const uint64_t add = (uint64_t)c->n_exc * (uint64_t)sub;
        if (c->count_pending + add > COUNT_COPY_MAX) {
          ENS_HIP(launch_fold_counts(c->n_units / ROW_UNITS, c->count_stride, c->d_blocks.p,
                                     c->d_site_count.p, c->d_counts64.p, m->stream));
          c->count_pending = 0;
        }
        c->count_pending += add;
      }
    }
    ENS_HIP(launch_ens_keys(m->n_total, m->d_nets, m->d_exc_net, m->d_ebase, m->d_keys, m->d_ids,
                            m->d_live, m->stream));
    size_t have = m->sort_bytes;
    ENS_HIP(hipcub::DeviceRadixSort::SortPairs(m->d_sort_tmp, have, m->d_keys,
                                               m->d_keys + m->n_total, m->d_ids,
                                               m->d_ids + m->n_total, (int)m->n_total, 0,
                                               m->end_bit, m->stream));
    a.hop_budget = (uint64_t)sub;
    ENS_HIP(launch_hop_ens(a, count, sample, m->stream));
    done += sub;
    ++launches;
    if (done < hops && live_check_due(launches)) {
      uint32_t live = 0;
      ENS_HIP(hipMemcpyAsync(&live, m->d_live, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             m->stream));
      ENS_HIP(hipStreamSynchronize(m->stream));
      if (live == 0u) break;
    }
  } while (done < hops);
  return KMC_OK;
}

}  // namespace

extern "C" {

int kmc_autotune_schedule(kmc_ctx* c, int64_t trial_hops, int32_t n_cand, const int64_t* cand,
                          int64_t* best, double* hops_per_s) {
  if (!c || trial_hops < 1 || n_cand < 1 || !cand)
    return fail(c, KMC_ERR_ARG, "kmc_autotune_schedule: bad argument");
  if (c->n_exc == 0) return fail(c, KMC_ERR_STATE, "kmc_autotune_schedule: call kmc_inject first");
  for (int i = 0; i < n_cand; ++i)
    if (cand[i] < 0) return fail(c, KMC_ERR_ARG, "kmc_autotune_schedule: negative candidate");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  state_copy bak;
  KMC_TRY(save_state(c, bak, false));
  uint32_t* counts = nullptr;
  hipError_t he = hipMalloc((void**)&counts, (size_t)KMC_COUNT_COPIES * c->count_stride * 4);
  if (he != hipSuccess) {
    (void)hipFree(bak.buf);
    return hip_fail(c, he, "kmc_autotune_schedule");
  }
  hop_args a{};
  a.v = view_of(c);
  a.v.n_bins = 0;           // trials record no samples
  a.v.site_count = counts;  // and count into scratch
  a.t_max = INFINITY;
  int64_t pick = cand[0];
  double best_rate = -1.0;
  int st = KMC_OK;
  for (int i = 0; i < n_cand && st == KMC_OK; ++i) {
    st = save_state(c, bak, true);
    if (st == KMC_OK && hipEventRecord(c->ev0, c->stream) != hipSuccess) st = KMC_ERR_HIP;
    if (st == KMC_OK) st = run_hops(c, a, trial_hops, cand[i], true, false, false, false);
    if (st == KMC_OK && hipEventRecord(c->ev1, c->stream) != hipSuccess) st = KMC_ERR_HIP;
    float ms = 0.f;
    if (st == KMC_OK && (hipEventSynchronize(c->ev1) != hipSuccess ||
                         hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess))
      st = KMC_ERR_HIP;
    const double rate = (double)c->n_exc * (double)trial_hops / (ms * 1e-3);
    if (hops_per_s) hops_per_s[i] = rate;
    if (st == KMC_OK && rate > best_rate) {
      best_rate = rate;
      pick = cand[i];
    }
  }
  if (st == KMC_OK) st = save_state(c, bak, true);
  if (st == KMC_OK && hipStreamSynchronize(c->stream) != hipSuccess) st = KMC_ERR_HIP;
  (void)hipFree(counts);
  (void)hipFree(bak.buf);
  if (st != KMC_OK) return fail(c, st, "kmc_autotune_schedule: a trial failed");
  c->resort_hops = pick;
  if (best) *best = pick;
  return KMC_OK;
}

const char* kmc_ensemble_last_error(void) { return g_ens_err.c_str(); }

int kmc_ensemble_destroy(kmc_ensemble* m) {
  if (!m) return KMC_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* bufs[] = {m->d_nets, m->d_wave_net, m->d_exc_net, m->d_pbase, m->d_ebase,
                  m->d_keys, m->d_ids,      m->d_sort_tmp, m->d_live};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (m->ev0) (void)hipEventDestroy(m->ev0);
  if (m->ev1) (void)hipEventDestroy(m->ev1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return KMC_OK;
}

int kmc_ensemble_create(int32_t n_networks, kmc_ctx* const* networks, kmc_ensemble** out) {
  if (!networks || !out || n_networks < 1 || n_networks > 65535)
    return efail(KMC_ERR_ARG, "kmc_ensemble_create: need 1 <= n_networks <= 65535");
  *out = nullptr;
  const int dev = networks[0] ? networks[0]->cfg.device : 0;
  std::vector<int64_t> pbase(n_networks), ebase(n_networks);
  int64_t np = 0, ne = 0;
  for (int k = 0; k < n_networks; ++k) {
    kmc_ctx* c = networks[k];
    if (!c) return efail(KMC_ERR_ARG, "kmc_ensemble_create: null member");
    if (c->cfg.device != dev) return efail(KMC_ERR_ARG, "kmc_ensemble_create: members on different devices");
    if (c->csr_state != 2 || c->n_exc == 0)
      return efail(KMC_ERR_STATE, "kmc_ensemble_create: every member needs its CSR and excitons");
    pbase[k] = np;
    ebase[k] = ne;
    np += (c->n_exc + 255) / 256 * 256;  // network ranges start on workgroup boundaries
    ne += c->n_exc;
  }
  if (ne >= ((int64_t)1 << 31) || np >= ((int64_t)1 << 32))
    return efail(KMC_ERR_CAPACITY, "kmc_ensemble_create: more than 2^31 excitons");
  std::vector<uint16_t> wave_net(np / 64), exc_net(ne);
  for (int k = 0; k < n_networks; ++k) {
    const int64_t E = networks[k]->n_exc;
    for (int64_t w = pbase[k] / 64; w < (pbase[k] + (E + 255) / 256 * 256) / 64; ++w) wave_net[w] = (uint16_t)k;
    for (int64_t e = 0; e < E; ++e) exc_net[ebase[k] + e] = (uint16_t)k;
  }
  kmc_ensemble* m = new kmc_ensemble();
  m->device = dev;
  m->ctxs.assign(networks, networks + n_networks);
  for (kmc_ctx* c : m->ctxs) {
    m->n_exc.push_back(c->n_exc);
    m->max_exc = std::max<int64_t>(m->max_exc, c->n_exc);
  }
  m->views.resize(n_networks);
  m->n_total = ne;
  m->n_positions = np;
  int bits = 0;
  while ((1 << bits) < n_networks) ++bits;
  m->end_bit = 32 + std::max(bits, 1);
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&m->ev0);
  if (e == hipSuccess) e = hipEventCreate(&m->ev1);
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_nets, n_networks * sizeof(net_view));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_wave_net, wave_net.size() * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_exc_net, exc_net.size() * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_pbase, n_networks * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_ebase, n_networks * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_keys, 2 * ne * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_ids, 2 * ne * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, m->sort_bytes, m->d_keys, m->d_keys + ne,
                                           m->d_ids, m->d_ids + ne, (int)ne, 0, m->end_bit,
                                           m->stream);
  if (e == hipSuccess) e = hipMalloc(&m->d_sort_tmp, std::max<size_t>(m->sort_bytes, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_live, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemsetAsync(m->d_live, 0, sizeof(uint32_t), m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m->d_wave_net, wave_net.data(), wave_net.size() * 2, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m->d_exc_net, exc_net.data(), exc_net.size() * 2, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m->d_pbase, pbase.data(), n_networks * 8, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m->d_ebase, ebase.data(), n_networks * 8, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    kmc_ensemble_destroy(m);
    return efail(KMC_ERR_HIP, std::string("kmc_ensemble_create: ") + hipGetErrorString(e));
  }
  *out = m;
  return KMC_OK;
}

int kmc_ensemble_set_hop_schedule(kmc_ensemble* m, int64_t resort_hops) {
  if (!m || resort_hops < 0) return efail(KMC_ERR_ARG, "kmc_ensemble_set_hop_schedule: bad argument");
  m->resort_hops = resort_hops;
  return KMC_OK;
}

int kmc_ensemble_run(kmc_ensemble* m, int64_t hops_per_exciton, double t_max, int32_t count_sites,
                     double* kernel_ms) {
  if (!m || hops_per_exciton < 0 || std::isnan(t_max))
    return efail(KMC_ERR_ARG, "kmc_ensemble_run: bad argument");
  ENS_HIP(hipSetDevice(m->device));
  bool sample = false;
  for (kmc_ctx* c : m->ctxs) sample = sample || c->n_bins > 0;
  int s = ens_upload(m, [](size_t, net_view&) {});
  if (s != KMC_OK) return s;
  ENS_HIP(hipEventRecord(m->ev0, m->stream));
  s = ens_run_hops(m, hops_per_exciton, m->resort_hops, count_sites != 0, sample, t_max, true);
  if (s != KMC_OK) return s;
  ENS_HIP(hipEventRecord(m->ev1, m->stream));
  ENS_HIP(hipStreamSynchronize(m->stream));
  if (kernel_ms) {
    float ms = 0.f;
    ENS_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    *kernel_ms = ms;
  }
  return KMC_OK;
}

int kmc_ensemble_autotune(kmc_ensemble* m, int64_t trial_hops, int32_t n_cand, const int64_t* cand,
                          int64_t* best, double* hops_per_s) {
  if (!m || trial_hops < 1 || n_cand < 1 || !cand)
    return efail(KMC_ERR_ARG, "kmc_ensemble_autotune: bad argument");
  for (int i = 0; i < n_cand; ++i)
    if (cand[i] < 0) return efail(KMC_ERR_ARG, "kmc_ensemble_autotune: negative candidate");
  ENS_HIP(hipSetDevice(m->device));
  const size_t n = m->ctxs.size();
  std::vector<state_copy> bak(n);
  std::vector<uint32_t*> counts(n, nullptr);
  auto release = [&]() {
    for (size_t k = 0; k < n; ++k) {
      if (bak[k].buf) (void)hipFree(bak[k].buf);
      if (counts[k]) (void)hipFree(counts[k]);
    }
  };
  int st = KMC_OK;
  for (size_t k = 0; k < n && st == KMC_OK; ++k) {
    kmc_ctx* c = m->ctxs[k];
    st = save_state(c, bak[k], false);
    if (st == KMC_OK && hipMalloc((void**)&counts[k], (size_t)KMC_COUNT_COPIES * c->count_stride * 4) != hipSuccess)
      st = KMC_ERR_ALLOC;
    if (st == KMC_OK && hipStreamSynchronize(c->stream) != hipSuccess) st = KMC_ERR_HIP;
  }
  if (st == KMC_OK)
    st = ens_upload(m, [&](size_t k, net_view& v) {
      v.n_bins = 0;
      v.site_count = counts[k];
    });
  int64_t pick = cand[0];
  double best_rate = -1.0;
  for (int i = 0; i < n_cand && st == KMC_OK; ++i) {
    for (size_t k = 0; k < n && st == KMC_OK; ++k) st = save_state(m->ctxs[k], bak[k], true);
    for (size_t k = 0; k < n && st == KMC_OK; ++k)
      if (hipStreamSynchronize(m->ctxs[k]->stream) != hipSuccess) st = KMC_ERR_HIP;
    float ms = 0.f;
    if (st == KMC_OK && hipEventRecord(m->ev0, m->stream) != hipSuccess) st = KMC_ERR_HIP;
    if (st == KMC_OK) st = ens_run_hops(m, trial_hops, cand[i], true, false, INFINITY, false);
    if (st == KMC_OK && (hipEventRecord(m->ev1, m->stream) != hipSuccess ||
                         hipEventSynchronize(m->ev1) != hipSuccess ||
                         hipEventElapsedTime(&ms, m->ev0, m->ev1) != hipSuccess))
      st = KMC_ERR_HIP;
    const double rate = (double)m->n_total * (double)trial_hops / (ms * 1e-3);
    if (hops_per_s) hops_per_s[i] = rate;
    if (st == KMC_OK && rate > best_rate) {
      best_rate = rate;
      pick = cand[i];
    }
  }
  for (size_t k = 0; k < n && st == KMC_OK; ++k) st = save_state(m->ctxs[k], bak[k], true);
  for (size_t k = 0; k < n && st == KMC_OK; ++k)
    if (hipStreamSynchronize(m->ctxs[k]->stream) != hipSuccess) st = KMC_ERR_HIP;
  release();
  if (st != KMC_OK) return efail(st, "kmc_ensemble_autotune: a trial failed");
  m->resort_hops = pick;
  if (best) *best = pick;
  return KMC_OK;
}

}  // extern "C"

// =============================================================================================
// Contact (steady-state) mode: init() / step() / save_metrics() / repopulate_contacts()
// (src/monte_carlo/monte_carlo.h:157-195, 274-316, 343-355, 443-684), velocity 0.
// =============================================================================================
namespace {

contact_slots contact_buf(kmc_ctx* c, int b) {
  return contact_slots{c->c_site[b].p, c->c_t[b].p,    c->c_hops[b].p, c->c_id[b].p,
                       c->c_y[b].p,    c->c_head[b].p, c->c_soff[b].p, c->c_tlast[b].p};
}

// The order in which a step's threads take the live slots.  Slots sort by the eighth of the film
// their site lies in (the eighths of the sorted slots go to the eight XCDs, so each XCD's L2
// serves one compact region), then by the site's chirality, then by the site's row.  The
// chirality sets the rates a particle meets along its tube, so the particles of one wave make
// similar numbers of hops in a step and fewer of its lanes idle while its last particle hops.
// A film whose keys would not fit 32 bits sorts by the row alone.
contact_key_layout contact_key_layout_of(const kmc_ctx* c, uint32_t lines) {
  const auto bits = [](uint64_t v) {  // bits that hold the values [0, v]
    int b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
  };
  contact_key_layout kl{};
  kl.lines = lines;
  uint8_t top = 0;
  for (const uint8_t ch : c->h_chir) top = std::max(top, ch);
  kl.chir_bits = top ? bits(top) : 0;
  kl.row_bits = bits(lines - 1u);
  if (4 + kl.chir_bits + kl.row_bits <= 32) {
    kl.regions = 1;
    kl.dead = (uint32_t)KMC_COUNT_COPIES << (kl.chir_bits + kl.row_bits);
    kl.key_bits = 4 + kl.chir_bits + kl.row_bits;
  } else {
    kl.regions = 0;
    kl.chir_bits = 0;
    kl.dead = lines;
    kl.key_bits = std::min(bits(lines), 32);
    kl.row_bits = kl.key_bits;
  }
  return kl;
}

contact_args contact_args_of(kmc_ctx* c) {
  contact_args a{};
  a.blocks = c->d_blocks.p;
  a.roff = c->d_roff.p;
  a.py = c->d_pos[1].p;
  a.seed = c->cfg.seed;
  a.s = contact_buf(c, c->c_buf);
  a.n_live = c->c_live.p;
  a.capacity = c->c_cap;
  a.n_seg = c->c_nseg;
  a.ymin = c->c_ymin;
  a.dy = c->c_dy;
  a.err = c->c_err.p;
  a.hops_made = reinterpret_cast<unsigned long long*>(c->c_made.p);
  if (c->c_vel > 0.0) {
    a.fv = view_of(c);
    a.fv.velocity = c->c_vel;
  }
  return a;
}

// Adds the elapsed time of event pair q to c_step_ms if it holds an unread step.
int contact_read_pair(kmc_ctx* c, int q) {
  if (!c->c_ev_busy[q]) return KMC_OK;
  float ms = 0.0f;
  KMC_HIP(c, hipEventSynchronize(c->c_ev[2 * q + 1]));
  KMC_HIP(c, hipEventElapsedTime(&ms, c->c_ev[2 * q], c->c_ev[2 * q + 1]));
  c->c_step_ms += (double)ms;
  c->c_ev_busy[q] = 0;
  return KMC_OK;
}

int contact_check_err(kmc_ctx* c, const char* who) {
  int32_t err = 0;
  KMC_HIP(c, hipMemcpy(&err, c->c_err.p, sizeof(err), hipMemcpyDeviceToHost));
  if (err & CONTACT_ERR_HOPS)
    return fail(c, KMC_ERR_CAPACITY, std::string(who) + ": a particle reached max_hops_per_step "
                                         "within one step (shorten dt)");
  if (err & CONTACT_ERR_CAPACITY)
    return fail(c, KMC_ERR_CAPACITY, std::string(who) + ": more particles than the capacity");
  if (err & CONTACT_ERR_IDS)
    return fail(c, KMC_ERR_CAPACITY, std::string(who) + ": more than 2^32 particles injected");
  if (err & CONTACT_ERR_FLY)
    return fail(c, KMC_ERR_CAPACITY, std::string(who) + ": a flight passed more than 2^22 sites "
                                         "within one step (shorten dt)");
  return KMC_OK;
}

}  // namespace

extern "C" {

int kmc_contact_init(kmc_ctx* c, int32_t n_seg, int64_t c1_pop, int64_t c2_pop, int64_t capacity) {
#pragma clang fp contract(off)
  const char* who = "kmc_contact_init";
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_contact_init: null context");
  if (c->csr_state != 2 || !c->d_blocks.p)
    return fail(c, KMC_ERR_STATE, "kmc_contact_init: CSR not complete");
  if (n_seg < 2 || n_seg > CONTACT_MAX_SEG)
    return fail(c, KMC_ERR_ARG, "kmc_contact_init: need 2 <= n_seg <= 1024");
  if (c1_pop < 0 || c2_pop < 0 || capacity <= 0 || capacity > (int64_t)0x7FFFFFFF)
    return fail(c, KMC_ERR_ARG, "kmc_contact_init: need populations >= 0, 0 < capacity < 2^31");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  if (c->velocity > 0.0) {
    KMC_TRY(check_links(c, who));
    for (uint8_t f : c->h_flags)
      if (f & KMC_SITE_DRAIN)
        return fail(c, KMC_ERR_STATE, "kmc_contact_init: with velocity > 0 a flight would "
                                      "re-inject at Kubo drain sites; clear KMC_SITE_DRAIN");
  }
  const std::vector<double>& X = c->h_pos[0];
  const std::vector<double>& Y = c->h_pos[1];
  const std::vector<double>& Z = c->h_pos[2];
  const int64_t N = c->n_sites;
  // the domain is the sites' bounding box (find_simulation_domain, monte_carlo.h:328-340)
  const double ymin = c->bb_lo[1], ymax = c->bb_hi[1];
  const double dy = (ymax - ymin) / double(n_seg);
  // get_area (monte_carlo.h:644-684), with its min-else-max update
  std::vector<double> xmax(n_seg, c->bb_lo[0]), xmin(n_seg, c->bb_hi[0]);
  std::vector<double> zmax(n_seg, c->bb_lo[2]), zmin(n_seg, c->bb_hi[2]);
  for (int64_t s = 0; s < N; ++s) {
    int i = (int)((Y[s] - ymin) / dy);
    i = i < 0 ? 0 : (i < n_seg ? i : n_seg - 1);
    if (xmin[i] > X[s]) xmin[i] = X[s];
    else if (xmax[i] < X[s]) xmax[i] = X[s];
    if (zmin[i] > Z[s]) zmin[i] = Z[s];
    else if (zmax[i] < Z[s]) zmax[i] = Z[s];
  }
  c->c_area.assign(n_seg, 0.0);
  for (int i = 0; i < n_seg; ++i) c->c_area[i] = (zmax[i] - zmin[i]) * (xmax[i] - xmin[i]);
  // site lists: contact_scats(i = 1, n_seg) (monte_carlo.h:494-516, closed bounds), then the
  // sections of create_particles (monte_carlo.h:274-316, [y1, y1 + dy))
  std::vector<int32_t> lists;
  std::vector<int64_t> off, cnt;
  for (int ci : {1, (int)n_seg}) {
    const double y1 = ymin + double(ci - 1) * dy;
    const double y2 = ymin + double(ci) * dy;
    off.push_back((int64_t)lists.size());
    for (int64_t s = 0; s < N; ++s)
      if (Y[s] >= y1 && Y[s] <= y2) lists.push_back((int32_t)s);
    cnt.push_back((int64_t)lists.size() - off.back());
  }
  const double dp = double(c2_pop - c1_pop) / (double(n_seg) - 1);
  std::vector<int64_t> n_init(n_seg);
  int64_t total = 0;
  for (int i = 0; i < n_seg; ++i) {
    n_init[i] = (int64_t)std::round(double(c1_pop) + double(i) * dp);
    const double y1 = ymin + double(i) * dy;
    const double y2 = y1 + dy;
    off.push_back((int64_t)lists.size());
    for (int64_t s = 0; s < N; ++s)
      if (y1 <= Y[s] && Y[s] < y2) lists.push_back((int32_t)s);
    cnt.push_back((int64_t)lists.size() - off.back());
    if (n_init[i] > 0 && cnt.back() == 0)
      return fail(c, KMC_ERR_ARG, "kmc_contact_init: section " + std::to_string(i) +
                                      " has no sites for its initial particles");
    total += n_init[i];
  }
  if ((c1_pop > 0 && cnt[0] == 0) || (c2_pop > 0 && cnt[1] == 0))
    return fail(c, KMC_ERR_ARG, "kmc_contact_init: a contact holding particles has no sites");
  if (total > capacity || c1_pop + c2_pop > capacity)
    return fail(c, KMC_ERR_CAPACITY, "kmc_contact_init: capacity below the populations");
  c->c_nseg = n_seg;
  c->c_pop[0] = c1_pop;
  c->c_pop[1] = c2_pop;
  c->c_cap = capacity;
  c->c_ymin = ymin;
  c->c_dy = dy;
  c->c_time = 0.0;
  c->c_vel = c->velocity;
  // repopulate_contacts (monte_carlo.h:443-455): [ymin, ymin + dy] and [ymin + (n-1) dy, ymax]
  c->c_rm_lo[0] = ymin;
  c->c_rm_hi[0] = ymin + dy;
  c->c_rm_lo[1] = ymin + double(n_seg - 1) * dy;
  c->c_rm_hi[1] = ymax;
  for (int k = 0; k < 2; ++k) {
    c->c_list_off[k] = off[k];
    c->c_list_n[k] = cnt[k];
  }
  KMC_TRY(dalloc(c, c->c_sites, std::max<size_t>(1, lists.size())));
  if (!lists.empty())
    KMC_HIP(c, hipMemcpy(c->c_sites.p, lists.data(), lists.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  for (int b = 0; b < 2; ++b) {
    KMC_TRY(dalloc(c, c->c_site[b], capacity));
    KMC_TRY(dalloc(c, c->c_t[b], capacity));
    KMC_TRY(dalloc(c, c->c_hops[b], capacity));
    KMC_TRY(dalloc(c, c->c_id[b], capacity));
    KMC_TRY(dalloc(c, c->c_y[b], capacity));
    KMC_TRY(dalloc(c, c->c_head[b], capacity));
    KMC_TRY(dalloc(c, c->c_soff[b], capacity));
    KMC_TRY(dalloc(c, c->c_tlast[b], capacity));
  }
  KMC_TRY(dalloc(c, c->c_keep, capacity));
  KMC_TRY(dalloc(c, c->c_off, capacity + 1));
  KMC_TRY(dalloc(c, c->c_scan_tmp, scan_tmp_elems(capacity)));
  KMC_TRY(dalloc(c, c->c_live, 1));
  KMC_TRY(dalloc(c, c->c_next, 1));
  KMC_TRY(dalloc(c, c->c_err, 1));
  KMC_TRY(dalloc(c, c->c_made, 1));
  KMC_HIP(c, hipMemsetAsync(c->c_made.p, 0, sizeof(uint64_t), c->stream));
  c->c_step_ms = 0.0;
  if (c->c_ev.empty()) {
    c->c_ev.assign(2 * CONTACT_EV_PAIRS, nullptr);
    for (hipEvent_t& e : c->c_ev) KMC_HIP(c, hipEventCreate(&e));
  }
  c->c_ev_busy.assign(CONTACT_EV_PAIRS, 0);
  if (c->c_live_ev.empty()) {
    c->c_live_ev.assign(CONTACT_LIVE_LAG, nullptr);
    for (hipEvent_t& e : c->c_live_ev) KMC_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (!c->c_live_host)
    KMC_HIP(c, hipHostMalloc((void**)&c->c_live_host, CONTACT_LIVE_LAG * sizeof(int64_t)));
  KMC_HIP(c, hipMemsetAsync(c->c_live.p, 0, sizeof(int64_t), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->c_next.p, 0, sizeof(uint64_t), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->c_err.p, 0, sizeof(int32_t), c->stream));
  c->c_buf = 0;
  // the initial linear profile, section by section (create_particles)
  contact_args a = contact_args_of(c);
  for (int i = 0; i < n_seg; ++i)
    if (n_init[i] > 0)
      KMC_HIP(c, launch_contact_inject(a, c->c_live.p, n_init[i], c->c_sites.p + off[2 + i],
                                       cnt[2 + i], c->c_next.p, 0.0, c->c_live.p, c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  KMC_TRY(contact_check_err(c, who));
  c->c_on = true;
  return KMC_OK;
}

int kmc_contact_run(kmc_ctx* c, double dt, int64_t n_steps, uint64_t max_hops_per_step,
                    int32_t* population, int32_t* crossings, int64_t* n_live) {
#pragma clang fp contract(off)
  const char* who = "kmc_contact_run";
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_contact_run: null context");
  if (!c->c_on) return fail(c, KMC_ERR_STATE, "kmc_contact_run: call kmc_contact_init first");
  if (!(dt > 0.0) || !std::isfinite(dt) || n_steps < 0 || max_hops_per_step == 0)
    return fail(c, KMC_ERR_ARG, "kmc_contact_run: need dt > 0, n_steps >= 0, max_hops > 0");
  if (c->csr_state != 2 || !c->d_blocks.p)
    return fail(c, KMC_ERR_STATE, "kmc_contact_run: CSR not complete");
  if (n_steps == 0) return KMC_OK;
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  const int64_t S = c->c_nseg;
  KMC_TRY(dalloc(c, c->c_pop_out, (size_t)n_steps * S));
  KMC_TRY(dalloc(c, c->c_cur_out, (size_t)n_steps * (S - 1)));
  KMC_TRY(dalloc(c, c->c_live_out, (size_t)n_steps));
  KMC_HIP(c, hipMemsetAsync(c->c_pop_out.p, 0, c->c_pop_out.n * sizeof(int32_t), c->stream));
  KMC_HIP(c, hipMemsetAsync(c->c_cur_out.p, 0, c->c_cur_out.n * sizeof(int32_t), c->stream));
  // errors are this call's: a hop cap reached leaves every particle consistent (the one that
  // stopped continues in the next step), so a retry with a larger cap may go on
  KMC_HIP(c, hipMemsetAsync(c->c_err.p, 0, sizeof(int32_t), c->stream));
  // each step's threads take the live slots in the order of contact_key_layout_of (a radix sort;
  // dead slots sort last): draws are keyed by particle id and hop count, so the order changes no
  // result
  const int64_t cap = c->c_cap;
  const contact_key_layout kl = contact_key_layout_of(c, (uint32_t)(c->n_units / ROW_UNITS));
  const int key_bits = kl.key_bits;
  KMC_TRY(dalloc(c, c->c_skeys, 2 * (size_t)cap));
  KMC_TRY(dalloc(c, c->c_sids, 2 * (size_t)cap));
  size_t sort_need = 0;
  KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_need, c->c_skeys.p,
                                                c->c_skeys.p + cap, c->c_sids.p, c->c_sids.p + cap,
                                                (int)cap, 0, key_bits, c->stream));
  KMC_TRY(dalloc(c, c->c_sort_tmp, std::max<size_t>(sort_need, 1)));
  // Step k covers the slots [0, bound): at most the live count, which a step raises by at most
  // the two contacts' injections.  The count is read exactly here, then after every step into
  // pinned memory, and step k uses the one after step k - CONTACT_LIVE_LAG.
  const int64_t inject = c->c_pop[0] + c->c_pop[1];
  int64_t live0 = 0;
  KMC_HIP(c, hipMemcpyAsync(c->c_live_host, c->c_live.p, sizeof(int64_t), hipMemcpyDeviceToHost,
                            c->stream));
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  live0 = c->c_live_host[0];
  for (int64_t k = 0; k < n_steps; ++k) {
    // monte_carlo::step, then save_metrics (main.cpp:95-97): every particle to the step's end
    int64_t bound = live0 + k * inject;
    if (k >= CONTACT_LIVE_LAG) {
      const int slot = (int)((k - CONTACT_LIVE_LAG) % CONTACT_LIVE_LAG);
      KMC_HIP(c, hipEventSynchronize(c->c_live_ev[slot]));
      bound = c->c_live_host[slot] + (CONTACT_LIVE_LAG - 1) * inject;
    }
    bound = std::min(bound, cap);
    contact_args a = contact_args_of(c);
    a.t_end = c->c_time + dt;
    a.max_hops = max_hops_per_step;
    a.pop = c->c_pop_out.p + k * S;
    a.cur = c->c_cur_out.p + k * (S - 1);
    KMC_HIP(c, launch_contact_keys(bound, c->c_live.p, c->c_site[c->c_buf].p, c->d_roff.p,
                                   c->d_chir.p, kl, c->c_skeys.p, c->c_sids.p, c->stream));
    size_t have = c->c_sort_tmp.n;
    if (bound > 0)
      KMC_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->c_sort_tmp.p, have, c->c_skeys.p,
                                                    c->c_skeys.p + cap, c->c_sids.p,
                                                    c->c_sids.p + cap, (int)bound, 0, key_bits,
                                                    c->stream));
    a.perm = c->c_sids.p + cap;
    const int q = (int)(k % CONTACT_EV_PAIRS);
    KMC_TRY(contact_read_pair(c, q));
    KMC_HIP(c, hipEventRecord(c->c_ev[2 * q], c->stream));
    KMC_HIP(c, launch_contact_step(a, c->c_vel > 0.0, bound, c->stream));
    KMC_HIP(c, hipEventRecord(c->c_ev[2 * q + 1], c->stream));
    c->c_ev_busy[q] = 1;
    c->c_time = a.t_end;  // _time += dt (monte_carlo.h:354)
    // repopulate_contacts (monte_carlo.h:443-491): contact 1, then contact n_seg; the kept
    // particles keep their order, the new ones follow them
    int64_t n_side = bound;  // slots the live ones lie in: contact 1's injection may add some
    for (int side = 0; side < 2; ++side) {
      a.s = contact_buf(c, c->c_buf);
      KMC_HIP(c, launch_contact_keep(a, n_side, c->c_rm_lo[side], c->c_rm_hi[side], c->c_keep.p,
                                     c->stream));
      KMC_HIP(c, exclusive_scan(c->c_keep.p, c->c_off.p, n_side, c->c_scan_tmp.p, c->stream));
      KMC_HIP(c, launch_contact_compact(n_side, contact_buf(c, c->c_buf),
                                        contact_buf(c, 1 - c->c_buf), c->c_keep.p, c->c_off.p,
                                        c->stream));
      c->c_buf = 1 - c->c_buf;
      a.s = contact_buf(c, c->c_buf);
      KMC_HIP(c, launch_contact_inject(a, c->c_off.p + n_side, c->c_pop[side],
                                       c->c_sites.p + c->c_list_off[side], c->c_list_n[side],
                                       c->c_next.p, c->c_time, c->c_live.p, c->stream));
      n_side = std::min(cap, n_side + c->c_pop[side]);
    }
    KMC_HIP(c, hipMemcpyAsync(c->c_live_out.p + k, c->c_live.p, sizeof(int64_t),
                              hipMemcpyDeviceToDevice, c->stream));
    const int slot = (int)(k % CONTACT_LIVE_LAG);
    KMC_HIP(c, hipMemcpyAsync(c->c_live_host + slot, c->c_live.p, sizeof(int64_t),
                              hipMemcpyDeviceToHost, c->stream));
    KMC_HIP(c, hipEventRecord(c->c_live_ev[slot], c->stream));
  }
  KMC_HIP(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < CONTACT_EV_PAIRS; ++q) KMC_TRY(contact_read_pair(c, q));
  KMC_TRY(contact_check_err(c, who));
  if (population)
    KMC_HIP(c, hipMemcpy(population, c->c_pop_out.p, c->c_pop_out.n * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
  if (crossings)
    KMC_HIP(c, hipMemcpy(crossings, c->c_cur_out.p, c->c_cur_out.n * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
  if (n_live)
    KMC_HIP(c, hipMemcpy(n_live, c->c_live_out.p, (size_t)n_steps * sizeof(int64_t),
                         hipMemcpyDeviceToHost));
  return KMC_OK;
}

int kmc_contact_stats(kmc_ctx* c, uint64_t* hops, double* step_kernel_ms) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_contact_stats: null context");
  if (!c->c_on) return fail(c, KMC_ERR_STATE, "kmc_contact_stats: call kmc_contact_init first");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  if (hops) {
    KMC_HIP(c, hipStreamSynchronize(c->stream));
    KMC_HIP(c, hipMemcpy(hops, c->c_made.p, sizeof(uint64_t), hipMemcpyDeviceToHost));
  }
  if (step_kernel_ms) *step_kernel_ms = c->c_step_ms;
  return KMC_OK;
}

int kmc_contact_sections(const kmc_ctx* c, double* area, double* ymin, double* dy, double* time) {
  if (!c) return fail(nullptr, KMC_ERR_ARG, "kmc_contact_sections: null context");
  if (!c->c_on) return fail(const_cast<kmc_ctx*>(c), KMC_ERR_STATE,
                            "kmc_contact_sections: call kmc_contact_init first");
  if (area) std::copy(c->c_area.begin(), c->c_area.end(), area);
  if (ymin) *ymin = c->c_ymin;
  if (dy) *dy = c->c_dy;
  if (time) *time = c->c_time;
  return KMC_OK;
}

int kmc_contact_particles(kmc_ctx* c, int64_t capacity, int64_t* n, uint32_t* id, int32_t* site,
                          double* t_next, uint64_t* hops, double* y) {
  if (!c || !n) return fail(c, KMC_ERR_ARG, "kmc_contact_particles: null argument");
  if (!c->c_on) return fail(c, KMC_ERR_STATE, "kmc_contact_particles: call kmc_contact_init first");
  KMC_HIP(c, hipSetDevice(c->cfg.device));
  int64_t live = 0;
  KMC_HIP(c, hipMemcpy(&live, c->c_live.p, sizeof(live), hipMemcpyDeviceToHost));
  *n = live;
  if (!id && !site && !t_next && !hops && !y) return KMC_OK;  // size query
  if (capacity < live) return fail(c, KMC_ERR_CAPACITY, "kmc_contact_particles: capacity below n");
  const int b = c->c_buf;
  if (live > 0) {
    if (id) KMC_HIP(c, hipMemcpy(id, c->c_id[b].p, live * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (site) KMC_HIP(c, hipMemcpy(site, c->c_site[b].p, live * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (t_next) KMC_HIP(c, hipMemcpy(t_next, c->c_t[b].p, live * sizeof(double), hipMemcpyDeviceToHost));
    if (hops) KMC_HIP(c, hipMemcpy(hops, c->c_hops[b].p, live * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (y) KMC_HIP(c, hipMemcpy(y, c->c_y[b].p, live * sizeof(double), hipMemcpyDeviceToHost));
  }
  return KMC_OK;
}

}  // extern "C"
