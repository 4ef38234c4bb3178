// kmc_kernels.h -- argument blocks and launchers of the KMC kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmc_internal.h"

namespace cntkmc {

struct site_arrays {
  int64_t n;
  const double *px, *py, *pz;
  const double *ox, *oy, *oz;
  const uint8_t* chir;
  const uint8_t* flags;
};

struct csr_arrays {
  site_hdr* hdr;
  double* cum;
  edge_rec* edge;
};

struct pair_geom_arrays {
  double *theta, *z, *a1, *a2;
};

struct exciton_arrays {
  int64_t n;
  int32_t* site;
  double* t_next;  // time of the pending hop
  uint64_t* hops;  // hops done so far (the generator's hop counter)
  double *dx, *dy, *dz;
  uint32_t* drains;
  uint64_t* hash;
  uint32_t* k_rec;  // MSD records written so far
  uint8_t* heading; // 1 = drifting toward the `right` link (particle.h:35)
  double* s_off;    // v > 0: distance travelled beyond the current site toward the next one
  double* t_last;   // v > 0: time of the last hop or injection (the flight starts there)
};

// Everything the hop loop needs of one network (one context).
struct net_view {
  exciton_arrays x;         // the network's excitons, indexed by exciton id
  const uint4* blocks;      // row blocks: 128-byte lines of ROW_UNITS 16-byte units
  const int64_t* roff;      // first line of every site's block
  const uint32_t* src_roff; // first lines of the source sites' blocks
  uint32_t n_src;
  uint64_t seed;
  double half_per[3];       // half the period on a periodic axis, +inf on an open one (the
                            // minimum-image displacement; kernel arguments, so they stay scalar)
  // MSD records of the sample window [bin0, n_bins): sample k (time (k + 1) dt_sample, k the
  // index on the whole grid) is row k - bin0 of [n_bins - bin0][x.n]
  int32_t bin0, n_bins;
  double dt_sample;
  double *rec_dx, *rec_dy, *rec_dz;
  uint32_t* rec_drain;
  // departure counts per block line: KMC_COUNT_COPIES arrays, count_stride apart
  uint32_t* site_count;
  int64_t count_stride;
  // drift along the tubes (particle::fly), used when velocity > 0
  double velocity;
  const uint2* links;       // per site: {left, right} block line | drain bit, or LINK_NONE
  int32_t* fly_err;         // set when a flight exceeds FLY_MAX_PASSES site passes
  // excitons still able to hop at the end of a launch (not trapped, t <= t_max), summed per
  // wave; zeroed by the sort-key kernel ahead of every launch
  uint32_t* live;
};

// Contact (steady-state) mode: particle slots [0, *n_live) of one buffer.
constexpr int CONTACT_MAX_SEG = 1024;
enum : int32_t {
  CONTACT_ERR_HOPS = 1, CONTACT_ERR_CAPACITY = 2, CONTACT_ERR_IDS = 4, CONTACT_ERR_FLY = 8
};
struct contact_slots {
  int32_t* site;
  double* t_next;
  uint64_t* hops;   // hops done by the particle (its draw counter)
  uint32_t* id;     // the particle's draw key, unique over the run
  double* y;        // FP64 y of the particle (its site's, plus the flight offset when v > 0)
  uint8_t* heading; // v > 0: 1 = toward `right`
  double* s_off;    // v > 0: distance travelled beyond the site toward the next one
  double* t_last;   // v > 0: time up to which the flight is taken
};
struct contact_args {
  const uint4* blocks;
  const int64_t* roff;
  const double* py;         // FP64 site y: sections and interface crossings
  uint64_t seed;
  contact_slots s;
  const int64_t* n_live;
  int64_t capacity;
  double t_end;             // the step ends at t_end (hops at t <= t_end happen in it)
  uint64_t max_hops;        // hops one particle may make in one step
  int32_t n_seg;
  double ymin, dy;
  int32_t* pop;             // [n_seg] particles per section after the step
  int32_t* cur;             // [n_seg - 1] net crossings of each interface during the step
  int32_t* err;             // CONTACT_ERR_* bits
  net_view fv;              // v > 0: links, periods, velocity and blocks for the flights
  unsigned long long* hops_made;  // hops made in contact mode so far (all particles)
  const uint32_t* perm;     // thread j steps slot perm[j] (the sorted slots); null: slot j
};

struct trace_view {         // trace [n_trace][tr_max] (single network only)
  int64_t n_trace = 0, tr_max = 0;
  int32_t* tr_site = nullptr;
  double* tr_time = nullptr;
  int64_t* tr_n = nullptr;
};

struct hop_args {
  net_view v;
  uint64_t hop_budget;
  double t_max;
  const uint32_t* perm;     // exciton processed at each position (NULL = identity)
  trace_view tv;
  // k_hop: positions [0, *n_alive) of perm (the idle excitons sorted after them are left out);
  // null: all v.x.n
  const uint32_t* n_alive;
};

struct ens_args {
  const net_view* nets;     // [n_networks]
  const uint16_t* wave_net; // network of every 64-position wave
  const int64_t* pbase;     // first position of each network (a multiple of 64)
  const int64_t* ebase;     // first ensemble-wide exciton number of each network
  const uint32_t* perm;     // ensemble-wide exciton numbers in sorted order
  uint64_t n_positions;
  uint64_t hop_budget;
  double t_max;
};

struct reduce_args {
  int64_t n_exc;
  int32_t bin0;             // grid index of row 0 of the records
  int32_t n_bins, n_batches;
  const uint32_t* k_rec;
  const double *rec_dx, *rec_dy, *rec_dz;
  const uint32_t* rec_drain;
  double* out;  // [n_bins][n_batches][5]
};

hipError_t exclusive_scan(const uint32_t* d_in, int64_t* d_out, int64_t n, int64_t* d_tmp,
                          hipStream_t st);
int64_t scan_tmp_elems(int64_t n);

hipError_t launch_cell_assign(const site_arrays& s, const cell_desc& c, int32_t* cell_of,
                              uint32_t* cell_count, hipStream_t st);
hipError_t launch_cell_scatter(int64_t n, const int32_t* cell_of, const int64_t* cell_start,
                               uint32_t* cell_fill, int32_t* cell_sites, hipStream_t st);
hipError_t launch_cell_sort(int64_t n_cells, const int64_t* cell_start, int32_t* cell_sites,
                            hipStream_t st);
hipError_t launch_count_pairs(const site_arrays& s, const cell_desc& c, const int64_t* cell_start,
                              const int32_t* cell_sites, uint32_t* deg, hipStream_t st);
hipError_t launch_fill_pairs(int mode, const site_arrays& s, const cell_desc& c,
                             const int64_t* cell_start, const int32_t* cell_sites,
                             const int64_t* row_ptr, const table_desc& t, const csr_arrays& out,
                             const pair_geom_arrays& geo, int32_t* err, hipStream_t st);
hipError_t launch_finish_pair_rates(int64_t n, const double* rates, const csr_arrays& out,
                                    int32_t* err, hipStream_t st);
hipError_t launch_inject(const exciton_arrays& x, const uint4* blocks, const uint32_t* src_roff,
                         uint32_t n_src, uint64_t seed, hipStream_t st);
hipError_t launch_block_sizes(int64_t n, const site_hdr* hdr, const int32_t* order, uint32_t* lines,
                              hipStream_t st);
hipError_t launch_place_blocks(int64_t n, const int32_t* order, const int64_t* first, int64_t* roff,
                               hipStream_t st);
// Which excitons a clock-bounded launch can leave out (k_exciton_keys); alive null: none.
struct idle_test {
  const double* t_next;
  double t_max;
  const uint32_t* k_rec;  // null: the launch records no MSD samples
  int32_t n_bins;         // the sample window's end on the grid
  double dt_sample;
  uint32_t* alive;        // += the excitons not idle (zero on entry)
  uint32_t* alive_next;   // set to zero for the next launch
};
// keys[e] = first line of exciton e's block (idle excitons: 0xFFFFFFFF, after every line),
// ids[e] = e; *live = 0 (the next launch's count)
hipError_t launch_exciton_keys(int64_t n, const int32_t* site, const int64_t* roff,
                               const idle_test& it, uint32_t* keys, uint32_t* ids, uint32_t* live,
                               hipStream_t st);
hipError_t launch_fill_blocks(int64_t n, const site_hdr* hdr, const double* cum,
                              const edge_rec* edge, const int64_t* roff, const double* px,
                              const double* py, const double* pz, uint4* blocks, hipStream_t st);
// Sort key of a live contact slot: with `regions`, (eighth of the film by row line, chirality of
// the site, row line) in bit fields [row_bits + chir_bits, +3], [row_bits, +chir_bits], [0,
// row_bits); without, the row line alone.  Dead slots key `dead`, which sorts last.
struct contact_key_layout {
  uint32_t lines;      // row lines of the film (the eighths split [0, lines))
  int32_t row_bits, chir_bits;
  int32_t regions;     // 0: the key is the row line alone
  uint32_t dead;
  int32_t key_bits;    // bits the radix sort looks at
};
// keys[i] = sort key of slot i (contact_key_layout), ids[i] = i
hipError_t launch_contact_keys(int64_t capacity, const int64_t* n_live, const int32_t* site,
                               const int64_t* roff, const uint8_t* chir,
                               const contact_key_layout& kl, uint32_t* keys, uint32_t* ids,
                               hipStream_t st);
// keys[i] = Morton code of site i's cell, ids[i] = i (block placement order, after a sort)
hipError_t launch_morton_keys(int64_t n, const int32_t* cell_of, int32_t nx, int32_t ny,
                              uint64_t* keys, int32_t* ids, hipStream_t st);
// out[i] = row word of site idx[i] (first line of its block | TGT_WIDE_BIT)
hipError_t launch_gather_offsets(int64_t n, const int32_t* idx, const int64_t* roff,
                                 const site_hdr* hdr, uint32_t* out, hipStream_t st);
// totals[site] += the per-XCD departure counts of each block line (the site's first line), which
// are then zeroed
hipError_t launch_fold_counts(int64_t n_lines, int64_t stride, const uint4* blocks,
                              uint32_t* copies, uint64_t* totals, hipStream_t st);
hipError_t launch_pair_classes(int64_t n, int64_t nnz, const site_hdr* hdr, const edge_rec* edge,
                               const uint8_t* chir, int32_t n_chir, uint8_t* cls, uint32_t* counts,
                               hipStream_t st);
hipError_t launch_pair_scatter(int64_t nnz, const uint8_t* cls, const uint8_t* mask,
                               const int64_t* starts, uint32_t* fill, int64_t* sel, hipStream_t st);
// counts[cls[p]] += 1 for every entry with mask[p] != 0
hipError_t launch_mask_count(int64_t nnz, const uint8_t* cls, const uint8_t* mask, uint32_t* counts,
                             hipStream_t st);
hipError_t launch_hop(const hop_args& a, bool count, bool sample, bool trace, bool fly,
                      hipStream_t st);
// links[i] = {line of left | drain bit, line of right | drain bit}, LINK_NONE for a tube end
hipError_t launch_links(int64_t n, const int32_t* left, const int32_t* right, const uint8_t* flags,
                        const int64_t* roff, const site_hdr* hdr, uint2* links, hipStream_t st);
hipError_t launch_hop_ens(const ens_args& a, bool count, bool sample, hipStream_t st);
hipError_t launch_ens_keys(int64_t n, const net_view* nets, const uint16_t* exc_net,
                           const int64_t* ebase, uint64_t* keys, uint32_t* ids, uint32_t* live,
                           hipStream_t st);
hipError_t launch_reduce(const reduce_args& a, hipStream_t st);
// the step of the live slots; n_max >= *a.n_live bounds the grid
hipError_t launch_contact_step(const contact_args& a, bool fly, int64_t n_max, hipStream_t st);
// keep[i] for the slots i < n (n >= *a.n_live; slots past the live count keep 0)
hipError_t launch_contact_keep(const contact_args& a, int64_t n, double y_lo, double y_hi,
                               uint32_t* keep, hipStream_t st);
hipError_t launch_contact_compact(int64_t capacity, const contact_slots& src,
                                  const contact_slots& dst, const uint32_t* keep,
                                  const int64_t* off, hipStream_t st);
hipError_t launch_contact_inject(const contact_args& a, const int64_t* base, int64_t count,
                                 const int32_t* sites, int64_t n_sites, uint64_t* next_id,
                                 double t_now, int64_t* n_live, hipStream_t st);

}  // namespace cntkmc
