/*
 * cntkmc.h -- C ABI of the MI355X kinetic Monte Carlo (KMC) exciton-hopping path.
 *
 * The reference (amirhosseindavoody/cnt_film_monte_carlo) has no plugin or FFI
 * layer: its hot path is reached only through the C++ classes of src/monte_carlo
 * compiled into one executable.  Every entry point below names the reference
 * function (file:line) whose job it takes over.  All entry points:
 *   - take plain pointers and sizes (caller-owned host arrays, copied in/out);
 *   - return an int status (KMC_OK = 0, negative = error) and never throw;
 *   - store a message for the last error, readable with kmc_last_error().
 * One context owns one HIP stream and all of its device buffers; use one
 * context per host thread.  Results do not depend on launch shapes: every
 * random draw comes from a counter-based generator keyed by (seed, exciton,
 * hop, stream).
 *
 * Units: SI (metres, seconds, 1/s), as in the reference.
 */
#ifndef CNTKMC_H
#define CNTKMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define KMC_OK 0
#define KMC_ERR_ARG (-1)      /* invalid argument (null pointer, bad size)   */
#define KMC_ERR_STATE (-2)    /* call out of order (e.g. run before inject)  */
#define KMC_ERR_HIP (-3)      /* a HIP runtime call failed                   */
#define KMC_ERR_ALLOC (-4)    /* device or host allocation failed            */
#define KMC_ERR_NODEVICE (-5) /* no HIP device available                     */
#define KMC_ERR_GEOMETRY (-6) /* degenerate geometry (cutoff vs period ...)  */
#define KMC_ERR_RATE (-7)     /* negative, NaN or infinite rate              */
#define KMC_ERR_CAPACITY (-8) /* output buffer too small                     */

/* ---- per-site flags ------------------------------------------------------ */
#define KMC_SITE_SOURCE 1u /* injection site (contact 1 / injection region)  */
#define KMC_SITE_DRAIN 2u  /* drain site: reaching it counts current and re-injects */

/* ---- rate modes for kmc_build_rates -------------------------------------- */
#define KMC_RATE_TABLE 0 /* nearest-node 4-D table (scattering_struct::get_rate) */
#define KMC_RATE_PAIR 1  /* one caller-supplied rate per CSR pair (exact / surrogate) */

/* ---- periodic axes -------------------------------------------------------- */
#define KMC_PERIODIC_X 1
#define KMC_PERIODIC_Y 2
#define KMC_PERIODIC_Z 4

typedef struct kmc_ctx kmc_ctx;

typedef struct kmc_config {
  int32_t device;        /* HIP device ordinal                                        */
  int32_t periodic_mask; /* KMC_PERIODIC_* bits; the reference has no periodic axis    */
  double cutoff;         /* max hopping radius [m] (input.json "max hopping radius")   */
  double min_distance;   /* pairs closer than this are skipped (scatterer.cpp:43: 0.4e-9) */
  double period_lo[3];   /* lower edge of the periodic cell per axis [m]                */
  double period_len[3];  /* period per periodic axis [m]; must exceed 2*cutoff          */
  uint64_t seed;         /* key of the counter-based generator                         */
} kmc_config;

/* Context life cycle.  Replaces mc::monte_carlo(json) (src/monte_carlo/monte_carlo.h:116-136). */
int kmc_create(const kmc_config* cfg, kmc_ctx** out);
int kmc_destroy(kmc_ctx* ctx);
const char* kmc_last_error(const kmc_ctx* ctx); /* ctx may be NULL: last global error */
int kmc_sync(kmc_ctx* ctx);                     /* wait for the context stream        */

/* Sites (scatterers).  Replaces create_scatterers (src/monte_carlo/monte_carlo.h:199-271) and
 * the per-site fields of mc::scatterer (src/monte_carlo/scatterer.h:21-40).  left/right are the
 * same-tube links (-1 = tube end).  chirality is an index into the rate-table set (0 when a
 * single table is used).  flags holds KMC_SITE_* bits.  Periodic coordinates must lie inside
 * [period_lo, period_lo + period_len).  Also computes the bounding box
 * (find_simulation_domain, src/monte_carlo/monte_carlo.h:328-340). */
int kmc_set_sites(kmc_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                  const double* ox, const double* oy, const double* oz, const int32_t* left,
                  const int32_t* right, const uint8_t* chirality, const uint8_t* flags);

/* Job (a): cell list.  Replaces create_scatterer_buckets (src/monte_carlo/monte_carlo.h:368-416):
 * cell edge = cutoff, n = ceil(extent/cutoff)+1 per open axis, buckets in site-index order. */
int kmc_build_cells(kmc_ctx* ctx);
/* Cell-list shape and contents (for inspection and tests). */
int kmc_get_cells(kmc_ctx* ctx, int32_t dims[3], int64_t* n_cells, int64_t* cell_start,
                  int32_t* cell_sites);

/* 4-D nearest-node rate tables.  Replaces scattering_struct (src/monte_carlo/scattering_struct.h:
 * 11-52).  rates is [n_chir*n_chir][n_theta][n_z][n_a1][n_a2] in row-major order, i.e. the order
 * scattering_struct::save writes (src/monte_carlo/scattering_struct.h:79-87); table (ci*n_chir+cj)
 * serves donor chirality ci and acceptor chirality cj. */
int kmc_set_rate_table(kmc_ctx* ctx, int32_t n_chir, int32_t n_theta, const double* theta,
                       int32_t n_z, const double* z_shift, int32_t n_a1, const double* axis_shift_1,
                       int32_t n_a2, const double* axis_shift_2, const double* rates);

/* Job (b): CSR out-rate table with FP64 cumulative sums.  Replaces scatterer::find_neighbors and
 * set_max_rate (src/monte_carlo/scatterer.cpp:34-83, src/monte_carlo/scatterer.h:89-93,
 * src/monte_carlo/monte_carlo.h:426-440).  Rows list neighbours in the reference order (stencil
 * x-outer, y, z-inner; ascending site index per bucket).  mode KMC_RATE_TABLE finishes the table;
 * mode KMC_RATE_PAIR stops after the pair geometry, to be finished by kmc_set_pair_rates. */
int kmc_build_rates(kmc_ctx* ctx, int32_t mode);
int kmc_get_nnz(kmc_ctx* ctx, int64_t* nnz);
/* Pair geometry (theta, z_shift, axis_shift_1, axis_shift_2) per CSR entry, as computed in
 * scatterer.cpp:44-61, plus source/target chirality.  Any pointer may be NULL. */
int kmc_get_pair_geometry(kmc_ctx* ctx, double* theta, double* z_shift, double* axis_shift_1,
                          double* axis_shift_2, uint8_t* chir_src, uint8_t* chir_dst);
/* Finish a KMC_RATE_PAIR build from host rates (one per CSR entry, in CSR order). */
int kmc_set_pair_rates(kmc_ctx* ctx, const double* rates);
/* Finish a KMC_RATE_PAIR build from rates already in device memory (HIP pointer, nnz doubles). */
int kmc_set_pair_rates_device(kmc_ctx* ctx, const double* d_rates);
/* CSR read-back (oracle feed).  disp is [nnz][3] float, the pair displacement target - source
 * (minimum image on periodic axes).  gamma[i] = total out-rate of site i.  NULL = skip. */
int kmc_get_csr(kmc_ctx* ctx, int64_t* row_ptr, int32_t* col, double* cum, float* disp,
                double* gamma);
/* Install an externally built CSR instead of kmc_build_rates (same layout as kmc_get_csr). */
int kmc_set_csr(kmc_ctx* ctx, int64_t n_sites, const int64_t* row_ptr, const int32_t* col,
                const double* cum, const float* disp);

/* Excitons.  Replaces kubo_create_particles (src/monte_carlo/monte_carlo.cpp:308-316) and the
 * particle constructor (src/monte_carlo/particle.h:48-52): each exciton starts on a source site
 * drawn uniformly, with its first dwell time drawn at that site. */
int kmc_inject(kmc_ctx* ctx, int64_t n_excitons);

/* MSD sampling grid: n_bins sample times t_k = (k+1) * t_sample_max / n_bins.  The displacement
 * of every exciton is recorded at each t_k it passes (kubo_save_avg_dispalcement_squared,
 * src/monte_carlo/monte_carlo.cpp:382-409, samples on a fixed clock).  Call before kmc_run. */
int kmc_set_samples(kmc_ctx* ctx, int32_t n_bins, double t_sample_max);

/* Trace buffer: record (site, hop time) of every hop of excitons [0, n_trace) for up to
 * max_hops hops (track_particle, src/monte_carlo/monte_carlo.h:786-818). */
int kmc_set_trace(kmc_ctx* ctx, int64_t n_trace, int64_t max_hops);

/* Job (c): persistent hop kernel.  Replaces the kubo_step loop
 * (src/monte_carlo/monte_carlo.cpp:319-342) with particle::step / update_state / ff_time
 * (src/monte_carlo/particle.cpp:57-80, scatterer.cpp:9-31, scatterer.h:74-80), event driven:
 * every exciton performs up to hops_per_exciton further hops, stopping early at the first hop
 * whose time exceeds t_max (pass INFINITY for a pure hop budget).  Re-injection on reaching a
 * drain site keeps the displacement and the pending dwell time (kubo_step, monte_carlo.cpp:330-336).
 * count_sites != 0 adds per-site departure counts.  kernel_ms (may be NULL) receives the device
 * time of the hop kernel, measured with HIP events on the context stream. */
int kmc_run(kmc_ctx* ctx, int64_t hops_per_exciton, double t_max, int32_t count_sites,
            double* kernel_ms);

/* Per-exciton state read-back: current site, time of the next hop, hops done, displacement
 * [E][3], drain (current) count, FNV-1a hash of the visited-site sequence.  NULL = skip. */
int kmc_get_excitons(kmc_ctx* ctx, int32_t* site, double* t_next, uint64_t* hops, double* disp,
                     uint32_t* drains, uint64_t* seq_hash);
/* Per-site departure counts (hops that left each site). */
int kmc_get_hop_counts(kmc_ctx* ctx, uint64_t* counts);
/* Raw MSD records: disp [n_bins][E][3] doubles, drains [n_bins][E], n_recorded [E]. */
int kmc_get_samples(kmc_ctx* ctx, double* disp, uint32_t* drains, uint32_t* n_recorded);
/* Trace read-back: sites/times [n_trace][max_hops], n_hops [n_trace]. */
int kmc_get_trace(kmc_ctx* ctx, int32_t* sites, double* times, int64_t* n_hops);

/* Job (d): wavefront-reduced MSD / current histograms.  Excitons are split into n_batches
 * contiguous id ranges; out is [n_bins][n_batches][5] = {count, sum dx^2, sum dy^2, sum dz^2,
 * sum drains} over the excitons whose record k exists.  Replaces the serial averages of
 * kubo_save_avg_dispalcement_squared (monte_carlo.cpp:382-409) and save_currents
 * (src/monte_carlo/monte_carlo.h:584-643). */
int kmc_reduce_samples(kmc_ctx* ctx, int32_t n_batches, double* out, double* kernel_ms);

/* Diffusion coefficient per axis from reduced samples (host arithmetic, no device work).
 * Least-squares slope through the origin of MSD(t) over bins k in [k_lo, k_hi) with a full
 * count, D = slope / 2 per axis, and the 95% confidence half-width from the batch spread
 * (Student t, n_batches - 1 degrees of freedom).  Post-processing of the reference:
 * visualization/monte_carlo_results.py:459-501 (MSD / t per axis). */
int kmc_estimate_diffusion(const double* reduced, int32_t n_bins, int32_t n_batches,
                           double t_sample_max, int32_t k_lo, int32_t k_hi, double D[3],
                           double ci95[3]);

/* ---- synthetic CNT networks (host only; no device work) -------------------- */
typedef struct kmc_netgen_params {
  uint64_t seed;
  int64_t n_tubes;
  double len_min, len_max;   /* tube length range [m] (uniform); equal for fixed length   */
  double box_lo[3], box_hi[3]; /* tube centres uniform in this box [m]                     */
  double site_spacing;       /* segment length [m] (10 nm)                                */
  int32_t n_chir;            /* chirality classes drawn uniformly per tube               */
  int32_t wrap_mask;         /* KMC_PERIODIC_* axes wrapped into [box_lo, box_hi)          */
} kmc_netgen_params;

/* Straight tubes with random in-plane orientation, one site per segment centre.  Pass NULL
 * outputs (or capacity 0) to get the site count in *n_out.  Orientation vectors satisfy
 * dot(o,o) == 1.0 exactly in FP64 so same-tube pairs take the parallel branch of
 * scatterer.cpp:48-52. */
int kmc_generate_network(const kmc_netgen_params* p, int64_t capacity, double* x, double* y,
                         double* z, double* ox, double* oy, double* oz, int32_t* left,
                         int32_t* right, uint8_t* chirality, int32_t* tube_id, int64_t* n_out);

/* Förster / Wong closed-form table (create_forster_scatt_table, src/monte_carlo/monte_carlo.cpp:
 * 156-200) on the grid given; rates is [n_theta][n_z][n_a1][n_a2].  Host only. */
int kmc_forster_table(double gamma_0, int32_t n_theta, const double* theta, int32_t n_z,
                      const double* z_shift, int32_t n_a1, const double* axis_shift_1,
                      int32_t n_a2, const double* axis_shift_2, double* rates);
/* arma::linspace as used for the table grids (monte_carlo.cpp:66-75).  Host only. */
int kmc_linspace(double start, double end, int32_t n, double* out);

/* Library version string. */
const char* kmc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CNTKMC_H */
