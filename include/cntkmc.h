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

/* The nearest-node lookup of scattering_struct::get_rate (src/monte_carlo/scattering_struct.h:
 * 38-52) on the host for n geometries, over tables laid out as kmc_set_rate_table takes them:
 * out[k] = table (ci[k] * n_chir + cj[k]) at the node nearest to (theta, z_shift, a1, a2)[k] on
 * each axis, the first such node on ties (index_min of |grid - x|).  nodes (optional, [n][4])
 * receives the four node indices.  The device build of the CSR uses the same lookup.  Host only. */
int kmc_table_rate(int32_t n_chir, int32_t n_theta, const double* theta, int32_t n_z,
                   const double* z_shift, int32_t n_a1, const double* axis_shift_1, int32_t n_a2,
                   const double* axis_shift_2, const double* rates, int64_t n, const uint8_t* ci,
                   const uint8_t* cj, const double* q_theta, const double* q_z_shift,
                   const double* q_axis_shift_1, const double* q_axis_shift_2, double* out,
                   int32_t* nodes);

/* The dwell of ff_time (src/monte_carlo/scatterer.h:74-80), -(1/gamma) log(r / RAND_MAX), for n
 * draws r in [1, 2^31 - 1] and positive total rates gamma, evaluated on the host with the hop
 * kernels' own logarithm (an IEEE reciprocal where the device refines its hardware estimate).
 * For inspection and tests of that arithmetic.  Host only. */
int kmc_ff_time(int64_t n, const double* gamma, const uint32_t* r, double* out);

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
/* The same grid recorded window_bins bins at a time, so that the records take
 * window_bins x n_excitons x 28 bytes instead of n_bins x n_excitons x 28 (the reference keeps
 * one running average per output line, monte_carlo.cpp:382-409).  The first window holds bins
 * [0, window_bins); kmc_reduce_samples, kmc_reduce_samples_noalloc and kmc_get_samples cover
 * the open window only (kmc_get_sample_window tells which bins).  Run until every exciton's
 * clock has passed the window's last sample time ((first + n) * t_sample_max / n_bins, e.g.
 * kmc_run with that t_max), reduce, then kmc_next_sample_window.  Sample times, records and
 * reductions are those of kmc_set_samples, bit for bit.  No reference counterpart. */
int kmc_set_samples_window(kmc_ctx* ctx, int32_t n_bins, double t_sample_max,
                           int32_t window_bins);
/* The open window: its first bin on the grid and its number of bins (the last window of the
 * grid may be shorter than window_bins). */
int kmc_get_sample_window(kmc_ctx* ctx, int32_t* first_bin, int32_t* n_window);
/* Opens the next window and clears the records.  KMC_ERR_STATE if some exciton has not yet
 * recorded every bin of the open window (its records would be lost) or the open window is the
 * grid's last. */
int kmc_next_sample_window(kmc_ctx* ctx);

/* Trace buffer: record (site, hop time) of every hop of excitons [0, n_trace) for up to
 * max_hops hops (track_particle, src/monte_carlo/monte_carlo.h:786-818). */
int kmc_set_trace(kmc_ctx* ctx, int64_t n_trace, int64_t max_hops);

/* Job (c): the hop kernel, in sub-launches (kmc_set_hop_schedule).  Replaces the kubo_step loop
 * (src/monte_carlo/monte_carlo.cpp:319-342) with particle::step / update_state / ff_time
 * (src/monte_carlo/particle.cpp:57-80, scatterer.cpp:9-31, scatterer.h:74-80), event driven:
 * every exciton performs up to hops_per_exciton further hops, stopping early at the first hop
 * whose time exceeds t_max (pass INFINITY for a pure hop budget).  Re-injection on reaching a
 * drain site keeps the displacement and the pending dwell time (kubo_step, monte_carlo.cpp:330-336).
 * count_sites != 0 adds per-site departure counts.  kernel_ms (may be NULL) receives the device
 * time of the whole pass (all sub-launches and the exciton sorts between them), measured with
 * HIP events on the context stream.  The run ends early once no exciton can hop any more (every
 * one past t_max or trapped), so a budget far beyond the hops that t_max allows (the reference's
 * while (time < max_time) loop, src/main.cpp:64-76) returns after the hops that happen; the
 * excitons already past t_max are left out of the later sub-launches.  Neither changes any
 * trajectory. */
int kmc_run(kmc_ctx* ctx, int64_t hops_per_exciton, double t_max, int32_t count_sites,
            double* kernel_ms);
/* After kmc_run: the hop launches it made and the excitons still able to hop (t <= t_max and
 * not trapped; -1 before the first run).  No reference counterpart (its loop tests the clock,
 * src/main.cpp:64-76). */
int kmc_get_run_stats(kmc_ctx* ctx, int64_t* launches, int64_t* live_excitons);
/* After a kmc_run given kernel_ms: the device time of its hop-kernel launches alone (summed
 * over the sub-launches, sorts and copies between them left out), from HIP events recorded on
 * the context stream around each launch.  KMC_ERR_STATE if the last run was not timed.  No
 * reference counterpart (measurement only). */
int kmc_get_hop_time(kmc_ctx* ctx, double* hop_kernel_ms);
/* Launch schedule of kmc_run (no reference counterpart; results do not depend on it): the hop
 * budget runs as sub-launches of at most resort_hops hops, each preceded by a sort of the
 * excitons by the memory block of their current site; 0 = one launch (default 625). */
int kmc_set_hop_schedule(kmc_ctx* ctx, int64_t resort_hops);
/* Exciton velocity along its tube [m/s] (particle::fly, src/monte_carlo/particle.cpp:9-54;
 * "exciton velocity [m/s]", src/monte_carlo/monte_carlo.h:161).  0 (default): pure hopping.
 * v > 0: during every dwell the exciton drifts at speed v along the left/right links of its
 * tube (FP32 site positions, minimum image on periodic axes), reversing at a tube end; a site it
 * reaches becomes its current site, whose rates the next hop uses; a drain site reached in
 * flight counts a drain and re-injects (the flight goes on from the source site).  An isolated
 * site (no links) does not move the exciton.  Each flight is bounded at 2^22 site passes
 * (kmc_run fails beyond).  Linked sites must not coincide (KMC_ERR_GEOMETRY).  Set it after
 * kmc_set_sites; it applies to the excitons of the next kmc_inject.  Ensembles need velocity 0. */
int kmc_set_velocity(kmc_ctx* ctx, double velocity);
/* Flight state (v > 0): distance beyond the current site toward the next site along the
 * heading, heading (1 = toward `right`), time of the last hop or injection.  The exciton's
 * displacement is the kmc_get_excitons displacement (that of its current site) plus the offset
 * along that segment.  NULL = skip. */
int kmc_get_flight(kmc_ctx* ctx, double* offset, uint8_t* heading, double* t_last);

/* ---- contact (steady-state) mode ---------------------------------------------------------
 * The reference's second workflow: init(), step(), save_metrics() and repopulate_contacts()
 * (src/monte_carlo/monte_carlo.h:157-195, 274-316, 343-355, 443-684; driven by main.cpp:91-99).
 * The y range of the sites' bounding box is cut into n_seg sections; section 0 is a contact
 * held at c1_pop particles, section n_seg - 1 one held at c2_pop, by removing every particle
 * whose y lies in them ([ymin, ymin + dy] and [ymin + (n_seg - 1) dy, ymax]) after each step
 * and placing that many new ones on random sites of the section (contact_scats: closed
 * bounds).  kmc_contact_init makes the initial linear profile of create_particles
 * (round(c1_pop + i (c2_pop - c1_pop) / (n_seg - 1)) particles on sites with y in
 * [ymin + i dy, ymin + i dy + dy)) and the section areas of get_area.  Needs a complete CSR;
 * capacity bounds the live particles.  With a velocity set (kmc_set_velocity, before this
 * call) the particles drift along their tubes as in particle::fly (particle.cpp:9-54): up to
 * each hop, and to the end of every step, where the flight is taken; a particle's y is then
 * its site's plus the offset along the tube, and the sites may carry no KMC_SITE_DRAIN flag.
 * Draws are keyed by a particle id (unique over the run) and the particle's hop count; the
 * reference's rand() order is not reproduced. */
int kmc_contact_init(kmc_ctx* ctx, int32_t n_seg, int64_t c1_pop, int64_t c2_pop,
                     int64_t capacity);
/* n_steps steps of dt.  In each, every particle makes the hops due by the step's end
 * (particle::step; more than max_hops_per_step hops of one particle in one step is an error,
 * KMC_ERR_CAPACITY, as is a flight passing more than 2^22 sites); then population[step][n_seg] receives the particles per section
 * (save_population_profile, section = (int)((y - ymin) / dy) clamped) and
 * crossings[step][n_seg - 1] the net crossings of the interfaces ymin + (k + 1) dy between
 * the particles' positions at the start and the end of the step (save_currents: +1 upward,
 * -1 downward); then the contacts are repopulated and n_live[step] receives the particle
 * count.  Densities and currents are these over (area dy) and (interface area dt), as the
 * reference writes them.  Any output may be NULL.  The call returns once every step is done;
 * while queueing them the host stays at most two steps ahead of the device (it reads each
 * step's live count back to bound the slots the following steps cover). */
int kmc_contact_run(kmc_ctx* ctx, double dt, int64_t n_steps, uint64_t max_hops_per_step,
                    int32_t* population, int32_t* crossings, int64_t* n_live);
/* Since kmc_contact_init: the hops made by all particles (also those since removed at the
 * contacts) and the device time of the step kernels (particle::step of every particle, HIP
 * events around each step's launch), without the repopulation.  No counterpart in the
 * reference, which times nothing; a throughput measurement of the contact mode.  Either
 * output may be NULL. */
int kmc_contact_stats(kmc_ctx* ctx, uint64_t* hops, double* step_kernel_ms);
/* area[n_seg] (get_area: x extent times z extent of the section's sites), ymin, dy and the
 * simulated time; any output may be NULL. */
int kmc_contact_sections(const kmc_ctx* ctx, double* area, double* ymin, double* dy,
                         double* time);
/* The live particles in slot order: id, site, time of the next hop, hops made and y (FP64).
 * With every array NULL only *n is set. */
int kmc_contact_particles(kmc_ctx* ctx, int64_t capacity, int64_t* n, uint32_t* id,
                          int32_t* site, double* t_next, uint64_t* hops, double* y);
/* Autotuner of that schedule (no reference counterpart): runs trial_hops hops from the current
 * state once per candidate resort_hops value (counts into scratch, no samples), restores the
 * state, installs the fastest candidate.  best and hops_per_s[n_cand] may be NULL. */
int kmc_autotune_schedule(kmc_ctx* ctx, int64_t trial_hops, int32_t n_candidates,
                          const int64_t* candidates, int64_t* best, double* hops_per_s);

/* Ensembles (no reference counterpart): several independent networks, each a finished context
 * (CSR built, excitons injected, optional MSD grid; no trace), run by one hop kernel launch per
 * sub-launch.  Every network keeps its own seed, exciton ids, sample grid and count arrays, so its
 * trajectories are bit-identical to kmc_run on that context alone; results stay in the member
 * contexts (read them with kmc_get_excitons, kmc_reduce_samples, ...).  Members must not be
 * re-injected or rebuilt while the ensemble exists.  Errors: kmc_ensemble_last_error. */
typedef struct kmc_ensemble kmc_ensemble;
int kmc_ensemble_create(int32_t n_networks, kmc_ctx* const* networks, kmc_ensemble** out);
int kmc_ensemble_destroy(kmc_ensemble* ens);
const char* kmc_ensemble_last_error(void);
int kmc_ensemble_set_hop_schedule(kmc_ensemble* ens, int64_t resort_hops);
int kmc_ensemble_run(kmc_ensemble* ens, int64_t hops_per_exciton, double t_max,
                     int32_t count_sites, double* kernel_ms);
int kmc_ensemble_autotune(kmc_ensemble* ens, int64_t trial_hops, int32_t n_candidates,
                          const int64_t* candidates, int64_t* best, double* hops_per_s);

/* Per-exciton state read-back: current site, time of the next hop, hops done, displacement
 * [E][3], drain (current) count, FNV-1a hash of the visited-site sequence.  NULL = skip. */
int kmc_get_excitons(kmc_ctx* ctx, int32_t* site, double* t_next, uint64_t* hops, double* disp,
                     uint32_t* drains, uint64_t* seq_hash);
/* Per-site departure counts (hops that left each site) since kmc_inject, 64-bit: the device
 * keeps 32-bit per-XCD copies that are folded into 64-bit totals before any copy could wrap. */
int kmc_get_hop_counts(kmc_ctx* ctx, uint64_t* counts);
/* Raw MSD records of the open sample window (all n_bins without kmc_set_samples_window):
 * disp [n][E][3] doubles, drains [n][E], and n_recorded [E], each exciton's count of samples
 * recorded on the whole grid. */
int kmc_get_samples(kmc_ctx* ctx, double* disp, uint32_t* drains, uint32_t* n_recorded);
/* Trace read-back: sites/times [n_trace][max_hops], n_hops [n_trace]. */
int kmc_get_trace(kmc_ctx* ctx, int32_t* sites, double* times, int64_t* n_hops);

/* Job (d): wavefront-reduced MSD / current histograms.  Excitons are split into n_batches
 * contiguous id ranges; out is [n][n_batches][5] = {count, sum dx^2, sum dy^2, sum dz^2,
 * sum drains} over the excitons whose record k exists, for the n bins of the open sample window
 * (all n_bins without kmc_set_samples_window).  Replaces the serial averages of
 * kubo_save_avg_dispalcement_squared (monte_carlo.cpp:382-409) and save_currents
 * (src/monte_carlo/monte_carlo.h:584-643). */
int kmc_reduce_samples(kmc_ctx* ctx, int32_t n_batches, double* out, double* kernel_ms);
/* kmc_reduce_samples without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The partial sums go to a static 256 KB device array per
 * GPU (held until the process ends), so n_bins * n_batches * 5 must not exceed 32768
 * (KMC_ERR_CAPACITY otherwise).  Calls on one GPU share that array: calls that overlap on one
 * GPU (from several threads or contexts) race.  Same output as kmc_reduce_samples, bit for bit. */
int kmc_reduce_samples_noalloc(kmc_ctx* ctx, int32_t n_batches, double* out, double* kernel_ms);

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

/* ---- site ingest (host only; no device work) --------------------------------------------
 * Mesh reader: create_scatterers (src/monte_carlo/monte_carlo.h:199-271).  dir holds the six
 * ASCII matrices single_cnt.{pos,orient}.{x,y,z}.dat (one row per tube, one column per point,
 * numbers separated by blanks or commas; positions in nm).  Site s = i * n_points + j of tube i,
 * point j: position x 1e-9 m, orientation as read, links left = s - 1 (j > 0), right = s + 1
 * (j + 1 < n_points), else -1.  Call with x == NULL (or capacity 0) for n_tubes and n_points.
 * Errors (missing file, ragged rows, shape mismatch) read with kmc_last_error(NULL). */
int kmc_read_mesh(const char* dir, int64_t capacity, int64_t* n_tubes, int64_t* n_points,
                  double* x, double* y, double* z, double* ox, double* oy, double* oz,
                  int32_t* left, int32_t* right);
/* trim_scats (src/monte_carlo/monte_carlo.h:722-782): removes the sites outside
 * [xlim[0], xlim[1]] x [ylim[0], ylim[1]] x [zlim[0], zlim[1]] in place, in the reference's
 * order (a site outside swaps with the last kept site, links re-pointed; links to removed
 * sites become -1).  *n: site count in, kept count out.  chirality and flags may be NULL;
 * when given they follow the permutation.  origin (may be NULL) receives the original index
 * of the site in every slot. */
int kmc_trim_sites(int64_t* n, double* x, double* y, double* z, double* ox, double* oy,
                   double* oz, int32_t* left, int32_t* right, uint8_t* chirality, uint8_t* flags,
                   int64_t* origin, const double xlim[2], const double ylim[2],
                   const double zlim[2]);
/* Kubo regions of kubo_init (src/monte_carlo/monte_carlo.cpp:296-302): the bounding box
 * (find_simulation_domain, monte_carlo.h:328-340) cut into n_sections (odd, >= 3) slabs per
 * axis; sites in the centre cell (closed bounds) get KMC_SITE_SOURCE (injection_region,
 * monte_carlo.cpp:203-228), sites outside [edge 1, edge n-1] on any axis get KMC_SITE_DRAIN
 * (get_removal_domain, :231-251: reaching one re-injects, kubo_step :330-336).  flags is
 * overwritten; domain_lo / domain_hi (may be NULL) receive the box. */
int kmc_kubo_regions(int64_t n, const double* x, const double* y, const double* z,
                     int32_t n_sections, uint8_t* flags, double domain_lo[3], double domain_hi[3]);

/* Förster / Wong closed-form table (create_forster_scatt_table, src/monte_carlo/monte_carlo.cpp:
 * 156-200) on the grid given; rates is [n_theta][n_z][n_a1][n_a2].  Host only. */
int kmc_forster_table(double gamma_0, int32_t n_theta, const double* theta, int32_t n_z,
                      const double* z_shift, int32_t n_a1, const double* axis_shift_1,
                      int32_t n_a2, const double* axis_shift_2, double* rates);
/* arma::linspace as used for the table grids (monte_carlo.cpp:66-75).  Host only. */
int kmc_linspace(double start, double end, int32_t n, double* out);

/* Rate-table files of scattering_struct::save (src/monte_carlo/scattering_struct.h:55-94):
 * <dir>/<name>.theta.dat, .z_shift.dat, .axis_shift_1.dat, .axis_shift_2.dat (one value per
 * line) and <name>.rates.dat (the lines "sizes:", the axis names, "n_theta,n_z,n_a1,n_a2", a
 * blank line, then one rate per line, theta outermost, axis_shift_2 innermost).  name NULL or
 * "" means "scat_table", the reference's.  Values are written with 17 significant digits, so a
 * save and a load give back the same doubles; the reference itself writes 6 (rates) and
 * Armadillo's 4-digit vector format (axes), which the loader reads too.  rates is
 * [n_theta][n_z][n_a1][n_a2], as kmc_set_rate_table takes one table.  Host only. */
int kmc_save_rate_table(const char* dir, const char* name, int64_t n_theta, int64_t n_z,
                        int64_t n_a1, int64_t n_a2, const double* theta, const double* z_shift,
                        const double* axis_shift_1, const double* axis_shift_2,
                        const double* rates);
/* Reads the files above.  With theta == NULL: fills dims[4] = {n_theta, n_z, n_a1, n_a2} and
 * returns (size query).  Otherwise dims must hold the file's sizes (KMC_ERR_CAPACITY if not)
 * and every array is filled.  The reference has no reader (its tables are never reloaded);
 * this one accepts what visualization/monte_carlo_results.py:261-282 reads.  Errors (missing
 * file, bad header, counts that disagree with the header) read with kmc_last_error(NULL). */
int kmc_load_rate_table(const char* dir, const char* name, int64_t dims[4], double* theta,
                        double* z_shift, double* axis_shift_1, double* axis_shift_2,
                        double* rates);

/* ---- CNT exciton model (host only; no device work) ------------------------------------
 * Restates cnt::calculate_exciton_dispersion (src/exciton_transfer/cnt.cpp:1056-1081):
 * get_parameters (:16-117), get_atom_coordinates (:120-209), electron_energy in the
 * K2-extended zone (:301-358), find_valleys (:362-416), find_relev_ik_range (:419-517),
 * calculate_vq (:520-631), calculate_polarization (:634-732), calculate_dielectric
 * (:735-780) and the A2 singlet exciton of calculate_A_excitons (:783-1053). */
typedef struct kmc_cnt kmc_cnt;

typedef struct kmc_cnt_info {
  int32_t n, m, length_cells;           /* chirality (n,m), length in cnt unit cells        */
  int32_t Nu, Q, M, t1, t2;             /* cnt.h:35-53                                       */
  int32_t nk, n_mu;                     /* K2-extended zone: ik in [0,nk), mu in [0,n_mu)   */
  int32_t nk_relev, nk_cm, nk_c;        /* exciton basis sizes (cnt.cpp:876-878)             */
  int32_t ik_cm0;                       /* ik_cm of dispersion row 0 (= -nk_relev)          */
  double radius, length_m, Au, a_l;     /* [m], [m], [m^2], [m]                              */
  double ch_vec[2], t_vec[2], aCC[2], dk_l[2], K1[2], K2[2]; /* rotated frame (cnt.cpp:44-58) */
  double e_min;                         /* lowest A2 singlet energy [J]                      */
} kmc_cnt_info;

/* Build the model of a (n,m) tube of length_cells unit cells (threads <= 0: all host cores). */
int kmc_cnt_create(int32_t n, int32_t m, int32_t length_cells, int32_t threads, kmc_cnt** out);
/* The same model with its two large sums on the GPU `device`: the Ohno-potential transform v(q)
 * (calculate_vq, cnt.cpp:520-631) and the polarization PI(q) (calculate_polarization,
 * cnt.cpp:634-732), ~95 % of the host build's time; the rest runs on `threads` host cores.  The
 * sums are taken in another order and with the device sincos / hypot, so v(q), PI(q) and what
 * follows from them agree with kmc_cnt_create to rounding, not bitwise.  KMC_ERR_HIP when the
 * device cannot be used. */
int kmc_cnt_create_on_device(int32_t n, int32_t m, int32_t length_cells, int32_t threads,
                             int32_t device, kmc_cnt** out);
int kmc_cnt_destroy(kmc_cnt* c);
const char* kmc_cnt_last_error(void);
int kmc_cnt_get_info(const kmc_cnt* c, kmc_cnt_info* info);
/* A2 singlet energies [nk_cm][nk_relev] (ascending per ik_cm) [J]. */
int kmc_cnt_dispersion(const kmc_cnt* c, double* energy);
/* Electron energies [2][nk][n_mu] (band 0 valence, 1 conduction) and wavefunctions
 * [n_mu][nk][band][atom] as interleaved complex doubles. */
int kmc_cnt_electrons(const kmc_cnt* c, double* energy, double* wavefunc);
/* v(q) [nq][2 n_mu - 1][4] complex, PI and epsilon [nq][2 n_mu - 1], nq = 2 nk - 1. */
int kmc_cnt_interaction(const kmc_cnt* c, double* vq, double* PI, double* eps);
/* A2 singlet wavefunctions [nk_cm][nk_relev][nk_c] complex and their electron-state indices
 * [nk_cm][nk_c][4] = {ik_c, mu_c, ik_v, mu_v} (cnt.h:136-146). */
int kmc_cnt_exciton_states(const kmc_cnt* c, double* psi, int32_t* idx);
/* The Hermitian A2 singlet kernel at one ik_cm index, column-major complex nk_relev^2. */
int kmc_cnt_singlet_kernel(const kmc_cnt* c, int32_t ik_cm_idx, double* kernel);
/* A and B atom positions of one unit cell [Nu][2] (unrolled) and rolled unit cells [Nu][3]. */
int kmc_cnt_atoms(const kmc_cnt* c, double* pos_a, double* pos_b, double* pos_u_3d);
/* Eigen-decomposition of a Hermitian matrix (column-major complex n x n): ascending
 * eigenvalues and eigenvectors in columns; stands in for arma::eig_sym (cnt.cpp:950-965). */
int kmc_hermitian_eig(int32_t n, const double* a, double* evals, double* vecs);

/* ---- exact first-order transfer rate ---------------------------------------------------
 * exciton_transfer::first_order (src/exciton_transfer/exciton_transfer.cpp:380-441) for a
 * donor / acceptor tube pair at temperature T [K] and Lorentzian broadening [J] (the reference
 * hard-codes 300 K and 4 meV, exciton_transfer.h:41-42).  Matched states collapse into
 * coefficients C_ab per pair of centre-of-mass momenta; rate = sum_ab C_ab |J_ab|^2 with the
 * Coulomb sum J_ab of calculate_J (exciton_transfer.cpp:307-377) evaluated on the device. */
typedef struct kmc_transfer kmc_transfer;

typedef struct kmc_transfer_info {
  int32_t n_donor_states, n_acceptor_states, n_pairs; /* relevant states, matched pairs     */
  int32_t n_mom_donor, n_mom_acceptor;                /* distinct ik_cm per side (<= 16)      */
  int32_t n_terms;                                    /* nonzero C_ab (<= 256)                */
  int32_t n_atoms_donor, n_atoms_acceptor;            /* A-site unit cells per tube           */
  double partition_Z, min_energy;                     /* exciton_transfer.cpp:385, 394-397    */
} kmc_transfer_info;

int kmc_transfer_create(int32_t device, const kmc_cnt* donor, const kmc_cnt* acceptor,
                        double temperature, double broadening, kmc_transfer** out);
int kmc_transfer_destroy(kmc_transfer* t);
const char* kmc_transfer_last_error(void);
int kmc_transfer_get_info(const kmc_transfer* t, kmc_transfer_info* info);
/* The n_terms nonzero terms: donor / acceptor centre-of-mass momenta (ik_cm) and C_ab [1/s/m^2]. */
int kmc_transfer_terms(const kmc_transfer* t, int32_t* ik_cm_donor, int32_t* ik_cm_acceptor,
                       double* coef);
/* Clamp (z_shift, axis_shift_1, axis_shift_2) into [lo, hi] before evaluation (default: none).
 * With the table grid's box this is the continuous counterpart of the nearest-node clamp of
 * scattering_struct::get_rate (scattering_struct.h:42-49). */
int kmc_transfer_set_clamp(kmc_transfer* t, const double lo[3], const double hi[3]);
/* Rates for n geometries (host arrays); kernel_ms (may be NULL) = device time. */
int kmc_transfer_rates(kmc_transfer* t, int64_t n, const double* theta, const double* z_shift,
                       const double* axis_shift_1, const double* axis_shift_2, double* rate,
                       double* kernel_ms);
/* Same on device arrays; with d_sel != NULL geometry g is entry d_sel[g] (read and written). */
int kmc_transfer_rates_device(kmc_transfer* t, int64_t n, const double* d_theta,
                              const double* d_z, const double* d_a1, const double* d_a2,
                              const int64_t* d_sel, double* d_rate);
/* Exact 4-D rate table on the grid given: create_davoody_scatt_table
 * (src/monte_carlo/monte_carlo.cpp:63-150) for one donor/acceptor pair.  rates (host) is
 * [n_theta][n_z][n_a1][n_a2] row-major, as kmc_set_rate_table takes it.  elapsed_ms = wall time
 * of the build (upload, kernel, download). */
int kmc_transfer_table(kmc_transfer* t, int32_t n_theta, const double* theta, int32_t n_z,
                       const double* z_shift, int32_t n_a1, const double* axis_shift_1,
                       int32_t n_a2, const double* axis_shift_2, double* rates, double* elapsed_ms);

/* Per-pair exact rates for a KMC_RATE_PAIR build: CSR entries are grouped by chirality class
 * (donor * n_chir + acceptor) on the device and class k is evaluated with transfers[k] (NULL:
 * rate 0) at the pair's (theta, z_shift, axis_shift_1, axis_shift_2); then the FP64 prefix sums
 * finish the CSR.  This replaces both the table fill of create_davoody_scatt_table
 * (src/monte_carlo/monte_carlo.cpp:108-134) and its nearest-node lookup: every pair gets
 * first_order at its own geometry (clamped as set by kmc_transfer_set_clamp).  elapsed_ms
 * (may be NULL) = wall time of the whole rate build. */
int kmc_build_rates_exact(kmc_ctx* ctx, int32_t n_chir, kmc_transfer* const* transfers,
                          double* elapsed_ms);
/* Chirality classes of the CSR entries: class_start[n_chir^2 + 1] (host) and the device index
 * list (entries of class k at d_sel[class_start[k] .. class_start[k+1])). */
int kmc_pair_classes(kmc_ctx* ctx, int32_t n_chir, int64_t* class_start, const int64_t** d_sel);
/* Per-pair rates of the last pair-rate build (host copy, nnz doubles). */
int kmc_get_pair_rates(kmc_ctx* ctx, double* rates);

/* ---- Learned rate surrogate (no reference counterpart) --------------------------------------
 * An MLP of KMC_MLP_HIDDEN hidden layers of width KMC_MLP_WIDTH with SiLU maps the features of a
 * pair geometry to a normalised log10 transfer rate.  It stands beside
 * exciton_transfer::first_order (src/exciton_transfer/exciton_transfer.cpp:380-441), which stays
 * the fallback for pairs outside the trained domain or in regions whose held-out residual
 * exceeded the tolerance.  Matrix operands are BF16 (weights, features, activations, gradients
 * of pre-activations); accumulation, biases, the output layer and the master weights are FP32.
 *
 * Features (KMC_MLP_IN slots, 13 used), from theta, the clamped (z, a1, a2) and T:
 *   0 cos(theta)  1 sin(theta)  2..4 z, a1, a2 mapped linearly from the clamp box to [-1, 1],
 *   with (a1, a2) replaced by (-a1, -a2) when a2 < 0 (or a2 == 0 and a1 < 0): reversing both
 *   tubes (the C2 axes of a nanotube) leaves the exact rate unchanged to about 1e-4, so both
 *   hop directions along a tube get one prediction.  The a1/a2 box must be symmetric.
 *   5..7 donor chirality one-hot  8..10 acceptor chirality one-hot
 *   11 ln(T / sqrt(t_lo t_hi)) / ln(sqrt(t_hi / t_lo))  (maps [t_lo, t_hi] to [-1, 1])
 *   12 1/T mapped linearly from [1/t_hi, 1/t_lo] to [-1, 1]  (Boltzmann factors are linear
 *      in 1/T in the log rate)
 * Output y: log10(rate / s^-1) = y_shift + y_scale * y.
 * Parameters (KMC_MLP_PARAMS floats, FP32): W0 [16][128], b0, W1 [128][128], b1, W2, b2, W3, b3,
 * w4 [128], b4 [1]; W_l is [in][out] row-major. */
#define KMC_MLP_IN 16
#define KMC_MLP_WIDTH 128
#define KMC_MLP_HIDDEN 4
#define KMC_MLP_CLASSES 3
#define KMC_MLP_PARAMS 51841

#define KMC_MLP_FLAG_DOMAIN 1   /* chirality or temperature outside the trained domain */
#define KMC_MLP_FLAG_RESIDUAL 2 /* bin whose held-out residual exceeded the tolerance   */

typedef struct kmc_mlp kmc_mlp;
typedef struct kmc_mlp_domain {
  double box_lo[3], box_hi[3]; /* z, a1, a2 clamp box (m) */
  double t_lo, t_hi;           /* trained temperature range (K) */
  double y_shift, y_scale;     /* output normalisation */
} kmc_mlp_domain;

int kmc_mlp_create(int32_t device, const kmc_mlp_domain* domain, kmc_mlp** out);
int kmc_mlp_destroy(kmc_mlp* mlp);
const char* kmc_mlp_last_error(void);
int kmc_mlp_set_params(kmc_mlp* mlp, const float* params);
int kmc_mlp_get_params(kmc_mlp* mlp, float* params);
/* Residual map: flags[dims[0]][dims[1]][dims[2]][dims[3]][dims[4]] over (donor class * 3 +
 * acceptor class, T bin, z bin, theta bin, a bin), bins uniform in T, clamped z and theta over
 * the domain and in (max(|a1|, |a2|) / max |a|)^2 (rings of equal area in the (a1, a2) square);
 * nonzero = send the pair to the exact routine.  dims[0] must be 9.  NULL clears the map. */
int kmc_mlp_set_residual_map(kmc_mlp* mlp, const int32_t dims[5], const uint8_t* flags);
/* Feature rows [n][KMC_MLP_IN] (host); class < 0 or >= 3 gives an all-zero one-hot. */
int kmc_mlp_features(const kmc_mlp* mlp, int64_t n, const double* theta, const double* z_shift,
                     const double* axis_shift_1, const double* axis_shift_2, const int32_t* cls_d,
                     const int32_t* cls_a, const double* temperature, float* features);
/* Network output y for n feature rows (host arrays), through the training kernels. */
int kmc_mlp_forward(kmc_mlp* mlp, int64_t n, const float* features, float* y);
/* Mean squared error of y against targets and its gradient w.r.t. every parameter (host). */
int kmc_mlp_loss_grad(kmc_mlp* mlp, int64_t n, const float* features, const float* targets,
                      double* loss, float* grad);
/* One Adam step (step >= 1, bias-corrected) with the gradient given (host). */
int kmc_mlp_adam_step(kmc_mlp* mlp, const float* grad, double lr, double beta1, double beta2,
                      double eps, int64_t step);
/* Train on n rows for `epochs` epochs of shuffled mini-batches (batch a multiple of 128; the
 * last partial batch of an epoch is dropped), Adam with a cosine learning rate from lr to
 * lr_min.  epoch_loss[epochs] (may be NULL) gets the mean batch loss of each epoch. */
int kmc_mlp_train(kmc_mlp* mlp, int64_t n, const float* features, const float* targets,
                  int32_t epochs, int32_t batch, double lr, double lr_min, uint64_t seed,
                  double* epoch_loss, double* elapsed_ms);
/* Fused inference kernel on device arrays: rates of n geometries of one chirality pair (cls_d,
 * cls_a) at temperature T.  d_flag (may be NULL) gets 0 or a KMC_MLP_FLAG_* value; flagged
 * entries still receive the network's rate. */
int kmc_mlp_rates_device(kmc_mlp* mlp, int64_t n, const double* d_theta, const double* d_z,
                         const double* d_a1, const double* d_a2, int32_t cls_d, int32_t cls_a,
                         double temperature, double* d_rate, uint8_t* d_flag);
/* Same on host arrays; kernel_ms (may be NULL) = device time of the inference kernel. */
int kmc_mlp_rates(kmc_mlp* mlp, int64_t n, const double* theta, const double* z_shift,
                  const double* axis_shift_1, const double* axis_shift_2, int32_t cls_d,
                  int32_t cls_a, double temperature, double* rate, uint8_t* flag,
                  double* kernel_ms);
/* Surrogate-built 4-D table (layout of kmc_transfer_table).  Flagged nodes, and the coaxial
 * nodes (theta = 0 and axis shift 1 = 0, the geometry of every same-tube hop), are recomputed by
 * `fallback` (the exact transfer of the same pair and temperature; NULL leaves the network's
 * value).  n_fallback (may be NULL) = nodes recomputed; elapsed_ms = wall time of the build. */
int kmc_mlp_table(kmc_mlp* mlp, int32_t cls_d, int32_t cls_a, double temperature,
                  kmc_transfer* fallback, int32_t n_theta, const double* theta, int32_t n_z,
                  const double* z_shift, int32_t n_a1, const double* axis_shift_1, int32_t n_a2,
                  const double* axis_shift_2, double* rates, int64_t* n_fallback,
                  double* elapsed_ms);
/* Per-pair surrogate rates for a KMC_RATE_PAIR build, then the CSR.  chir_class[n_chir] maps
 * the context's chirality ids to surrogate classes (-1: unknown, always exact); flagged pairs
 * are recomputed by transfers[ci * n_chir + cj] (entries may be NULL only for classes with no
 * flagged pair).  n_fallback (may be NULL) = pairs recomputed exactly. */
int kmc_build_rates_mlp(kmc_ctx* ctx, kmc_mlp* mlp, int32_t n_chir, const int32_t* chir_class,
                        double temperature, kmc_transfer* const* transfers, int64_t* n_fallback,
                        double* elapsed_ms);

/* Library version string. */
const char* kmc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CNTKMC_H */
