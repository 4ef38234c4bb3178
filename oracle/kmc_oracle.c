/*
 * kmc_oracle.c -- CPU oracle of the KMC exciton-hopping path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load this library, and only as the checker or the CPU baseline; the product
 * (cnt_film_monte_carlo_amd/) never links, loads or calls it.
 *
 * A plain-C, FP64 restatement of the reference algorithm (amirhosseindavoody/
 * cnt_film_monte_carlo), written independently of the HIP product.  Each function cites the
 * reference lines it follows.  Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 *
 * Draw mapping shared with the product (the counter-based generator replaces std::rand()):
 *   Philox4x32-10, key = {seed lo, seed hi}, counter = {c lo, c hi, exciton, stream}.
 *   hop h    : stream 0 at counter h >> 1 (one call for hops 2P and 2P + 1) -> words 2(h & 1)
 *              selection and 2(h & 1) + 1 dwell; a dwell word of 0 moves to streams 2+k at
 *              counter h; a drain at hop h re-injects on word 0 of stream 0x10000000, counter h
 *   injection: stream 1 -> w0 site, w1 dwell, w2 heading, w3 dwell spare; a zero dwell moves
 *              to the spare word, then to streams 0x80000000+k
 *   a 31-bit draw is w >> 1; redraws take k < 16 streams, words in order; after that r = 1.
 *
 * Parity pins (see tests/): Philox against the published Random123 known-answer vectors;
 * geometry against closed forms; cell lists and neighbour sets against brute force; the
 * selection against a linear scan; the dwell against Exp(Gamma) statistics; MSD against the
 * 1-D chain known answer D = k a^2.  Armadillo's summation order (dot, norm, linspace,
 * index_min) cannot be pinned here (Armadillo is absent): parity unpinned for those ulps.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_RAND_MAX 2147483647.0

/* ------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64 multiplies with Weyl key
 * increments.
 * ---------------------------------------------------------------------------------------- */
static void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t x0 = ctr_in[0], x1 = ctr_in[1], x2 = ctr_in[2], x3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  int round;
  for (round = 0; round < 10; ++round) {
    uint32_t hi0, lo0, hi1, lo1, y0, y1, y2, y3;
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    mulhilo32(0xD2511F53u, x0, &hi0, &lo0);
    mulhilo32(0xCD9E8D57u, x2, &hi1, &lo1);
    y0 = hi1 ^ x1 ^ k0;
    y1 = lo1;
    y2 = hi0 ^ x3 ^ k1;
    y3 = lo0;
    x0 = y0; x1 = y1; x2 = y2; x3 = y3;
  }
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

static void draws(uint64_t seed, uint32_t exciton, uint64_t hop, uint32_t stream, uint32_t w[4]) {
  uint32_t ctr[4], key[2];
  ctr[0] = (uint32_t)(hop & 0xFFFFFFFFu);
  ctr[1] = (uint32_t)(hop >> 32);
  ctr[2] = exciton;
  ctr[3] = stream;
  key[0] = (uint32_t)(seed & 0xFFFFFFFFu);
  key[1] = (uint32_t)(seed >> 32);
  orc_philox4x32_10(ctr, key, w);
}

/* a draw for ff_time, which loops while rand() == 0 (scatterer.h:76-77) */
static uint32_t dwell_draw(uint64_t seed, uint32_t e, uint64_t hop, uint32_t w1, uint32_t w3,
                           uint32_t redraw_base) {
  uint32_t k, i, w[4];
  if ((w1 >> 1) != 0u) return w1 >> 1;
  if ((w3 >> 1) != 0u) return w3 >> 1;
  for (k = 0; k < 16u; ++k) {
    draws(seed, e, hop, redraw_base + k, w);
    for (i = 0; i < 4; ++i)
      if ((w[i] >> 1) != 0u) return w[i] >> 1;
  }
  return 1u;
}

/* The selection and dwell words of hop h of exciton (or contact particle) e: hops 2P and 2P + 1
 * share the Philox call of counter P = h >> 1 on stream 0, the even hop taking words 0 and 1,
 * the odd hop words 2 and 3. */
static void hop_words(uint64_t seed, uint32_t e, uint64_t h, uint32_t* sel, uint32_t* dw) {
  uint32_t w[4];
  draws(seed, e, h >> 1, 0u, w);
  *sel = w[2 * (unsigned)(h & 1u)];
  *dw = w[2 * (unsigned)(h & 1u) + 1];
}

/* ff_time's draw for the dwell word dw of hop h: zero moves to streams 2 + k at counter h */
static uint32_t hop_dwell_draw(uint64_t seed, uint32_t e, uint64_t h, uint32_t dw) {
  uint32_t k, i, w[4];
  if ((dw >> 1) != 0u) return dw >> 1;
  for (k = 0; k < 16u; ++k) {
    draws(seed, e, h, 2u + k, w);
    for (i = 0; i < 4; ++i)
      if ((w[i] >> 1) != 0u) return w[i] >> 1;
  }
  return 1u;
}

/* the source-site word of a re-injection after a drain at hop h (stream 0x10000000) */
static uint32_t reinject_word(uint64_t seed, uint32_t e, uint64_t h) {
  uint32_t w[4];
  draws(seed, e, h, 0x10000000u, w);
  return w[0];
}

/* ------------------------------------------------------------------------------------------
 * Armadillo-order small-vector arithmetic (op_dot::direct_dot_arma, vec_norm_2_direct_std):
 * two accumulators over even / odd elements, summed at the end.
 * ---------------------------------------------------------------------------------------- */
static double arma_dot(const double* a, const double* b, int n) {
  double v1 = 0.0, v2 = 0.0;
  int i, j;
  for (i = 0, j = 1; j < n; i += 2, j += 2) {
    v1 += a[i] * b[i];
    v2 += a[j] * b[j];
  }
  if (i < n) v1 += a[i] * b[i];
  return v1 + v2;
}

static double arma_norm(const double* a, int n) {
  double acc1 = 0.0, acc2 = 0.0;
  int i, j;
  for (i = 0, j = 1; j < n; i += 2, j += 2) {
    acc1 += a[i] * a[i];
    acc2 += a[j] * a[j];
  }
  if (i < n) acc1 += a[i] * a[i];
  return sqrt(acc1 + acc2);
}

/* arma::linspace (used by create_*_scatt_table, monte_carlo.cpp:66-75, 158-167) */
int orc_linspace(double start, double end, int n, double* out) {
  int i;
  double delta;
  if (n < 1) return -1;
  if (n == 1) {
    out[0] = end;
    return 0;
  }
  delta = (end - start) / (double)(n - 1);
  for (i = 0; i < n - 1; ++i) out[i] = start + (double)i * delta;
  out[n - 1] = end;
  return 0;
}

/* arma::normalise: divide by the norm unless it is zero */
static void arma_normalise(const double* v, double* out) {
  double nv = arma_norm(v, 3);
  double d = (nv != 0.0) ? nv : 1.0;
  out[0] = v[0] / d;
  out[1] = v[1] / d;
  out[2] = v[2] / d;
}

/* create_forster_scatt_table (src/monte_carlo/monte_carlo.cpp:156-200) */
int orc_forster_table(double gamma_0, int n_th, const double* th, int n_z, const double* zs,
                      int n_a1, const double* a1, int n_a2, const double* a2, double* rates) {
  int it, iz, i1, i2;
  size_t idx = 0;
  for (it = 0; it < n_th; ++it)
    for (iz = 0; iz < n_z; ++iz)
      for (i1 = 0; i1 < n_a1; ++i1)
        for (i2 = 0; i2 < n_a2; ++i2) {
          double r1[3], r2[3], dR[3], n1[3], n2[3], nd[3], af;
          r1[0] = a1[i1]; r1[1] = 0.0; r1[2] = 0.0;
          r2[0] = a2[i2] * cos(th[it]); r2[1] = a2[i2] * sin(th[it]); r2[2] = zs[iz];
          dR[0] = r1[0] - r2[0]; dR[1] = r1[1] - r2[1]; dR[2] = r1[2] - r2[2];
          arma_normalise(r1, n1);
          arma_normalise(r2, n2);
          arma_normalise(dR, nd);
          af = cos(th[it]) - 3 * arma_dot(n1, nd, 3) * arma_dot(n2, nd, 3);
          rates[idx++] = gamma_0 * pow(af, 2) * pow(1.e-9 / arma_norm(dR, 3), 6);
        }
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * Network: sites, cell list, rate table
 * ---------------------------------------------------------------------------------------- */
typedef struct orc_net {
  int64_t n;
  double *p[3], *o[3];
  uint8_t *chir, *flags;
  double cutoff, min_dist;
  int periodic;
  double period_lo[3], period_len[3];
  /* cells */
  int dims[3];
  double lo[3], cell[3];
  int64_t n_cells;
  int64_t* cell_start;
  int32_t* cell_sites;
  /* table */
  int t_nchir, t_n[4];
  double* t_grid[4];
  double* t_rates;
  /* source list */
  int64_t n_src;
  int32_t* src;
} orc_net;

static int axis_cell(const orc_net* w, int a, double p) {
  if (w->periodic & (1 << a)) {
    int i = (int)floor((p - w->lo[a]) / w->cell[a]);
    if (i < 0) i = 0;
    if (i >= w->dims[a]) i = w->dims[a] - 1;
    return i;
  }
  /* int ix = (s.pos(0) - xmin) / radius;  (monte_carlo.h:390) */
  return (int)((p - w->lo[a]) / w->cell[a]);
}

void orc_net_free(orc_net* w) {
  int a;
  if (!w) return;
  for (a = 0; a < 3; ++a) {
    free(w->p[a]);
    free(w->o[a]);
  }
  for (a = 0; a < 4; ++a) free(w->t_grid[a]);
  free(w->chir); free(w->flags); free(w->cell_start); free(w->cell_sites); free(w->t_rates);
  free(w->src);
  free(w);
}

/* create_scatterer_buckets (src/monte_carlo/monte_carlo.h:368-416), plus a periodic axis
 * option (period split into floor(period / cutoff) equal cells). */
orc_net* orc_net_create(int64_t n, const double* x, const double* y, const double* z,
                        const double* ox, const double* oy, const double* oz,
                        const uint8_t* chir, const uint8_t* flags, double cutoff, double min_dist,
                        int periodic, const double* period_lo, const double* period_len) {
  orc_net* w = (orc_net*)calloc(1, sizeof(orc_net));
  const double* pin[3];
  const double* oin[3];
  int64_t i, *fill;
  int a;
  if (!w || n <= 0) return NULL;
  pin[0] = x; pin[1] = y; pin[2] = z;
  oin[0] = ox; oin[1] = oy; oin[2] = oz;
  w->n = n;
  for (a = 0; a < 3; ++a) {
    w->p[a] = (double*)malloc(n * sizeof(double));
    w->o[a] = (double*)malloc(n * sizeof(double));
    memcpy(w->p[a], pin[a], n * sizeof(double));
    memcpy(w->o[a], oin[a], n * sizeof(double));
  }
  w->chir = (uint8_t*)calloc(n, 1);
  w->flags = (uint8_t*)calloc(n, 1);
  if (chir) memcpy(w->chir, chir, n);
  if (flags) memcpy(w->flags, flags, n);
  w->cutoff = cutoff;
  w->min_dist = min_dist;
  w->periodic = periodic;
  for (a = 0; a < 3; ++a) {
    w->period_lo[a] = period_lo ? period_lo[a] : 0.0;
    w->period_len[a] = period_len ? period_len[a] : 0.0;
  }
  /* domain: bounding box of the sites (find_simulation_domain, monte_carlo.h:328-340) */
  for (a = 0; a < 3; ++a) {
    double mn = w->p[a][0], mx = w->p[a][0];
    for (i = 0; i < n; ++i) {
      mn = mn > w->p[a][i] ? w->p[a][i] : mn;
      mx = mx < w->p[a][i] ? w->p[a][i] : mx;
    }
    if (periodic & (1 << a)) {
      int m = (int)floor(w->period_len[a] / cutoff);
      w->dims[a] = m < 1 ? 1 : m;
      w->lo[a] = w->period_lo[a];
      w->cell[a] = w->period_len[a] / w->dims[a];
    } else {
      w->dims[a] = (int)ceil((mx - mn) / cutoff) + 1; /* monte_carlo.h:377 */
      w->lo[a] = mn;
      w->cell[a] = cutoff;
    }
  }
  w->n_cells = (int64_t)w->dims[0] * w->dims[1] * w->dims[2];
  w->cell_start = (int64_t*)calloc(w->n_cells + 1, sizeof(int64_t));
  w->cell_sites = (int32_t*)malloc(n * sizeof(int32_t));
  fill = (int64_t*)calloc(w->n_cells, sizeof(int64_t));
  /* bucket contents in site order (monte_carlo.h:389-395) */
  for (i = 0; i < n; ++i) {
    int64_t c = (int64_t)axis_cell(w, 0, w->p[0][i]) +
                (int64_t)axis_cell(w, 1, w->p[1][i]) * w->dims[0] +
                (int64_t)axis_cell(w, 2, w->p[2][i]) * w->dims[0] * w->dims[1];
    w->cell_start[c + 1]++;
  }
  for (i = 0; i < w->n_cells; ++i) w->cell_start[i + 1] += w->cell_start[i];
  for (i = 0; i < n; ++i) {
    int64_t c = (int64_t)axis_cell(w, 0, w->p[0][i]) +
                (int64_t)axis_cell(w, 1, w->p[1][i]) * w->dims[0] +
                (int64_t)axis_cell(w, 2, w->p[2][i]) * w->dims[0] * w->dims[1];
    w->cell_sites[w->cell_start[c] + fill[c]++] = (int32_t)i;
  }
  free(fill);
  w->src = (int32_t*)malloc(n * sizeof(int32_t));
  for (i = 0; i < n; ++i)
    if (w->flags[i] & 1u) w->src[w->n_src++] = (int32_t)i;
  return w;
}

int orc_net_cells(const orc_net* w, int32_t dims[3], int64_t* n_cells, int64_t* cell_start,
                  int32_t* cell_sites) {
  if (dims) {
    dims[0] = w->dims[0]; dims[1] = w->dims[1]; dims[2] = w->dims[2];
  }
  if (n_cells) *n_cells = w->n_cells;
  if (cell_start) memcpy(cell_start, w->cell_start, (w->n_cells + 1) * sizeof(int64_t));
  if (cell_sites) memcpy(cell_sites, w->cell_sites, w->n * sizeof(int32_t));
  return 0;
}

int orc_net_set_table(orc_net* w, int n_chir, int n_th, const double* th, int n_z,
                      const double* zs, int n_a1, const double* a1, int n_a2, const double* a2,
                      const double* rates) {
  const double* g[4];
  int n[4], a;
  size_t nr;
  g[0] = th; g[1] = zs; g[2] = a1; g[3] = a2;
  n[0] = n_th; n[1] = n_z; n[2] = n_a1; n[3] = n_a2;
  w->t_nchir = n_chir;
  for (a = 0; a < 4; ++a) {
    free(w->t_grid[a]);
    w->t_n[a] = n[a];
    w->t_grid[a] = (double*)malloc(n[a] * sizeof(double));
    memcpy(w->t_grid[a], g[a], n[a] * sizeof(double));
  }
  nr = (size_t)n_chir * n_chir * n_th * n_z * n_a1 * n_a2;
  free(w->t_rates);
  w->t_rates = (double*)malloc(nr * sizeof(double));
  memcpy(w->t_rates, rates, nr * sizeof(double));
  return 0;
}

/* tmp = abs(grid - x); tmp.index_min()  (scattering_struct.h:42-49) */
static unsigned arma_index_min_abs(const double* grid, int n, double x) {
  double best = INFINITY;
  unsigned idx = 0;
  int i;
  for (i = 0; i < n; ++i) {
    double v = fabs(grid[i] - x);
    if (v < best) {
      best = v;
      idx = (unsigned)i;
    }
  }
  return idx;
}

/* scattering_struct::get_rate (scattering_struct.h:38-52), per chirality pair */
static double get_rate(const orc_net* w, int ci, int cj, double theta, double zsh, double a1,
                       double a2) {
  unsigned ith = arma_index_min_abs(w->t_grid[0], w->t_n[0], theta);
  unsigned iz = arma_index_min_abs(w->t_grid[1], w->t_n[1], zsh);
  unsigned i1 = arma_index_min_abs(w->t_grid[2], w->t_n[2], a1);
  unsigned i2 = arma_index_min_abs(w->t_grid[3], w->t_n[3], a2);
  size_t tab = (size_t)ci * w->t_nchir + cj;
  size_t idx = (((tab * w->t_n[0] + ith) * w->t_n[1] + iz) * w->t_n[2] + i1) * w->t_n[3] + i2;
  return w->t_rates[idx];
}

/* get_rate over caller-given tables (scattering_struct.h:38-52) for n queries; nodes [n][4]
 * receives the node indices (optional).  Tables as orc_net_set_table takes them. */
int orc_table_rate(int n_chir, int n_th, const double* th, int n_z, const double* zs, int n_a1,
                   const double* a1, int n_a2, const double* a2, const double* rates, int64_t n,
                   const uint8_t* ci, const uint8_t* cj, const double* q_th, const double* q_z,
                   const double* q_a1, const double* q_a2, double* out, int32_t* nodes) {
  int64_t k;
  for (k = 0; k < n; ++k) {
    const unsigned i0 = arma_index_min_abs(th, n_th, q_th[k]);
    const unsigned i1 = arma_index_min_abs(zs, n_z, q_z[k]);
    const unsigned i2 = arma_index_min_abs(a1, n_a1, q_a1[k]);
    const unsigned i3 = arma_index_min_abs(a2, n_a2, q_a2[k]);
    const size_t tab = (size_t)(ci ? ci[k] : 0) * (size_t)n_chir + (size_t)(cj ? cj[k] : 0);
    out[k] = rates[(((tab * n_th + i0) * n_z + i1) * n_a1 + i2) * n_a2 + i3];
    if (nodes) {
      nodes[4 * k] = (int32_t)i0;
      nodes[4 * k + 1] = (int32_t)i1;
      nodes[4 * k + 2] = (int32_t)i2;
      nodes[4 * k + 3] = (int32_t)i3;
    }
  }
  return 0;
}

/* check_neighbor geometry (src/monte_carlo/scatterer.cpp:44-61) */
static void orc_geometry(const double* pi, const double* pj, const double* a1, const double* a2,
                         double* theta, double* zsh, double* ash1, double* ash2) {
  double dR[3], cosT, y1, y2, s2, v[3];
  int k;
  for (k = 0; k < 3; ++k) dR[k] = pi[k] - pj[k];
  cosT = arma_dot(a1, a2, 3);
  if (cosT == 1) {
    double d;
    *ash1 = 0;
    *ash2 = arma_dot(dR, a1, 3);
    *theta = 0;
    d = arma_dot(dR, a1, 3);
    for (k = 0; k < 3; ++k) v[k] = dR[k] - d * a1[k];
    *zsh = arma_norm(v, 3);
  } else {
    *theta = acos(cosT);
    y1 = arma_dot(a1, dR, 3);
    y2 = arma_dot(a2, dR, 3);
    s2 = 1 - pow(cosT, 2);
    *ash1 = (y1 + y2 * cosT) / s2;
    *ash2 = (y2 + y1 * cosT) / s2;
    for (k = 0; k < 3; ++k) v[k] = (*ash1 * a1[k] + pi[k]) - (*ash2 * a2[k] + pj[k]);
    *zsh = arma_norm(v, 3);
  }
}

/* image of coordinate q seen from p on a periodic axis of length L */
static double orc_image(double p, double q, double L) {
  double d = p - q;
  if (d > 0.5 * L) return q + L;
  if (d < -0.5 * L) return q - L;
  return q;
}

typedef struct {
  int32_t j;
  double pj[3];
} orc_cand;

/* Enumerate the neighbours of site i in reference order: stencil cells x-outer, y, z-inner
 * (monte_carlo.h:402-411), buckets in site order, keep min_dist < |dR| < cutoff
 * (scatterer.cpp:40-43).  Returns the count; fills out (capacity cap) when non-NULL. */
static int64_t orc_neighbors(const orc_net* w, int64_t i, orc_cand* out, int64_t cap) {
  int st[3][3], ns[3], a, b, c, d, k;
  int64_t cnt = 0, q;
  double pi[3];
  for (a = 0; a < 3; ++a) pi[a] = w->p[a][i];
  for (a = 0; a < 3; ++a) {
    int ia = axis_cell(w, a, pi[a]);
    ns[a] = 0;
    for (d = -1; d <= 1; ++d) {
      int v = ia + d;
      if (w->periodic & (1 << a)) {
        int dup = 0;
        v = ((v % w->dims[a]) + w->dims[a]) % w->dims[a];
        for (k = 0; k < ns[a]; ++k) dup |= (st[a][k] == v);
        if (!dup) st[a][ns[a]++] = v;
      } else if (v > -1 && v < w->dims[a]) {
        st[a][ns[a]++] = v;
      }
    }
  }
  for (a = 0; a < ns[0]; ++a)
    for (b = 0; b < ns[1]; ++b)
      for (c = 0; c < ns[2]; ++c) {
        int64_t cell = (int64_t)st[0][a] + (int64_t)st[1][b] * w->dims[0] +
                       (int64_t)st[2][c] * w->dims[0] * w->dims[1];
        for (q = w->cell_start[cell]; q < w->cell_start[cell + 1]; ++q) {
          int32_t j = w->cell_sites[q];
          double pj[3], dR[3], dist;
          for (k = 0; k < 3; ++k) {
            pj[k] = w->p[k][j];
            if (w->periodic & (1 << k)) pj[k] = orc_image(pi[k], pj[k], w->period_len[k]);
            dR[k] = pi[k] - pj[k];
          }
          dist = arma_norm(dR, 3);
          if ((dist < w->cutoff) && (dist > w->min_dist)) {
            if (out && cnt < cap) {
              out[cnt].j = j;
              for (k = 0; k < 3; ++k) out[cnt].pj[k] = pj[k];
            }
            ++cnt;
          }
        }
      }
  return cnt;
}

int64_t orc_net_count_pairs(const orc_net* w, int64_t* row_ptr) {
  int64_t i;
  row_ptr[0] = 0;
  for (i = 0; i < w->n; ++i) row_ptr[i + 1] = row_ptr[i] + orc_neighbors(w, i, NULL, 0);
  return row_ptr[w->n];
}

/* find_neighbors with inclusive FP64 prefix sum (scatterer.cpp:34-83).  mode 0: table
 * rates; mode 1: pair_rates[nnz] given in CSR order.  Any output may be NULL. */
int orc_net_fill(const orc_net* w, int mode, const int64_t* row_ptr, const double* pair_rates,
                 int32_t* col, double* cum, float* disp, double* g_th, double* g_z,
                 double* g_a1, double* g_a2) {
  int64_t i;
  int64_t maxdeg = 0;
  for (i = 0; i < w->n; ++i)
    if (row_ptr[i + 1] - row_ptr[i] > maxdeg) maxdeg = row_ptr[i + 1] - row_ptr[i];
#pragma omp parallel for schedule(dynamic, 256)
  for (i = 0; i < w->n; ++i) {
    orc_cand* nb = (orc_cand*)malloc((maxdeg + 1) * sizeof(orc_cand));
    int64_t m = orc_neighbors(w, i, nb, maxdeg + 1), q;
    double pi[3], ai[3];
    int k;
    for (k = 0; k < 3; ++k) {
      pi[k] = w->p[k][i];
      ai[k] = w->o[k][i];
    }
    for (q = 0; q < m; ++q) {
      int64_t pos = row_ptr[i] + q;
      int32_t j = nb[q].j;
      double aj[3], th, zs, s1, s2, rate;
      for (k = 0; k < 3; ++k) aj[k] = w->o[k][j];
      orc_geometry(pi, nb[q].pj, ai, aj, &th, &zs, &s1, &s2);
      if (col) col[pos] = j;
      if (disp)
        for (k = 0; k < 3; ++k) disp[3 * pos + k] = (float)(nb[q].pj[k] - pi[k]);
      if (g_th) g_th[pos] = th;
      if (g_z) g_z[pos] = zs;
      if (g_a1) g_a1[pos] = s1;
      if (g_a2) g_a2[pos] = s2;
      if (cum) {
        rate = (mode == 0) ? get_rate(w, w->chir[i], w->chir[j], th, zs, s1, s2) : pair_rates[pos];
        cum[pos] = (q == 0) ? rate : rate + cum[pos - 1];
      }
    }
    free(nb);
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * Excitons
 * ---------------------------------------------------------------------------------------- */
/* scatterer::update_state binary search (scatterer.cpp:17-30); cum = the row, n >= 1 */
static unsigned orc_select(const double* cum, unsigned n, double dice) {
  unsigned left = (unsigned)-1;
  unsigned right = n - 1;
  while (left + 1 < right) {
    unsigned mid = (left + right) / 2;
    if (cum[mid] <= dice)
      left = mid;
    else
      right = mid;
  }
  return right;
}

int orc_select_index(const double* cum, int n, double dice) {
  return (int)orc_select(cum, (unsigned)n, dice);
}

/* ff_time (scatterer.h:74-80) with the cached inverse rate of set_max_rate (scatterer.h:92) */
static double orc_dwell(const double* row, int64_t n, uint32_t r) {
  double max_rate, inv;
  if (n == 0 || !(row[n - 1] > 0.0)) return INFINITY; /* no neighbour, or 1/Gamma = inf */
  max_rate = row[n - 1];
  inv = 1. / max_rate;
  return -inv * log((double)r / (double)ORC_RAND_MAX);
}

/* Mismatches between x / 2147483647 (IEEE division) and the divide-free form the device uses,
 * q0 = x INV, rem = fma(-q0, D, x), q = fma(rem, INV, q0) with INV = RN(1 / D): over every
 * integer r in [lo, hi) (the dwell draws), then over n_random doubles with random significands
 * and exponents 2^30..2^70 (the selection dice Gamma r / D), from an xorshift stream. */
int64_t orc_div_rand_max_mismatches(uint32_t lo, uint32_t hi, int64_t n_random, uint64_t seed) {
  const double D = 2147483647.0, INV = 1.0 / 2147483647.0;
  int64_t bad = 0, i;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (i = (int64_t)lo; i < (int64_t)hi; ++i) {
    const double x = (double)i;
    const double q0 = x * INV;
    const double q = fma(fma(-q0, D, x), INV, q0);
    if (q != x / D) ++bad;
  }
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (i = 0; i < n_random; ++i) {
    uint64_t s = seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull), bits;
    double x, q0, q;
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    bits = (s & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 + 30 + (int)(i % 41)) << 52);
    memcpy(&x, &bits, sizeof x);
    q0 = x * INV;
    q = fma(fma(-q0, D, x), INV, q0);
    if (q != x / D) ++bad;
  }
  return bad;
}

double orc_dwell_time(double max_rate, uint32_t r) {
  double inv = 1. / max_rate;
  return -inv * log((double)r / (double)ORC_RAND_MAX);
}

#define FNV_OFF 0xcbf29ce484222325ull
#define FNV_PR 0x100000001b3ull

/* kubo_create_particles + particle constructor (monte_carlo.cpp:308-316, particle.h:48-52) */
int orc_inject(int64_t n_sites, const int64_t* row_ptr, const double* cum, int64_t n_src,
               const int32_t* src, uint64_t seed, int64_t E, int32_t* site, double* t_next,
               uint64_t* hops, double* disp, uint32_t* drains, uint64_t* hash, uint32_t* k_rec) {
  int64_t e;
  (void)n_sites;
  if (n_src <= 0) return -1;
#pragma omp parallel for
  for (e = 0; e < E; ++e) {
    uint32_t w[4], r;
    int32_t s;
    draws(seed, (uint32_t)e, 0, 1u, w);
    s = src[(w[0] >> 1) % (uint32_t)n_src];
    r = dwell_draw(seed, (uint32_t)e, 0, w[1], w[3], 0x80000000u);
    site[e] = s;
    t_next[e] = 0.0 + orc_dwell(cum + row_ptr[s], row_ptr[s + 1] - row_ptr[s], r);
    hops[e] = 0;
    disp[3 * e] = disp[3 * e + 1] = disp[3 * e + 2] = 0.0;
    drains[e] = 0;
    hash[e] = (FNV_OFF ^ (uint64_t)(uint32_t)s) * FNV_PR;
    if (k_rec) k_rec[e] = 0;
  }
  return 0;
}

/* Flight state at injection (particle.h:50-51: heading = rand() % 2, drawn after the dwell):
 * word 2 of the injection draw; offset 0, last event at t = 0. */
int orc_inject_flight(uint64_t seed, int64_t E, uint8_t* heading, double* s_off, double* t_last) {
  int64_t e;
  for (e = 0; e < E; ++e) {
    uint32_t w[4];
    draws(seed, (uint32_t)e, 0, 1u, w);
    heading[e] = (uint8_t)((w[2] >> 1) & 1u);
    s_off[e] = 0.0;
    t_last[e] = 0.0;
  }
  return 0;
}

typedef struct orc_run_args {
  /* CSR */
  int64_t n_sites;
  const int64_t* row_ptr;
  const int32_t* col;
  const double* cum;
  const float* pos; /* [n_sites][3] FP32 site positions: displacement = pos[target] - pos[source] */
  const uint8_t* flags;
  int64_t n_src;
  const int32_t* src;
  uint64_t seed;
  /* state [E] */
  int64_t E;
  int32_t* site;
  double* t_next;
  uint64_t* hops;
  double* dsp; /* [E][3] */
  uint32_t* drains;
  uint64_t* hash;
  uint32_t* k_rec;
  /* budget */
  int64_t hop_budget;
  double t_max;
  /* samples [T][E] */
  int32_t n_bins;
  double t_sample_max;
  double* rec_disp; /* [T][E][3] */
  uint32_t* rec_drain;
  /* counts [n_sites], may be NULL */
  uint64_t* site_count;
  /* trace [n_trace][tr_max] */
  int64_t n_trace, tr_max;
  int32_t* tr_site;
  double* tr_time;
  int64_t* tr_n;
  /* faithful mode: recompute neighbours per hop from this network (NULL = use the CSR) */
  const orc_net* net;
  int n_threads;
  double per_len[3]; /* period per axis (0 = open): displacements take the minimum image */
  /* drift along the tubes (particle::fly) when velocity > 0 */
  double velocity;
  const int32_t* left; /* [n_sites] same-tube links, -1 = tube end */
  const int32_t* right;
  uint8_t* heading;    /* [E] 1 = toward right */
  double* s_off;       /* [E] distance beyond the current site toward the next one */
  double* t_last;      /* [E] time of the last hop or injection */
  int32_t* fly_err;    /* set to 1 when a flight passes more than ORC_FLY_MAX_PASSES sites */
} orc_run_args;

#define ORC_FLY_MAX_PASSES (1u << 22)

/* The exciton between events when v > 0: current site, distance travelled from it toward the
 * next site along the heading, displacement of the current site's position, drains, hash. */
typedef struct orc_flight {
  int32_t s;
  double off;
  unsigned heading;
  double d[3];
  uint32_t nd;
  uint64_t hs;
} orc_flight;

/* One coordinate of a hop's displacement: the difference of two FP32 positions in FP64 (exact
 * for sites this close), shifted by one period on a periodic axis when it exceeds half of it. */
static double orc_min_image(double d, double period) {
  if (period > 0.0) {
    if (d > 0.5 * period) d -= period;
    else if (d < -0.5 * period) d += period;
  }
  return d;
}

/* Next site along the heading; at a tube end the other link, reversing the heading
 * (particle.cpp:18-37). */
static int32_t orc_fly_next(const orc_run_args* a, int32_t s, unsigned* heading) {
  int32_t l = a->left[s], r = a->right[s], nx;
  if (*heading) nx = r > -1 ? r : l;
  else nx = l > -1 ? l : r;
  *heading = (nx == r) ? 1u : 0u;
  return nx;
}

static void orc_fly_segment(const orc_run_args* a, int32_t s, int32_t n, double sd[3]) {
  int k;
  for (k = 0; k < 3; ++k)
    sd[k] = orc_min_image((double)a->pos[3 * n + k] - (double)a->pos[3 * s + k], a->per_len[k]);
}

/* particle::fly (particle.cpp:9-54) over a whole dwell tau: pass sites while the remaining
 * distance takes less than tau; a drain site reached counts a drain and re-injects at a source
 * site drawn from stream 0x40000000 + j (the j-th drain of this flight) at hop counter h, and the
 * flight goes on from there (kubo_step's removal, monte_carlo.cpp:330-336, at the moment the
 * site is reached).  Returns 0 when the pass bound is exceeded. */
static int orc_fly(const orc_run_args* a, orc_flight* f, double tau, uint32_t e, uint64_t h) {
  uint32_t j = 0, pass;
  if (a->left[f->s] < 0 && a->right[f->s] < 0) return 1; /* particle.cpp:11-12 */
  for (pass = 0; pass < ORC_FLY_MAX_PASSES; ++pass) {
    int32_t nx = orc_fly_next(a, f->s, &f->heading);
    double sd[3], len, dist;
    orc_fly_segment(a, f->s, nx, sd);
    len = arma_norm(sd, 3);
    dist = len - f->off;
    if (!(dist / a->velocity < tau)) {
      f->off = f->off + a->velocity * tau;
      return 1;
    }
    tau = tau - dist / a->velocity;
    f->d[0] = f->d[0] + sd[0];
    f->d[1] = f->d[1] + sd[1];
    f->d[2] = f->d[2] + sd[2];
    f->off = 0.0;
    f->s = nx;
    f->hs = (f->hs ^ ((uint64_t)(uint32_t)nx | (2ull << 32))) * FNV_PR;
    if (a->flags[nx] & 2u) {
      uint32_t w[4];
      ++f->nd;
      draws(a->seed, e, h, 0x40000000u + j, w);
      ++j;
      f->s = a->src[(w[0] >> 1) % (uint32_t)a->n_src];
      f->hs = (f->hs ^ ((uint64_t)(uint32_t)f->s | (1ull << 32))) * FNV_PR;
    }
    if (a->left[f->s] < 0 && a->right[f->s] < 0) return 1;
  }
  return 0;
}

/* The exciton's position relative to its current site: off along the segment to the next. */
static void orc_fly_offset(const orc_run_args* a, const orc_flight* f, double o[3]) {
  unsigned hd = f->heading;
  int32_t nx;
  double sd[3], q;
  o[0] = o[1] = o[2] = 0.0;
  if (!(f->off != 0.0)) return;
  nx = orc_fly_next(a, f->s, &hd);
  orc_fly_segment(a, f->s, nx, sd);
  q = f->off / arma_norm(sd, 3);
  o[0] = sd[0] * q;
  o[1] = sd[1] * q;
  o[2] = sd[2] * q;
}

/* One exciton through its budget: particle::step (particle.cpp:57-80) in event form,
 * update_state (scatterer.cpp:9-31), ff_time (scatterer.h:74-80), re-injection on a drain
 * site keeping delta_pos and ff_time (monte_carlo.cpp:330-336). */
static void orc_one(const orc_run_args* a, int64_t e, orc_cand* nb, double* rowbuf,
                    int32_t* colbuf, int64_t cap) {
  int32_t s = a->site[e];
  double t = a->t_next[e];
  uint64_t h = a->hops[e];
  double dx = a->dsp[3 * e], dy = a->dsp[3 * e + 1], dz = a->dsp[3 * e + 2];
  uint32_t nd = a->drains[e];
  uint64_t hs = a->hash[e];
  uint32_t k = a->k_rec ? a->k_rec[e] : 0;
  int64_t tr = (a->tr_n && e < a->n_trace) ? a->tr_n[e] : 0;
  const uint64_t h_end = h + (uint64_t)a->hop_budget;
  const double dts = a->n_bins > 0 ? a->t_sample_max / a->n_bins : 0.0;
  const int fly = a->velocity > 0.0;
  double tl = fly ? a->t_last[e] : 0.0, off = fly ? a->s_off[e] : 0.0;
  unsigned hd = fly ? a->heading[e] : 0u;

  for (;;) {
    const double* row;
    const int32_t* rcol;
    int64_t n;
    uint32_t w_sel, w_dw, r;
    unsigned idx;
    double dice;
    int32_t j;
    int drain;
    double t_hop;

    while ((int32_t)k < a->n_bins) {
      double tk = (double)(k + 1) * dts;
      size_t rr;
      if (!(tk < t)) break;
      rr = (size_t)k * a->E + e;
      if (fly) { /* the flight from the last event up to tk */
        orc_flight g;
        double o[3];
        g.s = s; g.off = off; g.heading = hd; g.d[0] = dx; g.d[1] = dy; g.d[2] = dz;
        g.nd = nd; g.hs = hs;
        if (!orc_fly(a, &g, tk - tl, (uint32_t)e, h)) {
          *a->fly_err = 1;
          break;
        }
        orc_fly_offset(a, &g, o);
        a->rec_disp[3 * rr] = g.d[0] + o[0];
        a->rec_disp[3 * rr + 1] = g.d[1] + o[1];
        a->rec_disp[3 * rr + 2] = g.d[2] + o[2];
        a->rec_drain[rr] = g.nd;
      } else {
        a->rec_disp[3 * rr] = dx;
        a->rec_disp[3 * rr + 1] = dy;
        a->rec_disp[3 * rr + 2] = dz;
        a->rec_drain[rr] = nd;
      }
      ++k;
    }
    if (fly) { /* fly through the dwell, then hop from the site reached (particle.cpp:62-75) */
      orc_flight g;
      if (h >= h_end || !(t <= a->t_max) || !(t < INFINITY)) break;
      g.s = s; g.off = off; g.heading = hd; g.d[0] = dx; g.d[1] = dy; g.d[2] = dz;
      g.nd = nd; g.hs = hs;
      if (!orc_fly(a, &g, t - tl, (uint32_t)e, h)) {
        *a->fly_err = 1;
        break;
      }
      s = g.s; off = g.off; hd = g.heading; dx = g.d[0]; dy = g.d[1]; dz = g.d[2];
      nd = g.nd; hs = g.hs;
    }
    /* current site row: cached CSR or recomputed (find_neighbors on every hop, scatterer.cpp:12) */
    if (a->net) {
      const orc_net* w0 = a->net;
      int64_t q;
      double pi[3], ai[3];
      int kk;
      n = orc_neighbors(w0, s, nb, cap);
      for (kk = 0; kk < 3; ++kk) {
        pi[kk] = w0->p[kk][s];
        ai[kk] = w0->o[kk][s];
      }
      for (q = 0; q < n; ++q) {
        double aj[3], th, zs, s1, s2, rate;
        int32_t jj = nb[q].j;
        for (kk = 0; kk < 3; ++kk) aj[kk] = w0->o[kk][jj];
        orc_geometry(pi, nb[q].pj, ai, aj, &th, &zs, &s1, &s2);
        rate = get_rate(w0, w0->chir[s], w0->chir[jj], th, zs, s1, s2);
        rowbuf[q] = (q == 0) ? rate : rate + rowbuf[q - 1];
        colbuf[q] = jj;
      }
      row = rowbuf;
      rcol = colbuf;
    } else {
      row = a->cum + a->row_ptr[s];
      rcol = a->col + a->row_ptr[s];
      n = a->row_ptr[s + 1] - a->row_ptr[s];
    }
    if (fly) {
      if (n == 0 || !(row[n - 1] > 0.0)) { /* no rates: the event leaves it in place */
        tl = t;
        t = INFINITY;
        ++h;
        continue;
      }
    } else if (h >= h_end || n == 0 || !(row[n - 1] > 0.0) || !(t <= a->t_max)) {
      break;
    }

    hop_words(a->seed, (uint32_t)e, h, &w_sel, &w_dw);
    dice = row[n - 1] * (double)(w_sel >> 1) / ORC_RAND_MAX;
    idx = orc_select(row, (unsigned)n, dice);
    j = rcol[idx];
    if (a->site_count) {
#pragma omp atomic
      a->site_count[s] += 1;
    }
    dx = dx + orc_min_image((double)a->pos[3 * j] - (double)a->pos[3 * s], a->per_len[0]);
    dy = dy + orc_min_image((double)a->pos[3 * j + 1] - (double)a->pos[3 * s + 1], a->per_len[1]);
    dz = dz + orc_min_image((double)a->pos[3 * j + 2] - (double)a->pos[3 * s + 2], a->per_len[2]);
    hs = (hs ^ (uint64_t)(uint32_t)j) * FNV_PR;
    r = hop_dwell_draw(a->seed, (uint32_t)e, h, w_dw);
    t_hop = t;
    {
      double tj;
      if (a->net) {
        /* the dwell uses the new site's cached 1/Gamma (set_max_rate) */
        int64_t nj = orc_neighbors(a->net, j, nb, cap), q;
        double gj = 0.0;
        const orc_net* w0 = a->net;
        double pj[3], aj[3];
        int kk;
        for (kk = 0; kk < 3; ++kk) {
          pj[kk] = w0->p[kk][j];
          aj[kk] = w0->o[kk][j];
        }
        for (q = 0; q < nj; ++q) {
          double ak[3], th, zs, s1, s2, rate;
          int32_t jj = nb[q].j;
          for (kk = 0; kk < 3; ++kk) ak[kk] = w0->o[kk][jj];
          orc_geometry(pj, nb[q].pj, aj, ak, &th, &zs, &s1, &s2);
          rate = get_rate(w0, w0->chir[j], w0->chir[jj], th, zs, s1, s2);
          gj = (q == 0) ? rate : rate + gj;
        }
        tj = (nj == 0 || !(gj > 0.0)) ? INFINITY : orc_dwell_time(gj, r);
      } else {
        tj = orc_dwell(a->cum + a->row_ptr[j], a->row_ptr[j + 1] - a->row_ptr[j], r);
      }
      t = t + tj;
    }
    s = j;
    drain = (a->flags[j] & 2u) != 0;
    if (drain) {
      ++nd;
      s = a->src[(reinject_word(a->seed, (uint32_t)e, h) >> 1) % (uint32_t)a->n_src];
      hs = (hs ^ ((uint64_t)(uint32_t)s | (1ull << 32))) * FNV_PR;
    }
    if (fly) {
      off = 0.0;
      tl = t_hop;
    }
    if (a->tr_site && e < a->n_trace && tr < a->tr_max) {
      a->tr_site[(size_t)e * a->tr_max + tr] = s;
      a->tr_time[(size_t)e * a->tr_max + tr] = t_hop;
      ++tr;
    }
    ++h;
  }
  a->site[e] = s;
  a->t_next[e] = t;
  a->hops[e] = h;
  a->dsp[3 * e] = dx;
  a->dsp[3 * e + 1] = dy;
  a->dsp[3 * e + 2] = dz;
  a->drains[e] = nd;
  a->hash[e] = hs;
  if (a->k_rec) a->k_rec[e] = k;
  if (a->tr_n && e < a->n_trace) a->tr_n[e] = tr;
  if (fly) {
    a->s_off[e] = off;
    a->t_last[e] = tl;
    a->heading[e] = (uint8_t)hd;
  }
}

int orc_run(const orc_run_args* a) {
  int64_t e, cap = 0;
  if (a->net) {
    int64_t i;
    for (i = 0; i < a->net->n; ++i) {
      int64_t m = orc_neighbors(a->net, i, NULL, 0);
      if (m > cap) cap = m;
    }
  }
  cap += 1;
#ifdef _OPENMP
  if (a->n_threads > 0) omp_set_num_threads(a->n_threads);
#endif
#pragma omp parallel
  {
    orc_cand* nb = (orc_cand*)malloc(cap * sizeof(orc_cand));
    double* rowbuf = (double*)malloc(cap * sizeof(double));
    int32_t* colbuf = (int32_t*)malloc(cap * sizeof(int32_t));
#pragma omp for schedule(dynamic, 16)
    for (e = 0; e < a->E; ++e) orc_one(a, e, nb, rowbuf, colbuf, cap);
    free(nb);
    free(rowbuf);
    free(colbuf);
  }
  return 0;
}

/* Serial per-bin, per-batch sums of the MSD records (kubo_save_avg_dispalcement_squared,
 * monte_carlo.cpp:382-409): out[T][B][5] = {count, sum dx^2, sum dy^2, sum dz^2, sum drains}. */
int orc_reduce_samples(int64_t E, int32_t n_bins, int32_t n_batches, const uint32_t* k_rec,
                       const double* rec_disp, const uint32_t* rec_drain, double* out) {
  int32_t k, b;
  for (k = 0; k < n_bins; ++k)
    for (b = 0; b < n_batches; ++b) {
      int64_t e0 = (int64_t)b * E / n_batches, e1 = (int64_t)(b + 1) * E / n_batches, e;
      double acc[5] = {0, 0, 0, 0, 0};
      for (e = e0; e < e1; ++e) {
        if ((uint32_t)k < k_rec[e]) {
          size_t r = (size_t)k * E + e;
          double x = rec_disp[3 * r], y = rec_disp[3 * r + 1], z = rec_disp[3 * r + 2];
          acc[0] += 1.0;
          acc[1] += pow(x, 2);
          acc[2] += pow(y, 2);
          acc[3] += pow(z, 2);
          acc[4] += (double)rec_drain[r];
        }
      }
      memcpy(out + ((size_t)k * n_batches + b) * 5, acc, sizeof(acc));
    }
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * Contact (steady-state) mode, velocity 0: monte_carlo::init / step / save_metrics /
 * repopulate_contacts (monte_carlo.h:157-195, 274-316, 343-355, 443-684; main.cpp:91-99),
 * written literally over one particle list.  Draw mapping of this mode: a particle carries an
 * id (0, 1, 2, ... in creation order); its hops draw (seed, id, its hop count, stream 0) with
 * dwell redraws on streams 2+k, its creation draws (seed, id, 0, 0x20000000) -- w0 the site
 * (rand() % size), w1 the dwell -- with redraws on 0x20000001+k.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
  uint32_t id;
  int32_t site;
  double t_next;
  uint64_t hops;
  double y;         /* FP64 y: the site's, plus the flight offset when v > 0 */
  unsigned heading; /* v > 0: the particle's heading, drawn at creation (particle.h:51) */
  double off;       /* v > 0: distance beyond the site toward the next one */
  double t_last;    /* v > 0: time up to which the flight is taken */
} orc_cparticle;

static int orc_new_particle(uint64_t seed, uint32_t id, const int32_t* list, int64_t n_list,
                            const int64_t* row_ptr, const double* cum, const double* y, double t_now,
                            orc_cparticle* p) {
  uint32_t w[4], r;
  int32_t s;
  if (n_list <= 0) return -1;
  draws(seed, id, 0, 0x20000000u, w);
  s = list[(w[0] >> 1) % (uint32_t)n_list];
  r = dwell_draw(seed, id, 0, w[1], w[3], 0x20000001u);
  p->id = id;
  p->site = s;
  p->t_next = t_now + orc_dwell(cum + row_ptr[s], row_ptr[s + 1] - row_ptr[s], r);
  p->hops = 0;
  p->y = y[s];
  p->heading = (w[2] >> 1) & 1u;
  p->off = 0.0;
  p->t_last = t_now;
  return 0;
}

/* velocity > 0 adds particle::fly (particle.cpp:9-54) with the links left / right over the FP32
 * positions pos32 [n][3] (per_len: periods, 0 = open): up to each hop, and to the end of every
 * step, where the flight is taken.  Returns 0, -1 for a bad argument or an empty list, -2 when
 * the particles exceed capacity, -3 when a flight passes more than ORC_FLY_MAX_PASSES sites. */
int orc_contact(int64_t n_sites, const double* x, const double* y, const double* z,
                const int64_t* row_ptr, const int32_t* col, const double* cum, uint64_t seed,
                int n_seg, int64_t c1_pop, int64_t c2_pop, int64_t capacity, double dt,
                int64_t n_steps, double velocity, const int32_t* left, const int32_t* right,
                const float* pos32, const double* per_len, int32_t* pop, int32_t* cur,
                int64_t* n_live, double* area, int64_t* n_final, uint32_t* p_id, int32_t* p_site,
                double* p_t, uint64_t* p_hops, double* p_y, uint64_t* hops_made) {
  double lo[3], hi[3], ymin, ymax, dy, dp, T = 0.0;
  double *xmn, *xmx, *zmn, *zmx;
  int32_t *c1, *cn, *sec;
  int64_t n1 = 0, nn = 0, s, k, np_ = 0, step, i;
  uint32_t next_id = 0;
  orc_cparticle *P, *Q;
  int a, side, ret = 0;
  orc_run_args ra;
  uint8_t* no_flags;
  uint64_t made_total = 0; /* hops made by all particles, removed ones included */
  if (n_sites <= 0 || n_seg < 2 || capacity <= 0) return -1;
  if (velocity > 0.0 && (!left || !right || !pos32 || !per_len)) return -1;
  /* the flight's view of the network: links, FP32 positions, periods; no drain sites */
  memset(&ra, 0, sizeof(ra));
  no_flags = calloc((size_t)n_sites, 1);
  ra.n_sites = n_sites;
  ra.pos = pos32;
  ra.flags = no_flags;
  ra.seed = seed;
  ra.velocity = velocity;
  ra.left = left;
  ra.right = right;
  for (a = 0; a < 3; ++a) ra.per_len[a] = per_len ? per_len[a] : 0.0;
  /* find_simulation_domain (monte_carlo.h:328-340) */
  for (a = 0; a < 3; ++a) {
    const double* v = a == 0 ? x : (a == 1 ? y : z);
    lo[a] = hi[a] = v[0];
    for (s = 1; s < n_sites; ++s) {
      if (v[s] < lo[a]) lo[a] = v[s];
      if (v[s] > hi[a]) hi[a] = v[s];
    }
  }
  ymin = lo[1];
  ymax = hi[1];
  dy = (ymax - ymin) / (double)n_seg;
  /* get_area (monte_carlo.h:644-684) */
  xmn = malloc(sizeof(double) * n_seg); xmx = malloc(sizeof(double) * n_seg);
  zmn = malloc(sizeof(double) * n_seg); zmx = malloc(sizeof(double) * n_seg);
  for (i = 0; i < n_seg; ++i) { xmx[i] = lo[0]; xmn[i] = hi[0]; zmx[i] = lo[2]; zmn[i] = hi[2]; }
  for (s = 0; s < n_sites; ++s) {
    int j = (int)((y[s] - ymin) / dy);
    j = j < 0 ? 0 : (j < n_seg ? j : n_seg - 1);
    if (xmn[j] > x[s]) xmn[j] = x[s];
    else if (xmx[j] < x[s]) xmx[j] = x[s];
    if (zmn[j] > z[s]) zmn[j] = z[s];
    else if (zmx[j] < z[s]) zmx[j] = z[s];
  }
  for (i = 0; i < n_seg; ++i) area[i] = (zmx[i] - zmn[i]) * (xmx[i] - xmn[i]);
  free(xmn); free(xmx); free(zmn); free(zmx);
  /* contact_scats(1) and contact_scats(n_seg) (monte_carlo.h:494-516) */
  c1 = malloc(sizeof(int32_t) * n_sites);
  cn = malloc(sizeof(int32_t) * n_sites);
  sec = malloc(sizeof(int32_t) * n_sites);
  for (s = 0; s < n_sites; ++s) {
    if (y[s] >= ymin + 0.0 * dy && y[s] <= ymin + 1.0 * dy) c1[n1++] = (int32_t)s;
    if (y[s] >= ymin + (double)(n_seg - 1) * dy && y[s] <= ymin + (double)n_seg * dy)
      cn[nn++] = (int32_t)s;
  }
  P = malloc(sizeof(orc_cparticle) * capacity);
  Q = malloc(sizeof(orc_cparticle) * capacity);
  /* create_particles (monte_carlo.h:274-316): the linear profile, section by section */
  dp = (double)(c2_pop - c1_pop) / ((double)n_seg - 1);
  for (i = 0; i < n_seg && ret == 0; ++i) {
    int64_t ns = 0, want = (int64_t)round((double)c1_pop + (double)i * dp);
    const double y1 = ymin + (double)i * dy;
    const double y2 = y1 + dy;
    for (s = 0; s < n_sites; ++s)
      if (y1 <= y[s] && y[s] < y2) sec[ns++] = (int32_t)s;
    for (k = 0; k < want && ret == 0; ++k) {
      if (np_ >= capacity) { ret = -2; break; }
      if (orc_new_particle(seed, next_id++, sec, ns, row_ptr, cum, y, 0.0, &P[np_++]) != 0) ret = -1;
    }
  }
  for (step = 0; step < n_steps && ret == 0; ++step) {
    const double t_end = T + dt;
    int32_t* pp = pop + step * n_seg;
    int32_t* cc = cur + step * (n_seg - 1);
    memset(pp, 0, sizeof(int32_t) * n_seg);
    memset(cc, 0, sizeof(int32_t) * (n_seg - 1));
    /* monte_carlo::step -> particle::step (particle.cpp:57-80) */
    for (k = 0; k < np_; ++k) {
      orc_cparticle* p = &P[k];
      const double y_old = p->y;
      double y_new;
      int j, q;
      orc_flight f;
      f.s = p->site;
      f.off = p->off;
      f.heading = p->heading;
      f.d[0] = f.d[1] = f.d[2] = 0.0;
      f.nd = 0;
      f.hs = 0;
      while (p->t_next <= t_end) {
        int64_t b, n;
        uint32_t w_sel, w_dw, r;
        double dice;
        if (velocity > 0.0) {
          /* the flight up to the hop; the hop is from the site reached */
          if (!orc_fly(&ra, &f, p->t_next - p->t_last, p->id, p->hops)) { ret = -3; break; }
          p->site = f.s;
        }
        b = row_ptr[p->site];
        n = row_ptr[p->site + 1] - b;
        if (n == 0 || !(cum[b + n - 1] > 0.0)) {                /* no rates: stays */
          if (velocity > 0.0) {
            p->t_last = p->t_next;
            p->t_next = INFINITY;
            p->hops++;
          }
          break;
        }
        hop_words(seed, p->id, p->hops, &w_sel, &w_dw);
        dice = (cum[b + n - 1] * (double)(w_sel >> 1)) / ORC_RAND_MAX;
        p->site = col[b + orc_select(cum + b, (unsigned)n, dice)];
        r = hop_dwell_draw(seed, p->id, p->hops, w_dw);
        if (velocity > 0.0) {
          f.s = p->site;
          f.off = 0.0;
          p->t_last = p->t_next;
        }
        p->t_next = p->t_next + orc_dwell(cum + row_ptr[p->site],
                                          row_ptr[p->site + 1] - row_ptr[p->site], r);
        p->hops++;
        made_total++;
      }
      if (velocity > 0.0) {
        double o[3];
        /* the rest of the step, taken (particle::fly(dt), particle.cpp:77) */
        if (!orc_fly(&ra, &f, t_end - p->t_last, p->id, p->hops)) ret = -3;
        p->t_last = t_end;
        p->site = f.s;
        p->off = f.off;
        p->heading = f.heading;
        orc_fly_offset(&ra, &f, o);
        y_new = y[p->site] + o[1];
      } else {
        y_new = y[p->site];
      }
      p->y = y_new;
      /* save_population_profile (monte_carlo.h:565-571) */
      j = (int)((y_new - ymin) / dy);
      j = j < 0 ? 0 : (j < n_seg ? j : n_seg - 1);
      pp[j]++;
      /* save_currents (monte_carlo.h:598-626): every interface */
      for (q = 1; q < n_seg; ++q) {
        const double yq = ymin + dy * (double)q;
        if (y_old < yq && y_new >= yq) cc[q - 1]++;
        else if (y_old >= yq && y_new < yq) cc[q - 1]--;
      }
    }
    T = t_end;
    /* repopulate_contacts (monte_carlo.h:443-491): contact 1, then contact n_seg; the kept
     * particles in their order, the new ones after them */
    for (side = 0; side < 2 && ret == 0; ++side) {
      const double ylo = side == 0 ? ymin : ymin + (double)(n_seg - 1) * dy;
      const double yhi = side == 0 ? ymin + dy : ymax;
      const int64_t want = side == 0 ? c1_pop : c2_pop;
      int64_t nk = 0;
      for (k = 0; k < np_; ++k) {
        const double yy = P[k].y;
        if (!(yy >= ylo && yy <= yhi)) Q[nk++] = P[k];
      }
      for (k = 0; k < want; ++k) {
        if (nk >= capacity) { ret = -2; break; }
        if (orc_new_particle(seed, next_id++, side == 0 ? c1 : cn, side == 0 ? n1 : nn, row_ptr,
                             cum, y, T, &Q[nk++]) != 0) { ret = -1; break; }
      }
      memcpy(P, Q, sizeof(orc_cparticle) * nk);
      np_ = nk;
    }
    n_live[step] = np_;
  }
  *n_final = np_;
  for (k = 0; k < np_; ++k) {
    p_id[k] = P[k].id;
    p_site[k] = P[k].site;
    p_t[k] = P[k].t_next;
    p_hops[k] = P[k].hops;
    p_y[k] = P[k].y;
  }
  free(c1); free(cn); free(sec); free(P); free(Q); free(no_flags);
  if (hops_made) *hops_made = made_total;
  return ret;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------------------------
 * Exact first-order transfer rate: a literal restatement of exciton_transfer::first_order
 * (src/exciton_transfer/exciton_transfer.cpp:380-441) with get_relevant_states (:226-270),
 * match_states (exciton_transfer.h:167-196), calculate_Q (:273-304) and the per-pair double
 * loop of calculate_J (:307-377).  The exciton model arrays come from the product's CNT model
 * (they are pinned separately: closed forms and LAPACK eigenvalues).
 * ---------------------------------------------------------------------------------------- */
#include <complex.h>

typedef struct orc_tube {
  int Nu, L, nk, n_mu, nr, nk_cm, nk_c, ik_cm0;
  const double* energy;   /* [nk_cm][nr] */
  const double* psi;      /* complex [nk_cm][nr][nk_c] */
  const int32_t* idx;     /* [nk_cm][nk_c][4] */
  const double* wf;       /* complex [n_mu][nk][band][atom] */
  double dk_l[2], aCC[2], t_vec[2], t_vec_3d[3];
  double Au, radius, length_m;
  const double* pos_u_2d; /* [Nu][2] */
  const double* pos_u_3d; /* [Nu][3] */
} orc_tube;

typedef struct {
  const orc_tube* t;
  int ik_cm_idx, n, ik_cm;
  double energy;
} orc_state;

static int cmp_state(const void* a, const void* b) {
  const orc_state* x = (const orc_state*)a;
  const orc_state* y = (const orc_state*)b;
  if (x->energy < y->energy) return -1;
  if (x->energy > y->energy) return 1;
  /* ties keep collection order (stable) */
  return (x->n * 100000 + x->ik_cm_idx) - (y->n * 100000 + y->ik_cm_idx);
}

static int orc_relevant(const orc_tube* t, double min_energy, double T, double kb, orc_state* out) {
  const double thr = min_energy + fabs(log(1.e-3) * kb * T);
  int cnt = 0, i_n, ic;
  for (i_n = 0; i_n < t->nr; i_n++)
    for (ic = 0; ic < t->nk_cm; ic++) {
      double e = t->energy[(size_t)ic * t->nr + i_n];
      if (e <= thr) {
        out[cnt].t = t;
        out[cnt].ik_cm_idx = ic;
        out[cnt].n = i_n;
        out[cnt].ik_cm = t->ik_cm0 + ic;
        out[cnt].energy = e;
        cnt++;
      }
    }
  qsort(out, cnt, sizeof(orc_state), cmp_state);
  return cnt;
}

static double complex orc_q_partial(const orc_state* s) {
  const orc_tube* t = s->t;
  double kx = s->ik_cm * t->dk_l[0], ky = s->ik_cm * t->dk_l[1];
  double complex ef0 = cexp(I * (kx * 0.0 + ky * 0.0));
  double complex ef1 = cexp(I * (kx * t->aCC[0] + ky * t->aCC[1]));
  double complex q = 0;
  int k;
  for (k = 0; k < t->nk_c; k++) {
    const int32_t* id = t->idx + ((size_t)s->ik_cm_idx * t->nk_c + k) * 4;
    const double* wc = t->wf + 2 * ((((size_t)id[1] * t->nk + id[0]) * 2 + 1) * 2);
    const double* wv = t->wf + 2 * ((((size_t)id[3] * t->nk + id[2]) * 2 + 0) * 2);
    double complex ccA = wc[0] + I * wc[1], ccB = wc[2] + I * wc[3];
    double complex cvA = wv[0] + I * wv[1], cvB = wv[2] + I * wv[3];
    const double* ps = t->psi + 2 * (((size_t)s->ik_cm_idx * t->nr + s->n) * t->nk_c + k);
    double complex acc = ccA * conj(cvA) * ef0 + ccB * conj(cvB) * ef1;
    q += (ps[0] + I * ps[1]) * acc;
  }
  return q;
}

static void orc_ru_3d(const orc_tube* t, double shift, double z_shift, double angle, double* out) {
  int n = t->Nu * t->L, i, a;
  double ymax, ymin, c = cos(angle), s = sin(angle);
  for (i = 0; i < t->L; i++)
    for (a = 0; a < t->Nu; a++) {
      int k = i * t->Nu + a;
      out[3 * k] = t->pos_u_3d[3 * a] + i * t->t_vec_3d[0];
      out[3 * k + 1] = t->pos_u_3d[3 * a + 1] + i * t->t_vec_3d[1];
      out[3 * k + 2] = t->pos_u_3d[3 * a + 2] + i * t->t_vec_3d[2];
    }
  ymax = out[1];
  ymin = out[1];
  for (i = 0; i < n; i++) {
    if (out[3 * i + 1] > ymax) ymax = out[3 * i + 1];
    if (out[3 * i + 1] < ymin) ymin = out[3 * i + 1];
  }
  for (i = 0; i < n; i++) {
    double x, y;
    out[3 * i + 1] -= (ymax + ymin) / 2.;
    out[3 * i + 1] += shift;
    out[3 * i + 2] += z_shift;
    x = out[3 * i] * c - out[3 * i + 1] * s;
    y = out[3 * i] * s + out[3 * i + 1] * c;
    out[3 * i] = x;
    out[3 * i + 1] = y;
  }
}

static void orc_ru_2d(const orc_tube* t, double* out) {
  int i, a;
  for (i = 0; i < t->L; i++)
    for (a = 0; a < t->Nu; a++) {
      int k = i * t->Nu + a;
      out[2 * k] = t->pos_u_2d[2 * a] + i * t->t_vec[0];
      out[2 * k + 1] = t->pos_u_2d[2 * a + 1] + i * t->t_vec[1];
    }
}

double orc_first_order(const orc_tube* d, const orc_tube* a, double T, double broad, double z_shift,
                       double ash1, double ash2, double theta, int* n_pairs_out) {
  const double pi = 3.141592, inv_pi = 1. / 3.141592;
  const double eV = 1.6 * pow(10, -19.0);
  const double hb = 6.5 * pow(10, -16.0) * eV;
  const double kb = 1.3865 * pow(10, -23.0);
  const double eps0 = 8.85 * pow(10, -12);
  const double q0 = 1.6 * pow(10, -19);
  double min_energy = d->energy[0], Z = 0, rate = 0, l0;
  size_t i;
  int nd, na, p, f, n_pairs = 0;
  orc_state *ds = (orc_state*)malloc(sizeof(orc_state) * d->nk_cm * d->nr);
  orc_state *as = (orc_state*)malloc(sizeof(orc_state) * a->nk_cm * a->nr);
  double *i3 = (double*)malloc(sizeof(double) * 3 * d->Nu * d->L);
  double *f3 = (double*)malloc(sizeof(double) * 3 * a->Nu * a->L);
  double *i2 = (double*)malloc(sizeof(double) * 2 * d->Nu * d->L);
  double *f2 = (double*)malloc(sizeof(double) * 2 * a->Nu * a->L);
  double complex* fexp = (double complex*)malloc(sizeof(double complex) * a->Nu * a->L);
  for (i = 0; i < (size_t)d->nk_cm * d->nr; i++) min_energy = d->energy[i] < min_energy ? d->energy[i] : min_energy;
  nd = orc_relevant(d, min_energy, T, kb, ds);
  na = orc_relevant(a, min_energy, T, kb, as);
  for (p = 0; p < nd; p++) Z += exp(-ds[p].energy / (kb * T));
  orc_ru_3d(d, ash1, 0, 0, i3);
  orc_ru_3d(a, ash2, z_shift, theta, f3);
  orc_ru_2d(d, i2);
  orc_ru_2d(a, f2);
  l0 = inv_pi * broad / (0 * 0 + broad * broad);
  for (p = 0; p < nd; p++)
    for (f = 0; f < na; f++) {
      double de = ds[p].energy - as[f].energy, coeff, M;
      double complex Q, J = 0;
      int ii, jj, ni = d->Nu * d->L, nf = a->Nu * a->L;
      if (!(inv_pi * broad / (de * de + broad * broad) > 1.e-2 * l0)) continue;
      n_pairs++;
      coeff = (pow(q0, 2) * d->Au * a->Au) /
              (16 * pow(pi, 3) * eps0 * d->radius * a->radius * sqrt(d->length_m * a->length_m));
      Q = coeff * conj(orc_q_partial(&ds[p])) * orc_q_partial(&as[f]);
      for (jj = 0; jj < nf; jj++) {
        double kx = as[f].ik_cm * a->dk_l[0], ky = as[f].ik_cm * a->dk_l[1];
        fexp[jj] = cexp(+I * (kx * f2[2 * jj] + ky * f2[2 * jj + 1]));
      }
      for (ii = 0; ii < ni; ii++) {
        double kx = ds[p].ik_cm * d->dk_l[0], ky = ds[p].ik_cm * d->dk_l[1];
        double complex iexp = cexp(-I * (kx * i2[2 * ii] + ky * i2[2 * ii + 1]));
        for (jj = 0; jj < nf; jj++) {
          double dv[3];
          dv[0] = i3[3 * ii] - f3[3 * jj];
          dv[1] = i3[3 * ii + 1] - f3[3 * jj + 1];
          dv[2] = i3[3 * ii + 2] - f3[3 * jj + 2];
          J += iexp * fexp[jj] / arma_norm(dv, 3);
        }
      }
      M = cabs(Q * J) / sqrt(d->length_m * a->length_m);
      rate += (2 * pi / hb) * (exp(-ds[p].energy / (kb * T)) / Z) * pow(M, 2) *
              (inv_pi * broad / (de * de + broad * broad));
    }
  free(ds); free(as); free(i3); free(f3); free(i2); free(f2); free(fexp);
  if (n_pairs_out) *n_pairs_out = n_pairs;
  return rate;
}
