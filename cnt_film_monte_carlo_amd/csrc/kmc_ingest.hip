// kmc_ingest.hip -- host-side site ingest of the KMC path: the mesh-file reader
// (create_scatterers), the trim to a box (trim_scats) and the Kubo injection / removal regions
// (injection_region, get_removal_domain).  No device work.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/cntkmc.h"

namespace cntkmc {
void set_last_error(const std::string& msg);
}

using namespace cntkmc;

namespace {

int ingest_fail(int code, const std::string& msg) {
  set_last_error(msg);
  return code;
}

// One ASCII matrix (rows = tubes, columns = points): numbers separated by blanks, tabs or
// commas, one row per line, blank lines skipped.  Every row must have the same length.
int read_matrix(const std::string& path, std::vector<double>& v, int64_t& rows, int64_t& cols) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: cannot open " + path);
  std::string text;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  v.clear();
  rows = 0;
  cols = -1;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t end = text.find('\n', pos);
    if (end == std::string::npos) end = text.size();
    std::string line = text.substr(pos, end - pos);
    pos = end + 1;
    for (char& ch : line)
      if (ch == ',' || ch == '\t' || ch == '\r') ch = ' ';
    const char* p = line.c_str();
    int64_t n = 0;
    while (true) {
      while (*p == ' ') ++p;
      if (!*p) break;
      char* q = nullptr;
      const double d = std::strtod(p, &q);
      if (q == p)
        return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: bad number in " + path + " row " +
                                            std::to_string(rows + 1));
      v.push_back(d);
      ++n;
      p = q;
    }
    if (n == 0) continue;
    if (cols >= 0 && n != cols)
      return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: ragged rows in " + path + " (row " +
                                          std::to_string(rows + 1) + " has " + std::to_string(n) +
                                          " values, row 1 has " + std::to_string(cols) + ")");
    cols = n;
    ++rows;
  }
  if (rows == 0) return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: no data in " + path);
  return KMC_OK;
}

}  // namespace

extern "C" {

int kmc_read_mesh(const char* dir, int64_t capacity, int64_t* n_tubes, int64_t* n_points,
                  double* x, double* y, double* z, double* ox, double* oy, double* oz,
                  int32_t* left, int32_t* right) {
  if (!dir || !n_tubes || !n_points) return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: null argument");
  static const char* kFiles[6] = {"single_cnt.pos.x.dat",    "single_cnt.pos.y.dat",
                                  "single_cnt.pos.z.dat",    "single_cnt.orient.x.dat",
                                  "single_cnt.orient.y.dat", "single_cnt.orient.z.dat"};
  std::vector<double> m[6];
  int64_t r0 = 0, c0 = 0;
  for (int k = 0; k < 6; ++k) {
    int64_t r = 0, c = 0;
    const int s = read_matrix(std::string(dir) + "/" + kFiles[k], m[k], r, c);
    if (s != KMC_OK) return s;
    if (k == 0) {
      r0 = r;
      c0 = c;
    } else if (r != r0 || c != c0) {
      return ingest_fail(KMC_ERR_ARG, std::string("kmc_read_mesh: ") + kFiles[k] +
                                          " has a different shape from " + kFiles[0]);
    }
  }
  *n_tubes = r0;
  *n_points = c0;
  const int64_t n = r0 * c0;
  if (n > (int64_t)0x7FFFFFFF) return ingest_fail(KMC_ERR_CAPACITY, "kmc_read_mesh: more than 2^31 sites");
  if (!x || capacity <= 0) return KMC_OK;  // size query
  if (capacity < n) return ingest_fail(KMC_ERR_CAPACITY, "kmc_read_mesh: capacity below n_tubes x n_points");
  if (!y || !z || !ox || !oy || !oz || !left || !right)
    return ingest_fail(KMC_ERR_ARG, "kmc_read_mesh: null output array");
  // site n = i * n_points + j; positions in nm scaled to m; links along a row
  // (create_scatterers, src/monte_carlo/monte_carlo.h:210-262)
  double* out[6] = {x, y, z, ox, oy, oz};
  for (int64_t i = 0; i < r0; ++i) {
    for (int64_t j = 0; j < c0; ++j) {
      const int64_t s = i * c0 + j;
      for (int k = 0; k < 3; ++k) out[k][s] = m[k][s] * 1.e-9;
      for (int k = 3; k < 6; ++k) out[k][s] = m[k][s];
      left[s] = j > 0 ? (int32_t)(s - 1) : -1;
      right[s] = j + 1 < c0 ? (int32_t)(s + 1) : -1;
    }
  }
  return KMC_OK;
}

int kmc_trim_sites(int64_t* n, double* x, double* y, double* z, double* ox, double* oy,
                   double* oz, int32_t* left, int32_t* right, uint8_t* chirality, uint8_t* flags,
                   int64_t* origin, const double xlim[2], const double ylim[2],
                   const double zlim[2]) {
  if (!n || *n < 0 || !x || !y || !z || !ox || !oy || !oz || !left || !right || !xlim || !ylim ||
      !zlim)
    return ingest_fail(KMC_ERR_ARG, "kmc_trim_sites: null argument");
  const int64_t N = *n;
  if (origin)
    for (int64_t i = 0; i < N; ++i) origin[i] = i;
  for (int64_t i = 0; i < N; ++i)
    if (left[i] < -1 || left[i] >= N || right[i] < -1 || right[i] >= N)
      return ingest_fail(KMC_ERR_ARG, "kmc_trim_sites: link out of range at site " + std::to_string(i));
  // trim_scats (src/monte_carlo/monte_carlo.h:722-782): a site outside the box swaps with the
  // last kept one (the neighbours of both are re-pointed first), then its neighbours lose
  // their link to it; the site moved into slot i is examined next.
  auto swap_sites = [&](int64_t i, int64_t j) {
    const int32_t il = left[i], ir = right[i], jl = left[j], jr = right[j];
    if (il > -1) right[il] = (int32_t)j;
    if (ir > -1) left[ir] = (int32_t)j;
    if (jl > -1) right[jl] = (int32_t)i;
    if (jr > -1) left[jr] = (int32_t)i;
    std::swap(x[i], x[j]);
    std::swap(y[i], y[j]);
    std::swap(z[i], z[j]);
    std::swap(ox[i], ox[j]);
    std::swap(oy[i], oy[j]);
    std::swap(oz[i], oz[j]);
    std::swap(left[i], left[j]);
    std::swap(right[i], right[j]);
    if (chirality) std::swap(chirality[i], chirality[j]);
    if (flags) std::swap(flags[i], flags[j]);
    if (origin) std::swap(origin[i], origin[j]);
  };
  int64_t j = N;
  for (int64_t i = 0; i < j;) {
    if (x[i] < xlim[0] || y[i] < ylim[0] || z[i] < zlim[0] || x[i] > xlim[1] || y[i] > ylim[1] ||
        z[i] > zlim[1]) {
      --j;
      swap_sites(i, j);
      if (left[j] > -1) right[left[j]] = -1;
      if (right[j] > -1) left[right[j]] = -1;
    } else {
      ++i;
    }
  }
  *n = j;
  return KMC_OK;
}

int kmc_kubo_regions(int64_t n, const double* x, const double* y, const double* z,
                     int32_t n_sections, uint8_t* flags, double domain_lo[3],
                     double domain_hi[3]) {
  if (n < 1 || !x || !y || !z || !flags)
    return ingest_fail(KMC_ERR_ARG, "kmc_kubo_regions: need n >= 1 and the site arrays");
  if (n_sections < 3 || n_sections % 2 == 0)
    return ingest_fail(KMC_ERR_ARG, "kmc_kubo_regions: n_sections must be odd and >= 3");
  const double* p[3] = {x, y, z};
  // find_simulation_domain (src/monte_carlo/monte_carlo.h:328-340)
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = p[k][0];
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = lo[k] > p[k][i] ? p[k][i] : lo[k];
      hi[k] = hi[k] < p[k][i] ? p[k][i] : hi[k];
    }
  // section edges e_k(i) = i * d_k + lo_k, d_k = (hi_k - lo_k) / n (monte_carlo.cpp:206-216)
  std::vector<double> e[3];
  for (int k = 0; k < 3; ++k) {
    const double d = (hi[k] - lo[k]) / double(n_sections);
    for (int i = 0; i <= n_sections; ++i) e[k].push_back(double(i) * d + lo[k]);
  }
  const int c = n_sections / 2;
  for (int64_t i = 0; i < n; ++i) {
    uint8_t f = 0;
    bool inj = true, out = false;
    for (int k = 0; k < 3; ++k) {
      // injection_region: the centre cell, closed bounds (monte_carlo.cpp:220-224)
      inj = inj && (e[k][c] <= p[k][i] && p[k][i] <= e[k][c + 1]);
      // get_removal_domain: [e(1), e(n-1)] (monte_carlo.cpp:246-248); outside re-injects
      out = out || (p[k][i] < e[k][1] || e[k][n_sections - 1] < p[k][i]);
    }
    if (inj) f |= KMC_SITE_SOURCE;
    if (out) f |= KMC_SITE_DRAIN;
    flags[i] = f;
  }
  if (domain_lo)
    for (int k = 0; k < 3; ++k) domain_lo[k] = lo[k];
  if (domain_hi)
    for (int k = 0; k < 3; ++k) domain_hi[k] = hi[k];
  return KMC_OK;
}

}  // extern "C"
