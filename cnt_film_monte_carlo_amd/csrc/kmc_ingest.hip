// kmc_ingest.hip -- host-side file and site ingest of the KMC path: the mesh-file reader
// (create_scatterers), the trim to a box (trim_scats), the Kubo injection / removal regions
// (injection_region, get_removal_domain) and the rate-table files (scattering_struct::save and
// a reader for them).  No device work.
#include <algorithm>
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

// Every number of a text file, in order (blanks, tabs, commas and line ends separate them).
int read_numbers(const std::string& path, const std::string& text, size_t from,
                 std::vector<double>& v, const char* who) {
  v.clear();
  const char* p = text.c_str() + from;
  const char* end = text.c_str() + text.size();
  while (p < end) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == ',')) ++p;
    if (p >= end) break;
    char* q = nullptr;
    const double d = std::strtod(p, &q);
    if (q == p)
      return ingest_fail(KMC_ERR_ARG, std::string(who) + ": bad number in " + path + " after " +
                                          std::to_string(v.size()) + " values");
    v.push_back(d);
    p = q;
  }
  return KMC_OK;
}

int slurp(const std::string& path, std::string& text, const char* who) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return ingest_fail(KMC_ERR_ARG, std::string(who) + ": cannot open " + path);
  text.clear();
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  return KMC_OK;
}

const char* const kAxisFiles[4] = {".theta.dat", ".z_shift.dat", ".axis_shift_1.dat",
                                   ".axis_shift_2.dat"};

}  // namespace

extern "C" {

int kmc_save_rate_table(const char* dir, const char* name, int64_t n_theta, int64_t n_z,
                        int64_t n_a1, int64_t n_a2, const double* theta, const double* z_shift,
                        const double* axis_shift_1, const double* axis_shift_2,
                        const double* rates) {
  const char* who = "kmc_save_rate_table";
  if (!dir || !theta || !z_shift || !axis_shift_1 || !axis_shift_2 || !rates)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": null argument");
  if (n_theta < 1 || n_z < 1 || n_a1 < 1 || n_a2 < 1)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": empty axis");
  const std::string base = std::string(dir) + "/" + (name && *name ? name : "scat_table");
  const int64_t dims[4] = {n_theta, n_z, n_a1, n_a2};
  const double* axes[4] = {theta, z_shift, axis_shift_1, axis_shift_2};
  // one value per line; 17 significant digits, so that reading back gives the same doubles
  for (int k = 0; k < 4; ++k) {
    FILE* f = std::fopen((base + kAxisFiles[k]).c_str(), "wb");
    if (!f) return ingest_fail(KMC_ERR_ARG, std::string(who) + ": cannot write " + base + kAxisFiles[k]);
    for (int64_t i = 0; i < dims[k]; ++i) std::fprintf(f, "%.17g\n", axes[k][i]);
    if (std::fclose(f) != 0)
      return ingest_fail(KMC_ERR_ARG, std::string(who) + ": write failed: " + base + kAxisFiles[k]);
  }
  // the four header lines and the theta-outer, axis_shift_2-inner order of
  // scattering_struct::save (src/monte_carlo/scattering_struct.h:75-91)
  FILE* f = std::fopen((base + ".rates.dat").c_str(), "wb");
  if (!f) return ingest_fail(KMC_ERR_ARG, std::string(who) + ": cannot write " + base + ".rates.dat");
  std::fprintf(f, "sizes:\ntheta, z_shift, axis_shift_1, axis_shift_2\n%lld,%lld,%lld,%lld\n\n",
               (long long)n_theta, (long long)n_z, (long long)n_a1, (long long)n_a2);
  const int64_t total = n_theta * n_z * n_a1 * n_a2;
  for (int64_t i = 0; i < total; ++i) std::fprintf(f, "%.17g\n", rates[i]);
  if (std::fclose(f) != 0)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": write failed: " + base + ".rates.dat");
  return KMC_OK;
}

int kmc_load_rate_table(const char* dir, const char* name, int64_t dims[4], double* theta,
                        double* z_shift, double* axis_shift_1, double* axis_shift_2,
                        double* rates) {
  const char* who = "kmc_load_rate_table";
  if (!dir || !dims) return ingest_fail(KMC_ERR_ARG, std::string(who) + ": null argument");
  const std::string base = std::string(dir) + "/" + (name && *name ? name : "scat_table");
  std::string text;
  int st = slurp(base + ".rates.dat", text, who);
  if (st != KMC_OK) return st;
  // header: "sizes:", the axis names, "n_theta,n_z,n_a1,n_a2", a blank line
  // (scattering_struct.h:76-79; the reference's reader skips four rows,
  // visualization/monte_carlo_results.py:264-270)
  size_t pos = 0;
  std::string line[4];
  for (int k = 0; k < 4; ++k) {
    const size_t e = text.find('\n', pos);
    if (e == std::string::npos)
      return ingest_fail(KMC_ERR_ARG, std::string(who) + ": short header in " + base + ".rates.dat");
    line[k] = text.substr(pos, e - pos);
    pos = e + 1;
  }
  long long d[4];
  if (line[0].rfind("sizes:", 0) != 0 ||
      std::sscanf(line[2].c_str(), "%lld , %lld , %lld , %lld", &d[0], &d[1], &d[2], &d[3]) != 4)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": bad header in " + base + ".rates.dat");
  for (int k = 0; k < 4; ++k)
    if (d[k] < 1 || d[k] > (1 << 24))
      return ingest_fail(KMC_ERR_ARG, std::string(who) + ": bad size in " + base + ".rates.dat");
  const bool fill = theta != nullptr;
  if (fill)
    for (int k = 0; k < 4; ++k)
      if (dims[k] != d[k])
        return ingest_fail(KMC_ERR_CAPACITY, std::string(who) + ": table sizes differ from dims");
  for (int k = 0; k < 4; ++k) dims[k] = d[k];
  if (!fill) return KMC_OK;  // size query
  if (!z_shift || !axis_shift_1 || !axis_shift_2 || !rates)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": null output array");
  std::vector<double> v;
  st = read_numbers(base + ".rates.dat", text, pos, v, who);
  if (st != KMC_OK) return st;
  const int64_t total = d[0] * d[1] * d[2] * d[3];
  if ((int64_t)v.size() != total)
    return ingest_fail(KMC_ERR_ARG, std::string(who) + ": " + base + ".rates.dat holds " +
                                        std::to_string(v.size()) + " rates, the sizes give " +
                                        std::to_string(total));
  std::copy(v.begin(), v.end(), rates);
  double* axes[4] = {theta, z_shift, axis_shift_1, axis_shift_2};
  for (int k = 0; k < 4; ++k) {
    st = slurp(base + kAxisFiles[k], text, who);
    if (st == KMC_OK) st = read_numbers(base + kAxisFiles[k], text, 0, v, who);
    if (st != KMC_OK) return st;
    if ((int64_t)v.size() != d[k])
      return ingest_fail(KMC_ERR_ARG, std::string(who) + ": " + base + kAxisFiles[k] + " holds " +
                                          std::to_string(v.size()) + " values, the header gives " +
                                          std::to_string(d[k]));
    std::copy(v.begin(), v.end(), axes[k]);
  }
  return KMC_OK;
}

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
