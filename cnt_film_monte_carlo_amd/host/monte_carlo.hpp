// monte_carlo.hpp -- C++ host layer over the C ABI of include/cntkmc.h.
//
// The reference drives its hot path through the methods of mc::monte_carlo
// (src/monte_carlo/monte_carlo.h, monte_carlo.cpp) from src/main.cpp.  This header keeps those
// method names and their meaning, so a driver written against the reference reads the same,
// while every step runs in libcntkmc.so on the GPU:
//
//   reference (mc::monte_carlo)                          here (cntkmc::monte_carlo)
//   create_scatterers(path)        monte_carlo.h:199-271  create_scatterers(dir)    kmc_read_mesh
//   trim_scats(x, y, z, list)      monte_carlo.h:722-782  trim_scats(x, y, z)       kmc_trim_sites
//   kubo_init()                    monte_carlo.cpp:254-305 kubo_init(n_sections)    kmc_kubo_regions,
//                                                                   kmc_set_sites, kmc_build_cells
//   create_forster_scatt_table     monte_carlo.cpp:156-200 create_forster_scatt_table
//   set_scat_table / set_max_rate  monte_carlo.h:419-440  set_scat_table, set_max_rate
//   kubo_create_particles()        monte_carlo.cpp:308-316 kubo_create_particles(n)  kmc_inject
//   kubo_step(dt)                  monte_carlo.cpp:319-342 kubo_step(dt)             kmc_run
//   kubo_save_avg_dispalcement_squared  monte_carlo.cpp:382-409  (same name)   kmc_get_excitons
//   time(), kubo_max_time()        monte_carlo.h                  time(), kubo_max_time()
//   init() / step(dt) / repopulate_contacts()  monte_carlo.h:157-195, 343-355, 443-455
//                                               contact_init, contact_step  kmc_contact_*
//
// Errors: the C ABI returns status codes; this layer turns them into cntkmc::error (a
// std::runtime_error carrying the code and kmc_last_error's message).
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "cntkmc.h"

namespace cntkmc {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const kmc_ctx* ctx, const char* what) {
  if (rc < 0) {
    const std::string m = kmc_last_error(ctx) ? kmc_last_error(ctx) : "";
    // the library's messages usually start with the entry point's name already
    throw error(rc, m.rfind(what, 0) == 0 ? m : std::string(what) + ": " + m);
  }
}

// Site arrays of a film (the reference's std::vector<scatterer>, as structure of arrays).
struct scatterers {
  std::vector<double> x, y, z, ox, oy, oz;
  std::vector<int32_t> left, right;
  std::vector<uint8_t> chirality, flags;
  int64_t size() const { return (int64_t)x.size(); }
  void resize(int64_t n) {
    for (auto* v : {&x, &y, &z, &ox, &oy, &oz}) v->assign(n, 0.0);
    left.assign(n, -1);
    right.assign(n, -1);
    chirality.assign(n, 0);
    flags.assign(n, 0);
  }
};

// The table grid of input.json ("theta [degrees]", "zshift [m]", "axis shift 1/2 [m]"), built
// with arma::linspace and the reference's truncated pi (src/helper/constants.h:10), as
// create_*_scatt_table does (monte_carlo.cpp:64-75).
struct table_grid {
  std::vector<double> theta, z_shift, axis_shift_1, axis_shift_2;
  static table_grid reference_default() {
    table_grid g;
    auto lin = [](double a, double b, int n) {
      std::vector<double> v(n);
      check(kmc_linspace(a, b, n, v.data()), nullptr, "kmc_linspace");
      return v;
    };
    g.theta = lin(0.0, 180.0, 21);
    for (double& t : g.theta) t *= 3.141592 / 180.0;
    g.z_shift = lin(1.5e-9, 10e-9, 11);
    g.axis_shift_1 = lin(-10e-9, 10e-9, 11);
    g.axis_shift_2 = lin(-10e-9, 10e-9, 11);
    return g;
  }
  size_t nodes() const {
    return theta.size() * z_shift.size() * axis_shift_1.size() * axis_shift_2.size();
  }
};

class monte_carlo {
 public:
  explicit monte_carlo(const kmc_config& cfg) { check(kmc_create(&cfg, &ctx_), nullptr, "kmc_create"); }
  ~monte_carlo() {
    if (ctx_) kmc_destroy(ctx_);
  }
  monte_carlo(const monte_carlo&) = delete;
  monte_carlo& operator=(const monte_carlo&) = delete;

  kmc_ctx* context() const { return ctx_; }
  scatterers& scatterer_list() { return scat_; }
  double time() const { return time_; }
  double kubo_max_time() const { return max_time_; }
  void set_kubo_max_time(double t) { max_time_ = t; }
  const double* domain_lo() const { return dom_lo_; }
  const double* domain_hi() const { return dom_hi_; }

  // create_scatterers: the six single_cnt.{pos,orient}.{x,y,z}.dat matrices of a mesh directory.
  void create_scatterers(const std::string& dir) {
    int64_t nt = 0, np = 0;
    check(kmc_read_mesh(dir.c_str(), 0, &nt, &np, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr),
          nullptr, "kmc_read_mesh");
    scat_.resize(nt * np);
    check(kmc_read_mesh(dir.c_str(), nt * np, &nt, &np, scat_.x.data(), scat_.y.data(),
                        scat_.z.data(), scat_.ox.data(), scat_.oy.data(), scat_.oz.data(),
                        scat_.left.data(), scat_.right.data()),
          nullptr, "kmc_read_mesh");
  }

  // Sites from the seeded generator of straight tubes (the synthetic films).
  void generate_scatterers(const kmc_netgen_params& p) {
    int64_t n = 0;
    check(kmc_generate_network(&p, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, nullptr, &n),
          nullptr, "kmc_generate_network");
    scat_.resize(n);
    check(kmc_generate_network(&p, n, scat_.x.data(), scat_.y.data(), scat_.z.data(),
                               scat_.ox.data(), scat_.oy.data(), scat_.oz.data(),
                               scat_.left.data(), scat_.right.data(), scat_.chirality.data(),
                               nullptr, &n),
          nullptr, "kmc_generate_network");
  }

  void trim_scats(const double xlim[2], const double ylim[2], const double zlim[2]) {
    int64_t n = scat_.size();
    check(kmc_trim_sites(&n, scat_.x.data(), scat_.y.data(), scat_.z.data(), scat_.ox.data(),
                         scat_.oy.data(), scat_.oz.data(), scat_.left.data(), scat_.right.data(),
                         scat_.chirality.data(), scat_.flags.data(), nullptr, xlim, ylim, zlim),
          nullptr, "kmc_trim_sites");
    for (auto* v : {&scat_.x, &scat_.y, &scat_.z, &scat_.ox, &scat_.oy, &scat_.oz}) v->resize(n);
    scat_.left.resize(n);
    scat_.right.resize(n);
    scat_.chirality.resize(n);
    scat_.flags.resize(n);
  }

  // kubo_init: injection region and removal domain from n_sections slabs per axis, then the
  // sites and their cell list on the device (create_scatterer_buckets).
  void kubo_init(int32_t n_sections) {
    check(kmc_kubo_regions(scat_.size(), scat_.x.data(), scat_.y.data(), scat_.z.data(),
                           n_sections, scat_.flags.data(), dom_lo_, dom_hi_),
          nullptr, "kmc_kubo_regions");
    upload_scatterers();
  }

  // The sites as they are (flags set by the caller), then the cell list.
  void upload_scatterers() {
    check(kmc_set_sites(ctx_, scat_.size(), scat_.x.data(), scat_.y.data(), scat_.z.data(),
                        scat_.ox.data(), scat_.oy.data(), scat_.oz.data(), scat_.left.data(),
                        scat_.right.data(), scat_.chirality.data(), scat_.flags.data()),
          ctx_, "kmc_set_sites");
    check(kmc_build_cells(ctx_), ctx_, "kmc_build_cells");
  }

  // create_forster_scatt_table ("forster": gamma_0 = 1e15, "wong": 1e13) on a grid.
  std::vector<double> create_forster_scatt_table(double gamma_0, const table_grid& g) {
    std::vector<double> r(g.nodes());
    check(kmc_forster_table(gamma_0, (int32_t)g.theta.size(), g.theta.data(),
                            (int32_t)g.z_shift.size(), g.z_shift.data(),
                            (int32_t)g.axis_shift_1.size(), g.axis_shift_1.data(),
                            (int32_t)g.axis_shift_2.size(), g.axis_shift_2.data(), r.data()),
          nullptr, "kmc_forster_table");
    return r;
  }

  // set_scat_table: n_chir^2 tables [theta][z][a1][a2], donor-major.
  void set_scat_table(const table_grid& g, const std::vector<double>& rates, int32_t n_chir = 1) {
    if (rates.size() != (size_t)n_chir * n_chir * g.nodes())
      throw error(KMC_ERR_ARG, "set_scat_table: rates do not match the grid");
    check(kmc_set_rate_table(ctx_, n_chir, (int32_t)g.theta.size(), g.theta.data(),
                             (int32_t)g.z_shift.size(), g.z_shift.data(),
                             (int32_t)g.axis_shift_1.size(), g.axis_shift_1.data(),
                             (int32_t)g.axis_shift_2.size(), g.axis_shift_2.data(), rates.data()),
          ctx_, "kmc_set_rate_table");
  }

  // set_max_rate: every site's neighbours, rates and FP64 cumulative sums (the CSR).
  void set_max_rate() { check(kmc_build_rates(ctx_, KMC_RATE_TABLE), ctx_, "kmc_build_rates"); }

  void kubo_create_particles(int64_t n_particles) {
    check(kmc_inject(ctx_, n_particles), ctx_, "kmc_inject");
    n_particles_ = n_particles;
    time_ = 0.0;
  }

  // kubo_step: every particle makes the hops due by time() + dt; re-injection at a drain site
  // (outside the removal domain) keeps the displacement.
  void kubo_step(double dt, bool count_sites = true) {
    const double t_end = time_ + dt;
    check(kmc_run(ctx_, std::numeric_limits<int64_t>::max() / 2, t_end, count_sites ? 1 : 0,
                  nullptr),
          ctx_, "kmc_run");
    time_ = t_end;
  }

  // kubo_save_avg_dispalcement_squared: <dx^2>, <dy^2>, <dz^2> over the particles at time(),
  // appended to <file> as "time,x,y,z" lines (the reference's particle_dispalcement.avg.
  // squared.dat, with its header), summed in particle order.
  std::vector<double> kubo_save_avg_dispalcement_squared(const std::string& file = "") {
    std::vector<double> disp((size_t)n_particles_ * 3);
    check(kmc_get_excitons(ctx_, nullptr, nullptr, nullptr, disp.data(), nullptr, nullptr), ctx_,
          "kmc_get_excitons");
    double avg[3] = {0.0, 0.0, 0.0};
    for (int64_t e = 0; e < n_particles_; ++e)
      for (int k = 0; k < 3; ++k) avg[k] += disp[3 * e + k] * disp[3 * e + k];
    for (double& a : avg) a /= double(n_particles_);
    if (!file.empty()) {
      if (!msd_out_.is_open()) {
        msd_out_.open(file, std::ios::out);
        msd_out_ << "# this file contains the average of dx^2, dy^2, and dz^2 of the particle "
                    "ensemble over time\n# number of particles: "
                 << n_particles_ << "\n\ntime,x,y,z\n";
        msd_out_.precision(17);
      }
      msd_out_ << std::scientific << time_ << "," << avg[0] << "," << avg[1] << "," << avg[2]
               << "\n";
    }
    return {avg[0], avg[1], avg[2]};
  }

  // Contact (steady-state) mode: init() with the initial linear profile, then step(dt) +
  // save_metrics(dt) + repopulate_contacts() per step.
  void contact_init(int32_t n_seg, int64_t c1_pop, int64_t c2_pop, int64_t capacity) {
    check(kmc_contact_init(ctx_, n_seg, c1_pop, c2_pop, capacity), ctx_, "kmc_contact_init");
    n_seg_ = n_seg;
    time_ = 0.0;
  }
  // One step: population per section and net interface crossings (save_metrics' inputs).
  int64_t contact_step(double dt, std::vector<int32_t>* population = nullptr,
                       std::vector<int32_t>* crossings = nullptr,
                       uint64_t max_hops_per_step = 1u << 20) {
    if (population) population->assign(n_seg_, 0);
    if (crossings) crossings->assign(n_seg_ > 1 ? n_seg_ - 1 : 0, 0);
    int64_t live = 0;
    check(kmc_contact_run(ctx_, dt, 1, max_hops_per_step,
                          population ? population->data() : nullptr,
                          crossings ? crossings->data() : nullptr, &live),
          ctx_, "kmc_contact_run");
    time_ += dt;
    return live;
  }

  // Per-site departure counts and the exciton state (read-back).
  std::vector<uint64_t> hop_counts() {
    std::vector<uint64_t> c((size_t)scat_.size());
    check(kmc_get_hop_counts(ctx_, c.data()), ctx_, "kmc_get_hop_counts");
    return c;
  }

 private:
  kmc_ctx* ctx_ = nullptr;
  scatterers scat_;
  double time_ = 0.0, max_time_ = 0.0;
  double dom_lo_[3] = {0, 0, 0}, dom_hi_[3] = {0, 0, 0};
  int64_t n_particles_ = 0;
  int32_t n_seg_ = 0;
  std::ofstream msd_out_;
};

}  // namespace cntkmc
