// kubo_main.cpp -- the Green-Kubo run of the reference's driver (src/main.cpp:64-76) as a C++
// program over libcntkmc.so, through the monte_carlo host layer:
//
//   sim.kubo_init();  sim.kubo_create_particles();
//   while (sim.time() < sim.kubo_max_time()) { sim.kubo_step(dt); sim.kubo_save_avg_...(); }
//
// on the film_5k_sites rules: 250 straight (7,5) tubes of 200 nm, random in-plane angle, centres
// in 500 x 500 x 50 nm, 10 nm segments (5,000 sites), cutoff 15 nm, Wong/Förster rates
// (gamma_0 = 1e13) on the input.json grid, the Kubo regions of 5 sections (input.json:65), no
// periodic axis (the reference has none).
//
// usage: kubo_main OUT_DIR [n_particles=2000] [n_steps=40] [max_time=2e-10] [seed=7]
// writes OUT_DIR/particle_dispalcement.avg.squared.dat (one line per step) and OUT_DIR/state.bin:
// int64 E, int64 N, then int32 site[E], f64 t_next[E], u64 hops[E], f64 disp[E][3],
// u32 drains[E], u64 hash[E], u64 counts[N].
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "monte_carlo.hpp"

namespace {

template <typename T>
void put(std::FILE* f, const std::vector<T>& v) {
  if (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
    throw cntkmc::error(KMC_ERR_ARG, "state.bin: short write");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT_DIR [n_particles] [n_steps] [max_time] [seed]\n", argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  const int64_t n_particles = argc > 2 ? std::atoll(argv[2]) : 2000;
  const int n_steps = argc > 3 ? std::atoi(argv[3]) : 40;
  const double max_time = argc > 4 ? std::atof(argv[4]) : 2e-10;
  const uint64_t seed = argc > 5 ? std::strtoull(argv[5], nullptr, 10) : 7;
  try {
    kmc_config cfg{};
    cfg.device = 0;
    cfg.periodic_mask = 0;
    cfg.cutoff = 15e-9;         // "max hopping radius [m]"
    cfg.min_distance = 0.4e-9;  // scatterer.cpp:43
    cfg.seed = seed;
    cntkmc::monte_carlo sim(cfg);

    kmc_netgen_params p{};
    p.seed = 1001;
    p.n_tubes = 250;
    p.len_min = p.len_max = 200e-9;
    p.box_lo[0] = p.box_lo[1] = p.box_lo[2] = 0.0;
    p.box_hi[0] = 500e-9;
    p.box_hi[1] = 500e-9;
    p.box_hi[2] = 50e-9;
    p.site_spacing = 10e-9;
    p.n_chir = 1;
    p.wrap_mask = 0;
    sim.generate_scatterers(p);

    sim.kubo_init(5);
    const cntkmc::table_grid grid = cntkmc::table_grid::reference_default();
    sim.set_scat_table(grid, sim.create_forster_scatt_table(1e13, grid));
    sim.set_max_rate();
    sim.kubo_create_particles(n_particles);
    sim.set_kubo_max_time(max_time);

    const double dt = max_time / n_steps;
    const std::string msd = out + "/particle_dispalcement.avg.squared.dat";
    int steps = 0;
    std::vector<double> last;
    while (sim.time() < sim.kubo_max_time() && steps < n_steps) {
      sim.kubo_step(dt);
      last = sim.kubo_save_avg_dispalcement_squared(msd);
      ++steps;
    }

    const int64_t E = n_particles, N = sim.scatterer_list().size();
    std::vector<int32_t> site(E);
    std::vector<double> t(E), disp(3 * E);
    std::vector<uint64_t> hops(E), hash(E);
    std::vector<uint32_t> drains(E);
    cntkmc::check(kmc_get_excitons(sim.context(), site.data(), t.data(), hops.data(), disp.data(),
                                   drains.data(), hash.data()),
                  sim.context(), "kmc_get_excitons");
    const std::vector<uint64_t> counts = sim.hop_counts();
    std::FILE* f = std::fopen((out + "/state.bin").c_str(), "wb");
    if (!f) throw cntkmc::error(KMC_ERR_ARG, "cannot open " + out + "/state.bin");
    put(f, std::vector<int64_t>{E, N});
    put(f, site);
    put(f, t);
    put(f, hops);
    put(f, disp);
    put(f, drains);
    put(f, hash);
    put(f, counts);
    std::fclose(f);
    uint64_t total = 0;
    for (uint64_t h : hops) total += h;
    std::printf("kubo_main: %lld sites, %lld particles, %d steps to %.6e s, %llu hops, "
                "<dx^2> %.6e <dy^2> %.6e <dz^2> %.6e\n",
                (long long)N, (long long)E, steps, sim.time(), (unsigned long long)total,
                last.empty() ? 0.0 : last[0], last.empty() ? 0.0 : last[1],
                last.empty() ? 0.0 : last[2]);
  } catch (const cntkmc::error& e) {
    std::fprintf(stderr, "kubo_main: %s (status %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
