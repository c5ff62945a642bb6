// Exercises the CPU oracle (oracle/ymc_oracle.cpp, OpenMP over chains) for sanitizer builds: seeding,
// MWC fills, single steps, and transport runs with every tally (maps, radial bins, totals, per-packet
// records, absorption grid) on a small two-table grid with an inclusion, pencil and uniform light.
// Built and run by tests/test_sanitizers.py with -fsanitize=address,undefined.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/ymc3d.h"

extern "C" {
int32_t ymc_oracle_init_rng(const uint32_t *table, uint32_t table_len, uint64_t xinit, uint64_t n, uint64_t *x,
                            uint32_t *a);
void ymc_oracle_mwc_fill(const uint64_t *x0, const uint32_t *a, uint32_t n_streams, uint32_t n_draws, int32_t kind,
                         float *out, uint64_t *x_final);
uint64_t ymc_oracle_build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes, ymc_chain *out,
                                 uint64_t cap);
int32_t ymc_oracle_run_abs(const ymc_geometry *geom, const ymc_optprops *props, const ymc_params *prm,
                           const ymc_chain *chains, const uint64_t *x0, const uint32_t *a0, uint64_t n_chains,
                           ymc_tally *out, ymc_photon_record *records, uint64_t n_records, int32_t threads,
                           uint64_t *absorption);
}

int main() {
  const uint32_t table[9] = {4294967118u, 4294966893u, 4294966830u, 4294966695u, 4294966224u,
                             4294966050u, 4294965885u, 4294965738u, 4294965612u};
  const int nx = 12, ny = 10, nz = 8, T = 3, L = 2;
  std::vector<uint8_t> rows((size_t)nx * ny * nz, 1);
  for (int k = 2; k < 5; ++k)
    for (int j = 3; j < 7; ++j)
      for (int i = 4; i < 8; ++i) rows[((size_t)k * ny + j) * nx + i] = 2;
  const ymc_geometry geo{nx, ny, nz, 0.01f, 0.01f, 0.01f, rows.data()};
  const float mua[T * L] = {0, 1, 20, 0, 2, 10}, mus[T * L] = {0, 100, 100, 0, 80, 120};
  const float g[T * L] = {0, 0.9f, 0.9f, 0, 0.8f, 0.85f}, n[T * L] = {1, 1.4f, 1.4f, 1, 1.37f, 1.45f};
  const ymc_optprops props{T, L, mua, mus, g, n};
  const uint64_t budgets[L] = {3000, 1700};
  const uint64_t nch = ymc_oracle_build_chains(budgets, L, 256, nullptr, 0);
  std::vector<ymc_chain> ch(nch);
  ymc_oracle_build_chains(budgets, L, 256, ch.data(), nch);
  std::vector<uint64_t> x(nch);
  std::vector<uint32_t> a(nch);
  if (ymc_oracle_init_rng(table, 9, 99, nch, x.data(), a.data()) != 0) return 1;
  std::vector<float> draws(16 * nch);
  std::vector<uint64_t> xf(nch);
  ymc_oracle_mwc_fill(x.data(), a.data(), (uint32_t)nch, 16, 1, draws.data(), xf.data());
  for (int illum = 0; illum < 2; ++illum) {
    ymc_params p{};
    p.seed = 99;
    p.illumination = illum;
    p.pencil_x = p.pencil_y = -1;
    p.radial_bins = 8;
    p.radial_dr = 0.01f;
    p.radial_x0 = p.radial_y0 = -1;
    p.kill_zero_step = 2;
    p.flags = YMC_FLAG_COUNTERS | YMC_FLAG_RECORD | YMC_FLAG_ABSORPTION;
    const size_t mapn = (size_t)L * nx * ny;
    std::vector<uint64_t> Rw(mapn), Rw2(mapn), Tw(mapn), Tw2(mapn), Rr(L * 8), Rr2(L * 8);
    std::vector<uint32_t> Rn(mapn), Tn(mapn), Rrn(L * 8);
    std::vector<ymc_totals> tot(L);
    ymc_tally t{Rw.data(), Rw2.data(), Rn.data(), Tw.data(), Tw2.data(), Tn.data(),
                Rr.data(), Rr2.data(), Rrn.data(), tot.data()};
    std::vector<ymc_photon_record> rec(4700);
    std::vector<uint64_t> absorb((size_t)L * nx * ny * nz);
    if (ymc_oracle_run_abs(&geo, &props, &p, ch.data(), x.data(), a.data(), nch, &t, rec.data(), rec.size(), 4,
                           absorb.data()) != 0)
      return 1;
    if (tot[0].photons != 3000 || tot[1].photons != 1700) return 1;
  }
  std::printf("oracle driver ok\n");
  return 0;
}
