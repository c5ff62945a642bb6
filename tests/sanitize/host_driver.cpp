// Exercises the engine's host utilities (ymc3d_amd/csrc/host_util.cpp) for sanitizer builds:
// the multiplier search, safe radius and safe box (all three run std::thread pools), stream
// seeding, chain schedules, the parametric voxeliser and the file readers and writers.  Built and
// run by tests/test_sanitizers.py with -fsanitize=address,undefined and with -fsanitize=thread;
// exits non-zero on a wrong result, and the sanitizers report (and abort) on their own findings.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "host_util.h"

#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                   \
    }                                                             \
  } while (0)

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  using namespace ymc;
  // multipliers: two threaded segment batches
  std::vector<uint32_t> tab(40000);
  CHECK(generate_multipliers((uint32_t)tab.size(), tab.data()) == YMC_OK);
  CHECK(tab[0] == 4294967118u);
  for (size_t i = 1; i < tab.size(); ++i) CHECK(tab[i] < tab[i - 1]);
  // seeding and chains
  std::vector<uint64_t> x(5000);
  std::vector<uint32_t> a(5000);
  CHECK(seed_streams(tab.data(), 1025, 12345, x.size(), x.data(), a.data()) == YMC_OK);
  const uint64_t budgets[3] = {100000, 7, 0};
  uint64_t n = 0;
  CHECK(build_chains(budgets, 3, 4096, nullptr, 0, &n) == YMC_OK);
  std::vector<ymc_chain> ch(n);
  CHECK(build_chains(budgets, 3, 4096, ch.data(), n, &n) == YMC_OK);
  uint64_t total = 0;
  for (const ymc_chain &c : ch) total += c.count;
  CHECK(total == 100007);
  // safe radius and safe box on random layered grids with inclusions, odd sizes included
  std::mt19937 rng(7);
  const int shapes[][3] = {{33, 17, 9}, {64, 64, 32}, {1, 1, 1}, {5, 70, 3}};
  for (const auto &s : shapes) {
    const int nx = s[0], ny = s[1], nz = s[2];
    const size_t nv = (size_t)nx * ny * nz;
    std::vector<uint8_t> rows(nv), rad(nv), zu(nv), zd(nv), rxy(nv);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) rows[((size_t)k * ny + j) * nx + i] = (uint8_t)(1 + (k * 3) / std::max(nz, 1));
    for (int q = 0; q < 20 && nv > 1; ++q) rows[rng() % nv] = (uint8_t)(4 + rng() % 3);
    safe_radius(rows.data(), nx, ny, nz, rad.data());
    safe_box(rows.data(), nx, ny, nz, rad.data(), 15, 31, zu.data(), zd.data(), rxy.data());
    for (size_t v = 0; v < nv; ++v) CHECK(rad[v] <= kSafeMax && zu[v] <= 15 && zd[v] <= 15 && rxy[v] <= 31);
  }
  // parametric voxeliser
  ymc_layered g{};
  g.top_voxels = 4;
  g.row_top = 1;
  g.row_bottom = 2;
  g.row_vessel = 3;
  g.vessel_x = 10.5f;
  g.vessel_z = 8.0f;
  g.vessel_r = 3.0f;
  std::vector<uint8_t> lay(24 * 20 * 16);
  CHECK(layered_voxelize(&g, 24, 20, 16, lay.data()) == YMC_OK);
  // files
  std::vector<int32_t> vals(6 * 5 * 4);
  for (size_t q = 0; q < vals.size(); ++q) vals[q] = (q % 7 == 0) ? 255 : (int32_t)(q % 3);
  const std::string gp = dir + "/g.bin", op = dir + "/o.bid", pp = dir + "/p.bin";
  CHECK(write_geometry(gp.c_str(), 6, 5, 4, 0.01, 0.02, 0.03, vals.data()) == YMC_OK);
  int32_t nx, ny, nz;
  double dx, dy, dz;
  uint8_t *t = nullptr;
  CHECK(read_geometry(gp.c_str(), &nx, &ny, &nz, &dx, &dy, &dz, &t) == YMC_OK);
  CHECK(nx == 6 && ny == 5 && nz == 4 && t[0] == 0 && t[1] == 2);
  std::free(t);
  const double rec[8] = {1, 100, 0.9, 1.4, 20, 100, 0.9, 1.4};
  CHECK(write_optprops(op.c_str(), 2, rec) == YMC_OK);
  int32_t T = 0;
  float *mua, *mus, *gg, *nn;
  CHECK(read_optprops(op.c_str(), &T, &mua, &mus, &gg, &nn) == YMC_OK);
  CHECK(T == 3 && mua[2] == 20.0f && nn[0] == 1.0f);
  std::free(mua);
  std::free(mus);
  std::free(gg);
  std::free(nn);
  std::vector<double> prop(6 * 5, 0.5);
  CHECK(write_property(pp.c_str(), 2, 6, 5, 1, 0.01, 0.02, 0, prop.data()) == YMC_OK);
  std::printf("host driver ok\n");
  return 0;
}
