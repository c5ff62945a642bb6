// cli.cpp - command-line driver, the counterpart of the reference's main (src/mc3d_main.cu:15-98).
//
//   ymc3d geometry.bin optprops.bid definition.txt outname [options]
//
// Positional arguments are the reference's (README.md:36).  The third argument, a dummy in the
// reference (src/mc3d_main.cu:24), is read as a definition file of `key value` lines when it
// exists (keys: photons, seed, illumination uniform|pencil, pencil_x, pencil_y); "-" skips it.
// Options override the file:
//   --photons N          packets to launch (default 30000000, src/mc3d_main.cu:57)
//   --seed S             RNG seed (default 1; the reference used time(NULL))
//   --pencil [X Y]       pencil beam (default at the grid centre); --uniform for uniform light
//   --tally NX NY DX DY  reflectance grid (default: the geometry's lateral grid)
//   --multipliers FILE   safeprimes_base32.txt-style multiplier table
//   --fresnel exact      textbook Fresnel formula instead of the reference's
//   --device D           HIP device
//   --absorption         also tally absorbed weight per voxel and write <outname>_abs.bin and
//                        <outname>_flu.bin (save_abs_map, src/mc3d_io.cu:229-255)
// Outputs: <outname>_drs.bin (save_property layout, reflectance normalised by the packet count,
// src/mc3d_io.cu:218-227), <outname>_drt.bin (same for transmittance) and
// <outname>_simparam.dat (src/mc3d_main.cu:85-87).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ymc3d.h"

namespace {

int die(const char *what) {
  std::fprintf(stderr, "ymc3d: %s: %s\n", what, ymc_last_error());
  return 1;
}

bool load_multipliers(const char *path, const std::string &exe_dir, std::vector<uint32_t> &table) {
  const uint32_t cap = (1u << 20) + 1;
  table.resize(cap);
  uint32_t n = 0;
  if (path) {
    if (ymc_multipliers_read_text(path, table.data(), cap, &n) != YMC_OK) return false;
    table.resize(n);
    return true;
  }
  // the reference looks for safeprimes_base32.txt in the working directory (src/mc3d_rng.cu:16)
  if (ymc_multipliers_read_text("safeprimes_base32.txt", table.data(), cap, &n) == YMC_OK) {
    table.resize(n);
    return true;
  }
  std::string bin = exe_dir + "/../data/multipliers.bin";
  if (FILE *f = std::fopen(bin.c_str(), "rb")) {
    size_t got = std::fread(table.data(), 4, cap, f);
    std::fclose(f);
    if (got >= 2) {
      table.resize(got);
      return true;
    }
  }
  return ymc_multipliers_generate(cap, table.data()) == YMC_OK;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "\nUsage: %s geometry.bin tissue.bid definition.txt OutputName [options]\n", argv[0]);
    return 1;
  }
  const char *geom_path = argv[1], *prop_path = argv[2], *def_path = argv[3], *out_name = argv[4];
  uint64_t photons = 30000000ull, seed = 1;
  int illum = YMC_ILLUM_UNIFORM, device = 0, fresnel = YMC_FRESNEL_YMC3D;
  float px = -1, py = -1;
  int tnx = 0, tny = 0;
  float tdx = 0, tdy = 0;
  const char *mult_path = nullptr;
  bool absorption = false;

  if (std::strcmp(def_path, "-") != 0) {
    if (FILE *f = std::fopen(def_path, "r")) {
      char key[64], val[128];
      while (std::fscanf(f, "%63s %127s", key, val) == 2) {
        if (!std::strcmp(key, "photons")) photons = std::strtoull(val, nullptr, 10);
        else if (!std::strcmp(key, "seed")) seed = std::strtoull(val, nullptr, 10);
        else if (!std::strcmp(key, "illumination")) illum = std::strcmp(val, "pencil") ? YMC_ILLUM_UNIFORM : YMC_ILLUM_PENCIL;
        else if (!std::strcmp(key, "pencil_x")) px = std::strtof(val, nullptr);
        else if (!std::strcmp(key, "pencil_y")) py = std::strtof(val, nullptr);
      }
      std::fclose(f);
    }
  }
  for (int i = 5; i < argc; ++i) {
    std::string a = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) {
        std::fprintf(stderr, "ymc3d: %s needs %d value(s)\n", a.c_str(), k);
        std::exit(1);
      }
    };
    if (a == "--photons") need(1), photons = std::strtoull(argv[++i], nullptr, 10);
    else if (a == "--seed") need(1), seed = std::strtoull(argv[++i], nullptr, 10);
    else if (a == "--uniform") illum = YMC_ILLUM_UNIFORM;
    else if (a == "--pencil") {
      illum = YMC_ILLUM_PENCIL;
      if (i + 2 < argc && argv[i + 1][0] != '-') {
        px = std::strtof(argv[++i], nullptr);
        py = std::strtof(argv[++i], nullptr);
      }
    } else if (a == "--tally") {
      need(4);
      tnx = std::atoi(argv[++i]);
      tny = std::atoi(argv[++i]);
      tdx = std::strtof(argv[++i], nullptr);
      tdy = std::strtof(argv[++i], nullptr);
    } else if (a == "--multipliers") need(1), mult_path = argv[++i];
    else if (a == "--fresnel") need(1), fresnel = std::strcmp(argv[++i], "exact") ? YMC_FRESNEL_YMC3D : YMC_FRESNEL_EXACT;
    else if (a == "--device") need(1), device = std::atoi(argv[++i]);
    else if (a == "--absorption") absorption = true;
    else {
      std::fprintf(stderr, "ymc3d: unknown option %s\n", a.c_str());
      return 1;
    }
  }

  int32_t nx, ny, nz;
  double dx, dy, dz;
  uint8_t *tissue = nullptr;
  if (ymc_read_geometry(geom_path, &nx, &ny, &nz, &dx, &dy, &dz, &tissue) != YMC_OK) return die("geometry");
  int32_t T;
  float *mua, *mus, *g, *n;
  if (ymc_read_optprops(prop_path, &T, &mua, &mus, &g, &n) != YMC_OK) return die("optical properties");
  std::fprintf(stderr, "\nRead geometry: %s and tissue definition: %s.\n", geom_path, prop_path);
  std::printf("#mua, mus, g, n. Rows: tissue type 1, tissue type 2, ...\n");
  for (int i = 1; i < T; ++i) std::printf("%g %g %g %g\n", mua[i], mus[i], g[i], n[i]);

  std::string exe = argv[0];
  std::string exe_dir = exe.find('/') == std::string::npos ? "." : exe.substr(0, exe.rfind('/'));
  std::vector<uint32_t> table;
  if (!load_multipliers(mult_path, exe_dir, table)) return die("multiplier table");

  ymc_engine *e = nullptr;
  if (ymc_engine_create(device, table.data(), (uint32_t)table.size(), &e) != YMC_OK) return die("engine");
  ymc_geometry geo{nx, ny, nz, (float)dx, (float)dy, (float)dz, tissue};
  ymc_optprops op{T, 1, mua, mus, g, n};
  if (ymc_engine_set_geometry(e, &geo) != YMC_OK) return die("geometry upload");
  if (ymc_engine_set_optprops(e, &op) != YMC_OK) return die("optical-property upload");
  ymc_params prm{};
  prm.num_photons = photons;
  prm.seed = seed;
  prm.illumination = illum;
  prm.pencil_x = px;
  prm.pencil_y = py;
  prm.tally_nx = tnx;
  prm.tally_ny = tny;
  prm.tally_dx = tdx;
  prm.tally_dy = tdy;
  prm.radial_x0 = prm.radial_y0 = -1;
  prm.fresnel_mode = fresnel;
  prm.kill_zero_step = 1;
  if (absorption) prm.flags |= YMC_FLAG_ABSORPTION;
  if (ymc_engine_prepare(e, &prm) != YMC_OK) return die("prepare");
  std::fprintf(stderr, "\nPerforming MC3D simulation on the GPU (%llu packets, seed %llu).\n",
               (unsigned long long)photons, (unsigned long long)seed);
  if (ymc_engine_run(e) != YMC_OK) return die("run");

  const int onx = tnx > 0 ? tnx : nx, ony = tny > 0 ? tny : ny;
  const double odx = tdx > 0 ? tdx : dx, ody = tdy > 0 ? tdy : dy;
  std::vector<uint64_t> rw((size_t)onx * ony), tw((size_t)onx * ony);
  ymc_totals tot{};
  ymc_tally out{};
  out.R_w = rw.data();
  out.T_w = tw.data();
  out.totals = &tot;
  if (ymc_engine_read_tally(e, &out) != YMC_OK) return die("read tally");
  double ms = 0;
  ymc_engine_last_kernel_ms(e, &ms);

  const double N = (double)tot.photons;
  double geomR = 0;
  std::vector<double> drs(rw.size()), drt(tw.size());
  for (size_t b = 0; b < rw.size(); ++b) {
    drs[b] = (double)rw[b] / YMC_FX_ONE / N;
    drt[b] = (double)tw[b] / YMC_FX_ONE / N;
    geomR += drs[b];
  }
  std::fprintf(stderr, "Total diffuse reflectance is %f.\n", (double)tot.R_w / YMC_FX_ONE / N);
  std::fprintf(stderr, "(diffuse reflectance over geometry area is %f)\n", geomR);
  std::fprintf(stderr, "Weight finishing below the surface is %f.\n", (double)tot.T_w / YMC_FX_ONE / N);
  std::fprintf(stderr, "Transport kernel: %.3f ms, %.4g packets/s.\n", ms, N / (ms * 1e-3));

  std::string base = out_name;
  if (ymc_write_property((base + "_drs.bin").c_str(), 2, onx, ony, 1, odx, ody, 0, drs.data()) != YMC_OK)
    return die("write _drs.bin");
  if (ymc_write_property((base + "_drt.bin").c_str(), 2, onx, ony, 1, odx, ody, 0, drt.data()) != YMC_OK)
    return die("write _drt.bin");
  if (absorption) {
    // Absorbed weight per launched packet in each voxel (_abs.bin) and that divided by the voxel's
    // mu_a (_flu.bin, 0 where mu_a is 0), as save_abs_map lays them out.  save_abs_map multiplies by
    // the packet count where a per-packet normalisation is meant; this writes per packet, like
    // _drs.bin.
    const size_t nv = (size_t)nx * ny * nz;
    std::vector<uint64_t> ag(nv);
    if (ymc_engine_read_absorption(e, ag.data(), nv) != YMC_OK) return die("read absorption");
    std::vector<double> abs_map(nv), flu(nv);
    double sum = 0;
    for (size_t v = 0; v < nv; ++v) {
      abs_map[v] = (double)ag[v] / YMC_FX_ONE / N;
      const float m = mua[tissue[v]];
      flu[v] = m > 0 ? abs_map[v] / m : 0.0;
      sum += abs_map[v];
    }
    std::fprintf(stderr, "Absorbed fraction in geometry is %f.\n", sum);
    if (ymc_write_property((base + "_abs.bin").c_str(), 3, nx, ny, nz, dx, dy, dz, abs_map.data()) != YMC_OK)
      return die("write _abs.bin");
    if (ymc_write_property((base + "_flu.bin").c_str(), 3, nx, ny, nz, dx, dy, dz, flu.data()) != YMC_OK)
      return die("write _flu.bin");
  }
  if (FILE *f = std::fopen((base + "_simparam.dat").c_str(), "w")) {
    std::fprintf(f, "%llu #num_photons\n", (unsigned long long)tot.photons);
    std::fclose(f);
  } else {
    std::fprintf(stderr, "ymc3d: cannot write %s_simparam.dat\n", out_name);
    return 1;
  }
  ymc_engine_destroy(e);
  ymc_free(tissue);
  ymc_free(mua);
  ymc_free(mus);
  ymc_free(g);
  ymc_free(n);
  return 0;
}
