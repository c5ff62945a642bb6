// host_util.cpp - host side of the engine: multiplier table, stream seeding, chain schedule,
// binary file formats and error reporting.  No device code.
#include "../../include/ymc3d.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "host_util.h"

namespace ymc {

static thread_local char g_last_error[512];

ymc_status fail(ymc_status st, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
  return st;
}

const char *last_error() { return g_last_error; }

// ------------------------------------------------------------------------------------------
// Multiplier generation.  A multiply-with-carry generator x' = a*x + c (base 2^32) has
// maximal period (a*2^32-2)/2 when a*2^32-1 and a*2^31-1 are both prime; this is the property
// the entries of GPU-MCML's safeprimes_base32.txt satisfy.  Candidates are scanned downward
// from 2^32-1 in segments: a sieve removes a with a small prime factor in either number, then
// a deterministic Miller-Rabin test (bases valid for every n < 2^64) decides the survivors.

namespace {

inline uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }

bool miller_rabin(uint64_t n) {
  if (n < 4) return n == 2 || n == 3;
  if ((n & 1) == 0) return false;
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    ++s;
  }
  static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  for (uint64_t b : bases) {
    uint64_t a = b % n;
    if (a == 0) continue;
    uint64_t x = 1, p = a, e = d;
    while (e) {
      if (e & 1) x = mul_mod(x, p, n);
      p = mul_mod(p, p, n);
      e >>= 1;
    }
    if (x == 1 || x == n - 1) continue;
    bool witness = true;
    for (int r = 1; r < s; ++r) {
      x = mul_mod(x, x, n);
      if (x == n - 1) {
        witness = false;
        break;
      }
    }
    if (witness) return false;
  }
  return true;
}

struct SievePrime {
  uint32_t p, r32, r31;  // a = r32 (mod p) <=> p | a*2^32-1 ; a = r31 (mod p) <=> p | a*2^31-1
};

std::vector<SievePrime> sieve_primes(uint32_t limit) {
  std::vector<char> comp(limit + 1, 0);
  std::vector<SievePrime> out;
  for (uint32_t p = 3; p <= limit; p += 2) {
    if (comp[p]) continue;
    for (uint64_t q = (uint64_t)p * p; q <= limit; q += 2 * p) comp[q] = 1;
    // inverse of 2 mod p is (p+1)/2; inverse of 2^k is its k-th power
    uint64_t inv2 = (p + 1) / 2, inv31 = 1;
    for (int k = 0; k < 31; ++k) inv31 = inv31 * inv2 % p;
    uint64_t inv32 = inv31 * inv2 % p;
    out.push_back({p, (uint32_t)inv32, (uint32_t)inv31});
  }
  return out;
}

// Valid multipliers in (hi - len, hi], descending.
void scan_segment(uint32_t hi, uint32_t len, const std::vector<SievePrime> &primes, std::vector<uint32_t> &out) {
  uint32_t lo = hi - len + 1;  // inclusive
  std::vector<char> bad(len, 0);
  for (const SievePrime &sp : primes) {
    for (uint32_t r : {sp.r32, sp.r31}) {
      // first a >= lo with a = r (mod p)
      uint32_t m = lo % sp.p;
      uint32_t off = (r + sp.p - m) % sp.p;
      for (uint64_t idx = off; idx < len; idx += sp.p) bad[idx] = 1;
    }
  }
  for (int64_t idx = (int64_t)len - 1; idx >= 0; --idx) {
    if (bad[idx]) continue;
    uint32_t a = lo + (uint32_t)idx;
    uint64_t n31 = ((uint64_t)a << 31) - 1;
    if (!miller_rabin(n31)) continue;
    uint64_t n32 = ((uint64_t)a << 32) - 1;
    if (!miller_rabin(n32)) continue;
    out.push_back(a);
  }
}

}  // namespace

ymc_status generate_multipliers(uint32_t count, uint32_t *out) {
  if (!out || count == 0) return fail(YMC_ERR_ARG, "generate_multipliers: bad arguments");
  const uint32_t kSeg = 1u << 20;
  static const std::vector<SievePrime> primes = sieve_primes(1u << 15);
  unsigned nthreads = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  uint32_t have = 0;
  uint64_t hi = 0xffffffffull;
  while (have < count) {
    // one batch of nthreads consecutive segments, scanned in parallel, merged in order
    std::vector<std::vector<uint32_t>> res(nthreads);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nthreads; ++t) {
      uint64_t seg_hi = hi - (uint64_t)t * kSeg;
      if (seg_hi < (1ull << 31) + kSeg) break;
      pool.emplace_back([&, t, seg_hi] { scan_segment((uint32_t)seg_hi, kSeg, primes, res[t]); });
    }
    if (pool.empty()) return fail(YMC_ERR_TABLE, "multiplier search ran below 2^31");
    for (auto &th : pool) th.join();
    for (size_t t = 0; t < pool.size() && have < count; ++t)
      for (uint32_t a : res[t]) {
        if (have == count) break;
        out[have++] = a;
      }
    hi -= (uint64_t)pool.size() * kSeg;
  }
  return YMC_OK;
}

ymc_status read_multipliers_text(const char *path, uint32_t *out, uint32_t cap, uint32_t *count) {
  if (!path || !out || !count) return fail(YMC_ERR_ARG, "read_multipliers_text: null argument");
  FILE *f = std::fopen(path, "r");
  if (!f) return fail(YMC_ERR_IO, "cannot open multiplier file %s", path);
  uint32_t n = 0;
  unsigned long long a, b, c;
  while (n < cap && std::fscanf(f, "%llu %llu %llu", &a, &b, &c) == 3) {
    if (a < 2 || a > 0xffffffffull) {
      std::fclose(f);
      return fail(YMC_ERR_TABLE, "multiplier file %s: bad value on line %u", path, n + 1);
    }
    out[n++] = (uint32_t)a;
  }
  std::fclose(f);
  *count = n;
  if (n < 2) return fail(YMC_ERR_TABLE, "multiplier file %s has fewer than 2 lines", path);
  return YMC_OK;
}

// ------------------------------------------------------------------------------------------
// Stream seeding, gpumcml_rng/gpumcml_rng.cu:59-118.  The first table entry (`begin`) drives a
// seeding MWC stream; each output stream i takes multiplier table[1 + i mod (len-1)] and draws
// its carry c = floor(u/2^32 * a) and value x from two consecutive seeding outputs, rejecting
// the forbidden states (x == 0, c >= a-1, low word 0xffffffff) exactly as the reference loop.

ymc_status seed_streams(const uint32_t *table, uint32_t table_len, uint64_t seed, uint64_t n, uint64_t *x_out,
                        uint32_t *a_out) {
  if (!table || table_len < 2 || (n && (!x_out || !a_out)))
    return fail(YMC_ERR_ARG, "seed_streams: need a table of >= 2 multipliers and output arrays");
  const uint32_t begin = table[0];
  const uint64_t nmult = table_len - 1;
  uint64_t s = seed;
  const uint32_t s_hi = (uint32_t)(s >> 32), s_lo = (uint32_t)s;
  if (s == 0 || s_hi >= begin - 1 || s_lo >= 0xffffffffu)
    return fail(YMC_ERR_SEED, "seed %llu is not a valid MWC state for multiplier %u", (unsigned long long)seed, begin);
  auto next = [&]() -> uint32_t {
    s = (uint64_t)(uint32_t)s * begin + (s >> 32);
    return (uint32_t)s;
  };
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t a = table[1 + i % nmult];
    uint64_t st;
    for (;;) {
      uint32_t u = next();
      uint64_t carry = (uint32_t)std::floor(((double)u / 4294967296.0) * (double)a);
      uint32_t v = next();
      st = (carry << 32) + v;
      bool bad = st == 0 || (uint32_t)(st >> 32) >= a - 1 || (uint32_t)st == 0xffffffffu;
      if (!bad) break;
    }
    x_out[i] = st;
    a_out[i] = a;
  }
  return YMC_OK;
}

// ------------------------------------------------------------------------------------------
// Chain schedule.  Chains are handed to lanes in order; their lengths shrink with the work
// left (guided scheduling) so the last chains are single packets and the end of the run is
// not held up by a long chain.  The schedule depends only on the budgets and chain_lanes, never
// on the device, which keeps results identical for any pool size.

ymc_status build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes, ymc_chain *out,
                        uint64_t cap, uint64_t *count) {
  if (!budgets || num_tables < 1 || chain_lanes < 1 || !count) return fail(YMC_ERR_ARG, "build_chains: bad arguments");
  std::vector<uint64_t> left(budgets, budgets + num_tables);
  uint64_t total = 0;
  for (uint64_t b : left) total += b;
  const uint64_t denom = 2ull * (uint64_t)chain_lanes;
  uint64_t n = 0;
  int w = 0;
  while (total) {
    while (left[w] == 0) w = (w + 1) % num_tables;
    uint64_t len = std::min<uint64_t>((total + denom - 1) / denom, 0xffffffffull);
    len = std::min(len, left[w]);
    if (out) {
      if (n >= cap) return fail(YMC_ERR_ARG, "build_chains: output capacity %llu too small", (unsigned long long)cap);
      out[n] = ymc_chain{(uint32_t)len, (uint32_t)w};
    }
    ++n;
    left[w] -= len;
    total -= len;
    w = (w + 1) % num_tables;
  }
  *count = n;
  return YMC_OK;
}

// ------------------------------------------------------------------------------------------
// Safe radius of every voxel: the largest r <= kSafeMax such that every voxel within Chebyshev
// distance r has the same row, lies inside the lateral grid (not beyond a mirror wall) and not
// above the top layer.  Below the grid the reference's texture clamps to the bottom layer
// (src/mc3d_gpu.cu:416-418), so the bottom neighbourhood is read through the same clamp.
// S_1(v): all 26 neighbours have v's row; S_r = erosion of S_{r-1} by the 3x3x3 cube, done as
// three separable 1-D erosions.

namespace {
void erode_axis(const uint8_t *in, uint8_t *out, int nx, int ny, int nz, int axis) {
  auto idx = [&](int i, int j, int k) { return ((size_t)k * ny + j) * nx + i; };
  unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthreads; ++t) {
    pool.emplace_back([&, t] {
      for (int k = (int)t; k < nz; k += (int)nthreads)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            uint8_t v = in[idx(i, j, k)];
            if (v) {
              if (axis == 0) {
                v = (i > 0 && i < nx - 1) ? (in[idx(i - 1, j, k)] & in[idx(i + 1, j, k)]) : 0;
              } else if (axis == 1) {
                v = (j > 0 && j < ny - 1) ? (in[idx(i, j - 1, k)] & in[idx(i, j + 1, k)]) : 0;
              } else {
                int kd = std::min(k + 1, nz - 1);  // clamp below the grid
                v = (k > 0) ? (in[idx(i, j, k - 1)] & in[idx(i, j, kd)]) : 0;
              }
            }
            out[idx(i, j, k)] = v;
          }
    });
  }
  for (auto &th : pool) th.join();
}
}  // namespace

void safe_radius(const uint8_t *rows, int nx, int ny, int nz, uint8_t *radius) {
  const size_t nv = (size_t)nx * ny * nz;
  auto idx = [&](int i, int j, int k) { return ((size_t)k * ny + j) * nx + i; };
  std::vector<uint8_t> cur(nv), tmp(nv), tmp2(nv);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        uint8_t ok = (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0) ? 1 : 0;
        const uint8_t t = rows[idx(i, j, k)];
        for (int dk = -1; dk <= 1 && ok; ++dk)
          for (int dj = -1; dj <= 1 && ok; ++dj)
            for (int di = -1; di <= 1 && ok; ++di) {
              int kk = std::min(k + dk, nz - 1);
              if (rows[idx(i + di, j + dj, kk)] != t) ok = 0;
            }
        cur[idx(i, j, k)] = ok;
        radius[idx(i, j, k)] = ok;
      }
  for (int r = 2; r <= kSafeMax; ++r) {
    erode_axis(cur.data(), tmp.data(), nx, ny, nz, 0);
    erode_axis(tmp.data(), tmp2.data(), nx, ny, nz, 1);
    erode_axis(tmp2.data(), cur.data(), nx, ny, nz, 2);
    bool any = false;
    for (size_t v = 0; v < nv; ++v) {
      radius[v] += cur[v];
      any |= cur[v] != 0;
    }
    if (!any) break;
  }
}

// ------------------------------------------------------------------------------------------
// Safe box of every voxel: extents zup (towards the top), zdn (down) and rxy (lateral, square)
// such that every voxel of [i-rxy, i+rxy] x [j-rxy, j+rxy] x [k-zup, k+zdn] has the voxel's row,
// lies inside the lateral grid and not above the top layer (below the grid, the clamp to the
// bottom layer).  Layered tissue is wide and thin, so a box reaches much further sideways than the
// cube of safe_radius, and a voxel at the top or bottom of a layer keeps the layer's height on
// one side.  Per voxel:
//   up, dn = vertical runs of the voxel's row above and below it (capped at zcap);
//   rxy(a, b) = largest s <= xycap such that the s-square of the voxel's slice has the voxel's row
//               everywhere, up >= a and dn >= b everywhere, found by binary search over square
//               range-min/max tables;
// candidates (r, r), (up, dn), (up, r) and (r, dn) for (zup, zdn), r the cube radius capped at
// zcap, each with its rxy; among those containing the cube (rxy >= r), the one with the largest
// harmonic mean of rxy+1, rxy+1 and (zup+zdn)/2+1 is kept, the first listed winning ties.

namespace {
// Square sparse tables of one slice: level m holds min/max over [i, i+2^m) x [j, j+2^m).
struct SquareTables {
  int nx, ny, levels;
  std::vector<uint8_t> upmin, dnmin, rowmin, rowmax;  // [levels][ny][nx]
  void build(const uint8_t *up, const uint8_t *dn, const uint8_t *row, int nx_, int ny_) {
    nx = nx_;
    ny = ny_;
    levels = 1;
    while ((1 << levels) <= std::max(nx, ny)) ++levels;
    const size_t plane = (size_t)nx * ny;
    for (auto *v : {&upmin, &dnmin, &rowmin, &rowmax}) v->assign(plane * levels, 0);
    std::copy(up, up + plane, upmin.begin());
    std::copy(dn, dn + plane, dnmin.begin());
    std::copy(row, row + plane, rowmin.begin());
    std::copy(row, row + plane, rowmax.begin());
    for (int m = 1; m < levels; ++m) {
      const int h = 1 << (m - 1), w = 1 << m;
      const size_t o = plane * m, po = plane * (m - 1);
      for (int j = 0; j + w <= ny; ++j)
        for (int i = 0; i + w <= nx; ++i) {
          const size_t a = (size_t)j * nx + i, b = a + h, c = a + (size_t)h * nx, d = c + h;
          auto mn = [&](const std::vector<uint8_t> &t) {
            return std::min(std::min(t[po + a], t[po + b]), std::min(t[po + c], t[po + d]));
          };
          upmin[o + a] = mn(upmin);
          dnmin[o + a] = mn(dnmin);
          rowmin[o + a] = mn(rowmin);
          rowmax[o + a] = std::max(std::max(rowmax[po + a], rowmax[po + b]), std::max(rowmax[po + c], rowmax[po + d]));
        }
    }
  }
  // square [i-s, i+s] x [j-s, j+s] (inside the slice): all rows == r, up >= a and dn >= b
  bool uniform(int i, int j, int s, uint8_t r, int a, int b) const {
    const int L = 2 * s + 1;
    int m = 0;
    while ((2 << m) <= L) ++m;
    const int w = 1 << m;
    const size_t plane = (size_t)nx * ny;
    const int x0 = i - s, x1 = i + s - w + 1, y0 = j - s, y1 = j + s - w + 1;
    const size_t q[4] = {(size_t)y0 * nx + x0, (size_t)y0 * nx + x1, (size_t)y1 * nx + x0, (size_t)y1 * nx + x1};
    for (size_t c : q) {
      const size_t o = plane * m + c;
      if (upmin[o] < a || dnmin[o] < b || rowmin[o] != r || rowmax[o] != r) return false;
    }
    return true;
  }
};
}  // namespace

void safe_box(const uint8_t *rows, int nx, int ny, int nz, const uint8_t *radius, int zcap, int xycap,
              uint8_t *zup, uint8_t *zdn, uint8_t *rxy) {
  const size_t plane = (size_t)nx * ny;
  std::vector<uint8_t> up(plane * nz), dn(plane * nz);
  for (size_t c = 0; c < plane; ++c) {
    int run = 0;
    for (int k = 0; k < nz; ++k) {
      run = (k > 0 && rows[(size_t)(k - 1) * plane + c] == rows[(size_t)k * plane + c]) ? run + 1 : 0;
      up[(size_t)k * plane + c] = (uint8_t)std::min(run, zcap);
    }
    run = zcap;  // the bottom layer repeats below the grid
    for (int k = nz - 1; k >= 0; --k) {
      if (k < nz - 1) run = rows[(size_t)(k + 1) * plane + c] == rows[(size_t)k * plane + c] ? run + 1 : 0;
      dn[(size_t)k * plane + c] = (uint8_t)std::min(run, zcap);
    }
  }
  unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthreads; ++t) {
    pool.emplace_back([&, t] {
      SquareTables tab;
      for (int k = (int)t; k < nz; k += (int)nthreads) {
        const size_t off = (size_t)k * plane;
        tab.build(&up[off], &dn[off], &rows[off], nx, ny);
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            const size_t v = off + (size_t)j * nx + i;
            const uint8_t row = rows[v];
            const int smax = std::min(std::min(std::min(i, nx - 1 - i), std::min(j, ny - 1 - j)), xycap);
            auto widest = [&](int a, int b) {
              int lo = 0, hi = smax;  // s = 0 holds: the voxel itself
              while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (tab.uniform(i, j, mid, row, a, b)) lo = mid; else hi = mid - 1;
              }
              return lo;
            };
            const int r = std::min<int>(radius[v], zcap), u = up[v], d = dn[v];
            const int cand[4][2] = {{r, r}, {u, d}, {u, r}, {r, d}};
            int bu = r, bd = r, bx = widest(r, r);
            auto score = [](int a, int b, int x) { return 1.0 / (2.0 / (x + 1) + 2.0 / (a + b + 2)); };
            double best = score(bu, bd, bx);
            for (int q = 1; q < 4; ++q) {
              const int a = cand[q][0], b = cand[q][1];
              if (a == bu && b == bd) continue;
              const int x = widest(a, b);
              if (x < r) continue;
              const double sc = score(a, b, x);
              if (sc > best) {
                best = sc;
                bu = a;
                bd = b;
                bx = x;
              }
            }
            zup[v] = (uint8_t)bu;
            zdn[v] = (uint8_t)bd;
            rxy[v] = (uint8_t)bx;
          }
      }
    });
  }
  for (auto &th : pool) th.join();
}

// Parametric layered geometry as a voxel grid: the same binary32 cylinder test as the device, each
// product and sum rounded on its own (separate statements, contraction off).
ymc_status layered_voxelize(const ymc_layered *g, int nx, int ny, int nz, uint8_t *rows) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  if (!g || !rows || nx <= 0 || ny <= 0 || nz <= 0) return fail(YMC_ERR_ARG, "layered_voxelize: bad arguments");
  const float r2 = g->vessel_r * g->vessel_r;
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      int row = k < g->top_voxels ? g->row_top : g->row_bottom;
      if (g->row_vessel >= 0) {
        float ci = (float)i + 0.5f;
        float ck = (float)k + 0.5f;
        float fx = ci - g->vessel_x;
        float fz = ck - g->vessel_z;
        float ax = fx * fx;
        float az = fz * fz;
        float d2 = ax + az;
        if (d2 <= r2) row = g->row_vessel;
      }
      for (int j = 0; j < ny; ++j) rows[((size_t)k * ny + j) * nx + i] = (uint8_t)row;
    }
  return YMC_OK;
}

// ------------------------------------------------------------------------------------------
// Binary formats, src/mc3d_io.cu.

namespace {
struct File {
  FILE *f;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};
}  // namespace

ymc_status read_geometry(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dx, double *dy,
                         double *dz, uint8_t **tissue) {
  if (!path || !nx || !ny || !nz || !dx || !dy || !dz || !tissue) return fail(YMC_ERR_ARG, "read_geometry: null argument");
  File F(std::fopen(path, "rb"));
  if (!F.f) return fail(YMC_ERR_IO, "cannot open geometry file %s", path);
  int32_t hdr[4];
  double d[3];
  if (std::fread(hdr, 4, 4, F.f) != 4 || std::fread(d, 8, 3, F.f) != 3)
    return fail(YMC_ERR_IO, "geometry file %s: short header", path);
  if (hdr[0] != 3) return fail(YMC_ERR_FORMAT, "geometry file %s: dimension %d is not 3", path, hdr[0]);
  if (hdr[1] <= 0 || hdr[2] <= 0 || hdr[3] <= 0) return fail(YMC_ERR_FORMAT, "geometry file %s: bad size", path);
  if (!(d[0] > 0) || !(d[1] > 0) || !(d[2] > 0)) return fail(YMC_ERR_FORMAT, "geometry file %s: bad voxel size", path);
  size_t nv = (size_t)hdr[1] * hdr[2] * hdr[3];
  std::vector<int32_t> raw(nv);
  if (std::fread(raw.data(), 4, nv, F.f) != nv) return fail(YMC_ERR_IO, "geometry file %s: short voxel data", path);
  uint8_t *t = (uint8_t *)std::malloc(nv);
  if (!t) return fail(YMC_ERR_NOMEM, "read_geometry: out of memory");
  for (size_t v = 0; v < nv; ++v) {
    int32_t val = raw[v];
    if (val < 0 || val > 255) {
      std::free(t);
      return fail(YMC_ERR_FORMAT, "geometry file %s: voxel value %d outside 0..255", path, val);
    }
    t[v] = (val == 255) ? (uint8_t)YMC_TISSUE_AIR : (uint8_t)(val + 1);  // src/mc3d_io.cu:156-163
  }
  *nx = hdr[1];
  *ny = hdr[2];
  *nz = hdr[3];
  *dx = d[0];
  *dy = d[1];
  *dz = d[2];
  *tissue = t;
  return YMC_OK;
}

ymc_status write_geometry(const char *path, int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                          const int32_t *vals) {
  if (!path || !vals || nx <= 0 || ny <= 0 || nz <= 0) return fail(YMC_ERR_ARG, "write_geometry: bad arguments");
  File F(std::fopen(path, "wb"));
  if (!F.f) return fail(YMC_ERR_IO, "cannot create %s", path);
  int32_t hdr[4] = {3, nx, ny, nz};
  double d[3] = {dx, dy, dz};
  size_t nv = (size_t)nx * ny * nz;
  if (std::fwrite(hdr, 4, 4, F.f) != 4 || std::fwrite(d, 8, 3, F.f) != 3 || std::fwrite(vals, 4, nv, F.f) != nv)
    return fail(YMC_ERR_IO, "short write to %s", path);
  return YMC_OK;
}

ymc_status read_optprops(const char *path, int32_t *num_types, float **mua, float **mus, float **g, float **n) {
  if (!path || !num_types || !mua || !mus || !g || !n) return fail(YMC_ERR_ARG, "read_optprops: null argument");
  File F(std::fopen(path, "rb"));
  if (!F.f) return fail(YMC_ERR_IO, "cannot open optical-properties file %s", path);
  int32_t T;
  if (std::fread(&T, 4, 1, F.f) != 1) return fail(YMC_ERR_IO, "optical-properties file %s: short header", path);
  if (T < 1 || T > 255) return fail(YMC_ERR_FORMAT, "optical-properties file %s: %d tissue types", path, T);
  std::vector<double> rec((size_t)T * 4);
  if (std::fread(rec.data(), 8, rec.size(), F.f) != rec.size())
    return fail(YMC_ERR_IO, "optical-properties file %s: short data", path);
  const int32_t rows = T + 1;
  float *arr[4];
  for (int q = 0; q < 4; ++q) {
    arr[q] = (float *)std::malloc(sizeof(float) * rows);
    if (!arr[q]) {
      for (int r = 0; r < q; ++r) std::free(arr[r]);
      return fail(YMC_ERR_NOMEM, "read_optprops: out of memory");
    }
  }
  // row 0: ambient medium (src/mc3d_io.cu:38-42)
  arr[0][0] = 0;
  arr[1][0] = 0;
  arr[2][0] = 0;
  arr[3][0] = 1.0f;
  for (int k = 1; k < rows; ++k)  // records in the order mua, mus, g, n (src/mc3d_io.cu:45-52)
    for (int q = 0; q < 4; ++q) arr[q][k] = (float)rec[(size_t)(k - 1) * 4 + q];
  *num_types = rows;
  *mua = arr[0];
  *mus = arr[1];
  *g = arr[2];
  *n = arr[3];
  return YMC_OK;
}

ymc_status write_optprops(const char *path, int32_t num_records, const double *records) {
  if (!path || !records || num_records < 1) return fail(YMC_ERR_ARG, "write_optprops: bad arguments");
  File F(std::fopen(path, "wb"));
  if (!F.f) return fail(YMC_ERR_IO, "cannot create %s", path);
  size_t nd = (size_t)num_records * 4;
  if (std::fwrite(&num_records, 4, 1, F.f) != 1 || std::fwrite(records, 8, nd, F.f) != nd)
    return fail(YMC_ERR_IO, "short write to %s", path);
  return YMC_OK;
}

ymc_status write_property(const char *path, int32_t dims, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                          double dz, const double *data) {
  if (!path || !data || (dims != 2 && dims != 3) || nx <= 0 || ny <= 0 || (dims == 3 && nz <= 0))
    return fail(YMC_ERR_ARG, "write_property: bad arguments");
  File F(std::fopen(path, "wb"));
  if (!F.f) return fail(YMC_ERR_IO, "cannot create %s", path);
  int32_t ih[4] = {dims, nx, ny, nz};
  double dh[3] = {dx, dy, dz};
  size_t cnt = (size_t)nx * ny * (dims == 3 ? nz : 1);
  if (std::fwrite(ih, 4, 1 + dims, F.f) != (size_t)(1 + dims) || std::fwrite(dh, 8, dims, F.f) != (size_t)dims ||
      std::fwrite(data, 8, cnt, F.f) != cnt)
    return fail(YMC_ERR_IO, "short write to %s", path);
  return YMC_OK;
}

}  // namespace ymc

// ============================================================================================
// C ABI wrappers for the host utilities.

extern "C" {

const char *ymc_last_error(void) { return ymc::last_error(); }
int32_t ymc_abi_version(void) { return YMC3D_ABI_VERSION; }

ymc_status ymc_multipliers_generate(uint32_t count, uint32_t *out) { return ymc::generate_multipliers(count, out); }
ymc_status ymc_multipliers_read_text(const char *path, uint32_t *out, uint32_t cap, uint32_t *count) {
  return ymc::read_multipliers_text(path, out, cap, count);
}
ymc_status ymc_seed_streams(const uint32_t *table, uint32_t table_len, uint64_t seed, uint64_t n, uint64_t *x_out,
                            uint32_t *a_out) {
  return ymc::seed_streams(table, table_len, seed, n, x_out, a_out);
}
ymc_status ymc_build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes, ymc_chain *out,
                            uint64_t cap, uint64_t *count) {
  return ymc::build_chains(budgets, num_tables, chain_lanes, out, cap, count);
}
ymc_status ymc_read_geometry(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dx, double *dy,
                             double *dz, uint8_t **tissue) {
  return ymc::read_geometry(path, nx, ny, nz, dx, dy, dz, tissue);
}
ymc_status ymc_write_geometry(const char *path, int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                              const int32_t *file_values) {
  return ymc::write_geometry(path, nx, ny, nz, dx, dy, dz, file_values);
}
ymc_status ymc_read_optprops(const char *path, int32_t *num_types, float **mua, float **mus, float **g, float **n) {
  return ymc::read_optprops(path, num_types, mua, mus, g, n);
}
ymc_status ymc_write_optprops(const char *path, int32_t num_records, const double *records) {
  return ymc::write_optprops(path, num_records, records);
}
ymc_status ymc_write_property(const char *path, int32_t dims, int32_t nx, int32_t ny, int32_t nz, double dx,
                              double dy, double dz, const double *data) {
  return ymc::write_property(path, dims, nx, ny, nz, dx, dy, dz, data);
}
void ymc_free(void *p) { std::free(p); }
ymc_status ymc_layered_voxelize(const ymc_layered *g, int32_t nx, int32_t ny, int32_t nz, uint8_t *rows) {
  return ymc::layered_voxelize(g, nx, ny, nz, rows);
}

}  // extern "C"
