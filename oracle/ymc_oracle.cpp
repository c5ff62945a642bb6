// ymc_oracle.cpp - CPU restatement of the reference photon-transport path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py load it; the product library (ymc3d_amd/) never does.
//
// Every function cites the reference lines it restates.  The arithmetic follows the reference
// expression by expression (operand order, float vs double promotion of the C constants); it is
// compiled with -ffp-contract=off and without fast-math so that IEEE single precision is what
// the C source says.  Where the reference's behaviour depends on its launch structure (frozen
// lanes drawing random numbers until the 45-step launch ends, src/mc3d_gpu.cu:175-179) the
// oracle uses the chain semantics of the engine instead: each chain of packets owns one MWC
// stream and runs its packets back to back, each packet until the step at which it freezes.
//
// Parity status of each part (how it is pinned) is listed in tests/test_oracle_*.py headers:
//   MWC step and float conversion ........ closed form (tests/test_oracle_rng.py)
//   multipliers .......................... defining primality property, independent checker
//   init_RNG seeding ..................... the begin-multiplier stream split into carry and value
//                                          in Python integers / binary64 (tests/test_oracle_rng.py)
//   grid coord / intersection ............ brute force on tiny grids
//   Fresnel reflectance / HG / specular .. closed forms and sampling moments
//   Fresnel direction update ............. Snell's law, law of reflection, total internal
//                                          reflection, one step at a time through ymc_oracle_step
//                                          (tests/test_oracle_fresnel_direction.py)
//   transport ............................ van de Hulst slab, energy identity, symmetry
#include "../include/ymc3d.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------------------------------
// RNG: multiply-with-carry, base 2^32 (gpumcml_rng/gpumcml_rng.cu:38-54).
// State x holds the value in the low word and the carry in the high word.

inline uint32_t mwc_step(uint64_t &x, uint32_t a) {
  x = (x & 0xffffffffull) * (uint64_t)a + (x >> 32);  // gpumcml_rng.cu:40
  return (uint32_t)x;
}

// Round-toward-zero conversion of a 32-bit integer to binary32: keep the 24 most significant
// bits, which makes the ordinary conversion exact (IEEE 754 definition of RZ).
inline float u32_to_float_rz(uint32_t u) {
  if (u < (1u << 24)) return (float)u;
  int lz = __builtin_clz(u);
  int sh = 8 - lz;  // bits below the 24-bit significand
  return (float)((u >> sh) << sh);
}

// rand_MWC_co: RZ(x mod 2^32) / 2^32 in [0,1) (gpumcml_rng.cu:38-46).  The division by 2^32 is
// an exact power-of-two scaling.
inline float mwc_co(uint64_t &x, uint32_t a) { return std::ldexp(u32_to_float_rz(mwc_step(x, a)), -32); }
// rand_MWC_oc: 1 - co in (0,1] (gpumcml_rng.cu:51-54).
inline float mwc_oc(uint64_t &x, uint32_t a) { return 1.0f - mwc_co(x, a); }

// ------------------------------------------------------------------------------------------
// Primality (deterministic Miller-Rabin for 64-bit n with the first 12 prime bases).

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }
inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod(r, b, m);
    b = mulmod(b, b, m);
    e >>= 1;
  }
  return r;
}
bool is_prime_u64(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : small) {
    if (n % p == 0) return n == p;
  }
  uint64_t d = n - 1;
  int r = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    ++r;
  }
  for (uint64_t a : small) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int i = 1; i < r; ++i) {
      x = mulmod(x, x, n);
      if (x == n - 1) {
        comp = false;
        break;
      }
    }
    if (comp) return false;
  }
  return true;
}

// ------------------------------------------------------------------------------------------
// Geometry helpers.

struct Geom {
  int nx, ny, nz;
  float dx, dy, dz;
  float Lx, Ly;  // length_x = sample_dx*num_x in float (src/mc3d_io.cu:143-145)
  const uint8_t *t;
};

// floor() of a float quotient converted to int with saturation (NaN -> 0), clamped to +-2^30
// so that the i+1 / i-1 arithmetic below cannot overflow.  For finite in-range values this is
// the plain `int i = floorf(x/dx)` of src/mc3d_gpu.cu:54.
inline int floor_to_int(float q) {
  float f = std::floor(q);
  if (!(f == f)) return 0;
  if (f > 1073741824.0f) return 1073741824;
  if (f < -1073741824.0f) return -1073741824;
  return (int)f;
}

// photontrack_get_grid_coord (src/mc3d_gpu.cu:53-63).
inline int grid_coord(float x, float ux, float dx) {
  int i = floor_to_int(x / dx);
  i += (x == (float)(i + 1) * dx);
  float bmin = (float)i * dx;
  i = ((x > bmin) || ((x == bmin) && (ux >= 0))) ? i : i - 1;
  return i;
}

// photontrack_get_tissue_type (src/mc3d_gpu.cu:24-27) with the texture's clamp addressing
// (src/mc3d_gpu.cu:416-418).
inline int tissue_at(const Geom &G, int i, int j, int k, float z, float uz) {
  int ic = std::min(std::max(i, 0), G.nx - 1);
  int jc = std::min(std::max(j, 0), G.ny - 1);
  int kc = std::min(std::max(k, 0), G.nz - 1);
  int tt = G.t[((size_t)kc * G.ny + jc) * G.nx + ic];
  return ((z <= 0) && (uz < 0)) ? YMC_TISSUE_AIR : tt;
}

// photontrack_intersection (src/mc3d_gpu.cu:29-49).
inline float intersection(const Geom &G, int i, int j, int k, float &x, float ux, float &y, float uy,
                          float &z, float uz) {
  float bx = (float)((ux >= 0) ? (i + 1) : i) * G.dx;
  float tx = (bx - x) / ux;
  float by = (float)((uy >= 0) ? (j + 1) : j) * G.dy;
  float ty = (by - y) / uy;
  float bz = (float)((uz >= 0) ? (k + 1) : k) * G.dz;
  float tz = (bz - z) / uz;
  float t = std::fmin(tx, std::fmin(ty, tz));
  x = (t == tx) ? bx : x + ux * t;
  y = (t == ty) ? by : y + uy * t;
  z = (t == tz) ? bz : z + uz * t;
  return t;
}

// photontrack_calc_specular_reflection (src/mc3d_gpu.cu:67-70).
inline float specular(float n0, float n1) {
  float r = (n0 - n1) / (n0 + n1);
  return r * r;
}

// photontrack_mirror (src/mc3d_gpu.cu:113-126).
inline bool mirror(float &x, float &ux, float L) {
  bool refl = (x >= L) || (x <= 0);
  x = (x >= L) ? 2 * L - x : x;
  x = (x <= 0) ? std::fabs(x) : x;
  ux = refl ? -1 * ux : ux;
  return refl;
}

// Fresnel reflectance of src/mc3d_gpu.cu:227-245.  mode 0 keeps the reference's R2 denominator
// (cos_ai*cos_at + sin_at*sin_at, :241); mode 1 uses cos(ai - at) = cos_ai*cos_at + sin_ai*sin_at.
// Outputs cos_at and ni/nt for the direction update.
inline float fresnel_R(float cos_ai_in, float ni, float nt, int mode, float *cos_at_out, float *ni_nt_out) {
  float cos_ai = std::fabs(cos_ai_in);
  cos_ai = (cos_ai > 1) ? 1 : cos_ai;
  float ai = std::acos(cos_ai);
  float ni_nt = ni / nt;
  float sin_ai = std::sin(ai);
  float sin_at = ni_nt * sin_ai;
  float cos_at = std::sqrt(1.0f - sin_at * sin_at);
  float R1 = (sin_ai * cos_at - cos_ai * sin_at) / (sin_ai * cos_at + cos_ai * sin_at);
  float R2 = (mode == YMC_FRESNEL_EXACT)
                 ? (cos_ai * cos_at - sin_ai * sin_at) / (cos_ai * cos_at + sin_ai * sin_at)
                 : (cos_ai * cos_at - sin_ai * sin_at) / (cos_ai * cos_at + sin_at * sin_at);
  // float R = 0.5*R1*R1*(1.0f + R2*R2): 0.5 is a double literal, so the product is double.
  double h = 0.5 * (double)R1;
  h = h * (double)R1;
  float q = 1.0f + R2 * R2;
  float R = (float)(h * (double)q);
  R = (sin_at > 1) ? 1 : R;  // critical angle (:244)
  R = (ni == nt) ? 0 : R;    // (:245)
  *cos_at_out = cos_at;
  *ni_nt_out = ni_nt;
  return R;
}

// Henyey-Greenstein cos(theta) from a (0,1] draw (src/mc3d_gpu.cu:282-283).
inline float hg_costheta(float g, float xi_oc) {
  float ct = (1 - g * g) / (1 - g + 2 * g * xi_oc);
  ct = 1.0f / (2 * g) * (1 + g * g - ct * ct);
  return ct;
}

// Fixed-point weight (scale 2^32, round to nearest even).  W*2^32 is exact in binary32.
inline uint64_t fx(float w) {
  float s = std::ldexp(w, 32);
  s = std::nearbyint(s);
  if (!(s > 0)) return 0;
  return (uint64_t)s;
}
inline uint64_t fx_sq(uint64_t f) { return (uint64_t)(((unsigned __int128)f * f + (1ull << 31)) >> 32); }

struct Props {
  const float *mua, *mus, *g, *n;  // one table, num_types rows
};

struct Photon {
  float x, y, z, ux, uy, uz, W, s;
  int i, j, k, t;
};

struct Acc {
  uint64_t A_w = 0, killed_w = 0, killed = 0, steps = 0, fresnel = 0;
  int64_t roulette_w = 0;
  uint64_t *abs = nullptr;  // absorption grid of the packet's table (optional)
};

// Launch of a new packet: photontrack_reinitialize (src/mc3d_gpu.cu:74-111) followed by the
// grid-coordinate and tissue lookup at the start of photontrack_step (:168-171).
inline void launch(Photon &p, uint64_t &rx, uint32_t ra, const Geom &G, const Props &P, bool uniform,
                   float px, float py) {
  if (uniform) {
    p.x = G.Lx * mwc_co(rx, ra);  // :88
    p.y = G.Ly * mwc_co(rx, ra);  // :89
  } else {
    p.x = px;  // :93
    p.y = py;  // :94
  }
  p.z = 0;
  float Rsp = specular(P.n[YMC_TISSUE_AIR], P.n[1]);  // n1 = tissue row 1 (:422, :500)
  p.ux = 0;
  p.uy = 0;
  p.uz = 1;
  p.W = 1.0f - Rsp;              // :101
  p.s = -std::log(mwc_oc(rx, ra));  // :102
  p.i = grid_coord(p.x, p.ux, G.dx);
  p.j = grid_coord(p.y, p.uy, G.dy);
  p.k = grid_coord(p.z, p.uz, G.dz);
  p.t = tissue_at(G, p.i, p.j, p.k, p.z, p.uz);
}

// Steps per launch of the reference (NUM_PHOTON_STEPS, src/mc3d_types.h:16).  A new packet always
// starts at the beginning of a launch (photontrack_reinitialize runs before photontrack_step,
// src/mc3d_gpu.cu:499-504), so the end-of-launch check of :343 sees the packet's steps 44, 89, ...
const int kLaunchSteps = 45;

// One iteration of the step loop (src/mc3d_gpu.cu:181-339) for a live packet, followed by the
// kill of :343-353: non-finite state is checked after every step (a NaN/inf state persists until
// the reference's check sees it), a zero step length (db == 0) where the reference checks it:
// kill_zero_step 1 = at the end of each 45-step launch (packet step index % 45 == 44),
// 2 = after every step, 0 = never.  `index` is the packet's 0-based step number.
// Two termination rules the reference does not have (it would not terminate in these cases):
// a negative step length (db < 0: the position has left the voxel its indices name, which
// happens when an interaction rounds onto a mirror wall and the index commit of :334-339 puts
// the packet outside the grid; the wall test of :114 then bounces it forever) and more than
// kMaxSteps steps.  Both are applied after every step and counted as kills.
// Returns true when the packet froze.
const uint64_t kMaxSteps = 1ull << 24;

bool step(Photon &p, uint64_t &rx, uint32_t ra, const Geom &G, const Props &P, int fresnel_mode,
          int kill_zero_step, uint64_t index, Acc &acc) {
  acc.steps++;
  float ni = P.n[p.t];
  float nt = ni;
  float mu_a = P.mua[p.t];
  float mu_s = P.mus[p.t];
  float g = P.g[p.t];
  float mu_t = mu_a + mu_s;

  float newx = p.x, newy = p.y, newz = p.z;
  float db = intersection(G, p.i, p.j, p.k, newx, p.ux, newy, p.uy, newz, p.uz);  // :194

  bool hit = db * mu_t <= p.s;          // :197
  float ds = hit ? db * mu_t : p.s;     // :198
  newx = hit ? newx : p.x + p.ux * ds / mu_t;  // :201-203
  newy = hit ? newy : p.y + p.uy * ds / mu_t;
  newz = hit ? newz : p.z + p.uz * ds / mu_t;
  p.s = p.s - ds;  // :207

  int nexti = grid_coord(newx, p.ux, G.dx);  // :209-211
  int nextj = grid_coord(newy, p.uy, G.dy);
  int nextk = grid_coord(newz, p.uz, G.dz);
  int next_t = tissue_at(G, nexti, nextj, nextk, newz, p.uz);  // :213
  nt = P.n[next_t];

  float trx = p.ux, tr_y = p.uy, trz = p.uz;
  bool refl = false;
  if (ni != nt) {  // :226-257
    acc.fresnel++;
    float cai = (p.i != nexti) ? p.ux : ((p.j != nextj) ? p.uy : p.uz);
    float cos_at, ni_nt;
    float R = fresnel_R(cai, ni, nt, fresnel_mode, &cos_at, &ni_nt);
    float eps = mwc_co(rx, ra);
    refl = eps < R;
    trx = refl ? p.ux * ((p.i == nexti) ? 1 : -1)
               : ((p.i != nexti) ? std::copysign(cos_at, p.ux * cos_at) : p.ux * ni_nt);
    tr_y = refl ? p.uy * ((p.j == nextj) ? 1 : -1)
                : (((p.j != nextj) && (p.i == nexti)) ? std::copysign(cos_at, p.uy * cos_at) : p.uy * ni_nt);
    trz = refl ? p.uz * ((p.k == nextk) ? 1 : -1)
               : (((p.k != nextk) && (p.j == nextj) && (p.i == nexti)) ? std::copysign(cos_at, p.uz * cos_at)
                                                                       : p.uz * ni_nt);
  }

  // Absorption (:262-263).
  float dW = (mu_t > 0) ? mu_a / mu_t * p.W : 0;
  float W0 = p.W;
  p.W = p.W - dW * (float)(!hit);
  acc.A_w += fx(W0) - fx(p.W);
  // Absorption grid (the disabled deposit of src/mc3d_gpu.cu:265-272): the absorbed weight goes
  // to the voxel the step started in, clamped to the grid like the texture lookup.
  if (acc.abs && !hit) {
    const int ic = std::min(std::max(p.i, 0), G.nx - 1);
    const int jc = std::min(std::max(p.j, 0), G.ny - 1);
    const int kc = std::min(std::max(p.k, 0), G.nz - 1);
    __atomic_fetch_add(&acc.abs[((size_t)kc * G.ny + jc) * G.nx + ic], fx(W0) - fx(p.W), __ATOMIC_RELAXED);
  }

  // Russian roulette (:275-277); SURVIVAL_CHANCE and W_THRESHOLD are double constants.
  float eps_r = mwc_co(rx, ra);
  float newW = ((double)eps_r < YMC_SURVIVAL_CHANCE) ? (float)((double)p.W / YMC_SURVIVAL_CHANCE) : 0;
  if (((double)p.W < YMC_W_THRESHOLD) && (!hit) && (mu_t != 0)) {
    acc.roulette_w += (int64_t)fx(newW) - (int64_t)fx(p.W);
    p.W = newW;
  }

  // Henyey-Greenstein scattering direction (:282-294).
  float costheta = hg_costheta(g, mwc_oc(rx, ra));
  float sintheta = std::sqrt(1 - costheta * costheta);
  float phi_divpi = 2 * mwc_co(rx, ra);
  float sinphi, cosphi;
  // sincospif(phi): sin and cos of pi*phi.
  sincosf(3.14159265358979323846f * phi_divpi, &sinphi, &cosphi);
  float ux = p.ux, uy = p.uy, uz = p.uz;
  float sx, sy, sz;
  if ((double)std::fabs(uz) <= YMC_UZ_THRESH) {
    float invsqrt = 1.0f / std::sqrt(1 - uz * uz);
    sx = sintheta * (ux * uz * cosphi - uy * sinphi) * invsqrt + ux * costheta;
    sy = sintheta * (uy * uz * cosphi + ux * sinphi) * invsqrt + uy * costheta;
    sz = -1 * std::sqrt(1 - uz * uz) * sintheta * cosphi + uz * costheta;
  } else {
    sx = sintheta * cosphi;
    sy = sintheta * sinphi;
    sz = std::copysign(costheta, costheta * uz);
  }

  // Scatter vs transmit/reflect direction (:297-299), small-component cleanup (:302-305),
  // renormalisation (:307-310).
  float nux = (!hit) ? sx : trx;
  float nuy = (!hit) ? sy : tr_y;
  float nuz = (!hit) ? sz : trz;
  const float zero_thresh = (float)(1 - YMC_UZ_THRESH);
  nux = (std::fabs(nux) < zero_thresh) ? 0 : nux;
  nuy = (std::fabs(nuy) < zero_thresh) ? 0 : nuy;
  nuz = (std::fabs(nuz) < zero_thresh) ? 0 : nuz;
  float nrm = 1.0f / std::sqrt(nux * nux + nuy * nuy + nuz * nuz);
  nux *= nrm;
  nuy *= nrm;
  nuz *= nrm;

  bool freeze = (p.W <= 0) || (mu_t == 0);  // :313
  if (!freeze) {
    p.ux = nux;
    p.uy = nuy;
    p.uz = nuz;
  }
  float eps_s = mwc_oc(rx, ra);  // :319-321
  float newS = (p.W > 0) ? -std::log(eps_s) : 0;
  p.s = (freeze || hit) ? p.s : newS;
  if (!freeze) {  // :324-326
    p.x = newx;
    p.y = newy;
    p.z = newz;
  }
  // :330 (C++ short-circuit: the y mirror is skipped once a reflection is recorded).
  refl = refl || mirror(p.x, p.ux, G.Lx) || mirror(p.y, p.uy, G.Ly);
  if (!((hit && refl) || freeze)) {  // :334-339
    p.i = nexti;
    p.j = nextj;
    p.k = nextk;
    p.t = next_t;
  }

  // :343-353
  bool zero_check = kill_zero_step == 2 || (kill_zero_step == 1 && index % kLaunchSteps == kLaunchSteps - 1);
  bool bad = !std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z) || !std::isfinite(p.s) ||
             !std::isfinite(p.ux) || !std::isfinite(p.uy) || !std::isfinite(p.uz) || (zero_check && db == 0) ||
             db < 0 || index + 1 >= kMaxSteps;
  if (bad) {
    acc.killed++;
    acc.killed_w += fx(p.W);
    freeze = true;
    p.x = p.y = p.z = 0;
    p.ux = p.uy = 0;
    p.uz = 1;
    p.W = 0;
    p.s = 0;
    return true;
  }
  return freeze;
}

struct Tallies {
  int tables, tnx, tny, nr;
  std::vector<uint64_t> Rw, Rw2, Tw, Tw2, Rrw, Rrw2;
  std::vector<uint32_t> Rn, Tn, Rrn;
  std::vector<ymc_totals> tot;
};

inline void add64(uint64_t *p, uint64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
inline void add64s(int64_t *p, int64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
inline void add32(uint32_t *p, uint32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

thread_local char g_err[256];

}  // namespace

// ============================================================================================
// Exported test entry points.

extern "C" {

const char *ymc_oracle_last_error(void) { return g_err; }

uint32_t ymc_oracle_mwc_step(uint64_t *x, uint32_t a) { return mwc_step(*x, a); }
float ymc_oracle_u32_to_float_rz(uint32_t u) { return u32_to_float_rz(u); }

// n_streams streams advanced n_draws times; out[d*n_streams+s] = co (kind 0) or oc (kind 1).
void ymc_oracle_mwc_fill(const uint64_t *x0, const uint32_t *a, uint32_t n_streams, uint32_t n_draws,
                         int32_t kind, float *out, uint64_t *x_final) {
  for (uint32_t s = 0; s < n_streams; ++s) {
    uint64_t x = x0[s];
    for (uint32_t d = 0; d < n_draws; ++d) {
      float v = kind ? mwc_oc(x, a[s]) : mwc_co(x, a[s]);
      out[(size_t)d * n_streams + s] = v;
    }
    if (x_final) x_final[s] = x;
  }
}

int32_t ymc_oracle_is_prime(uint64_t n) { return is_prime_u64(n) ? 1 : 0; }

// Defining property of a GPU-MCML multiplier: a*2^32-1 and a*2^31-1 both prime.
int32_t ymc_oracle_is_multiplier(uint32_t a) {
  uint64_t p1 = ((uint64_t)a << 32) - 1;
  uint64_t p2 = ((uint64_t)a << 31) - 1;
  return (is_prime_u64(p2) && is_prime_u64(p1)) ? 1 : 0;
}

// init_RNG (gpumcml_rng/gpumcml_rng.cu:59-118).  table[0] = begin (first line, :80); stream i
// reads table[1 + i % (table_len-1)] (:97; the reference repeats the last line past EOF, the
// engine cycles instead).  Returns 0 on success, 4 (YMC_ERR_SEED) for an invalid seed (:88).
int32_t ymc_oracle_init_rng(const uint32_t *table, uint32_t table_len, uint64_t xinit, uint64_t n,
                            uint64_t *x, uint32_t *a) {
  if (table_len < 2) return YMC_ERR_TABLE;
  uint32_t begin = table[0];
  if ((xinit == 0ull) | (((uint32_t)(xinit >> 32)) >= (begin - 1)) | (((uint32_t)xinit) >= 0xffffffffu))
    return YMC_ERR_SEED;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t fora = table[1 + i % (table_len - 1)];
    a[i] = fora;
    uint64_t xi = 0;
    while ((xi == 0) | (((uint32_t)(xi >> 32)) >= (fora - 1)) | (((uint32_t)xi) >= 0xffffffffu)) {
      xinit = (xinit & 0xffffffffull) * begin + (xinit >> 32);
      xi = (uint32_t)std::floor((((double)((uint32_t)xinit)) / (double)0x100000000ull) * fora);
      xi = xi << 32;
      xinit = (xinit & 0xffffffffull) * begin + (xinit >> 32);
      xi += (uint32_t)xinit;
    }
    x[i] = xi;
  }
  return 0;
}

int32_t ymc_oracle_grid_coord(float x, float ux, float dx) { return grid_coord(x, ux, dx); }

// Returns t; pos[3] is updated to the snapped position.
float ymc_oracle_intersection(int32_t i, int32_t j, int32_t k, float dx, float dy, float dz, float *pos,
                              const float *u) {
  Geom G{};
  G.dx = dx;
  G.dy = dy;
  G.dz = dz;
  return intersection(G, i, j, k, pos[0], u[0], pos[1], u[1], pos[2], u[2]);
}

float ymc_oracle_fresnel(float cos_ai, float ni, float nt, int32_t mode, float *cos_at) {
  float ca, r;
  float R = fresnel_R(cos_ai, ni, nt, mode, &ca, &r);
  if (cos_at) *cos_at = ca;
  return R;
}

float ymc_oracle_hg_costheta(float g, float xi_oc) { return hg_costheta(g, xi_oc); }
float ymc_oracle_specular(float n0, float n1) { return specular(n0, n1); }

// One step of photontrack_step (src/mc3d_gpu.cu:181-353) for one packet on a voxel grid with one
// optical-property table, exposed so that single events (a Fresnel crossing, an interaction) can
// be checked against physics on their own.  state[8] = x, y, z, ux, uy, uz, W, s and idx[4] =
// i, j, k, t (row) are read and written back; *rx is the packet's MWC state.  kill_zero_step and
// index as for the transport (index = the packet's 0-based step number).  Returns 1 when the
// packet froze (or was killed), 0 when it lives on; *fresnel receives 1 when the step took the
// Fresnel branch (:226-257).
int32_t ymc_oracle_step(const ymc_geometry *geom, const float *mua, const float *mus, const float *g,
                        const float *n, int32_t fresnel_mode, int32_t kill_zero_step, uint64_t index, float *state,
                        int32_t *idx, uint64_t *rx, uint32_t ra, int32_t *fresnel) {
  Geom G{};
  G.nx = geom->nx;
  G.ny = geom->ny;
  G.nz = geom->nz;
  G.dx = geom->dx;
  G.dy = geom->dy;
  G.dz = geom->dz;
  G.Lx = G.dx * (float)G.nx;
  G.Ly = G.dy * (float)G.ny;
  G.t = geom->tissue;
  Props P{mua, mus, g, n};
  Photon p{};
  p.x = state[0];
  p.y = state[1];
  p.z = state[2];
  p.ux = state[3];
  p.uy = state[4];
  p.uz = state[5];
  p.W = state[6];
  p.s = state[7];
  p.i = idx[0];
  p.j = idx[1];
  p.k = idx[2];
  p.t = idx[3];
  Acc acc;
  const bool froze = step(p, *rx, ra, G, P, fresnel_mode, kill_zero_step, index, acc);
  const float out[8] = {p.x, p.y, p.z, p.ux, p.uy, p.uz, p.W, p.s};
  std::memcpy(state, out, sizeof(out));
  idx[0] = p.i;
  idx[1] = p.j;
  idx[2] = p.k;
  idx[3] = p.t;
  if (fresnel) *fresnel = (int32_t)acc.fresnel;
  return froze ? 1 : 0;
}

// Guided chain schedule, restated from its definition: chain length
// ceil(remaining_total / (2*chain_lanes)), tables visited round-robin, each chain capped by
// what its table has left.
uint64_t ymc_oracle_build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes,
                                 ymc_chain *out, uint64_t cap) {
  std::vector<uint64_t> rem(budgets, budgets + num_tables);
  uint64_t total = 0;
  for (int w = 0; w < num_tables; ++w) total += rem[w];
  uint64_t n = 0;
  int w = 0;
  uint64_t denom = 2ull * (uint64_t)chain_lanes;
  while (total > 0) {
    while (rem[w] == 0) w = (w + 1) % num_tables;
    uint64_t chunk = (total + denom - 1) / denom;
    if (chunk > 0xffffffffull) chunk = 0xffffffffull;
    uint64_t c = std::min(chunk, rem[w]);
    if (out && n < cap) out[n] = ymc_chain{(uint32_t)c, (uint32_t)w};
    ++n;
    rem[w] -= c;
    total -= c;
    w = (w + 1) % num_tables;
  }
  return n;
}

// Full transport over explicit chains.  Chain c runs chains[c].count packets of table
// chains[c].table from stream (x0[c], a0[c]).  Tallies follow photon_detector
// (src/mc3d_photons.cu:151-171) on the tally grid; records (optional) are indexed by the global
// packet number (chains in order, packets in order).  threads <= 0 uses the OpenMP default.
// Same, with an optional absorption grid abs_grid [num_tables][nz][ny][nx] of fixed-point
// absorbed weight (zeroed here).
int32_t ymc_oracle_run_abs(const ymc_geometry *geom, const ymc_optprops *props, const ymc_params *prm,
                           const ymc_chain *chains, const uint64_t *x0, const uint32_t *a0, uint64_t n_chains,
                           ymc_tally *out, ymc_photon_record *records, uint64_t n_records, int32_t threads,
                           uint64_t *abs_grid);

int32_t ymc_oracle_run(const ymc_geometry *geom, const ymc_optprops *props, const ymc_params *prm,
                       const ymc_chain *chains, const uint64_t *x0, const uint32_t *a0, uint64_t n_chains,
                       ymc_tally *out, ymc_photon_record *records, uint64_t n_records, int32_t threads) {
  return ymc_oracle_run_abs(geom, props, prm, chains, x0, a0, n_chains, out, records, n_records, threads, nullptr);
}

int32_t ymc_oracle_run_abs(const ymc_geometry *geom, const ymc_optprops *props, const ymc_params *prm,
                           const ymc_chain *chains, const uint64_t *x0, const uint32_t *a0, uint64_t n_chains,
                           ymc_tally *out, ymc_photon_record *records, uint64_t n_records, int32_t threads,
                           uint64_t *abs_grid) {
  if (!geom || !props || !prm || !chains || !x0 || !a0 || !geom->tissue) {
    snprintf(g_err, sizeof g_err, "null argument");
    return YMC_ERR_ARG;
  }
  Geom G;
  G.nx = geom->nx;
  G.ny = geom->ny;
  G.nz = geom->nz;
  G.dx = geom->dx;
  G.dy = geom->dy;
  G.dz = geom->dz;
  G.Lx = G.dx * G.nx;
  G.Ly = G.dy * G.ny;
  G.t = geom->tissue;
  const int T = props->num_types;
  const int L = props->num_tables;
  for (size_t v = 0; v < (size_t)G.nx * G.ny * G.nz; ++v) {
    if (G.t[v] >= T) {
      snprintf(g_err, sizeof g_err, "voxel row %d >= num_types %d", (int)G.t[v], T);
      return YMC_ERR_FORMAT;
    }
  }
  Tallies tl;
  tl.tables = L;
  tl.tnx = prm->tally_nx > 0 ? prm->tally_nx : G.nx;
  tl.tny = prm->tally_ny > 0 ? prm->tally_ny : G.ny;
  const float tdx = prm->tally_dx > 0 ? prm->tally_dx : G.dx;
  const float tdy = prm->tally_dy > 0 ? prm->tally_dy : G.dy;
  tl.nr = prm->radial_bins > 0 ? prm->radial_bins : 0;
  size_t mapn = (size_t)L * tl.tnx * tl.tny;
  tl.Rw.assign(mapn, 0);
  tl.Rw2.assign(mapn, 0);
  tl.Tw.assign(mapn, 0);
  tl.Tw2.assign(mapn, 0);
  tl.Rn.assign(mapn, 0);
  tl.Tn.assign(mapn, 0);
  tl.Rrw.assign((size_t)L * tl.nr, 0);
  tl.Rrw2.assign((size_t)L * tl.nr, 0);
  tl.Rrn.assign((size_t)L * tl.nr, 0);
  tl.tot.assign(L, ymc_totals{});

  const bool uniform = prm->illumination == YMC_ILLUM_UNIFORM;
  const float px = prm->pencil_x >= 0 ? prm->pencil_x : G.Lx / 2;
  const float py = prm->pencil_y >= 0 ? prm->pencil_y : G.Ly / 2;
  const float rx0 = prm->radial_x0 >= 0 ? prm->radial_x0 : px;
  const float ry0 = prm->radial_y0 >= 0 ? prm->radial_y0 : py;
  const float rdr = prm->radial_dr;
  const int kill0 = prm->kill_zero_step;
  const int fmode = prm->fresnel_mode;

  std::vector<uint64_t> first(n_chains + 1, 0);
  for (uint64_t c = 0; c < n_chains; ++c) first[c + 1] = first[c] + chains[c].count;
  for (uint64_t c = 0; c < n_chains; ++c) {
    if (chains[c].table >= (uint32_t)L) {
      snprintf(g_err, sizeof g_err, "chain table out of range");
      return YMC_ERR_ARG;
    }
  }

  if (abs_grid) std::memset(abs_grid, 0, sizeof(uint64_t) * L * G.nx * G.ny * G.nz);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  volatile const int32_t *cancel = prm->cancel;
  int stopped = 0;

#pragma omp parallel for schedule(dynamic, 16) reduction(| : stopped)
  for (int64_t c = 0; c < (int64_t)n_chains; ++c) {
    const uint32_t w = chains[c].table;
    Props P{props->mua + (size_t)w * T, props->mus + (size_t)w * T, props->g + (size_t)w * T,
            props->n + (size_t)w * T};
    uint64_t rx = x0[c];
    const uint32_t ra = a0[c];
    Acc acc;
    if (abs_grid) acc.abs = abs_grid + (size_t)w * G.nx * G.ny * G.nz;
    uint64_t done = 0;
    for (uint32_t q = 0; q < chains[c].count; ++q) {
      if (cancel && *cancel) {
        stopped = 1;
        break;
      }
      Photon p;
      launch(p, rx, ra, G, P, uniform, px, py);
      uint64_t s0 = acc.steps;
      uint64_t k0 = acc.killed;
      for (uint64_t index = 0; !step(p, rx, ra, G, P, fmode, kill0, index, acc); ++index) {
      }
      ++done;
      bool killed = acc.killed != k0;
      uint64_t f = fx(p.W);
      uint64_t f2 = fx_sq(f);
      ymc_totals &tt = tl.tot[w];
      if (p.z <= 0) {
        if (f) {
          add64(&tt.R_w, f);
          add64(&tt.R_w2, f2);
          add64(&tt.R_n, 1);
          int ix = grid_coord(p.x, p.ux, tdx);
          int iy = grid_coord(p.y, p.uy, tdy);
          if (ix >= 0 && iy >= 0 && ix < tl.tnx && iy < tl.tny) {
            size_t b = ((size_t)w * tl.tny + iy) * tl.tnx + ix;
            add64(&tl.Rw[b], f);
            add64(&tl.Rw2[b], f2);
            add32(&tl.Rn[b], 1);
          }
          if (tl.nr > 0) {
            float ddx = p.x - rx0, ddy = p.y - ry0;
            float r = std::sqrt(ddx * ddx + ddy * ddy);
            int rb = floor_to_int(r / rdr);
            if (rb >= 0 && rb < tl.nr) {
              size_t b = (size_t)w * tl.nr + rb;
              add64(&tl.Rrw[b], f);
              add64(&tl.Rrw2[b], f2);
              add32(&tl.Rrn[b], 1);
            }
          }
        }
      } else if (f) {
        add64(&tt.T_w, f);
        add64(&tt.T_w2, f2);
        add64(&tt.T_n, 1);
        int ix = grid_coord(p.x, p.ux, tdx);
        int iy = grid_coord(p.y, p.uy, tdy);
        if (ix >= 0 && iy >= 0 && ix < tl.tnx && iy < tl.tny) {
          size_t b = ((size_t)w * tl.tny + iy) * tl.tnx + ix;
          add64(&tl.Tw[b], f);
          add64(&tl.Tw2[b], f2);
          add32(&tl.Tn[b], 1);
        }
      }
      uint64_t gi = first[c] + q;
      if (records && gi < n_records) {
        ymc_photon_record &r = records[gi];
        r.x = p.x;
        r.y = p.y;
        r.z = p.z;
        r.w = p.W;
        r.steps = (uint32_t)(acc.steps - s0);
        r.fate = killed ? 2 : (p.z <= 0 ? 0 : 1);
      }
    }
    ymc_totals &tt = tl.tot[w];
    add64(&tt.photons, done);
    add64(&tt.A_w, acc.A_w);
    add64s(&tt.roulette_w, acc.roulette_w);
    add64(&tt.killed_w, acc.killed_w);
    add64(&tt.killed, acc.killed);
    add64(&tt.steps, acc.steps);
    add64(&tt.fresnel, acc.fresnel);
  }

  if (out) {
    auto cp64 = [](uint64_t *dst, const std::vector<uint64_t> &src) {
      if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * 8);
    };
    auto cp32 = [](uint32_t *dst, const std::vector<uint32_t> &src) {
      if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * 4);
    };
    cp64(out->R_w, tl.Rw);
    cp64(out->R_w2, tl.Rw2);
    cp32(out->R_n, tl.Rn);
    cp64(out->T_w, tl.Tw);
    cp64(out->T_w2, tl.Tw2);
    cp32(out->T_n, tl.Tn);
    cp64(out->Rr_w, tl.Rrw);
    cp64(out->Rr_w2, tl.Rrw2);
    cp32(out->Rr_n, tl.Rrn);
    if (out->totals) std::memcpy(out->totals, tl.tot.data(), sizeof(ymc_totals) * L);
  }
  return stopped ? YMC_ERR_CANCELLED : YMC_OK;
}

int32_t ymc_oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
