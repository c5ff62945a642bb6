// transport.hip - persistent photon-transport kernel for gfx950 and the device engine.
//
// One lane carries one photon packet at a time.  Lanes take chains of packets from a global
// counter (a chain owns one multiply-with-carry stream, so a packet's random numbers do not
// depend on which lane runs it or when), run each packet step by step until it freezes, tally
// it with 64-bit fixed-point atomics and start the next packet of the chain.  The kernel stays
// resident until every chain is done.
//
// Per step the lane follows photontrack_step (src/mc3d_gpu.cu:181-339) with the reference's
// draw order (optional Fresnel draw, roulette, HG cos, HG phi, new step length), so a packet's
// trajectory matches the CPU oracle draw for draw.  What differs is how the arithmetic is
// issued on CDNA4:
//  - optical properties of every (table, tissue) row sit in LDS with derived constants
//    (mu_t, mu_a/mu_t, 1/mu_t, HG terms) precomputed on the host;
//  - the voxel grid is one 16-bit word per voxel in global memory (read through L2: the row and
//    the safe box below; 32-bit words for grids with 8 or more rows), clamped like the
//    reference's texture (src/mc3d_gpu.cu:416-418);
//  - the boundary distance uses cached reciprocals of the direction cosines, updated only
//    when the direction changes;
//  - the next voxel index comes from the face test of the DDA instead of three floor divisions
//    (the divisions return only for the rare packet whose x or y index lies outside the grid);
//  - the HG, roulette and step-length math runs only for lanes that interact, behind a
//    wave-uniform branch; lanes that only cross a voxel face still advance the stream by the
//    four draws the reference consumes on every step;
//  - plain voxel crossings inside a region of one tissue are taken together: each voxel word
//    carries a safe box (vertical and lateral half-extents of a box of voxels that all share its
//    tissue and lie inside the lateral walls and below the top surface), and a step takes the
//    safe box of its anchor voxel as its region: the flight runs to the box's faces or to the
//    interaction point, the crossings inside the box cost no voxel load and no Fresnel test, the
//    stream is advanced by their draws, and the step ends as the reference's step from the last
//    voxel of the flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/ymc3d.h"
#include "host_util.h"

namespace ymc {

// ------------------------------------------------------------------------------------------
// Device helpers.

// MWC step (gpumcml_rng/gpumcml_rng.cu:40) as one v_mad_u64_u32.
__device__ __forceinline__ uint32_t mwc_next(uint64_t &x, uint32_t a) {
  // Written as the instruction itself: in some inlined contexts the compiler otherwise loses
  // track of the zero upper half of the multiplier and issues a second 64-bit multiply-add.
  uint64_t r;
  asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(r) : "v"((uint32_t)x), "v"(a), "v"(x >> 32) : "vcc");
  x = r;
  return (uint32_t)x;
}

// Round-toward-zero u32 -> f32, scaled by 2^-32: rand_MWC_co (gpumcml_rng.cu:41).  Masking the
// bits below the 24-bit significand makes the conversion exact; the scale is a power of two.
// With lz leading zeros, the mask 0xffffff00 >> lz (arithmetic shift, amount mod 32) clears the
// 8 - lz low bits for lz < 8 and nothing otherwise.  v_ffbh_u32 gives lz, and 0xffffffff for
// u == 0, which shifts the mask to all ones: three instructions instead of __clz's clamp to 32
// and a variable shift pair.
__device__ __forceinline__ float u32_co(uint32_t u) {
  uint32_t lz;
  asm("v_ffbh_u32 %0, %1" : "=v"(lz) : "v"(u));
  const uint32_t m = u & (uint32_t)((int32_t)0xffffff00 >> (lz & 31u));
  return (float)m * 0x1p-32f;
}
__device__ __forceinline__ float u32_oc(uint32_t u) { return 1.0f - u32_co(u); }

// Product rounded on its own.  Device code is compiled with contraction allowed, and HIP's
// __fmul_rn is a plain operator, so a product feeding a subtraction can become one fma; the pragma
// keeps this one a separate, rounded multiply.
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}

// Stream skip.  An MWC step maps the state v = carry*2^32 + x (0 < v < p, p = a*2^32 - 1) to
// a*x + carry, which is v*2^-32 mod p (a*2^32 = 1 mod p).  k steps multiply v by 2^-32k mod p, so
// the draws of n skipped plain crossings (4 each) take one modular product per base-4 digit of n;
// the lowest digit (at most 3 crossings, 12 steps) is cheaper stepped.
// The product is Montgomery's with R = 2^64: p = -1 mod 2^32, so reducing by one 32-bit limb,
// z -> (z + (z mod 2^32)*p) / 2^32 = a*(z mod 2^32) + z/2^32, is an MWC step itself.
// mwc_mulmod returns v*t*2^-64 mod p for 0 <= v, t < p.
__device__ __forceinline__ uint64_t mwc_mulmod(uint64_t v, uint64_t t, uint32_t a) {
  const uint64_t p = ((uint64_t)a << 32) - 1;
  const uint32_t v0 = (uint32_t)v, v1 = (uint32_t)(v >> 32), t0 = (uint32_t)t, t1 = (uint32_t)(t >> 32);
  // z = v*t in four 32-bit limbs: z0, z1, z23
  const uint64_t A = (uint64_t)v0 * t0;
  const uint64_t B = (uint64_t)v0 * t1 + (A >> 32);
  const uint64_t C = (uint64_t)v1 * t0 + (uint32_t)B;
  const uint64_t D = (uint64_t)v1 * t1 + (B >> 32);
  const uint32_t z0 = (uint32_t)A, z1 = (uint32_t)C;
  const uint64_t z23 = D + (C >> 32);
  // first limb reduction: w = a*z0 + z/2^32 (96 bits: w_hi:w_lo)
  const uint64_t s1 = (z23 << 32) | z1;
  const uint64_t w_lo = (uint64_t)a * z0 + s1;
  const uint32_t w_hi = (uint32_t)(z23 >> 32) + (w_lo < s1 ? 1u : 0u);
  // second: u = a*(w mod 2^32) + w/2^32 (65 bits: c:u).  Bound: w < p^2/2^32 + p + 1, and
  // a*(w mod 2^32) <= p + 1 - a, so u < p^2/2^64 + (p + 1)/2^32 + p + 1 - a; with p^2/2^64 < p - p/2^32
  // (a < 2^32) that is u < 2p + 1 + 2^-32 - a < 2p, since every multiplier has a >= 2.  One
  // conditional subtraction makes the result canonical (the 65th bit set means u >= 2^64 > p).
  const uint64_t s2 = ((uint64_t)w_hi << 32) | (w_lo >> 32);
  const uint64_t u = (uint64_t)a * (uint32_t)w_lo + s2;
  const bool c = u < s2;
  return (c || u >= p) ? u - p : u;
}

// Per-lane skip table: entry k - 1 (k = 1..7) holds E(4k) and entry 7 holds E(32), where E(n) =
// 2^-128n * 2^64 mod p skips the draws of n crossings (4n steps) in one mwc_mulmod.  E(4) is the
// state 14 MWC steps after v = 1 (2^-32*14 = 2^-128*4 * 2^64); the others follow from
// mwc_mulmod(E(a), E(b)) = E(a + b).  A skip of n = d + 4k + 32m crossings (d, k, m its digits in
// bases 4, 8, 2) steps d crossings and takes one product for k and one for m: flights of fewer than 32
// crossings, nearly all of them, need one product.  tab points at the lane's first entry; entries are
// STRIDE apart (one column per lane of the workgroup).
constexpr int kSkipEntries = 8;
constexpr int kSkipMax = 63;  // crossings one skip covers
constexpr int kMaxAxis = 1 << 24;  // voxels per grid dimension, below which indices convert to float exactly
constexpr float kMaxVesselRadius = 16384.0f;  // voxels; bounds the rounding of the box's cylinder gap
template <int STRIDE>
__device__ __forceinline__ void build_skip(uint64_t *tab, uint32_t a) {
  uint64_t e4 = 1;
  for (int k = 0; k < 14; ++k) mwc_next(e4, a);
  const uint64_t e8 = mwc_mulmod(e4, e4, a);
  const uint64_t e12 = mwc_mulmod(e8, e4, a);
  const uint64_t e16 = mwc_mulmod(e8, e8, a);
  const uint64_t e20 = mwc_mulmod(e16, e4, a);
  const uint64_t e24 = mwc_mulmod(e12, e12, a);
  const uint64_t e28 = mwc_mulmod(e24, e4, a);
  const uint64_t e32 = mwc_mulmod(e16, e16, a);
  tab[0] = e4;
  tab[STRIDE] = e8;
  tab[2 * STRIDE] = e12;
  tab[3 * STRIDE] = e16;
  tab[4 * STRIDE] = e20;
  tab[5 * STRIDE] = e24;
  tab[6 * STRIDE] = e28;
  tab[7 * STRIDE] = e32;
}
// Advances the stream by the draws of n plain crossings, 0 <= n <= kSkipMax.
template <int STRIDE>
__device__ __forceinline__ void skip_crossings(uint64_t &rx, uint32_t ra, const uint64_t *tab, int n) {
  const int d = n & 3, k = (n >> 2) & 7, m = n >> 5;
  for (int q = 0; q < d; ++q) {
    mwc_next(rx, ra);
    mwc_next(rx, ra);
    mwc_next(rx, ra);
    mwc_next(rx, ra);
  }
  if (k) rx = mwc_mulmod(rx, tab[(k - 1) * STRIDE], ra);
  if (m) rx = mwc_mulmod(rx, tab[7 * STRIDE], ra);
}

__device__ __forceinline__ float rcp(float v) { return __builtin_amdgcn_rcpf(v); }
// -ln(v) for v in [2^-32, 1]: v_log_f32 (log2, no denormal range in this domain) times ln 2.
__device__ __forceinline__ float neg_ln(float v) { return __builtin_amdgcn_logf(v) * -0.693147180559945309f; }

// Fixed-point weight, scale 2^32, round to nearest even (0 <= w < 2^32 after scaling).
__device__ __forceinline__ uint64_t fx(float w) {
  float s = rintf(w * 0x1p32f);
  return s > 0.0f ? (uint64_t)s : 0ull;
}
// (f^2 + 2^31) >> 32 for f < 2^33 (weights below 2; packet weights never exceed 1).  With f =
// fh 2^32 + fl, fh <= 1: (f^2 + 2^31) >> 32 = fh (2^32 + 2 fl) + ((fl^2 + 2^31) >> 32) exactly, since
// the other terms of f^2 are multiples of 2^32; fl^2 + 2^31 < 2^64 is one 32x32 -> 64-bit
// multiply-add instead of a 128-bit product.
__device__ __forceinline__ uint64_t fx_sq(uint64_t f) {
  const uint32_t fl = (uint32_t)f;
  const uint64_t low = ((uint64_t)fl * fl + (1ull << 31)) >> 32;
  return (f >> 32) ? low + (1ull << 32) + 2ull * fl : low;
}

// min(max(v, lo), hi) for lo <= hi as one v_med3_i32: the compiler issues a max and a min unless it
// can prove lo <= hi.  Callers guarantee the order.
__device__ __forceinline__ int clamp_vv(int v, int lo, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
  return r;
}
// The same against [0, hi] with a wave-uniform hi >= 0.
__device__ __forceinline__ int clamp_0s(int v, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "s"(hi));
  return r;
}

// photontrack_get_grid_coord (src/mc3d_gpu.cu:53-63), IEEE division; used at launch and tally.
__device__ __forceinline__ int grid_coord(float x, float ux, float dx) {
  float q = floorf(x / dx);
  q = fminf(fmaxf(q, -1073741824.0f), 1073741824.0f);
  int i = (q == q) ? (int)q : 0;
  i += (x == (float)(i + 1) * dx);
  float bmin = (float)i * dx;
  i = ((x > bmin) || ((x == bmin) && (ux >= 0))) ? i : i - 1;
  return i;
}

constexpr int kBlock = 256;  // threads per workgroup of the transport kernel

// Per-row constants in LDS: three float4 per (table, tissue) row.
//   r0 = {mua, mus, g, n}
//   r1 = {mu_t, mua/mu_t (0 if mu_t == 0), 1/mu_t, 1/(2g)}
//   r2 = {1 - g*g, 1 - g, 2*g, 1 + g*g}
struct RowConst {
  float4 r0, r1, r2;
};

// Row idx of a RowConst table, addressed with a 24-bit multiply (full rate, where a 32-bit integer
// multiply issues at a quarter of it); row indices stay below 2^24 (checked when the tables are set).
__device__ __forceinline__ const RowConst &row_at(const RowConst *rows, uint32_t idx) {
  return *reinterpret_cast<const RowConst *>(reinterpret_cast<const char *>(rows) +
                                             __umul24(idx, (uint32_t)sizeof(RowConst)));
}

// Copies a row of constants (three float4) into a lane's LDS slot.
// The lane's current row of constants in LDS (parametric runs).  The step reads the refractive
// index, r1 and r2 of it, kept in three workgroup-wide arrays: 36 B per lane.
struct LaneRow {
  float4 *r1, *r2;
  float *n;
};
__device__ __forceinline__ void set_lane_row(const LaneRow &d, const RowConst &src) {
  const float n = src.r0.w;
  const float4 r1 = src.r1, r2 = src.r2;
  *d.n = n;
  *d.r1 = r1;
  *d.r2 = r2;
}

struct KArgs {
  // geometry: one word per voxel holding its row and safe box (extents zup, zdn, rxy).  Grids with
  // fewer than 8 rows: 16-bit words, row in bits 0-2, zup in 3-6, zdn in 7-10, rxy in 11-15
  // (vox16); otherwise 32-bit words, one byte each for row, zup, zdn, rxy (vox32).  One of the
  // two is null.
  const uint16_t *vox16;
  const uint32_t *vox32;
  // Field layout of the words the kernel reads (16-bit grid words as above; 32-bit grid words and
  // parametric-geometry words: bytes): row = w & w_row, zup and zdn are w_zbits wide at bits
  // w_zup and w_zdn, rxy = w >> w_rxy.
  uint32_t w_row, w_zup, w_zdn, w_zbits, w_rxy;
  int nx, ny, nz;
  float dx, dy, dz, Lx, Ly;
  float idx, idy, idz;  // 1/dx, 1/dy, 1/dz
  // parametric geometry per table (layered runs only)
  const ymc_layered *layered;
  // optical properties
  const RowConst *rows;  // [num_tables * num_types]
  const float *w0;       // [num_tables] launch weight 1 - Rsp
  int num_types, num_tables;
  // physics / illumination
  float px, py;
  float uz_thresh;  // largest float <= 0.99999 (the reference compares in double)
  float zero_thresh;
  float w_thresh;   // smallest float >= 1e-4 (W < 1e-4 in double <=> W < w_thresh in float)
  int fresnel_exact;
  int kill0;
  // tally
  int tnx, tny;
  float tdx, tdy;
  int nr;
  float rdr, rx0, ry0;
  // Reflectance and transmittance maps share one allocation per quantity: [2][tables][tny][tnx],
  // R in the first half and T in the second, so a packet's destination is an index offset.
  unsigned long long *map_w, *map_w2, *Rr_w, *Rr_w2;
  unsigned int *map_n, *Rr_n;
  unsigned long long map_half;  // tables * tny * tnx
  unsigned long long *tot;  // [num_tables][kTotFields]
  // chains
  const ymc_chain *chains;
  const uint64_t *chain_x;
  const uint32_t *chain_a;
  const uint64_t *chain_first;  // packet index of each chain's first packet (records only)
  unsigned long long *next_chain;
  unsigned long long n_chains;
  const volatile int *cancel;
  int poll_interval;
  int jump_cap;  // most plain crossings one box flight may take (at most kSkipMax with a skip table);
                 // -1 with box flights off, so that every box flight is redone in the voxel
  int skip_off;  // byte offset of the per-lane stream-skip table in dynamic LDS, or -1 (then the
                 // skipped draws are taken one by one)
  int relaunch_lanes;  // the launch phase runs when this many lanes of a wave wait for it,
  int relaunch_wait;   // or after this many iterations with a waiting lane
  ymc_photon_record *records;
  unsigned long long n_records;
  // absorption grid [tables][nz][ny][nx] (YMC_FLAG_ABSORPTION) or null
  unsigned long long *abs_grid;
  uint32_t nvox;
};

// Offsets into the per-table totals block (ymc_totals field order).
enum TotField {
  kPhotons = 0, kRw, kRw2, kTw, kTw2, kRn, kTn, kAw, kRoul, kKilledW, kKilled, kSteps, kFresnel, kTotFields
};
static_assert(sizeof(ymc_totals) == kTotFields * 8, "ymc_totals layout");

__device__ __forceinline__ void atomic_add_u64(unsigned long long *p, unsigned long long v) { atomicAdd(p, v); }

struct Packet {
  float x, y, z, ux, uy, uz, W, s;
  float iux, iuy, iuz;
  int i, j, k, t;
  uint32_t nstep;  // steps taken by this packet
  // Safe box of the anchor voxel (ai, aj, ak): [ai-rxy, ai+rxy] x [aj-rxy, aj+rxy] x
  // [ak-zup, ak+zdn], stored as oi = rxy - ai, oj = rxy - aj, ok = zup - ak and the extents
  // rxy2 = 2 rxy, rz2 = zup + zdn, so that voxel (i, j, k) lies in the box iff
  // (unsigned)(i + oi) <= rxy2, (unsigned)(j + oj) <= rxy2 and (unsigned)(k + ok) <= rz2.  No box:
  // oi = -2^30 (kNoCube), which no index passes (extents <= 510).  The two extents share one word
  // (rxy2 | rz2 << 16): as two fields the compiler kept them in a private array in scratch.
  int oi, oj, ok;
  uint32_t ext;
};

constexpr int kNoCube = -(1 << 30);

// Voxel word fields, in the layout the kernel arguments name (bit-field extracts, no repacking).
__device__ __forceinline__ int word_row(const KArgs &a, uint32_t w) { return (int)(w & a.w_row); }

// Voxel (i, j, k) with word w becomes the anchor; with box false the packet has no box (an x offset
// that no index passes, whatever the extents hold).  Every field is written on both paths, with
// selected values: a branch that wrote some fields on one side only led the compiler to keep the
// box fields in scratch.
__device__ __forceinline__ void set_anchor(Packet &p, const KArgs &a, int i, int j, int k, uint32_t w,
                                           bool box = true) {
  const int zup = (int)__builtin_amdgcn_ubfe(w, a.w_zup, a.w_zbits);
  const int zdn = (int)__builtin_amdgcn_ubfe(w, a.w_zdn, a.w_zbits);
  const int rxy = (int)(w >> a.w_rxy);
  p.oi = box ? rxy - i : kNoCube;  // the sentinel in oi alone marks "no box" (has_box)
  p.oj = rxy - j;
  p.ok = zup - k;
  p.ext = (uint32_t)(2 * rxy) | ((uint32_t)(zup + zdn) << 16);
}
// The packet's voxel lies in the anchor's box whenever the anchor has one: every commit of an index
// either makes the new voxel the anchor (launch, the end of a step) or stays inside the box (the
// crossings of a box flight), and a step that commits no index keeps both.  So "in the box" is "has a
// box".
__device__ __forceinline__ bool has_box(const Packet &p) { return p.oi != kNoCube; }

// Steps per launch of the reference (NUM_PHOTON_STEPS, src/mc3d_types.h:16).  Packets start at a
// launch boundary (src/mc3d_gpu.cu:499-504), so its end-of-launch zero-step check (:343) sees a
// packet's steps 44, 89, ...
constexpr uint32_t kLaunchSteps = 45;
// Step cap per packet (a packet that reaches it is killed; see the kill rule in step_packet).
constexpr uint32_t kMaxSteps = 1u << 24;

// Grids are limited to 2^31 voxels at upload, so the clamped index fits 32-bit arithmetic.
__device__ __forceinline__ uint32_t voxel_word(const KArgs &a, int i, int j, int k) {
  const uint32_t ic = (uint32_t)clamp_0s(i, a.nx - 1);
  const uint32_t jc = (uint32_t)clamp_0s(j, a.ny - 1);
  const uint32_t kc = (uint32_t)clamp_0s(k, a.nz - 1);
  const uint32_t idx = (kc * (uint32_t)a.ny + jc) * (uint32_t)a.nx + ic;
  return a.vox16 ? (uint32_t)a.vox16[idx] : a.vox32[idx];
}

// Index after a move: an axis steps when the new coordinate reached the face ahead (the face at
// b, in the direction of travel).
__device__ __forceinline__ int next_index(int i, float u, float n, float b) {
  const bool up = u >= 0.0f;
  return i + (int)(up & (n >= b)) - (int)(!up & (n <= b));
}

// Voxel word of a parametric geometry (ymc_layered): the row of voxel (i, j, k) of the grid the
// parameters describe (same clamp as the texture), and a safe cube as the box, bounded from the
// distances to the layer interface, the cylinder surface and the walls.  The cylinder test is
// evaluated with every binary32 operation rounded, exactly as the host voxeliser does.  (Boxes
// reaching further sideways, beside or above the cylinder, measured no faster on these
// geometries, whose long walks stay deep in the bottom layer where the cube is already large.)
__device__ __forceinline__ uint32_t layered_word(const KArgs &a, const ymc_layered &g, int i, int j, int k) {
  const int ic = clamp_0s(i, a.nx - 1);
  const int jc = clamp_0s(j, a.ny - 1);
  const int kc = clamp_0s(k, a.nz - 1);
  const bool top = kc < g.top_voxels;
  int row = top ? g.row_top : g.row_bottom;
  int r = top ? min(kc, g.top_voxels - 1 - kc) : kc - g.top_voxels;
  if (g.row_vessel >= 0) {
    // every operation rounded on its own, as on the host (no contraction into fma)
#pragma clang fp contract(off)
    const float fx = ((float)ic + 0.5f) - g.vessel_x;
    const float fz = ((float)kc + 0.5f) - g.vessel_z;
    const float d2 = fx * fx + fz * fz;
    if (d2 <= g.vessel_r * g.vessel_r) row = g.row_vessel;
    // voxels within Chebyshev distance r have centres within r*sqrt(2) in the xz plane.  The root
    // need not be correctly rounded: extents are capped at kSafeMax = 31, so the cylinder limits a
    // box only where gap < 44, i.e. d < vessel_r + 45 < 2^15 (radii are held below 2^14 voxels),
    // and there the roundings of d2 and the 1 ulp of v_sqrt_f32 together stay under 2^-22 d, below
    // 0.008 voxels: inside the 0.01-voxel margin.  The rows (the test above) use no root.
    const float gap = fabsf(__builtin_amdgcn_sqrtf(d2) - g.vessel_r) - 0.01f;
    r = min(r, gap > 0.0f ? (int)(gap * 0.70710678f) : 0);
  }
  r = min(r, min(min(ic, a.nx - 1 - ic), min(jc, a.ny - 1 - jc)));
  r = min(max(r, 0), kSafeMax);
  return (uint32_t)row | ((uint32_t)r << 8) | ((uint32_t)r << 16) | ((uint32_t)r << 24);
}

template <bool LAYERED>
__device__ __forceinline__ uint32_t geometry_word(const KArgs &a, const ymc_layered *g, int i, int j, int k) {
  if (LAYERED) return layered_word(a, *g, i, j, k);
  return voxel_word(a, i, j, k);
}

struct LaneAcc {
  uint32_t steps, fresnel, photons;
  uint64_t absorbed;
};

template <bool UNIFORM, bool LAYERED>
__device__ __forceinline__ void launch_packet(Packet &p, uint64_t &rx, uint32_t ra, const KArgs &a, float w0,
                                              const ymc_layered *lg) {
  if (UNIFORM) {
    p.x = a.Lx * u32_co(mwc_next(rx, ra));  // src/mc3d_gpu.cu:88
    p.y = a.Ly * u32_co(mwc_next(rx, ra));  // :89
  } else {
    p.x = a.px;
    p.y = a.py;
  }
  p.z = 0.0f;
  p.ux = 0.0f;
  p.uy = 0.0f;
  p.uz = 1.0f;
  p.iux = INFINITY;
  p.iuy = INFINITY;
  p.iuz = 1.0f;
  p.W = w0;                                        // :101
  p.s = neg_ln(u32_oc(mwc_next(rx, ra)));         // :102
  p.i = grid_coord(p.x, 0.0f, a.dx);
  p.j = grid_coord(p.y, 0.0f, a.dy);
  p.k = 0;                                         // grid_coord(0, +1) = 0
  const uint32_t w = geometry_word<LAYERED>(a, lg, p.i, p.j, 0);  // uz > 0: never the air override
  p.t = word_row(a, w);
  p.nstep = 0;
  set_anchor(p, a, p.i, p.j, 0, w);
}

// Voxel of coordinate c along one axis, clamped to [lo, hi].  Rounding can put it one voxel off
// next to a face; the step below checks the end point of a flight against the voxel this names, so
// any voxel of the box is safe here.  The clamp is in integers (the conversion saturates; NaN gives 0).
__device__ __forceinline__ int axis_voxel(float c, float inv_d, int lo, int hi) {
  // floor and conversion in one instruction (the compiler issues v_floor_f32 + v_cvt_i32_f32)
  int q;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(q) : "v"(c * inv_d));
  return clamp_vv(q, lo, hi);
}

// Outcome of one step.
enum StepResult : int { kLive = 0, kFrozen = 1, kDead = 2 };

// One step of a live packet, with the plain crossings of its flight before it.
//
// The step's region is the safe box of the anchor voxel or the packet's voxel.  Inside the box every
// voxel has the anchor's tissue, no mirror wall and no top surface, so each face the straight path
// crosses there is a plain crossing of the reference's step: hit, no Fresnel draw, new indices,
// s -= db*mu_t, four draws.  The flight runs to the region's faces ahead or to the interaction
// point, whichever comes first (photontrack_intersection and the hop decision, src/mc3d_gpu.cu:29-49,
// 196-207, on the region's faces); the plain crossings on the way are counted, the stream is advanced
// by their draws (with the lane's skip table, up to two modular products, when the kernel has one in
// LDS), and the step ends as the reference's step from the last voxel of the flight: the crossing out
// of the box (next voxel, Fresnel, ...) or the interaction.  With the voxel as the region this is the
// reference's step.
//   Positions are the reference's up to rounding (the end point is one fma from the flight's start
// instead of one per crossing) and the draws are the same, so packets stay in lockstep with
// step-by-step transport.  The region is the voxel when the packet has no box, sits on a face ahead
// (the reference takes that zero-length step on its own, subject to the zero-step kill) or the
// medium is clear, and a box flight is redone in the voxel when it would cross more than jump_cap
// faces or when its end point does not lie strictly inside the last voxel across every axis it does
// not leave the box by (a face met exactly at the end would be a step of its own in the reference).
// A zero-length crossing inside a flight (a crossing's end position rounding onto the next face) is
// taken like the other crossings, without the zero-step kill.  A packet frozen at the end of a flight
// (roulette) keeps the flight's start as its position, where the reference keeps the last face it
// crossed; its weight is zero, so no tally depends on it.
template <bool COUNTERS, bool LAYERED, bool RECORD>
__device__ __forceinline__ int step_packet(Packet &p, uint64_t &rx, uint32_t ra, const KArgs &a,
                                            const RowConst *rows, uint32_t rbase, unsigned long long *tot,
                                            LaneAcc &acc, const ymc_layered *lg, uint32_t table, float4 nrow,
                                            const uint64_t *skip, LaneRow cur) {
  // this table's rows start at rows[rbase]; parametric runs keep the packet's row in the lane's LDS
  // slot `cur` (their tables stay in global memory)
  float ni;
  float4 c1;
  if (LAYERED) {
    ni = *cur.n;
    c1 = *cur.r1;
  } else {
    const RowConst &row = row_at(rows, rbase + p.t);
    ni = row.r0.w;
    c1 = row.r1;
  }
  const float mu_t = c1.x;
  const bool upx = p.ux >= 0.0f, upy = p.uy >= 0.0f, upz = p.uz >= 0.0f;

  // Faces ahead of the packet's voxel (:33-35).
  const float vbx = (float)(upx ? p.i + 1 : p.i) * a.dx;
  const float vby = (float)(upy ? p.j + 1 : p.j) * a.dy;
  const float vbz = (float)(upz ? p.k + 1 : p.k) * a.dz;
  const bool on_face = (p.x == vbx) | (p.y == vby) | (p.z == vbz);
  // An x or y index outside the grid (a packet launched on the far wall) anchors a box that reaches
  // past the walls: such a packet steps in its voxel.
  const bool outside = ((uint32_t)p.i >= (uint32_t)a.nx) | ((uint32_t)p.j >= (uint32_t)a.ny);
  bool box = has_box(p) && (mu_t > 0.0f) && !on_face && !outside;
  // skip table present (a kernel argument, so the tests below are wave-uniform)
  const bool has_skip = a.skip_off >= 0;
  const int cap = a.jump_cap;

  float db, dbm, nx, ny, nz, bx, by, bz;
  int ei, ej, ek, n;
  bool hit;
  // The flight in a region; returns false for a box flight that fails its checks.
  auto flight = [&](const bool box) -> bool {
    // region [lo, hi] on each axis
    const int rxy2 = box ? (int)(p.ext & 0xffffu) : 0, rz2 = box ? (int)(p.ext >> 16) : 0;
    const int loi = box ? -p.oi : p.i, loj = box ? -p.oj : p.j, lok = box ? -p.ok : p.k;
    const int hii = loi + rxy2, hij = loj + rxy2, hik = lok + rz2;
    const float rbx = box ? (float)(upx ? hii + 1 : loi) * a.dx : vbx;
    const float rby = box ? (float)(upy ? hij + 1 : loj) * a.dy : vby;
    const float rbz = box ? (float)(upz ? hik + 1 : lok) * a.dz : vbz;
    // photontrack_intersection (:29-49) with cached reciprocals, on the region's faces
    const float tx = (rbx - p.x) * p.iux;
    const float ty = (rby - p.y) * p.iuy;
    const float tz = (rbz - p.z) * p.iuz;
    db = fminf(tx, fminf(ty, tz));
    dbm = mul_rn(db, mu_t);  // rounded product (no contraction into s - db*mu_t)
    hit = dbm <= p.s;        // :197
    const float hop = hit ? db : p.s * c1.z;  // distance moved (:198-206)
    const bool ex = hit && db == tx, ey = hit && db == ty, ez = hit && db == tz;  // faces reached
    nx = ex ? rbx : fmaf(p.ux, hop, p.x);
    ny = ey ? rby : fmaf(p.uy, hop, p.y);
    nz = ez ? rbz : fmaf(p.uz, hop, p.z);
    if (!box) {  // the reference's step in the packet's voxel
      ei = p.i;
      ej = p.j;
      ek = p.k;
      bx = vbx;
      by = vby;
      bz = vbz;
      n = 0;
      return true;
    }
    // the flight's last voxel and its faces
    ei = axis_voxel(nx, a.idx, loi, hii);
    ej = axis_voxel(ny, a.idy, loj, hij);
    ek = axis_voxel(nz, a.idz, lok, hik);
    // (float)(i + 1) is (float)i + 1 exactly: every grid dimension is below kMaxAxis = 2^24
    const float fei = (float)ei, fej = (float)ej, fek = (float)ek;
    const float lox = fei * a.dx, hix = (fei + 1.0f) * a.dx;
    const float loy = fej * a.dy, hiy = (fej + 1.0f) * a.dy;
    const float loz = fek * a.dz, hiz = (fek + 1.0f) * a.dz;
    bx = upx ? hix : lox;
    by = upy ? hiy : loy;
    bz = upz ? hiz : loz;
    const int di = upx ? ei - p.i : p.i - ei;
    const int dj = upy ? ej - p.j : p.j - ej;
    const int dk = upz ? ek - p.k : p.k - ek;
    n = di + dj + dk;  // plain crossings before the step
    const bool inside = (ex | ((nx > lox) & (nx < hix))) & (ey | ((ny > loy) & (ny < hiy))) &
                        (ez | ((nz > loz) & (nz < hiz)));
    return inside & ((di | dj | dk) >= 0) & (n <= cap) & (p.nstep + (uint32_t)n < kMaxSteps);
  };
  const bool ok = flight(box);
  if (__builtin_amdgcn_ballot_w64(!ok) != 0) {  // rare: redo failed box flights in the voxel
    if (!ok) flight(false);
  }
  if (COUNTERS) acc.steps += (uint32_t)n;
  p.nstep += (uint32_t)n;
  p.i = ei;
  p.j = ej;
  p.k = ek;
  if (has_skip) {
    skip_crossings<kBlock>(rx, ra, skip, n);
  } else {
    for (int q = 0; q < n; ++q) {
      mwc_next(rx, ra);
      mwc_next(rx, ra);
      mwc_next(rx, ra);
      mwc_next(rx, ra);
    }
  }
  const float ds = hit ? dbm : p.s;   // :198
  const float s_left = p.s - ds;      // :207

  // Next voxel (:209-213): an axis steps when the new coordinate reached the face ahead.  This
  // equals the reference's grid_coord of the new position while each index names a grid voxel
  // whose closed extent holds the position.  An x or y index outside the grid (a packet launched
  // on the far wall, then folded back inside by a mirror wall) takes grid_coord itself, which
  // brings the index back to the position as the reference does.
  int ni_ = next_index(p.i, p.ux, nx, bx);
  int nj_ = next_index(p.j, p.uy, ny, by);
  int nk_ = next_index(p.k, p.uz, nz, bz);
  if (__builtin_amdgcn_ballot_w64(outside) != 0) {
    if (outside) {
      ni_ = grid_coord(nx, p.ux, a.dx);
      nj_ = grid_coord(ny, p.uy, a.dy);
      nk_ = grid_coord(nz, p.uz, a.dz);
    }
  }
  const uint32_t nw = geometry_word<LAYERED>(a, lg, ni_, nj_, nk_);
  const bool air_override = nz <= 0.0f && p.uz < 0.0f;
  const int next_t = air_override ? YMC_TISSUE_AIR : word_row(a, nw);
  float nt;  // parametric geometries with up to 4 rows: from the lane's registers
  if (LAYERED && a.num_types <= 4)
    nt = next_t == 0 ? nrow.x : (next_t == 1 ? nrow.y : (next_t == 2 ? nrow.z : nrow.w));
  else
    nt = row_at(rows, rbase + next_t).r0.w;

  float nux = p.ux, nuy = p.uy, nuz = p.uz;
  bool refl = false;
  bool turned = !hit;  // direction will change this step
  if (ni != nt) {      // Fresnel (:226-257)
    if (COUNTERS) acc.fresnel++;
    const bool xi = p.i != ni_, yj = p.j != nj_, zk = p.k != nk_;
    float cos_ai = xi ? p.ux : (yj ? p.uy : p.uz);
    cos_ai = fminf(fabsf(cos_ai), 1.0f);
    const float ni_nt = ni * rcp(nt);
    const float sin_ai = __builtin_amdgcn_sqrtf((1.0f - cos_ai) * (1.0f + cos_ai));  // sin(acos(c)), well conditioned
    const float sin_at = ni_nt * sin_ai;
    const float cos_at = __builtin_amdgcn_sqrtf(1.0f - sin_at * sin_at);
    const float R1 = (sin_ai * cos_at - cos_ai * sin_at) * rcp(sin_ai * cos_at + cos_ai * sin_at);
    const float r2d = a.fresnel_exact ? (cos_ai * cos_at + sin_ai * sin_at) : (cos_ai * cos_at + sin_at * sin_at);
    const float R2 = (cos_ai * cos_at - sin_ai * sin_at) * rcp(r2d);
    float R = 0.5f * R1 * R1 * (1.0f + R2 * R2);
    R = (sin_at > 1.0f) ? 1.0f : R;
    const float eps = u32_co(mwc_next(rx, ra));
    refl = eps < R;
    const float sx = copysignf(cos_at, p.ux), sy = copysignf(cos_at, p.uy), sz = copysignf(cos_at, p.uz);
    if (hit) {
      nux = refl ? (xi ? -p.ux : p.ux) : (xi ? sx : p.ux * ni_nt);
      nuy = refl ? (yj ? -p.uy : p.uy) : ((yj && !xi) ? sy : p.uy * ni_nt);
      nuz = refl ? (zk ? -p.uz : p.uz) : ((zk && !yj && !xi) ? sz : p.uz * ni_nt);
    }
    turned = true;
  }

  // The four draws of every step: roulette, HG cos, HG phi, new step length (:275-321).
  const uint32_t u_roul = mwc_next(rx, ra);
  const uint32_t u_cos = mwc_next(rx, ra);
  const uint32_t u_phi = mwc_next(rx, ra);
  const uint32_t u_len = mwc_next(rx, ra);

  float W = p.W;
  float s_new = s_left;
  // Interactions: absorption, roulette, scattering, new step length.
  if (__builtin_amdgcn_ballot_w64(!hit) != 0) {
    if (!hit) {
      const float W_in = W;
      W = W - c1.y * W;  // dW = (mu_a/mu_t)*W (:262-263)
      if (COUNTERS) acc.absorbed += fx(W_in) - fx(W);
      if (a.abs_grid) {  // the disabled deposit of :265-272, at the voxel the step started in
        const uint32_t ic = (uint32_t)min(max(p.i, 0), a.nx - 1);
        const uint32_t jc = (uint32_t)min(max(p.j, 0), a.ny - 1);
        const uint32_t kc = (uint32_t)min(max(p.k, 0), a.nz - 1);
        atomicAdd(&a.abs_grid[(size_t)table * a.nvox + (kc * (uint32_t)a.ny + jc) * (uint32_t)a.nx + ic],
                  (unsigned long long)(fx(W_in) - fx(W)));
      }
      if (W < a.w_thresh && mu_t != 0.0f) {  // roulette (:275-277)
        const float e = u32_co(u_roul);
        // The reference compares and divides in binary64: e < 0.1 and (float)(W / 0.1).  The
        // binary64 constant is 0.1 (1 + 2^-54) exactly, so W / 0.1 = 10 W (1 - 2^-54 + ...), within
        // half an ulp of 10 W, which binary64 holds exactly (W has 24 significant bits): the quotient
        // rounds to 10 W and the float result is RN(10 W) = W * 10.0f.  No float lies in
        // [0.1, 0.1f), so e < 0.1 <=> e < 0.1f.  Both identities are checked over every float of
        // their range on the CPU.
        static_assert(YMC_SURVIVAL_CHANCE == 0.1, "the binary32 roulette below assumes a chance of 0.1");
        const float Wr = (e < 0.1f) ? W * 10.0f : 0.0f;
        if (COUNTERS) atomicAdd(&tot[kRoul], (unsigned long long)((long long)fx(Wr) - (long long)fx(W)));
        W = Wr;
      }
      const float4 c2 = LAYERED ? *cur.r2 : row_at(rows, rbase + p.t).r2;
      // HG (:282-294)
      const float xi = u32_oc(u_cos);
      float ct = c2.x * rcp(fmaf(c2.z, xi, c2.y));
      ct = c1.w * (c2.w - ct * ct);
      const float st = __builtin_amdgcn_sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
      const float rev = u32_co(u_phi);  // phi / (2 pi)
      const float sp = __builtin_amdgcn_sinf(rev);
      const float cp = __builtin_amdgcn_cosf(rev);
      const float ux = p.ux, uy = p.uy, uz = p.uz;
      if (fabsf(uz) <= a.uz_thresh) {
        const float one_m = 1.0f - uz * uz;
        const float inv = __builtin_amdgcn_rsqf(one_m);
        const float stv = st * inv;
        nux = stv * (ux * uz * cp - uy * sp) + ux * ct;
        nuy = stv * (uy * uz * cp + ux * sp) + uy * ct;
        nuz = -(one_m * inv) * st * cp + uz * ct;
      } else {
        nux = st * cp;
        nuy = st * sp;
        nuz = copysignf(ct, ct * uz);
      }
      s_new = (W > 0.0f) ? neg_ln(u32_oc(u_len)) : 0.0f;  // :319-321
    }
  }

  const bool freeze = (W <= 0.0f) || (mu_t == 0.0f);  // :313
  if (turned && !freeze) {
    // :302-310
    nux = (fabsf(nux) < a.zero_thresh) ? 0.0f : nux;
    nuy = (fabsf(nuy) < a.zero_thresh) ? 0.0f : nuy;
    nuz = (fabsf(nuz) < a.zero_thresh) ? 0.0f : nuz;
    const float nrm = __builtin_amdgcn_rsqf(nux * nux + nuy * nuy + nuz * nuz);
    p.ux = nux * nrm;
    p.uy = nuy * nrm;
    p.uz = nuz * nrm;
    p.iux = rcp(p.ux);
    p.iuy = rcp(p.uy);
    p.iuz = rcp(p.uz);
  }
  p.W = W;
  p.s = (freeze || hit) ? s_left : s_new;
  if (!freeze) {
    p.x = nx;
    p.y = ny;
    p.z = nz;
  }
  // Mirror walls (:113-126, :330), with the reference's short-circuit order.  Wall hits are rare, so
  // the wave skips the block unless one of its lanes hits (the compiler otherwise issues both folds,
  // masked, in every step).
  const bool wall_x = (p.x >= a.Lx) | (p.x <= 0.0f);
  const bool wall_y = (p.y >= a.Ly) | (p.y <= 0.0f);
  if (__builtin_amdgcn_ballot_w64(!refl & (wall_x | wall_y)) != 0 && !refl) {
    if (wall_x) {
      p.x = (p.x >= a.Lx) ? 2.0f * a.Lx - p.x : p.x;
      p.x = (p.x <= 0.0f) ? fabsf(p.x) : p.x;
      p.ux = -p.ux;
      p.iux = -p.iux;
      refl = true;
    } else if (wall_y) {
      p.y = (p.y >= a.Ly) ? 2.0f * a.Ly - p.y : p.y;
      p.y = (p.y <= 0.0f) ? fabsf(p.y) : p.y;
      p.uy = -p.uy;
      p.iuy = -p.iuy;
      refl = true;
    }
  }
  if (!((hit && refl) || freeze)) {  // :334-339
    p.i = ni_;
    p.j = nj_;
    p.k = nk_;
    if (LAYERED && next_t != p.t) set_lane_row(cur, row_at(rows, rbase + next_t));
    p.t = next_t;
    // the voxel just loaded becomes the anchor (no box when the packet is leaving through the top)
    set_anchor(p, a, ni_, nj_, nk_, nw, !air_override);
  }
  if (COUNTERS) acc.steps++;

  // Kill (:343-353): non-finite state after every step; a zero step length where the reference
  // checks it (kill0 1: end of each 45-step launch, 2: every step, 0: never).  Two rules the
  // reference lacks keep every packet finite: a negative step length (the position has left the
  // voxel its indices name - an interaction rounded onto a mirror wall and the index commit put
  // the packet outside the grid, after which the wall test bounces it forever) and a step cap.
  p.nstep++;
  bool zero_kill = false;
  if (db == 0.0f)  // rare: the launch-boundary test runs only for lanes that need it
    zero_kill = a.kill0 == 2 || (a.kill0 == 1 && (p.nstep % kLaunchSteps) == 0);
  // Non-finite state: v * 0 is +-0 for finite v and NaN for an infinity or a NaN, and a sum of +-0
  // stays +-0, so the sum below is NaN exactly when one of the seven is not finite (multiplies and
  // fmas with an inline zero, where seven isfinite tests were seven compares against an SGPR).
  const float nf = fmaf(p.x, 0.0f, p.y * 0.0f) + fmaf(p.z, 0.0f, p.s * 0.0f) +
                   fmaf(p.ux, 0.0f, fmaf(p.uy, 0.0f, p.uz * 0.0f));
  // bitwise, not short-circuit: every test is a compare, and no branch splits the reset below
  const bool bad = (nf != nf) | zero_kill | (db < 0.0f) | (p.nstep >= kMaxSteps);
  if (bad) {
    if (COUNTERS) {
      atomicAdd(&tot[kKilled], 1ull);
      atomicAdd(&tot[kKilledW], (unsigned long long)fx(p.W));
    }
    // A dead packet tallies nothing (zero weight); its position is cleared for the record only.
    // Clearing the other fields unconditionally cost a select each in every iteration.
    if (RECORD) {
      p.x = p.y = p.z = 0.0f;
      p.ux = p.uy = 0.0f;
      p.uz = 1.0f;
      p.s = 0.0f;
    }
    p.W = 0.0f;
    return kDead;
  }
  return freeze ? kFrozen : kLive;
}

// photon_detector (src/mc3d_photons.cu:151-171) on the tally grid, in fixed point.
template <bool RADIAL>
__device__ __forceinline__ void tally_packet(const Packet &p, const KArgs &a, uint32_t table,
                                             unsigned long long *tot) {
  const uint64_t f = fx(p.W);
  if (f == 0) return;
  const uint64_t f2 = fx_sq(f);
  const int ix = grid_coord(p.x, p.ux, a.tdx);
  const int iy = grid_coord(p.y, p.uy, a.tdy);
  const bool inside = ix >= 0 && iy >= 0 && ix < a.tnx && iy < a.tny;
  const bool refl = p.z <= 0.0f;  // src/mc3d_photons.cu:413
  const size_t bin = (refl ? 0 : a.map_half) + ((size_t)table * a.tny + (inside ? iy : 0)) * a.tnx + (inside ? ix : 0);
  atomicAdd(&tot[refl ? kRw : kTw], (unsigned long long)f);
  atomicAdd(&tot[refl ? kRw2 : kTw2], (unsigned long long)f2);
  atomicAdd(&tot[refl ? kRn : kTn], 1ull);
  if (inside) {
    atomicAdd(&a.map_w[bin], (unsigned long long)f);
    atomicAdd(&a.map_w2[bin], (unsigned long long)f2);
    atomicAdd(&a.map_n[bin], 1u);
  }
  if (RADIAL && refl) {
    const float ddx = p.x - a.rx0, ddy = p.y - a.ry0;
    const float r = sqrtf(ddx * ddx + ddy * ddy);
    const float qb = floorf(r / a.rdr);
    if (qb < (float)a.nr) {
      const size_t rb = (size_t)table * a.nr + (int)qb;
      atomicAdd(&a.Rr_w[rb], (unsigned long long)f);
      atomicAdd(&a.Rr_w2[rb], (unsigned long long)f2);
      atomicAdd(&a.Rr_n[rb], 1u);
    }
  }
}


// Byte offset of the per-table totals in the kernel's dynamic LDS (8-byte aligned).
__host__ __device__ inline size_t lds_tot_offset(int tables, int types) {
  size_t off = (size_t)tables * types * sizeof(RowConst) + (size_t)tables * sizeof(float);
  return (off + 7) & ~(size_t)7;
}
// Parametric runs hold each lane's current row of constants in LDS (LaneRow: 36 B per thread).
__host__ __device__ inline size_t lds_bytes(int tables, int types, bool layered) {
  return layered ? (size_t)kBlock * (2 * sizeof(float4) + sizeof(float))
                 : lds_tot_offset(tables, types) + (size_t)tables * kTotFields * 8;
}

// Occupancy.  The voxel-grid variants without counters, records or radial tally are held to 7 waves
// per SIMD (72 VGPRs): held to 8 (64 VGPRs) the box-flight step keeps a value in scratch, and the
// tuned pool of the 256x256x128 layered-skin workload runs 7 waves per SIMD either way.  The
// parametric variants without counters or records are held to 6 (80 VGPRs; two or three values go
// to scratch, one of them reloaded for a skip product), which with the skip table (25 KB of LDS per
// workgroup) ran the surrogate's Monte Carlo 3.5% faster than the compiler's 5 waves.  The other
// variants keep the compiler's choice: the counter, record and radial variants spill to scratch at 7.
template <bool UNIFORM, bool COUNTERS, bool RADIAL, bool RECORD, bool LAYERED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    (!COUNTERS && !RECORD && !LAYERED && !RADIAL) ? 7 : ((!COUNTERS && !RECORD && LAYERED) ? 6 : 1),
    8))) void transport_kernel(KArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  // Voxel-grid runs (few tables, many packets each) keep the optical-property rows, launch
  // weights and the workgroup's per-table totals in LDS; the totals are accumulated with LDS
  // atomics and added to the device totals once, when the block exits.  Parametric-geometry runs
  // (one table per sample, thousands of tables, few packets each) read rows, weights and
  // geometries from global memory (cached) and add their totals there directly.
  RowConst *s_rows = reinterpret_cast<RowConst *>(lds_raw);
  float *s_w0 = reinterpret_cast<float *>(s_rows + a.num_tables * a.num_types);
  unsigned long long *s_tot = reinterpret_cast<unsigned long long *>(lds_raw + lds_tot_offset(a.num_tables, a.num_types));
  if (!LAYERED) {
    for (int q = threadIdx.x; q < a.num_tables * kTotFields; q += blockDim.x) s_tot[q] = 0;
    const int nrow = a.num_tables * a.num_types;
    const float4 *src = reinterpret_cast<const float4 *>(a.rows);
    float4 *dst = reinterpret_cast<float4 *>(s_rows);
    for (int q = threadIdx.x; q < nrow * 3; q += blockDim.x) dst[q] = src[q];
    for (int q = threadIdx.x; q < a.num_tables; q += blockDim.x) s_w0[q] = a.w0[q];
    __syncthreads();
  }
  const RowConst *rows = LAYERED ? a.rows : s_rows;
  LaneRow s_cur{nullptr, nullptr, nullptr};  // the lane's row slot (parametric runs)
  if (LAYERED) {
    float4 *r1 = reinterpret_cast<float4 *>(lds_raw);
    s_cur = LaneRow{r1 + threadIdx.x, r1 + kBlock + threadIdx.x, reinterpret_cast<float *>(r1 + 2 * kBlock) + threadIdx.x};
  }
  uint64_t *s_skip = a.skip_off >= 0 ? reinterpret_cast<uint64_t *>(lds_raw + a.skip_off) + threadIdx.x : nullptr;
  unsigned long long *totals = LAYERED ? a.tot : s_tot;

  const uint32_t lane = __lane_id();
  Packet p;
  uint64_t rx = 0;
  uint32_t ra = 0;
  uint32_t left = 0, table = 0, done_in_chain = 0;
  uint32_t rbase = 0;  // first row of the lane's table, set with each chain
  unsigned long long chain = 0;
  bool alive = false, drained = false, stop = false;
  bool ended = false;  // the lane's packet has ended and waits for its tally
  int ended_res = kLive;
  uint32_t waited = 0;  // loop iterations in which some lane of the wave waited for the launch phase
  LaneAcc acc{0, 0, 0, 0};
  uint32_t iter = 0;
  // Parametric geometry of the lane's table and the refractive indices of its rows 0-3, loaded
  // with each chain: each step otherwise waits on these loads before it can name the next voxel's
  // row and decide on Fresnel.
  ymc_layered geo{};
  float4 nrow = make_float4(1.0f, 1.0f, 1.0f, 1.0f);

  // A packet that ended: tally, optional record, count.
  auto finish = [&](int res) {
    unsigned long long *tt = totals + (size_t)table * kTotFields;
    tally_packet<RADIAL>(p, a, table, tt);
    if (RECORD) {
      const unsigned long long gi = a.chain_first[chain] + done_in_chain;
      if (gi < a.n_records) {
        ymc_photon_record r;
        r.x = p.x;
        r.y = p.y;
        r.z = p.z;
        r.w = p.W;
        r.steps = p.nstep;
        r.fate = res == kDead ? 2 : (p.z <= 0.0f ? 0 : 1);
        a.records[gi] = r;
      }
    }
    --left;
    ++done_in_chain;
    acc.photons++;
  };

  for (;;) {
    // Cancel flag (host memory), read by the first lane of the wave every poll_interval steps.
    if (a.cancel && (iter % (uint32_t)a.poll_interval) == 0) {
      int c = 0;
      if (lane == (uint32_t)__builtin_amdgcn_readfirstlane(lane)) c = *a.cancel;
      c = __shfl(c, __builtin_amdgcn_readfirstlane(lane));
      stop = stop || (c != 0);
    }
    ++iter;
    // Launch phase: tally ended packets, take new chains, launch packets.  The phase costs the wave
    // as much for one lane as for many, so lanes whose packet ended wait (doing nothing) and the
    // phase runs for several at once: when relaunch_lanes lanes wait, when no lane of the wave is
    // alive, or after relaunch_wait iterations.  Packets draw from their chains' streams, so the
    // moment a lane takes one changes no result.
    const bool wants = ended || (!alive && !stop && !(drained && left == 0));
    const uint32_t nwant = (uint32_t)__popcll(__ballot(wants));
    const bool phase = nwant >= (uint32_t)a.relaunch_lanes ||
                       (nwant > 0 && (__ballot(alive) == 0 || waited >= (uint32_t)a.relaunch_wait));
    waited = phase ? 0 : waited + (nwant > 0 ? 1 : 0);
    if (phase && ended) {
      finish(ended_res);
      ended = false;
    }
    if (phase && !alive && !stop) {
      if (left == 0 && !drained) {
        // Chain finished: flush its counters, then take the next chain.  One atomic per wave.
        if (COUNTERS && acc.photons) {
          unsigned long long *tt = totals + (size_t)table * kTotFields;
          atomicAdd(&tt[kPhotons], (unsigned long long)acc.photons);
          atomicAdd(&tt[kSteps], (unsigned long long)acc.steps);
          atomicAdd(&tt[kFresnel], (unsigned long long)acc.fresnel);
          atomicAdd(&tt[kAw], (unsigned long long)acc.absorbed);
          acc = LaneAcc{0, 0, 0, 0};
        } else if (!COUNTERS && acc.photons) {
          atomicAdd(&totals[(size_t)table * kTotFields + kPhotons], (unsigned long long)acc.photons);
          acc.photons = 0;
        }
        const uint64_t want = __ballot(1);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
        const uint32_t leader = __builtin_amdgcn_readfirstlane(lane);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.next_chain, (unsigned long long)__popcll(want));
        base = __shfl(base, leader);
        const unsigned long long c = base + rank;
        if (c < a.n_chains) {
          chain = c;
          const ymc_chain ch = a.chains[c];
          left = ch.count;
          table = ch.table;
          rbase = __umul24(table, (uint32_t)a.num_types);
          rx = a.chain_x[c];
          ra = a.chain_a[c];
          if (LAYERED) {
            geo = a.layered[table];
            const RowConst *tr = a.rows + (size_t)table * a.num_types;
            const int T = a.num_types;
            nrow = make_float4(tr[0].r0.w, T > 1 ? tr[1].r0.w : 1.0f, T > 2 ? tr[2].r0.w : 1.0f, T > 3 ? tr[3].r0.w : 1.0f);
          }
          if (s_skip) build_skip<kBlock>(s_skip, ra);
          done_in_chain = 0;
        } else {
          drained = true;
        }
      }
      if (left > 0) {
        launch_packet<UNIFORM, LAYERED>(p, rx, ra, a, LAYERED ? a.w0[table] : s_w0[table], &geo);
        if (LAYERED) set_lane_row(s_cur, row_at(rows, rbase + p.t));
        alive = true;
      }
    }
    if (__ballot(alive) == 0) break;
    if (alive) {
      unsigned long long *tt = totals + (size_t)table * kTotFields;
      const int res =
          step_packet<COUNTERS, LAYERED, RECORD>(p, rx, ra, a, rows, rbase, tt, acc, &geo, table, nrow, s_skip, s_cur);
      if (res != kLive) {  // tallied in the next launch phase
        alive = false;
        ended = true;
        ended_res = res;
      }
    }
  }
  // Final flush: lane counters into the totals, then (LDS totals) the block's totals into the
  // device totals.
  if (acc.photons) {
    unsigned long long *tt = totals + (size_t)table * kTotFields;
    atomicAdd(&tt[kPhotons], (unsigned long long)acc.photons);
    if (COUNTERS) {
      atomicAdd(&tt[kSteps], (unsigned long long)acc.steps);
      atomicAdd(&tt[kFresnel], (unsigned long long)acc.fresnel);
      atomicAdd(&tt[kAw], (unsigned long long)acc.absorbed);
    }
  }
  if (!LAYERED) {
    __syncthreads();
    for (int q = threadIdx.x; q < a.num_tables * kTotFields; q += blockDim.x) {
      const unsigned long long v = s_tot[q];
      if (v) atomicAdd(&a.tot[q], v);
    }
  }
}

// Parametric-geometry test hook: the voxel word (row, rz, rxy) of every voxel, as the transport
// kernel computes it.
__global__ void layered_words_kernel(KArgs a, ymc_layered g, uint32_t *out) {
  const uint32_t n = (uint32_t)a.nx * a.ny * a.nz;
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int i = (int)(v % (uint32_t)a.nx), j = (int)((v / (uint32_t)a.nx) % (uint32_t)a.ny);
  const int k = (int)(v / ((uint32_t)a.nx * a.ny));
  out[v] = layered_word(a, g, i, j, k);
}

// Stream-skip test hook: each lane builds its skip table in LDS, as the transport kernel does at
// a chain start, and skips the draws of `crossings` plain crossings.
__global__ __launch_bounds__(kBlock) void rng_skip_kernel(const uint64_t *x, const uint32_t *a, uint32_t n,
                                                          int crossings, uint64_t *xf) {
  __shared__ uint64_t tab[kSkipEntries * kBlock];
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint64_t st = x[s];
  const uint32_t m = a[s];
  build_skip<kBlock>(tab + threadIdx.x, m);
  skip_crossings<kBlock>(st, m, tab + threadIdx.x, crossings);
  xf[s] = st;
}

// FP32 FMA rate probe: eight independent fma chains per lane, so a SIMD always has an FMA to issue.
// Built like the rest of the library (no packed FP32).  OPERANDS is the number of distinct VGPRs an
// FMA reads: 1 = v*v+v, 2 = v*m+c (m wave-uniform in an SGPR, c in a VGPR), 3 = v*a+b with per-lane
// a and b, the usual case in the transport kernel.
constexpr int kProbeChains = 8, kProbeUnroll = 16;
template <int OPERANDS>
__global__ __launch_bounds__(256) void fma_probe_kernel(float *out, int iters, float m, float c) {
  float v[kProbeChains];
#pragma unroll
  for (int q = 0; q < kProbeChains; ++q) v[q] = (float)(threadIdx.x + q) * 1e-3f;
  const float a = m + (float)threadIdx.x * 1e-9f, b = c * (float)(threadIdx.x & 7);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u)
#pragma unroll
      for (int q = 0; q < kProbeChains; ++q) {
        if (OPERANDS == 1) v[q] = fmaf(v[q], v[q], v[q]);
        else if (OPERANDS == 2) v[q] = fmaf(v[q], m, c);
        else v[q] = fmaf(v[q], a, b);
      }
  }
  float sum = 0.0f;
#pragma unroll
  for (int q = 0; q < kProbeChains; ++q) sum += v[q];
  out[blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

// RNG test hook: n_streams lanes, n_draws draws each.
__global__ void rng_fill_kernel(const uint64_t *x, const uint32_t *a, uint32_t n, uint32_t draws, int kind,
                                float *out, uint64_t *xf) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint64_t st = x[s];
  const uint32_t m = a[s];
  for (uint32_t d = 0; d < draws; ++d) {
    const uint32_t u = mwc_next(st, m);
    out[(size_t)d * n + s] = kind ? u32_oc(u) : u32_co(u);
  }
  xf[s] = st;
}

// ------------------------------------------------------------------------------------------
// Host side.

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t err_ = (expr);                                                                 \
    if (err_ != hipSuccess)                                                                   \
      return fail(YMC_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_), __FILE__, __LINE__); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ymc_status ensure(size_t bytes) {
    if (bytes <= n && p) return YMC_OK;
    release();
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(YMC_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    n = bytes;
    return YMC_OK;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

}  // namespace ymc

struct ymc_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cus = 0;
  std::vector<uint32_t> table;
  // geometry
  ymc::DevBuf vox;
  int nx = 0, ny = 0, nz = 0;
  float dx = 0, dy = 0, dz = 0;
  bool have_geom = false;
  bool vox16 = false;  // voxel grid stored as 16-bit words (else 32-bit)
  bool last_skip_table = false;  // the last run skipped jumped draws with the LDS table
  // parametric geometry (ymc_engine_set_layered) instead of a voxel grid
  bool layered = false;
  ymc::DevBuf layered_buf;
  std::vector<ymc_layered> layered_host;
  // optprops
  ymc::DevBuf rows, w0;
  int num_types = 0, num_tables = 0;
  bool have_props = false;
  std::vector<float> host_n;  // for validation
  // prepared run
  ymc::DevBuf chains, chain_x, chain_a, chain_first;
  uint64_t n_chains = 0;
  uint64_t n_packets = 0;
  ymc_params params{};
  bool prepared = false;
  // tallies
  ymc::DevBuf map_w, map_w2, map_n, Rr_w, Rr_w2, Rr_n, tot, counter, records, absorb;
  bool absorbing = false;  // the prepared run tallies the absorption grid
  int tnx = 0, tny = 0, nr = 0;
  float tdx = 0, tdy = 0;
  uint64_t n_records = 0;
  double last_ms = 0;
  int last_pool = 0;
  // Cancel flag the kernel polls: pinned, mapped host memory.  ymc_engine_run copies the
  // caller's flag (any host pointer) into it while the kernel runs.
  int *cancel_host = nullptr;
  int *cancel_dev = nullptr;
};

namespace ymc {

static float largest_float_le(double v) {
  float f = (float)v;
  if ((double)f > v) f = std::nextafter(f, -INFINITY);
  return f;
}
static float smallest_float_ge(double v) {
  float f = (float)v;
  if ((double)f < v) f = std::nextafter(f, INFINITY);
  return f;
}

// Largest absorption grid an engine allocates (tables x voxels x 8 bytes).
constexpr size_t kMaxAbsorptionBytes = size_t(32) << 30;
// Largest packet budget of one table (see upload_chains).
constexpr uint64_t kMaxTablePackets = uint64_t(1) << 31;

static ymc_status check_engine(ymc_engine *e) {
  if (!e) return fail(YMC_ERR_ARG, "null engine");
  if (hipSetDevice(e->device) != hipSuccess) return fail(YMC_ERR_DEVICE, "hipSetDevice(%d) failed", e->device);
  return YMC_OK;
}

template <int KEY>
static const void *kernel_ptr() {
  return reinterpret_cast<const void *>(
      &transport_kernel<(KEY & 16) != 0, (KEY & 8) != 0, (KEY & 4) != 0, (KEY & 2) != 0, (KEY & 1) != 0>);
}

template <int... K>
static const void *pick(int key, std::integer_sequence<int, K...>) {
  static const void *const table[] = {kernel_ptr<K>()...};
  return table[key];
}

static const void *select_kernel(bool uniform, bool counters, bool radial, bool record, bool layered) {
  const int key = (uniform ? 16 : 0) | (counters ? 8 : 0) | (radial ? 4 : 0) | (record ? 2 : 0) | (layered ? 1 : 0);
  return pick(key, std::make_integer_sequence<int, 32>{});
}

}  // namespace ymc

using namespace ymc;

extern "C" {

ymc_status ymc_engine_create(int32_t device, const uint32_t *table, uint32_t table_len, ymc_engine **out) {
  if (!out || !table || table_len < 2) return fail(YMC_ERR_ARG, "ymc_engine_create: need a multiplier table of >= 2 entries");
  for (uint32_t q = 0; q < table_len; ++q)
    if (table[q] < 2) return fail(YMC_ERR_ARG, "ymc_engine_create: multiplier %u at %u is below 2", table[q], q);
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(YMC_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(YMC_ERR_ARG, "device %d out of range (have %d)", device, ndev);
  ymc_engine *e = new ymc_engine();
  e->device = device;
  e->table.assign(table, table + table_len);
  hipError_t st = hipSetDevice(device);
  if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (st == hipSuccess) st = hipEventCreate(&e->ev0);
  if (st == hipSuccess) st = hipEventCreate(&e->ev1);
  if (st == hipSuccess) st = hipDeviceGetAttribute(&e->num_cus, hipDeviceAttributeMultiprocessorCount, device);
  if (st != hipSuccess) {
    ymc_status r = fail(YMC_ERR_DEVICE, "engine setup failed: %s", hipGetErrorString(st));
    ymc_engine_destroy(e);
    return r;
  }
  if (e->counter.ensure(64) != YMC_OK || e->tot.ensure(sizeof(ymc_totals))) {
    ymc_engine_destroy(e);
    return fail(YMC_ERR_NOMEM, "engine allocation failed");
  }
  st = hipHostMalloc(reinterpret_cast<void **>(&e->cancel_host), sizeof(int), hipHostMallocMapped | hipHostMallocCoherent);
  if (st == hipSuccess) st = hipHostGetDevicePointer(reinterpret_cast<void **>(&e->cancel_dev), e->cancel_host, 0);
  if (st != hipSuccess) {
    ymc_status r = fail(YMC_ERR_DEVICE, "pinned cancel flag: %s", hipGetErrorString(st));
    ymc_engine_destroy(e);
    return r;
  }
  *e->cancel_host = 0;
  *out = e;
  return YMC_OK;
}

void ymc_engine_destroy(ymc_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  if (e->cancel_host) (void)hipHostFree(e->cancel_host);
  delete e;  // DevBuf destructors free device memory
}

ymc_status ymc_engine_set_layered(ymc_engine *e, int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz,
                                  const ymc_layered *per_table, int32_t num_tables) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!per_table || num_tables < 1 || nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0) || !(dy > 0) || !(dz > 0) ||
      (size_t)nx * ny * nz >= (1ull << 31) || std::max(nx, std::max(ny, nz)) >= kMaxAxis)
    return fail(YMC_ERR_ARG, "ymc_engine_set_layered: invalid arguments");
  for (int w = 0; w < num_tables; ++w) {
    const ymc_layered &g = per_table[w];
    if (g.top_voxels < 0 || g.row_top < 0 || g.row_top > 255 || g.row_bottom < 0 || g.row_bottom > 255 ||
        g.row_vessel > 255 || !(g.vessel_r >= 0 && g.vessel_r < kMaxVesselRadius) || !std::isfinite(g.vessel_x) ||
        !std::isfinite(g.vessel_z))
      return fail(YMC_ERR_ARG, "layered geometry %d invalid", w);
  }
  if ((st = e->layered_buf.ensure(num_tables * sizeof(ymc_layered)))) return st;
  HIP_TRY(hipMemcpy(e->layered_buf.p, per_table, num_tables * sizeof(ymc_layered), hipMemcpyHostToDevice));
  e->layered_host.assign(per_table, per_table + num_tables);
  e->nx = nx;
  e->ny = ny;
  e->nz = nz;
  e->dx = dx;
  e->dy = dy;
  e->dz = dz;
  e->have_geom = true;
  e->layered = true;
  e->prepared = false;
  return YMC_OK;
}

ymc_status ymc_safe_box(const uint8_t *rows, int32_t nx, int32_t ny, int32_t nz, int32_t zcap, int32_t xycap,
                        uint8_t *zup, uint8_t *zdn, uint8_t *rxy) {
  if (!rows || !zup || !zdn || !rxy || nx <= 0 || ny <= 0 || nz <= 0 || zcap < 0 || zcap > 255 || xycap < 0 ||
      xycap > 255)
    return fail(YMC_ERR_ARG, "ymc_safe_box: bad arguments");
  std::vector<uint8_t> radius((size_t)nx * ny * nz);
  safe_radius(rows, nx, ny, nz, radius.data());
  safe_box(rows, nx, ny, nz, radius.data(), zcap, xycap, zup, zdn, rxy);
  return YMC_OK;
}

ymc_status ymc_safe_radius(const uint8_t *rows, int32_t nx, int32_t ny, int32_t nz, uint8_t *out) {
  if (!rows || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail(YMC_ERR_ARG, "ymc_safe_radius: bad arguments");
  safe_radius(rows, nx, ny, nz, out);
  return YMC_OK;
}

ymc_status ymc_engine_set_geometry(ymc_engine *e, const ymc_geometry *g) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!g || !g->tissue || g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || !(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0))
    return fail(YMC_ERR_ARG, "ymc_engine_set_geometry: invalid geometry");
  size_t nv = (size_t)g->nx * g->ny * g->nz;
  if (nv >= (1ull << 31)) return fail(YMC_ERR_ARG, "ymc_engine_set_geometry: more than 2^31 voxels");
  if (std::max(g->nx, std::max(g->ny, g->nz)) >= kMaxAxis)
    return fail(YMC_ERR_ARG, "ymc_engine_set_geometry: a dimension of 2^24 voxels or more");
  std::vector<uint8_t> radius(nv), zup(nv), zdn(nv), rxy(nv);
  safe_radius(g->tissue, g->nx, g->ny, g->nz, radius.data());
  uint8_t max_row = 0;
  for (size_t v = 0; v < nv; ++v) max_row = std::max(max_row, g->tissue[v]);
  e->vox16 = max_row < 8;  // rows fit 3 bits: 16-bit words, boxes capped to fit 4 + 4 + 5 bits
  safe_box(g->tissue, g->nx, g->ny, g->nz, radius.data(), e->vox16 ? 15 : 255, e->vox16 ? 31 : 255, zup.data(),
           zdn.data(), rxy.data());
  if (e->vox16) {
    std::vector<uint16_t> words(nv);
    for (size_t v = 0; v < nv; ++v)
      words[v] = (uint16_t)(g->tissue[v] | (zup[v] << 3) | (zdn[v] << 7) | (rxy[v] << 11));
    if ((st = e->vox.ensure(nv * 2))) return st;
    HIP_TRY(hipMemcpy(e->vox.p, words.data(), nv * 2, hipMemcpyHostToDevice));
  } else {
    std::vector<uint32_t> words(nv);
    for (size_t v = 0; v < nv; ++v)
      words[v] = (uint32_t)g->tissue[v] | ((uint32_t)zup[v] << 8) | ((uint32_t)zdn[v] << 16) | ((uint32_t)rxy[v] << 24);
    if ((st = e->vox.ensure(nv * 4))) return st;
    HIP_TRY(hipMemcpy(e->vox.p, words.data(), nv * 4, hipMemcpyHostToDevice));
  }
  e->nx = g->nx;
  e->ny = g->ny;
  e->nz = g->nz;
  e->dx = g->dx;
  e->dy = g->dy;
  e->dz = g->dz;
  // remember the largest row for validation against the property tables
  uint8_t mx = 0;
  for (size_t v = 0; v < nv; ++v) mx = std::max(mx, g->tissue[v]);
  e->host_n.assign(1, (float)mx);
  e->have_geom = true;
  e->layered = false;
  e->prepared = false;
  return YMC_OK;
}

ymc_status ymc_engine_set_optprops(ymc_engine *e, const ymc_optprops *p) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!p || p->num_types < 2 || p->num_tables < 1 || !p->mua || !p->mus || !p->g || !p->n)
    return fail(YMC_ERR_ARG, "ymc_engine_set_optprops: need >= 2 rows (ambient + tissue) and >= 1 table");
  const int T = p->num_types, L = p->num_tables;
  if ((size_t)T * L >= ((size_t)1 << 24))  // the kernel addresses rows with 24-bit multiplies
    return fail(YMC_ERR_ARG, "ymc_engine_set_optprops: %d tables x %d rows exceed 2^24 rows", L, T);
  std::vector<RowConst> rows((size_t)T * L);
  std::vector<float> w0(L);
  for (int w = 0; w < L; ++w) {
    for (int t = 0; t < T; ++t) {
      size_t q = (size_t)w * T + t;
      float mua = p->mua[q], mus = p->mus[q], g = p->g[q], n = p->n[q];
      if (!(mua >= 0) || !(mus >= 0) || !(n > 0) || !(g >= -1 && g <= 1))
        return fail(YMC_ERR_ARG, "optical properties of table %d row %d out of range", w, t);
      float mut = mua + mus;
      RowConst &r = rows[q];
      r.r0 = make_float4(mua, mus, g, n);
      r.r1 = make_float4(mut, mut > 0 ? mua / mut : 0.0f, mut > 0 ? 1.0f / mut : 0.0f, 1.0f / (2 * g));
      r.r2 = make_float4(1 - g * g, 1 - g, 2 * g, 1 + g * g);
    }
    // photontrack_calc_specular_reflection(n[air], n[1]) (src/mc3d_gpu.cu:67-70, :500)
    float n0 = p->n[(size_t)w * T + YMC_TISSUE_AIR], n1 = p->n[(size_t)w * T + 1];
    float r = (n0 - n1) / (n0 + n1);
    w0[w] = 1.0f - r * r;
  }
  if ((st = e->rows.ensure(rows.size() * sizeof(RowConst)))) return st;
  if ((st = e->w0.ensure(w0.size() * sizeof(float)))) return st;
  HIP_TRY(hipMemcpy(e->rows.p, rows.data(), rows.size() * sizeof(RowConst), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->w0.p, w0.data(), w0.size() * sizeof(float), hipMemcpyHostToDevice));
  e->num_types = T;
  e->num_tables = L;
  e->have_props = true;
  e->prepared = false;
  return YMC_OK;
}

static ymc_status prepare_common(ymc_engine *e, const ymc_params *prm) {
  if (!e->have_geom || !e->have_props) return fail(YMC_ERR_STATE, "set geometry and optical properties first");
  if (!prm) return fail(YMC_ERR_ARG, "null params");
  if (!e->layered && !e->host_n.empty() && e->host_n[0] >= (float)e->num_types)
    return fail(YMC_ERR_FORMAT, "geometry uses row %d but the tables have %d rows", (int)e->host_n[0], e->num_types);
  if (e->layered) {
    if ((int)e->layered_host.size() != e->num_tables)
      return fail(YMC_ERR_ARG, "layered geometry has %zu tables, optical properties %d", e->layered_host.size(),
                  e->num_tables);
    for (const ymc_layered &g : e->layered_host) {
      if (g.row_top < 0 || g.row_top >= e->num_types || g.row_bottom < 0 || g.row_bottom >= e->num_types ||
          g.row_vessel >= e->num_types)
        return fail(YMC_ERR_FORMAT, "layered geometry names a row outside the %d-row tables", e->num_types);
    }
  }
  if (prm->jump_mode < 0 || prm->jump_mode > 2) return fail(YMC_ERR_ARG, "jump_mode must be 0, 1 or 2");
  if (prm->jump_max_crossings < 0) return fail(YMC_ERR_ARG, "jump_max_crossings must be >= 0");
  if (prm->relaunch_lanes < 0 || prm->relaunch_lanes > 64 || prm->relaunch_wait < 0)
    return fail(YMC_ERR_ARG, "relaunch_lanes must be 0..64 and relaunch_wait >= 0");
  e->params = *prm;
  e->tnx = prm->tally_nx > 0 ? prm->tally_nx : e->nx;
  e->tny = prm->tally_ny > 0 ? prm->tally_ny : e->ny;
  e->tdx = prm->tally_dx > 0 ? prm->tally_dx : e->dx;
  e->tdy = prm->tally_dy > 0 ? prm->tally_dy : e->dy;
  e->nr = prm->radial_bins > 0 ? prm->radial_bins : 0;
  if (e->nr > 0 && !(prm->radial_dr > 0)) return fail(YMC_ERR_ARG, "radial_dr must be > 0");
  const size_t L = e->num_tables;
  const size_t mapn = L * e->tnx * e->tny;
  ymc_status st;
  if ((st = e->map_w.ensure(2 * mapn * 8)) || (st = e->map_w2.ensure(2 * mapn * 8)) ||
      (st = e->map_n.ensure(2 * mapn * 4)) || (st = e->Rr_w.ensure(L * e->nr * 8)) || (st = e->Rr_w2.ensure(L * e->nr * 8)) ||
      (st = e->Rr_n.ensure(L * e->nr * 4)) || (st = e->tot.ensure(L * sizeof(ymc_totals))))
    return st;
  return YMC_OK;
}

static ymc_status upload_chains(ymc_engine *e, const ymc_chain *chains, const uint64_t *x, const uint32_t *a,
                                uint64_t n) {
  ymc_status st;
  if ((st = e->chains.ensure(n * sizeof(ymc_chain))) || (st = e->chain_x.ensure(n * 8)) ||
      (st = e->chain_a.ensure(n * 4)))
    return st;
  HIP_TRY(hipMemcpy(e->chains.p, chains, n * sizeof(ymc_chain), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->chain_x.p, x, n * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->chain_a.p, a, n * 4, hipMemcpyHostToDevice));
  uint64_t total = 0;
  std::vector<uint64_t> first(n), per_table(e->num_tables, 0);
  for (uint64_t c = 0; c < n; ++c) {
    if (chains[c].table >= (uint32_t)e->num_tables) return fail(YMC_ERR_ARG, "chain %llu: table out of range", (unsigned long long)c);
    first[c] = total;
    total += chains[c].count;
    per_table[chains[c].table] += chains[c].count;
  }
  // A packet's weight never exceeds 1 (2^32 in fixed point) and its absorbed weight exceeds its
  // launch weight only by roulette gains of under 1e-3 each, so the u64 weight sums and the u32
  // per-bin counts of a table cannot wrap below this many packets.
  for (int w = 0; w < e->num_tables; ++w)
    if (per_table[w] > kMaxTablePackets)
      return fail(YMC_ERR_ARG, "table %d: %llu packets exceed the %llu a table's 64-bit tallies hold", w,
                  (unsigned long long)per_table[w], (unsigned long long)kMaxTablePackets);
  if (e->params.flags & YMC_FLAG_RECORD) {
    if ((st = e->chain_first.ensure(n * 8)) || (st = e->records.ensure(total * sizeof(ymc_photon_record)))) return st;
    HIP_TRY(hipMemcpy(e->chain_first.p, first.data(), n * 8, hipMemcpyHostToDevice));
    e->n_records = total;
  } else {
    e->n_records = 0;
  }
  e->absorbing = (e->params.flags & YMC_FLAG_ABSORPTION) != 0;
  if (e->absorbing) {
    const size_t bytes = (size_t)e->num_tables * e->nx * e->ny * e->nz * 8;
    if (bytes > kMaxAbsorptionBytes)
      return fail(YMC_ERR_ARG, "absorption grid of %zu bytes exceeds the %zu-byte limit", bytes, kMaxAbsorptionBytes);
    if ((st = e->absorb.ensure(bytes))) return st;
  }
  e->n_chains = n;
  e->n_packets = total;
  e->prepared = true;
  return YMC_OK;
}

ymc_status ymc_engine_prepare(ymc_engine *e, const ymc_params *prm) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if ((st = prepare_common(e, prm))) return st;
  const int L = e->num_tables;
  std::vector<uint64_t> budgets(L);
  for (int w = 0; w < L; ++w) {
    budgets[w] = prm->budgets ? prm->budgets[w] : prm->num_photons;
    if (budgets[w] > kMaxTablePackets)  // checked again per chain list in upload_chains
      return fail(YMC_ERR_ARG, "table %d: budget of %llu packets exceeds %llu", w, (unsigned long long)budgets[w],
                  (unsigned long long)kMaxTablePackets);
  }
  const int lanes = prm->chain_lanes > 0 ? prm->chain_lanes : 524288;
  uint64_t n = 0;
  if ((st = build_chains(budgets.data(), L, lanes, nullptr, 0, &n))) return st;
  std::vector<ymc_chain> ch(n);
  if ((st = build_chains(budgets.data(), L, lanes, ch.data(), n, &n))) return st;
  std::vector<uint64_t> x(n);
  std::vector<uint32_t> a(n);
  if ((st = seed_streams(e->table.data(), (uint32_t)e->table.size(), prm->seed ? prm->seed : 1, n, x.data(), a.data())))
    return st;
  return upload_chains(e, ch.data(), x.data(), a.data(), n);
}

ymc_status ymc_engine_prepare_chains(ymc_engine *e, const ymc_params *prm, const ymc_chain *chains, const uint64_t *x,
                                     const uint32_t *a, uint64_t n_chains) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!chains || !x || !a) return fail(YMC_ERR_ARG, "ymc_engine_prepare_chains: null chain arrays");
  // Streams must be MWC states of their multiplier: a >= 2 and carry*2^32 + value < a*2^32 - 1 (the
  // stream skip's modular arithmetic relies on both).
  for (uint64_t c = 0; c < n_chains; ++c)
    if (a[c] < 2 || x[c] >= (((uint64_t)a[c] << 32) - 1))
      return fail(YMC_ERR_ARG, "chain %llu: state 0x%llx is not an MWC state of multiplier %u", (unsigned long long)c,
                  (unsigned long long)x[c], a[c]);
  if ((st = prepare_common(e, prm))) return st;
  return upload_chains(e, chains, x, a, n_chains);
}

ymc_status ymc_engine_run(ymc_engine *e) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!e->prepared) return fail(YMC_ERR_STATE, "ymc_engine_run: call ymc_engine_prepare first");
  const ymc_params &prm = e->params;
  const size_t L = e->num_tables;
  const size_t mapn = L * e->tnx * e->tny;
  hipStream_t s = e->stream;
  HIP_TRY(hipMemsetAsync(e->map_w.p, 0, 2 * mapn * 8, s));
  HIP_TRY(hipMemsetAsync(e->map_w2.p, 0, 2 * mapn * 8, s));
  HIP_TRY(hipMemsetAsync(e->map_n.p, 0, 2 * mapn * 4, s));
  if (e->nr) {
    HIP_TRY(hipMemsetAsync(e->Rr_w.p, 0, L * e->nr * 8, s));
    HIP_TRY(hipMemsetAsync(e->Rr_w2.p, 0, L * e->nr * 8, s));
    HIP_TRY(hipMemsetAsync(e->Rr_n.p, 0, L * e->nr * 4, s));
  }
  HIP_TRY(hipMemsetAsync(e->tot.p, 0, L * sizeof(ymc_totals), s));
  HIP_TRY(hipMemsetAsync(e->counter.p, 0, 8, s));
  if (e->n_records) HIP_TRY(hipMemsetAsync(e->records.p, 0, e->n_records * sizeof(ymc_photon_record), s));
  const size_t nvox = (size_t)e->nx * e->ny * e->nz;
  if (e->absorbing) HIP_TRY(hipMemsetAsync(e->absorb.p, 0, L * nvox * 8, s));

  KArgs a{};
  a.vox16 = (e->vox16 && !e->layered) ? e->vox.as<uint16_t>() : nullptr;
  a.vox32 = (!e->vox16 && !e->layered) ? e->vox.as<uint32_t>() : nullptr;
  const bool w16 = a.vox16 != nullptr;
  a.w_row = w16 ? 7u : 0xffu;
  a.w_zup = w16 ? 3u : 8u;
  a.w_zdn = w16 ? 7u : 16u;
  a.w_zbits = w16 ? 4u : 8u;
  a.w_rxy = w16 ? 11u : 24u;
  a.layered = e->layered_buf.as<ymc_layered>();
  a.nx = e->nx;
  a.ny = e->ny;
  a.nz = e->nz;
  a.dx = e->dx;
  a.dy = e->dy;
  a.dz = e->dz;
  a.Lx = e->dx * (float)e->nx;
  a.idx = 1.0f / e->dx;
  a.idy = 1.0f / e->dy;
  a.idz = 1.0f / e->dz;
  a.Ly = e->dy * (float)e->ny;
  a.rows = e->rows.as<RowConst>();
  a.w0 = e->w0.as<float>();
  a.num_types = e->num_types;
  a.num_tables = e->num_tables;
  a.px = prm.pencil_x >= 0 ? prm.pencil_x : a.Lx / 2;
  a.py = prm.pencil_y >= 0 ? prm.pencil_y : a.Ly / 2;
  a.uz_thresh = largest_float_le(YMC_UZ_THRESH);
  a.zero_thresh = (float)(1 - YMC_UZ_THRESH);
  a.w_thresh = smallest_float_ge(YMC_W_THRESHOLD);
  a.fresnel_exact = prm.fresnel_mode == YMC_FRESNEL_EXACT;
  a.kill0 = prm.kill_zero_step;
  a.tnx = e->tnx;
  a.tny = e->tny;
  a.tdx = e->tdx;
  a.tdy = e->tdy;
  a.nr = e->nr;
  a.rdr = prm.radial_dr;
  a.rx0 = prm.radial_x0 >= 0 ? prm.radial_x0 : a.px;
  a.ry0 = prm.radial_y0 >= 0 ? prm.radial_y0 : a.py;
  a.map_w = e->map_w.as<unsigned long long>();
  a.map_w2 = e->map_w2.as<unsigned long long>();
  a.map_n = e->map_n.as<unsigned int>();
  a.map_half = mapn;
  a.Rr_w = e->Rr_w.as<unsigned long long>();
  a.Rr_w2 = e->Rr_w2.as<unsigned long long>();
  a.Rr_n = e->Rr_n.as<unsigned int>();
  a.tot = e->tot.as<unsigned long long>();
  a.chains = e->chains.as<ymc_chain>();
  a.chain_x = e->chain_x.as<uint64_t>();
  a.chain_a = e->chain_a.as<uint32_t>();
  a.chain_first = e->chain_first.as<uint64_t>();
  a.next_chain = e->counter.as<unsigned long long>();
  a.n_chains = e->n_chains;
  *e->cancel_host = (prm.cancel && *prm.cancel) ? 1 : 0;
  a.cancel = e->cancel_dev;
  a.poll_interval = prm.compact_interval > 0 ? prm.compact_interval : 1024;
  const bool jump_on = prm.jump_mode != 2;
  const int jump_cap = prm.jump_max_crossings > 0 ? prm.jump_max_crossings : kSkipMax;
  a.relaunch_lanes = prm.relaunch_lanes > 0 ? prm.relaunch_lanes : 12;
  a.relaunch_wait = prm.relaunch_wait > 0 ? prm.relaunch_wait : 16;
  a.records = e->records.as<ymc_photon_record>();
  a.n_records = e->n_records;
  a.abs_grid = e->absorbing ? e->absorb.as<unsigned long long>() : nullptr;
  a.nvox = (uint32_t)nvox;

  const bool uniform = prm.illumination == YMC_ILLUM_UNIFORM;
  const bool counters = (prm.flags & YMC_FLAG_COUNTERS) != 0;
  const bool record = (prm.flags & YMC_FLAG_RECORD) != 0;
  const void *kern = select_kernel(uniform, counters, e->nr > 0, record, e->layered);
  const size_t lds0 = lds_bytes(e->num_tables, e->num_types, e->layered);
  if (lds0 > 64 * 1024) return fail(YMC_ERR_ARG, "optical-property tables need %zu B of LDS (max 65536)", lds0);
  // The per-lane stream-skip table goes into LDS when it costs no occupancy.
  const size_t skip_off = (lds0 + 7) & ~(size_t)7;
  const size_t lds1 = skip_off + (size_t)kSkipEntries * kBlock * 8;
  int per_cu = 0, per_cu_skip = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds0));
  if (lds1 <= 64 * 1024) HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_skip, kern, kBlock, lds1));
  const bool skip_table = jump_on && per_cu_skip >= per_cu && per_cu_skip > 0;
  a.skip_off = skip_table ? (int)skip_off : -1;
  // Box flights off: a cap of -1 fails every box flight's check, and the step redoes it in the voxel.
  a.jump_cap = !jump_on ? -1 : (skip_table ? std::min(jump_cap, kSkipMax) : jump_cap);
  const size_t lds = skip_table ? lds1 : lds0;
  e->last_skip_table = skip_table;
  if (per_cu < 1) per_cu = 1;
  int blocks = per_cu * e->num_cus;
  if (prm.pool_lanes > 0) blocks = std::max(1, prm.pool_lanes / kBlock);
  e->last_pool = blocks * kBlock;
  void *args[] = {&a};
  HIP_TRY(hipEventRecord(e->ev0, s));
  HIP_TRY(hipLaunchKernel(kern, dim3(blocks), dim3(kBlock), args, lds, s));
  HIP_TRY(hipEventRecord(e->ev1, s));
  // Wait for the kernel, forwarding the caller's cancel flag to the pinned copy it polls.
  for (;;) {
    hipError_t q = hipStreamQuery(s);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) return fail(YMC_ERR_DEVICE, "transport kernel failed: %s", hipGetErrorString(q));
    if (prm.cancel && *prm.cancel && !*e->cancel_host) *e->cancel_host = 1;
    std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->last_ms = ms;
  if (prm.cancel && *prm.cancel) return fail(YMC_ERR_CANCELLED, "run cancelled; tallies hold the finished packets");
  return YMC_OK;
}

ymc_status ymc_engine_read_tally(ymc_engine *e, ymc_tally *out) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!out) return fail(YMC_ERR_ARG, "null tally");
  if (!e->prepared) return fail(YMC_ERR_STATE, "no run prepared");
  const size_t L = e->num_tables;
  const size_t mapn = L * e->tnx * e->tny;
  HIP_TRY(hipStreamSynchronize(e->stream));
  auto cp = [&](void *dst, const DevBuf &src, size_t bytes, size_t offset = 0) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpy(dst, static_cast<const char *>(src.p) + offset, bytes, hipMemcpyDeviceToHost);
  };
  HIP_TRY(cp(out->R_w, e->map_w, mapn * 8));
  HIP_TRY(cp(out->R_w2, e->map_w2, mapn * 8));
  HIP_TRY(cp(out->R_n, e->map_n, mapn * 4));
  HIP_TRY(cp(out->T_w, e->map_w, mapn * 8, mapn * 8));
  HIP_TRY(cp(out->T_w2, e->map_w2, mapn * 8, mapn * 8));
  HIP_TRY(cp(out->T_n, e->map_n, mapn * 4, mapn * 4));
  HIP_TRY(cp(out->Rr_w, e->Rr_w, L * e->nr * 8));
  HIP_TRY(cp(out->Rr_w2, e->Rr_w2, L * e->nr * 8));
  HIP_TRY(cp(out->Rr_n, e->Rr_n, L * e->nr * 4));
  HIP_TRY(cp(out->totals, e->tot, L * sizeof(ymc_totals)));
  return YMC_OK;
}

ymc_status ymc_engine_read_absorption(ymc_engine *e, uint64_t *out, uint64_t count) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!out) return fail(YMC_ERR_ARG, "null output");
  if (!e->prepared || !e->absorbing) return fail(YMC_ERR_STATE, "last run did not tally absorption (YMC_FLAG_ABSORPTION)");
  const uint64_t n = (uint64_t)e->num_tables * e->nx * e->ny * e->nz;
  if (count != n) return fail(YMC_ERR_ARG, "absorption grid has %llu entries, not %llu", (unsigned long long)n,
                              (unsigned long long)count);
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(out, e->absorb.p, n * 8, hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_engine_read_records(ymc_engine *e, ymc_photon_record *out, uint64_t cap) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!out) return fail(YMC_ERR_ARG, "null output");
  if (!e->n_records) return fail(YMC_ERR_STATE, "last run did not record packets (YMC_FLAG_RECORD)");
  uint64_t n = std::min(cap, e->n_records);
  HIP_TRY(hipMemcpy(out, e->records.p, n * sizeof(ymc_photon_record), hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_engine_last_kernel_ms(ymc_engine *e, double *ms) {
  if (!e || !ms) return fail(YMC_ERR_ARG, "null argument");
  *ms = e->last_ms;
  return YMC_OK;
}

ymc_status ymc_engine_pool_lanes(ymc_engine *e, int32_t *lanes) {
  if (!e || !lanes) return fail(YMC_ERR_ARG, "null argument");
  *lanes = e->last_pool;
  return YMC_OK;
}

ymc_status ymc_engine_rng_fill(ymc_engine *e, const uint64_t *x, const uint32_t *a, uint32_t n_streams,
                               uint32_t n_draws, int32_t kind, float *out, uint64_t *x_final) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!x || !a || !out || !x_final || n_streams == 0) return fail(YMC_ERR_ARG, "ymc_engine_rng_fill: bad arguments");
  DevBuf dx, da, dout, dxf;
  if ((st = dx.ensure((size_t)n_streams * 8)) || (st = da.ensure((size_t)n_streams * 4)) ||
      (st = dout.ensure((size_t)n_streams * n_draws * 4 + 4)) || (st = dxf.ensure((size_t)n_streams * 8)))
    return st;
  HIP_TRY(hipMemcpy(dx.p, x, (size_t)n_streams * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(da.p, a, (size_t)n_streams * 4, hipMemcpyHostToDevice));
  const uint32_t blk = 256, grid = (n_streams + blk - 1) / blk;
  hipLaunchKernelGGL(rng_fill_kernel, dim3(grid), dim3(blk), 0, e->stream, dx.as<uint64_t>(), da.as<uint32_t>(),
                     n_streams, n_draws, kind, dout.as<float>(), dxf.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(out, dout.p, (size_t)n_streams * n_draws * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(x_final, dxf.p, (size_t)n_streams * 8, hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_engine_layered_words(ymc_engine *e, const ymc_layered *g, int32_t nx, int32_t ny, int32_t nz,
                                    uint32_t *out) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!g || !out || nx <= 0 || ny <= 0 || nz <= 0 || (size_t)nx * ny * nz >= (1ull << 31) ||
      std::max(nx, std::max(ny, nz)) >= kMaxAxis || !(g->vessel_r >= 0 && g->vessel_r < kMaxVesselRadius))
    return fail(YMC_ERR_ARG, "ymc_engine_layered_words: bad arguments");
  const size_t n = (size_t)nx * ny * nz;
  DevBuf dout;
  if ((st = dout.ensure(n * 4))) return st;
  KArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  hipLaunchKernelGGL(layered_words_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, e->stream, a, *g,
                     dout.as<uint32_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(out, dout.p, n * 4, hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_engine_rng_skip(ymc_engine *e, const uint64_t *x, const uint32_t *a, uint32_t n_streams,
                               uint32_t crossings, uint64_t *x_final) {
  ymc_status st = check_engine(e);
  if (st) return st;
  if (!x || !a || !x_final || n_streams == 0) return fail(YMC_ERR_ARG, "ymc_engine_rng_skip: bad arguments");
  if (crossings > (uint32_t)kSkipMax) return fail(YMC_ERR_ARG, "ymc_engine_rng_skip: crossings must be <= %d", kSkipMax);
  for (uint32_t q = 0; q < n_streams; ++q)
    if (a[q] < 2 || x[q] >= (((uint64_t)a[q] << 32) - 1))
      return fail(YMC_ERR_ARG, "ymc_engine_rng_skip: stream %u is not an MWC state of its multiplier", q);
  DevBuf dx, da, dxf;
  if ((st = dx.ensure((size_t)n_streams * 8)) || (st = da.ensure((size_t)n_streams * 4)) ||
      (st = dxf.ensure((size_t)n_streams * 8)))
    return st;
  HIP_TRY(hipMemcpy(dx.p, x, (size_t)n_streams * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(da.p, a, (size_t)n_streams * 4, hipMemcpyHostToDevice));
  const uint32_t grid = (n_streams + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(rng_skip_kernel, dim3(grid), dim3(kBlock), 0, e->stream, dx.as<uint64_t>(), da.as<uint32_t>(),
                     n_streams, (int)crossings, dxf.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(x_final, dxf.p, (size_t)n_streams * 8, hipMemcpyDeviceToHost));
  return YMC_OK;
}

ymc_status ymc_probe_fma_rate(int32_t device, int32_t waves_per_simd, int32_t operands, int32_t iters,
                              double *tflops, double *ms) {
  if (!tflops || waves_per_simd < 1 || waves_per_simd > 8 || operands < 1 || operands > 3 || iters < 1)
    return fail(YMC_ERR_ARG, "ymc_probe_fma_rate: waves_per_simd 1..8, operands 1..3, iters >= 1");
  auto kern = operands == 1 ? fma_probe_kernel<1> : operands == 2 ? fma_probe_kernel<2> : fma_probe_kernel<3>;
  HIP_TRY(hipSetDevice(device));
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const uint32_t blocks = (uint32_t)cus * (uint32_t)waves_per_simd;  // 4 waves per block, one per SIMD
  DevBuf out;
  ymc_status st;
  if ((st = out.ensure((size_t)blocks * 256 * 4))) return st;
  struct Events {  // destroyed on every return
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HIP_TRY(hipEventCreate(&ev.e0));
  HIP_TRY(hipEventCreate(&ev.e1));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, nullptr, out.as<float>(), iters, 0.999999f, 1e-7f);
  HIP_TRY(hipEventRecord(ev.e0, nullptr));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, nullptr, out.as<float>(), iters, 0.999999f, 1e-7f);
  HIP_TRY(hipEventRecord(ev.e1, nullptr));
  HIP_TRY(hipEventSynchronize(ev.e1));
  HIP_TRY(hipGetLastError());
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, ev.e0, ev.e1));
  const double flops = 2.0 * blocks * 256.0 * iters * kProbeUnroll * kProbeChains;
  *tflops = flops / (t * 1e-3) / 1e12;
  if (ms) *ms = t;
  return YMC_OK;
}

ymc_status ymc_run_3dmc_gpu(int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz, const int32_t *tissue,
                            int32_t num_types, const float *mua, const float *mus, const float *g, const float *n,
                            int32_t num_photons, uint64_t seed, const uint32_t *table, uint32_t table_len, double *R,
                            double *R_tot, double *T, double *T_tot, float *A, int32_t *num_finished) {
  (void)A;
  if (!tissue || !R || !T || !R_tot || !T_tot || !num_finished || num_photons < 0)
    return fail(YMC_ERR_ARG, "ymc_run_3dmc_gpu: null argument");
  size_t nv = (size_t)nx * ny * nz;
  std::vector<uint8_t> vox(nv);
  for (size_t v = 0; v < nv; ++v) {
    if (tissue[v] < 0 || tissue[v] >= num_types || tissue[v] > 255)
      return fail(YMC_ERR_FORMAT, "voxel %zu: row %d outside the optical-property table", v, tissue[v]);
    vox[v] = (uint8_t)tissue[v];
  }
  ymc_engine *e = nullptr;
  ymc_status st = ymc_engine_create(0, table, table_len, &e);
  if (st) return st;
  struct Guard {
    ymc_engine *e;
    ~Guard() { ymc_engine_destroy(e); }
  } guard{e};
  ymc_geometry geo{nx, ny, nz, dx, dy, dz, vox.data()};
  ymc_optprops op{num_types, 1, mua, mus, g, n};
  if ((st = ymc_engine_set_geometry(e, &geo)) || (st = ymc_engine_set_optprops(e, &op))) return st;
  ymc_params prm{};
  prm.num_photons = (uint64_t)num_photons;
  prm.seed = seed;
  prm.illumination = YMC_ILLUM_UNIFORM;
  prm.pencil_x = prm.pencil_y = -1;
  prm.radial_x0 = prm.radial_y0 = -1;
  prm.kill_zero_step = 1;
  if ((st = ymc_engine_prepare(e, &prm)) || (st = ymc_engine_run(e))) return st;
  std::vector<uint64_t> rw(nv / nz), tw(nv / nz);
  ymc_totals tt{};
  ymc_tally out{};
  out.R_w = rw.data();
  out.T_w = tw.data();
  out.totals = &tt;
  if ((st = ymc_engine_read_tally(e, &out))) return st;
  for (size_t b = 0; b < rw.size(); ++b) {
    R[b] += (double)rw[b] / YMC_FX_ONE;  // accumulate like the reference's tally
    T[b] += (double)tw[b] / YMC_FX_ONE;
  }
  *R_tot = (double)tt.R_w / YMC_FX_ONE;
  *T_tot = (double)tt.T_w / YMC_FX_ONE;
  *num_finished = (int32_t)tt.photons;
  return YMC_OK;
}

ymc_status ymc_run(int32_t device, const uint32_t *table, uint32_t table_len, const ymc_geometry *geom,
                   const ymc_optprops *props, const ymc_params *params, ymc_tally *out, uint64_t *absorption) {
  if (!geom || !props || !params || !out) return fail(YMC_ERR_ARG, "ymc_run: null argument");
  if (((params->flags & YMC_FLAG_ABSORPTION) != 0) != (absorption != nullptr))
    return fail(YMC_ERR_ARG, "ymc_run: pass an absorption buffer exactly when YMC_FLAG_ABSORPTION is set");
  if (params->flags & YMC_FLAG_RECORD) return fail(YMC_ERR_ARG, "ymc_run: records need the engine API");
  ymc_engine *e = nullptr;
  ymc_status st = ymc_engine_create(device, table, table_len, &e);
  if (st) return st;
  struct Guard {
    ymc_engine *e;
    ~Guard() { ymc_engine_destroy(e); }
  } guard{e};
  if ((st = ymc_engine_set_geometry(e, geom)) || (st = ymc_engine_set_optprops(e, props)) ||
      (st = ymc_engine_prepare(e, params)))
    return st;
  st = ymc_engine_run(e);
  if (st != YMC_OK && st != YMC_ERR_CANCELLED) return st;
  ymc_status rs = ymc_engine_read_tally(e, out);
  if (rs) return rs;
  if (absorption) {
    const uint64_t n = (uint64_t)props->num_tables * geom->nx * geom->ny * geom->nz;
    if ((rs = ymc_engine_read_absorption(e, absorption, n))) return rs;
  }
  return st;
}

}  // extern "C"
