/*
 * ymc3d.h - C ABI of the MI355X photon-transport engine.
 *
 * The engine replaces the reference's host run loop
 *   void run_3dmc_gpu(geometry_t, opticalprops_t, int num_photons, double *R, double *R_tot,
 *                     double *T, double *T_tot, float *A, int *num_finished_photons)
 * (src/mc3d_gpu.h:16, defined src/mc3d_gpu.cu:388-601) together with the two kernels it
 * launches (photontrack_reinitialize src/mc3d_gpu.cu:74-111, photontrack_step :128-376), the
 * RNG seeding it calls (rng_init src/mc3d_rng.cu:12-37 -> init_RNG gpumcml_rng/gpumcml_rng.cu:59-118)
 * and the host tally (photon_detector src/mc3d_photons.cu:151-171).
 *
 * Conventions:
 *  - Plain C types only; every pointer argument is a host pointer unless the name says "dev".
 *  - Every entry point returns an ymc_status; nothing calls exit() or installs signal handlers.
 *  - The caller owns every output buffer; the engine owns every device allocation.
 *  - Weights are tallied as unsigned 64-bit fixed point with scale 2^32 (YMC_FX_ONE), so the
 *    sums do not depend on the order in which packets finish.
 *  - One engine is used by one host thread at a time; distinct engines may run concurrently.
 */
#ifndef YMC3D_H
#define YMC3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YMC3D_ABI_VERSION 1

/* Fixed-point scale of every weight sum: value = raw / 2^32. */
#define YMC_FX_ONE 4294967296.0

/* Row 0 of every optical-property table is the ambient medium (src/mc3d_io.cu:38-42). */
#define YMC_TISSUE_AIR 0

/* Physics constants of the reference (src/mc3d_types.h:11-13). */
#define YMC_SURVIVAL_CHANCE 0.1
#define YMC_W_THRESHOLD 0.0001
#define YMC_UZ_THRESH 0.99999

typedef enum ymc_status {
  YMC_OK = 0,
  YMC_ERR_ARG = 1,       /* null pointer, bad size or out-of-range parameter */
  YMC_ERR_IO = 2,        /* file missing, unreadable or short */
  YMC_ERR_FORMAT = 3,    /* header or content violates the file format */
  YMC_ERR_SEED = 4,      /* seed rejected by the init_RNG validity rule (gpumcml_rng.cu:88) */
  YMC_ERR_TABLE = 5,     /* multiplier table missing, too short or invalid */
  YMC_ERR_DEVICE = 6,    /* HIP runtime error (message via ymc_last_error) */
  YMC_ERR_NOMEM = 7,     /* host or device allocation failed */
  YMC_ERR_CANCELLED = 8, /* run stopped by the cancel flag; tallies hold the finished packets */
  YMC_ERR_STATE = 9      /* call order violated (e.g. run before geometry was set) */
} ymc_status;

enum { YMC_ILLUM_UNIFORM = 0, YMC_ILLUM_PENCIL = 1 };
enum { YMC_FRESNEL_YMC3D = 0, YMC_FRESNEL_EXACT = 1 };

/* Run flags. */
#define YMC_FLAG_COUNTERS 0x1u /* collect steps / Fresnel / absorption / roulette counters */
#define YMC_FLAG_RECORD 0x2u   /* write one ymc_photon_record per packet (test hook) */
/* Tally the absorbed weight per voxel: the deposit the reference has disabled
 * (src/mc3d_gpu.cu:265-272), written by save_abs_map (src/mc3d_io.cu:229-255). */
#define YMC_FLAG_ABSORPTION 0x4u

/* Voxel geometry.  Mirrors geometry_t (src/mc3d_types.h:38-43) after geometry_read_from_file
 * (src/mc3d_io.cu:107-169) has remapped the file values: file value v in 0..254 becomes row
 * v+1, file value 255 becomes row 0 (ambient).  Rows therefore fit in uint8. */
typedef struct ymc_geometry {
  int32_t nx, ny, nz;     /* voxel counts: each below 2^24, fewer than 2^31 in all */
  float dx, dy, dz;       /* voxel size, same length unit as 1/mu */
  const uint8_t *tissue;  /* nz*ny*nx rows, x fastest */
} ymc_geometry;

/* Parametric geometry evaluated per voxel on the device: a top layer over a semi-infinite bottom
 * layer with a cylinder parallel to y.  Voxel (i,j,k) has its centre at (i+0.5, j+0.5, k+0.5) in
 * voxel units; it belongs to the cylinder when (i+0.5-vessel_x)^2 + (k+0.5-vessel_z)^2 <=
 * vessel_r^2 (binary32, each operation rounded), else to the top layer when k < top_voxels, else to
 * the bottom layer.  Used for batched runs in which every table has its own geometry (the voxel
 * grid these parameters describe is what ymc_layered_voxelize writes).  No counterpart in the
 * reference, which reads voxel grids only (src/mc3d_io.cu:107-169). */
typedef struct ymc_layered {
  int32_t top_voxels;            /* thickness of the top layer in voxels */
  int32_t row_top, row_bottom;   /* optical-property rows of the layers */
  int32_t row_vessel;            /* row of the cylinder, or -1 for none */
  float vessel_x, vessel_z;      /* cylinder axis, voxel units */
  float vessel_r;                /* cylinder radius, voxel units, below 2^14 */
  float reserved;
} ymc_layered;

/* Optical properties.  Mirrors opticalprops_t (src/mc3d_types.h:28-36): num_types rows
 * including row 0 = ambient (mua=mus=g=0, n=1).  num_tables > 1 stacks independent
 * tables (wavelengths / property sets) for a batched run. */
typedef struct ymc_optprops {
  int32_t num_types;   /* rows per table, including row 0 */
  int32_t num_tables;  /* >= 1 */
  const float *mua;    /* [num_tables * num_types] absorption coefficient */
  const float *mus;    /* scattering coefficient */
  const float *g;      /* Henyey-Greenstein anisotropy */
  const float *n;      /* refractive index */
} ymc_optprops;

/* Run parameters.  Zero-initialise and set what you need; zeros select the defaults. */
typedef struct ymc_params {
  uint64_t num_photons;        /* packets per table (used when budgets == NULL); at most 2^31 per table,
                                  beyond which the 64-bit weight sums could wrap (YMC_ERR_ARG) */
  const uint64_t *budgets;     /* optional [num_tables] packet budgets, each at most 2^31 */
  uint64_t seed;               /* init_RNG seed; 0 selects 1 */
  const volatile int32_t *cancel; /* optional host flag polled by the device; nonzero stops launches */
  int32_t illumination;        /* YMC_ILLUM_UNIFORM (src/mc3d_gpu.cu:88-89) or _PENCIL (:92-94) */
  float pencil_x, pencil_y;    /* beam position; negative selects the grid centre */
  int32_t tally_nx, tally_ny;  /* reflectance/transmittance grid; 0 selects geometry nx, ny */
  float tally_dx, tally_dy;    /* bin size; 0 selects geometry dx, dy */
  int32_t radial_bins;         /* radial reflectance bins; 0 disables the radial tally */
  float radial_dr;             /* radial bin width */
  float radial_x0, radial_y0;  /* radial centre; negative selects the beam position */
  int32_t fresnel_mode;        /* YMC_FRESNEL_YMC3D (reference formula) or _EXACT */
  int32_t kill_zero_step;      /* zero-step kill of src/mc3d_gpu.cu:343: 1 = at the end of each 45-step
                                  launch as the reference (default), 2 = every step, 0 = never */
  int32_t chain_lanes;         /* nominal lane count of the chain schedule; 0 selects 524288.  The
                                  first chains hold ceil(N / (2*chain_lanes)) packets: a lane runs a
                                  chain's packets back to back before it returns to the work queue,
                                  so this sets how often finished lanes are refilled (the pool's
                                  compaction interval) */
  int32_t pool_lanes;          /* resident lanes (device); 0 selects full occupancy */
  int32_t compact_interval;    /* loop iterations between polls of the cancel flag; 0 selects 1024 */
  uint32_t flags;              /* YMC_FLAG_* */
  int32_t jump_mode;           /* safe-box flights: 0 selects default (on), 1 on, 2 off.  A step
                                  takes the plain crossings of its flight inside the anchor voxel's
                                  safe box (faces between voxels of one tissue) together, drawing
                                  what those crossings draw */
  int32_t jump_max_crossings;  /* most crossings one box flight may take; 0 selects 63.  At most 63
                                  when their draws are skipped with the per-lane LDS table, which
                                  the engine uses whenever it costs no occupancy (so the default's
                                  results do not depend on that choice) */
  int32_t relaunch_lanes;      /* a lane whose packet ended waits until this many lanes of its wave
                                  wait (tally, next chain, launch run for all at once); 0 selects 12,
                                  1 launches at once */
  int32_t relaunch_wait;       /* ... or until some lane has waited this many loop iterations;
                                  0 selects 16 */
} ymc_params;

/* Per-table totals (fixed point, scale YMC_FX_ONE). */
typedef struct ymc_totals {
  uint64_t photons;     /* packets finished */
  uint64_t R_w, R_w2;   /* reflected weight (z <= 0 at exit), anywhere, and its square */
  uint64_t T_w, T_w2;   /* weight that finished with z > 0 */
  uint64_t R_n, T_n;    /* packets with positive weight in R / T */
  uint64_t A_w;         /* absorbed weight (implicit capture, src/mc3d_gpu.cu:262-263) */
  int64_t roulette_w;   /* net weight added by Russian roulette (src/mc3d_gpu.cu:275-277) */
  uint64_t killed_w;    /* weight removed by the non-finite / zero-step kill (:343-353) */
  uint64_t killed;      /* packets killed that way */
  uint64_t steps;       /* photon steps (iterations of src/mc3d_gpu.cu:175-340) */
  uint64_t fresnel;     /* Fresnel events (src/mc3d_gpu.cu:226-257) */
} ymc_totals;

/* Tally destination; any pointer may be NULL.  Sizes: maps num_tables*tally_ny*tally_nx
 * (x fastest), radial num_tables*radial_bins, totals num_tables. */
typedef struct ymc_tally {
  uint64_t *R_w, *R_w2;   /* per-bin reflectance weight sums (src/mc3d_photons.cu:160-161) */
  uint32_t *R_n;
  uint64_t *T_w, *T_w2;   /* per-bin transmittance (src/mc3d_photons.cu:166-167) */
  uint32_t *T_n;
  uint64_t *Rr_w, *Rr_w2; /* radial reflectance */
  uint32_t *Rr_n;
  ymc_totals *totals;
} ymc_tally;

/* Final state of one packet (YMC_FLAG_RECORD). */
typedef struct ymc_photon_record {
  float x, y, z, w;  /* exit position and weight */
  uint32_t steps;    /* steps taken by this packet */
  int32_t fate;      /* 0 reflected (z<=0), 1 finished with z>0, 2 killed */
} ymc_photon_record;

/* One chain: `count` consecutive packets of table `table`, all drawn from one RNG stream. */
typedef struct ymc_chain {
  uint32_t count;
  uint32_t table;
} ymc_chain;

/* ------------------------------------------------------------------------------------------ */
/* Host utilities (no device needed).                                                         */

const char *ymc_last_error(void);
int32_t ymc_abi_version(void);

/* Multiplier table.  Entry 0 is the seeding multiplier `begin`, entries 1.. are the stream
 * multipliers.  Replaces the file read of init_RNG (gpumcml_rng/gpumcml_rng.cu:72-98).
 * ymc_multipliers_generate writes the `count` largest a < 2^32 with a*2^32-1 and a*2^31-1
 * both prime, in descending order. */
ymc_status ymc_multipliers_generate(uint32_t count, uint32_t *out);
/* Reads the first column of a safeprimes_base32.txt-style text file (up to cap lines). */
ymc_status ymc_multipliers_read_text(const char *path, uint32_t *out, uint32_t cap, uint32_t *count);

/* Stream seeding: init_RNG (gpumcml_rng/gpumcml_rng.cu:59-118) for n streams; stream i uses
 * multiplier table[1 + i % (table_len-1)].  Returns YMC_ERR_SEED for an invalid seed. */
ymc_status ymc_seed_streams(const uint32_t *table, uint32_t table_len, uint64_t seed,
                            uint64_t n, uint64_t *x_out, uint32_t *a_out);

/* Chain schedule: splits the budgets into chains in dispatch order (guided: chain length
 * ceil(remaining / (2*chain_lanes)), tables interleaved).  Call with out == NULL to get the
 * count.  The schedule depends only on the budgets and chain_lanes. */
ymc_status ymc_build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes,
                            ymc_chain *out, uint64_t cap, uint64_t *count);

/* Binary formats (src/mc3d_io.cu).  Geometry/optprops readers allocate with malloc; release
 * with ymc_free.  ymc_read_geometry applies the +1 / 255->0 remap of src/mc3d_io.cu:156-163. */
ymc_status ymc_read_geometry(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dx,
                             double *dy, double *dz, uint8_t **tissue);
ymc_status ymc_write_geometry(const char *path, int32_t nx, int32_t ny, int32_t nz, double dx,
                              double dy, double dz, const int32_t *file_values);
/* Reads T records of (mua, mus, g, n) doubles and prepends the ambient row
 * (src/mc3d_io.cu:16-66); outputs have T+1 entries. */
ymc_status ymc_read_optprops(const char *path, int32_t *num_types, float **mua, float **mus,
                             float **g, float **n);
ymc_status ymc_write_optprops(const char *path, int32_t num_records, const double *records);
/* save_property layout (src/mc3d_io.cu:175-216): int32 dims, int32 nx, ny[, nz], f64 dx, dy[, dz],
 * f64 data. */
ymc_status ymc_write_property(const char *path, int32_t dims, int32_t nx, int32_t ny, int32_t nz,
                              double dx, double dy, double dz, const double *data);
void ymc_free(void *p);

/* Writes the rows of a parametric geometry (see ymc_layered) as a voxel grid [nz][ny][nx]. */
ymc_status ymc_layered_voxelize(const ymc_layered *g, int32_t nx, int32_t ny, int32_t nz, uint8_t *rows);

/* Safe radius per voxel (host): the largest r <= 31 such that every voxel within Chebyshev
 * distance r has the same row, lies inside the lateral grid and not above the top layer (below
 * the grid the bottom layer repeats).  Exported as a test hook.  No counterpart in the
 * reference. */
ymc_status ymc_safe_radius(const uint8_t *rows, int32_t nx, int32_t ny, int32_t nz, uint8_t *out);

/* Safe box per voxel (host): extents zup, zdn <= zcap (above and below) and rxy <= xycap (lateral,
 * square) such that every voxel of [i-rxy, i+rxy] x [j-rxy, j+rxy] x [k-zup, k+zdn] has the same
 * row, lies inside the lateral grid and not above the top layer (below the grid the bottom layer
 * repeats).  Candidates for (zup, zdn): (r, r), (up, dn), (up, r), (r, dn), with r the safe radius
 * and up/dn the vertical runs of the row, each with its widest lateral square; of those
 * containing the safe-radius cube, the one with the largest harmonic mean of rxy+1, rxy+1 and
 * (zup+zdn)/2+1 is kept, the first listed winning ties.  The engine packs it into its voxel words
 * (caps 15/31 for 16-bit words, 255/255 for 32-bit) and a box flight stays inside it; exported as
 * a test hook.  No counterpart in the reference. */
ymc_status ymc_safe_box(const uint8_t *rows, int32_t nx, int32_t ny, int32_t nz, int32_t zcap, int32_t xycap,
                        uint8_t *zup, uint8_t *zdn, uint8_t *rxy);

/* ------------------------------------------------------------------------------------------ */
/* Device engine.                                                                             */

typedef struct ymc_engine ymc_engine;

/* Creates an engine on HIP device `device`.  The multiplier table is copied; table_len >= 2. */
ymc_status ymc_engine_create(int32_t device, const uint32_t *table, uint32_t table_len,
                             ymc_engine **out);
void ymc_engine_destroy(ymc_engine *e);

/* Uploads the voxel grid (the reference's 3-D texture, src/mc3d_gpu.cu:398-421). */
ymc_status ymc_engine_set_geometry(ymc_engine *e, const ymc_geometry *g);
/* Selects a parametric geometry instead of a voxel grid: grid size nx*ny*nz of voxels dx*dy*dz (the
 * same limits as ymc_geometry), geometry per table (num_tables entries, matching the
 * optical-property tables at run time). */
ymc_status ymc_engine_set_layered(ymc_engine *e, int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz,
                                  const ymc_layered *per_table, int32_t num_tables);
/* Uploads the optical-property tables (opticalprops_transfer_to_device, src/mc3d_io.cu:70-86). */
ymc_status ymc_engine_set_optprops(ymc_engine *e, const ymc_optprops *p);

/* Seeds the streams and uploads the chain schedule for `params` (host work, not timed). */
ymc_status ymc_engine_prepare(ymc_engine *e, const ymc_params *params);
/* Same, with an explicit chain list and stream states (test hook / external schedulers). */
ymc_status ymc_engine_prepare_chains(ymc_engine *e, const ymc_params *params,
                                     const ymc_chain *chains, const uint64_t *x, const uint32_t *a,
                                     uint64_t n_chains);
/* Runs the prepared photon budget: clears the device tallies, launches the persistent
 * transport kernel and waits for it.  Tallies stay on the device. */
ymc_status ymc_engine_run(ymc_engine *e);
/* Copies the device tallies into `out`. */
ymc_status ymc_engine_read_tally(ymc_engine *e, ymc_tally *out);
/* Copies the absorption grid of the last run made with YMC_FLAG_ABSORPTION: count =
 * num_tables * nz * ny * nx entries [table][k][j][i] of fixed-point absorbed weight (raw / 2^32),
 * deposited at the voxel an absorbing step started in, clamped to the grid like the tissue
 * lookup.  Its sum equals the totals' A_w. */
ymc_status ymc_engine_read_absorption(ymc_engine *e, uint64_t *out, uint64_t count);
/* Copies per-packet records of the last run made with YMC_FLAG_RECORD. */
ymc_status ymc_engine_read_records(ymc_engine *e, ymc_photon_record *out, uint64_t cap);
/* Milliseconds of the last ymc_engine_run transport kernel, from HIP events on its stream. */
ymc_status ymc_engine_last_kernel_ms(ymc_engine *e, double *ms);
/* Resident lanes used by the last run. */
ymc_status ymc_engine_pool_lanes(ymc_engine *e, int32_t *lanes);

/* RNG test hook: advances n_streams MWC streams by n_draws draws on the device and returns
 * every draw as the float of rand_MWC_co (kind 0, gpumcml_rng.cu:38-46) or rand_MWC_oc
 * (kind 1, :51-54) in out[draw * n_streams + stream], plus the final states. */
ymc_status ymc_engine_rng_fill(ymc_engine *e, const uint64_t *x, const uint32_t *a,
                               uint32_t n_streams, uint32_t n_draws, int32_t kind, float *out,
                               uint64_t *x_final);

/* Parametric-geometry test hook: the device's voxel word of every voxel of the grid g describes,
 * out[(k*ny + j)*nx + i] = row | zup << 8 | zdn << 16 | rxy << 24, the extents of the safe box a
 * box flight may stay in (see ymc_safe_box).  No counterpart in the reference. */
ymc_status ymc_engine_layered_words(ymc_engine *e, const ymc_layered *g, int32_t nx, int32_t ny, int32_t nz,
                                    uint32_t *out);

/* Stream-skip test hook: advances n_streams MWC streams (gpumcml_rng.cu:38-46) past the draws of
 * `crossings` plain voxel crossings (4 draws each, crossings <= 63) with the modular skip table a
 * box flight uses, and returns the final states; they equal those of 4*crossings draws taken one
 * by one.  No counterpart in the reference, whose crossings draw one at a time. */
ymc_status ymc_engine_rng_skip(ymc_engine *e, const uint64_t *x, const uint32_t *a,
                               uint32_t n_streams, uint32_t crossings, uint64_t *x_final);

/* FP32 FMA rate probe: times `iters` x 128 v_fma_f32 per lane (eight independent chains) at
 * `waves_per_simd` waves on every SIMD of `device`, each FMA reading `operands` distinct VGPRs
 * (1: v*v+v; 2: v*m+c with m in an SGPR; 3: v*a+b, all per lane), and returns TFLOP/s (2 per FMA)
 * and the kernel time.  It measures what plain, unpacked FP32 FMA reaches against the chip's
 * published FP32 peak, the transport kernel's roofline denominator.  No counterpart in the
 * reference. */
ymc_status ymc_probe_fma_rate(int32_t device, int32_t waves_per_simd, int32_t operands, int32_t iters,
                              double *tflops, double *ms);

/* One-shot drop-in for run_3dmc_gpu (src/mc3d_gpu.cu:388): same arguments as plain pointers.
 * tissue holds remapped int32 rows (as geometry_t.tissue_type); optprops arrays have
 * num_types entries with row 0 = ambient.  R and T receive raw weight sums per geometry
 * bin (accumulated, as the reference does), R_tot / T_tot the sums over all exits, A is
 * unused (the reference never fills it, src/mc3d_gpu.cu:265), num_finished the packet
 * count.  Uniform illumination and the reference constants; seed and multiplier table
 * as given. */
ymc_status ymc_run_3dmc_gpu(int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz,
                            const int32_t *tissue, int32_t num_types, const float *mua,
                            const float *mus, const float *g, const float *n, int32_t num_photons,
                            uint64_t seed, const uint32_t *table, uint32_t table_len, double *R,
                            double *R_tot, double *T, double *T_tot, float *A,
                            int32_t *num_finished);

/* One-shot run through a temporary engine on `device`: geometry, optical-property tables (one or
 * many, with per-table budgets in params->budgets), parameters and seed as for the engine API;
 * tallies into `out` (caller-sized: tables x tally grid, tables x radial bins, tables totals) and,
 * with YMC_FLAG_ABSORPTION, the absorption grid into `absorption` (tables x nz x ny x nx).  The
 * whole of run_3dmc_gpu (src/mc3d_gpu.cu:388-601) with every parameter explicit; returns
 * YMC_ERR_CANCELLED with the finished packets tallied when params->cancel is raised. */
ymc_status ymc_run(int32_t device, const uint32_t *table, uint32_t table_len, const ymc_geometry *geom,
                   const ymc_optprops *props, const ymc_params *params, ymc_tally *out, uint64_t *absorption);

/* ------------------------------------------------------------------------------------------ */
/* Reflectance surrogate: MLP with GELU hidden layers trained and served on the device with BF16
 * matrix-core products (FP32 accumulation) and Adam on FP32 master weights.  No counterpart in the
 * reference.  Parameters are one flat FP32 vector: for each layer l, W_l [in_l x out_l] row-major
 * followed by b_l [out_l]. */

typedef struct ymc_mlp ymc_mlp;

/* widths: n_widths layer widths, input first (e.g. 12, 256, 256, 256, 33). He initialisation from seed. */
ymc_status ymc_mlp_create(int32_t device, const int32_t *widths, int32_t n_widths, uint64_t seed, ymc_mlp **out);
void ymc_mlp_destroy(ymc_mlp *n);
uint64_t ymc_mlp_num_params(const ymc_mlp *n);
ymc_status ymc_mlp_set_params(ymc_mlp *n, const float *params, uint64_t count);  /* resets Adam state */
ymc_status ymc_mlp_get_params(ymc_mlp *n, float *params, uint64_t count);
/* Outputs y [rows x out] for inputs x [rows x in], in launches of 4096 rows. */
ymc_status ymc_mlp_forward(ymc_mlp *n, const float *x, int64_t rows, float *y);
/* The same inputs through `count` (1..8) networks of identical widths (each <= 256, <= 8 layers) on
 * one device, in one fused launch per 4096 rows: y [count][rows][out]. */
ymc_status ymc_mlp_forward_ensemble(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows, float *y);
/* The same fused ensemble pass, allocating nothing: no hipMalloc, hipFree or allocating library call.
 * No counterpart in the reference; it is for callers that cannot allocate device memory (tight serving
 * loops, code that forbids allocation).  Inputs and outputs are the caller's host buffers as for
 * ymc_mlp_forward_ensemble and the results are identical.  Rows are staged through a 4 MiB buffer
 * that the library holds on each device for the life of the process: each chunk of rows takes
 * in + count * out floats per row, chunks hold up to 4096 rows, and a pass needing more than
 * 32768 floats per row (1048576 / 32) is rejected with YMC_ERR_ARG.  All calls on one device share
 * that buffer: calls that overlap on one device race and must be serialised by the caller. */
ymc_status ymc_mlp_forward_ensemble_noalloc(ymc_mlp *const *nets, int32_t count, const float *x, int64_t rows,
                                            float *y);
/* Timing hook: device milliseconds per fused ensemble pass of `rows` (<= 4096) rows already resident
 * on the device (zero inputs), averaged over `reps` launches timed with HIP events on the first
 * network's stream. */
ymc_status ymc_mlp_time_forward(ymc_mlp *const *nets, int32_t count, int32_t rows, int32_t reps, double *ms_per_pass);
/* Mean-squared-error loss and its gradient (flat, like the parameters) for one batch (test hook). */
ymc_status ymc_mlp_gradients(ymc_mlp *n, const float *x, const float *y, int32_t rows, float *grad, float *loss);
/* Uploads a training set (x [rows x in], y [rows x out]). */
ymc_status ymc_mlp_set_data(ymc_mlp *n, const float *x, const float *y, int64_t rows);
/* Adam epochs over the training set in shuffled minibatches, learning rate cosine-decayed from lr to
 * lr_final; epoch_loss (optional, [epochs]) receives the last minibatch loss of each epoch. */
ymc_status ymc_mlp_train(ymc_mlp *n, int32_t epochs, int32_t batch, float lr, float lr_final, uint64_t seed,
                         float *epoch_loss);

#ifdef __cplusplus
}
#endif

#endif /* YMC3D_H */
