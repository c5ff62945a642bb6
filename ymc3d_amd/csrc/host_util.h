// host_util.h - internal declarations shared by the host translation units.
#pragma once

#include "../../include/ymc3d.h"

namespace ymc {

ymc_status fail(ymc_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Largest safe radius stored per voxel (see safe_radius).
constexpr int kSafeMax = 31;
void safe_radius(const uint8_t *rows, int nx, int ny, int nz, uint8_t *radius);
// Safe box of every voxel (see safe_box): extents zup, zdn <= zcap above and below the voxel and
// lateral half-extent rxy <= xycap, from the cube radius computed by safe_radius.
void safe_box(const uint8_t *rows, int nx, int ny, int nz, const uint8_t *radius, int zcap, int xycap,
              uint8_t *zup, uint8_t *zdn, uint8_t *rxy);

ymc_status layered_voxelize(const ymc_layered *g, int nx, int ny, int nz, uint8_t *rows);
ymc_status generate_multipliers(uint32_t count, uint32_t *out);
ymc_status read_multipliers_text(const char *path, uint32_t *out, uint32_t cap, uint32_t *count);
ymc_status seed_streams(const uint32_t *table, uint32_t table_len, uint64_t seed, uint64_t n, uint64_t *x_out,
                        uint32_t *a_out);
ymc_status build_chains(const uint64_t *budgets, int32_t num_tables, int32_t chain_lanes, ymc_chain *out,
                        uint64_t cap, uint64_t *count);
ymc_status read_geometry(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, double *dx, double *dy,
                         double *dz, uint8_t **tissue);
ymc_status write_geometry(const char *path, int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                          const int32_t *vals);
ymc_status read_optprops(const char *path, int32_t *num_types, float **mua, float **mus, float **g, float **n);
ymc_status write_optprops(const char *path, int32_t num_records, const double *records);
ymc_status write_property(const char *path, int32_t dims, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                          double dz, const double *data);

}  // namespace ymc
