/* ps3d_nc.h -- ps3d's field and field_stats files in the NetCDF classic data model,
 * written and read by host code without libnetcdf: the 64-bit offset format
 * (CDF-2) where a variable's record fits 4 GiB, else the 64-bit data format (CDF-5,
 * e.g. a 1024^3 field record of 8.6 GB).
 *
 * Field files carry dimensions t (unlimited), z (nz+1), y, x, coordinates x, y, z
 * and record variables t, x/y/z_velocity, x/y/z_vorticity and optionally pressure
 * and horizontal_divergence, stored in C order (t, z, y, x).  Host arrays passed
 * here use the library's layout: volume[(ix*ny + iy)*(nz+1) + iz], components of
 * vel/vor contiguous one after another (3 volumes).
 *
 * ps3d writes NetCDF-4 files with a physical_quantities group; the classic format
 * has no groups, so those quantities are not written.
 * Every entry returns 0 or a ps3d_status code; ps3d_nc_last_error() explains.
 */
#ifndef PS3D_NC_H
#define PS3D_NC_H

#include "ps3d_gpu.h"

/* One field_stats record: t, the PS3D_NSTAT statistics (enum ps3d_stat), then the
 * extensions helicity, dt and prefactor. */
#define PS3D_NC_STATS_ROW (PS3D_NSTAT + 4)

#ifdef __cplusplus
extern "C" {
#endif

const char *ps3d_nc_last_error(void);

/* Flags of ps3d_nc_create_fields. */
enum ps3d_nc_flags {
    PS3D_NC_DERIVED = 1,   /* add pressure and horizontal_divergence                 */
    PS3D_NC_CDF5 = 2       /* 64-bit data format even when every record fits CDF-2   */
};

/* = create_netcdf_field_file (src/fields/field_netcdf.f90:65-139) with the axes of
 * write_netcdf_axis_3d (src/netcdf/netcdf_writer.f90:390-432): creates (or
 * truncates) `path`, writes the x, y, z coordinates (lower + i*extent/n).  The
 * format is CDF-2 unless a variable's record reaches 4 GiB (then CDF-5). */
int ps3d_nc_create_fields(const char *path, int nx, int ny, int nz, const double *lower,
                          const double *extent, int flags);

/* = write_netcdf_fields (field_netcdf.f90:177-254): appends one record at time t.
 * vel, vor: 3 volumes each; pres, delta: 1 volume each, required when the file was
 * created with derived fields, ignored otherwise. */
int ps3d_nc_append_fields(const char *path, double t, int nx, int ny, int nz, const double *vel,
                          const double *vor, const double *pres, const double *delta);

/* = read_netcdf_fields (field_netcdf.f90:276-433): the vorticity (3 volumes) and,
 * when vel is not NULL, the velocity of record `step` (counted from 1; -1 or a
 * step past the last reads the last record); t (may be NULL) receives its time.
 * The file's grid must equal (nx, ny, nz). */
int ps3d_nc_read_fields(const char *path, int step, double *t, int nx, int ny, int nz, double *vel,
                        double *vor);

/* = get_netcdf_box (src/netcdf/netcdf_reader.f90:319-337), as setup_domain_and_parameters
 * uses it (src/utils/utils.f90:91-134): the grid and box of a field file from its
 * global attributes ncells, extent and origin. */
int ps3d_nc_grid(const char *path, int *ncells, double *extent, double *origin);

/* Number of records of a field or field_stats file and the time of the last one
 * (t_last may be NULL; it is 0 when there are no records): what ps3d reads to
 * resume an existing file, with get_num_steps and get_time
 * (field_netcdf.f90:79-89, src/netcdf/netcdf_reader.f90:85-143). */
int ps3d_nc_records(const char *path, int *nrecords, double *t_last);

/* = create_netcdf_field_stats_file (src/fields/field_diagnostics_netcdf.f90:124-188):
 * the record variable t and ps3d's 40 statistics in ps3d's order with its names,
 * long names and units (36-75, 453-798), then the extension variables helicity,
 * dt and prefactor; global attributes file_type, ncells, extent and origin
 * (write_netcdf_box, src/netcdf/netcdf_writer.f90:355-369). */
int ps3d_nc_create_stats(const char *path, int nx, int ny, int nz, const double *lower, const double *extent);

/* Assembles a PS3D_NC_STATS_ROW record from the adapt record diag (PS3D_NDIAG, from
 * ps3d_gpu_adapt or a step's row of ps3d_gpu_advance_n / _until) and the field
 * statistics of the same state (PS3D_NSTAT, ps3d_gpu_field_stats): t from diag, the
 * entries ps3d's adapt sets (vortmax .. z_vormean, gmax, umax, vmax, wmax, usgmax,
 * lsgmax, rolling_mean_gmax) from diag, the others from stats. */
int ps3d_nc_stats_row(const double *diag, const double *stats, double *row);

/* = write_netcdf_field_stats (field_diagnostics_netcdf.f90:211-252): appends nrows
 * records of PS3D_NC_STATS_ROW doubles. */
int ps3d_nc_append_stats(const char *path, const double *rows, int nrows);

#ifdef __cplusplus
}
#endif

#endif
