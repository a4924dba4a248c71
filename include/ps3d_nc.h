/* ps3d_nc.h -- ps3d's field and field_stats files (NetCDF classic, 64-bit offset
 * format), written and read by host code without libnetcdf.
 *
 * Field files carry dimensions t (unlimited), z (nz+1), y, x, coordinates x, y, z
 * and record variables t, x/y/z_velocity, x/y/z_vorticity and optionally pressure
 * and horizontal_divergence, stored in C order (t, z, y, x).  Host arrays passed
 * here use the library's layout: volume[(ix*ny + iy)*(nz+1) + iz], components of
 * vel/vor contiguous one after another (3 volumes).
 *
 * ps3d writes NetCDF-4 files with a physical_quantities group; the classic format
 * has no groups, so those quantities are not written.  Fields larger than 4 GiB per
 * variable and record (beyond the 64-bit offset format's limit) are rejected.
 * Every entry returns 0 or a ps3d_status code; ps3d_nc_last_error() explains.
 */
#ifndef PS3D_NC_H
#define PS3D_NC_H

#ifdef __cplusplus
extern "C" {
#endif

const char *ps3d_nc_last_error(void);

/* = create_netcdf_field_file (src/fields/field_netcdf.f90:65-139) with the axes of
 * write_netcdf_axis_3d (src/netcdf/netcdf_writer.f90:390-432): creates (or
 * truncates) `path`, writes the x, y, z coordinates (lower + i*extent/n).
 * with_derived != 0 adds pressure and horizontal_divergence. */
int ps3d_nc_create_fields(const char *path, int nx, int ny, int nz, const double *lower,
                          const double *extent, int with_derived);

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
 * one record variable per entry of the diagnostics record (enum ps3d_diag), named
 * as ps3d names the quantity where it has one (ke, en, vortmax, vortrms, vorch,
 * x/y/z_vormean, gmax, umax, vmax, wmax, usgmax, lsgmax, rolling_mean_gmax). */
int ps3d_nc_create_stats(const char *path);

/* = write_netcdf_field_stats (field_diagnostics_netcdf.f90:211-252): appends nrows
 * diagnostics records (PS3D_NDIAG doubles each, as ps3d_gpu_advance_n returns). */
int ps3d_nc_append_stats(const char *path, const double *rows, int nrows);

#ifdef __cplusplus
}
#endif

#endif
