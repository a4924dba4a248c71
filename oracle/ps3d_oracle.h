/*
 * ps3d_oracle.h -- CPU FP64 restatement of ps3d's pseudo-spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or the timed CPU
 * baseline).  The product library never links it.
 *
 * Layout: every volume is Fortran (0:nz, 0:ny-1, 0:nx-1), z fastest, as in
 * src/fields/fields.f90:59-65; vector fields are 3 such volumes back to back.
 */
#ifndef PS3D_ORACLE_H
#define PS3D_ORACLE_H
#include "../include/ps3d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* stafft semantics on x(m, n), m fastest (src/fft/stafft.f90:196-550).
 * Return 0, or PS3D_E_FACTOR when n has a prime factor other than 2, 3, 5. */
int orc_forfft(int m, int n, double *x);
int orc_revfft(int m, int n, double *x);
int orc_dct(int m, int n, double *x);   /* x(m, 0:n) */
int orc_dst(int m, int n, double *x);   /* x(m, 1:n) */
/* Cyclic Jacobi eigenvalues (src/utils/jacobi.f90:270-305): a[9] column-major
 * (upper triangle used), d[3] out, descending. */
void orc_jacobi_eigenvalues(double *a, double *d);

orc_t *orc_create(const ps3d_params *p);
void orc_destroy(orc_t *o);
const char *orc_last_error(void);
void orc_set_threads(int n);

/* Same operator set and meaning as ps3d_gpu_op (enum ps3d_op). */
int orc_op(orc_t *o, int op, const double *in, double *out);

int orc_set_vorticity(orc_t *o, const double *vor);
int orc_vor2vel(orc_t *o);
int orc_source(orc_t *o);
int orc_adapt(orc_t *o, double t, double *dt, double *diag);
int orc_step(orc_t *o, double *t, double dt);
int orc_advance_n(orc_t *o, double *t, int nsteps, double *diag_series);
int orc_derived(orc_t *o);
/* field_stats record (40 values, NC_KE .. NC_ROMAX) after vor2vel + derived */
int orc_field_stats(orc_t *o, const double *diag, double *stats);
/* impl_rk4 step with the tendency held at src (3 volumes) and vdiss (ny*nx)
 * given, svor (3 volumes) updated in place: test_implicit_rk's setting. */
int orc_rk4_check(orc_t *o, double dt, const double *vdiss, const double *src, double *svor);
/* cn2 step with the tendency held at src and vdiss (ny*nx) given; svor updated
 * in place (the same setting for cn2.f90:92-181). */
int orc_cn2_check(orc_t *o, double dt, const double *vdiss, const double *src, double *svor);
int orc_get(orc_t *o, int which, double *out);
int orc_set(orc_t *o, int which, const double *in);
/* Tables for checks: 0 k2l2 (ny*nx), 1 k2l2i, 2 vhdis, 3 filt (npts), 4 green,
 * 5 phim, 6 phip, 7 thetam, 8 thetap, 9 dthetam, 10 dthetap, 11 gamtop (nz+1),
 * 12 gambot, 13 rkx (nx), 14 rky (ny), 15 rkz (nz+1). Returns element count. */
long orc_table(orc_t *o, int which, double *out);
double orc_viscosity(orc_t *o);

#ifdef __cplusplus
}
#endif
#endif
