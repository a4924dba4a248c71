/*
 * ps3d_gpu.h -- C ABI of the MI355X-native pseudo-spectral time step.
 *
 * One handle owns one grid resident in the HBM of the current HIP device.
 * All host arrays use ps3d's own field layout: Fortran (0:nz, 0:ny-1, 0:nx-1),
 * z fastest, FP64 (src/fields/fields.f90:59-65).  Vector fields are three such
 * volumes back to back, i.e. Fortran (0:nz, 0:ny-1, 0:nx-1, 3).
 *
 * Every entry point returns 0 on success and a nonzero ps3d_status on error;
 * ps3d_gpu_last_error() then gives a message (thread-local).  The reference has
 * no return codes: its errors stop the program through mpi_stop / mpi_exit_on_error
 * (src/mpi/mpi_utils.f90:15-42) or a bare `stop` (src/fft/stafft.f90:86-96).  A
 * Fortran caller maps a nonzero code onto the same mpi_stop message.
 *
 * A handle is single-threaded: calls on one handle must not overlap.  Each handle
 * runs its kernels on its own HIP stream; independent handles may be used
 * concurrently (the batched ensemble relies on this).
 */
#ifndef PS3D_GPU_H
#define PS3D_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum ps3d_status {
    PS3D_OK = 0,
    PS3D_E_ARG = 1,         /* invalid argument / parameter                      */
    PS3D_E_FACTOR = 2,      /* grid length not supported by the transforms       */
    PS3D_E_HIP = 3,         /* HIP runtime error                                 */
    PS3D_E_NOMEM = 4,       /* device allocation failed                          */
    PS3D_E_STATE = 5,       /* call out of order (e.g. step before set_vorticity) */
    PS3D_E_UNSUPPORTED = 6  /* option not implemented                            */
};

/* filtering (src/utils/options.f90:74): */
enum ps3d_filtering { PS3D_FILTER_HOU_LI = 0, PS3D_FILTER_23RULE = 1 };
/* stepper (src/utils/options.f90:24, src/ps3d.f90:90-99): */
enum ps3d_stepper { PS3D_STEPPER_IMPL_RK4 = 0, PS3D_STEPPER_CN2 = 1 };
/* vor_visc%pretype (src/stepper/advance.f90:381-410): */
enum ps3d_pretype {
    PS3D_PRE_CONSTANT = 0, PS3D_PRE_VORCH = 1, PS3D_PRE_BFMAX = 2,
    PS3D_PRE_ROLL_MEAN_MAX_STRAIN = 3, PS3D_PRE_MAX_STRAIN = 4, PS3D_PRE_US_MAX_STRAIN = 5
};
/* vor_visc%length_scale (src/inversion/inversion_utils.f90:157-184): */
enum ps3d_length_scale { PS3D_LS_KOLMOGOROV = 0, PS3D_LS_GEOPHYSICAL = 1 };

/* Run parameters: the &PS3D namelist entries the time step reads
 * (src/utils/options.f90:28-112) plus the grid box of the field file
 * (src/utils/utils.f90:91-134) and the planetary vorticity f_cor
 * (src/utils/physics.f90).  dt_fixed is an extension: <= 0 keeps ps3d's adaptive
 * step (src/stepper/advance.f90:325-333); > 0 uses min(dt_fixed, limit - t). */
typedef struct ps3d_params {
    int nx, ny, nz;
    double lower[3];
    double extent[3];
    int filtering;            /* enum ps3d_filtering                 */
    int stepper;              /* enum ps3d_stepper                   */
    int nnu;                  /* vor_visc%nnu                        */
    double prediss;           /* vor_visc%prediss                    */
    int pretype;              /* enum ps3d_pretype                   */
    int roll_mean_win_size;   /* vor_visc%roll_mean_win_size         */
    int length_scale;         /* enum ps3d_length_scale              */
    double alpha;             /* time%alpha                          */
    double limit;             /* time%limit                          */
    double dt_fixed;          /* extension, see above                */
    double f_cor[3];          /* planetary vorticity (physics.f90)   */
} ps3d_params;

/* Defaults of src/utils/options.f90 (filtering "Hou & Li", impl-diff-rk4,
 * pretype roll-mean-max-strain, window 1000, Kolmogorov, alpha 0.1) with nnu=3,
 * prediss=10 as recommended at options.f90:52-53. */
void ps3d_params_default(ps3d_params *p);

/* Which field a get/set call addresses (host layout as described above). */
enum ps3d_field {
    PS3D_SVOR = 0,    /* 3 volumes: mixed-spectral vorticity (state)            */
    PS3D_VOR = 1,     /* 3 volumes: physical vorticity (after vor2vel)          */
    PS3D_VEL = 2,     /* 3 volumes: physical velocity (after vor2vel)           */
    PS3D_SVEL = 3,    /* 3 volumes: semi-spectral velocity (after vor2vel)      */
    PS3D_SVORTS = 4,  /* 3 volumes: mixed-spectral tendency (after source)      */
    PS3D_PRES = 5,    /* 1 volume : physical pressure (after derived)           */
    PS3D_DELTA = 6    /* 1 volume : physical horizontal divergence              */
};

/* Diagnostics record written by adapt / advance: PS3D_NDIAG doubles. */
enum ps3d_diag {
    PS3D_D_T = 0, PS3D_D_DT, PS3D_D_KE, PS3D_D_EN, PS3D_D_HEL,
    PS3D_D_VORTMAX, PS3D_D_VORTRMS, PS3D_D_VORCH,
    PS3D_D_VORMEAN_X, PS3D_D_VORMEAN_Y, PS3D_D_VORMEAN_Z,
    PS3D_D_GGMAX, PS3D_D_UMAX, PS3D_D_VMAX, PS3D_D_WMAX,
    PS3D_D_USGGMAX, PS3D_D_LSGGMAX, PS3D_D_RGMAX, PS3D_D_PREFACTOR,
    PS3D_NDIAG
};

/* ps3d's field_stats record, in the order of its variables NC_KE .. NC_ROMAX
 * (src/fields/field_diagnostics_netcdf.f90:36-75); PS3D_NSTAT doubles. */
enum ps3d_stat {
    PS3D_ST_KE = 0, PS3D_ST_EN, PS3D_ST_OMAX, PS3D_ST_ORMS, PS3D_ST_OCHAR,
    PS3D_ST_OXMEAN, PS3D_ST_OYMEAN, PS3D_ST_OZMEAN,
    PS3D_ST_KEXY, PS3D_ST_KEZ, PS3D_ST_ENXY, PS3D_ST_ENZ,
    PS3D_ST_OXMIN, PS3D_ST_OYMIN, PS3D_ST_OZMIN, PS3D_ST_OXMAX, PS3D_ST_OYMAX, PS3D_ST_OZMAX,
    PS3D_ST_HEMAX, PS3D_ST_GMAX, PS3D_ST_BFMAX, PS3D_ST_UMAX, PS3D_ST_VMAX, PS3D_ST_WMAX,
    PS3D_ST_USOXMAX, PS3D_ST_LSOXMAX, PS3D_ST_USOYMAX, PS3D_ST_LSOYMAX, PS3D_ST_USOZMAX, PS3D_ST_LSOZMAX,
    PS3D_ST_USUHMAX, PS3D_ST_USGMAX, PS3D_ST_LSGMAX, PS3D_ST_USZRMS, PS3D_ST_USDELRMS,
    PS3D_ST_RGMAX, PS3D_ST_RBFMAX, PS3D_ST_RIMIN, PS3D_ST_ROMIN, PS3D_ST_ROMAX,
    PS3D_NSTAT
};

/* Single operators, for kernel checks (ps3d_gpu_op). Each reads `in` and writes
 * `out` (one volume each, host layout) on the handle's grid. */
enum ps3d_op {
    PS3D_OP_FFTXYP2S = 0,        /* src/fft/sta3dfft.f90:136-194                 */
    PS3D_OP_FFTXYS2P = 1,        /* src/fft/sta3dfft.f90:202-260                 */
    PS3D_OP_DIFFX = 2,           /* src/fft/sta3dfft.f90:304-337                 */
    PS3D_OP_DIFFY = 3,           /* src/fft/sta3dfft.f90:345-377                 */
    PS3D_OP_FFTSINE = 4,         /* src/fft/sta3dfft.f90:264-278                 */
    PS3D_OP_FFTCOSINE = 5,       /* src/fft/sta3dfft.f90:282-296                 */
    PS3D_OP_CENTRAL_DIFFZ = 6,   /* src/inversion/inversion_utils.f90:653-673    */
    PS3D_OP_COMBINE_SEMI = 7,    /* src/inversion/inversion_utils.f90:617-647    */
    PS3D_OP_DECOMPOSE_SEMI = 8,  /* src/inversion/inversion_utils.f90:563-592    */
    PS3D_OP_COMBINE_PHYS = 9,    /* src/inversion/inversion_utils.f90:599-611    */
    PS3D_OP_DECOMPOSE_PHYS = 10  /* src/inversion/inversion_utils.f90:549-557    */
};

typedef struct ps3d_handle ps3d_handle;

const char *ps3d_gpu_last_error(void);
/* Library version string. */
const char *ps3d_gpu_version(void);

/* = init_inversion + initialise_fft (inversion_utils.f90:222-372, sta3dfft.f90:53-110)
 * on the current HIP device. Rejects lengths with a prime factor other than 2, 3
 * and 5 (factorisen, stafft.f90:128-187; nx, ny, nz even) with PS3D_E_FACTOR.  The handle owns a non-blocking stream;
 * its device memory is cleared and the set-up complete when this returns. */
int ps3d_gpu_create(const ps3d_params *p, ps3d_handle **h);
/* = finalise_inversion (inversion_utils.f90:459-480). */
int ps3d_gpu_destroy(ps3d_handle *h);

/* = setup_fields (src/utils/utils.f90:136-184): decompose the physical vorticity
 * (3 volumes) into svor, store ini_vor_mean, vor2vel, and fix the viscosity from
 * the kinetic energy and enstrophy (init_diffusion). */
int ps3d_gpu_set_vorticity(ps3d_handle *h, const double *vor);

/* Spectral resampling between two grids of the same box (no counterpart in the
 * reference).  which = PS3D_SVOR: dst's initial state is src's vorticity, truncated
 * to a coarser grid (the sharp spectral cutoff filter) or zero-padded to a finer one,
 * then set up as ps3d_gpu_set_vorticity does; which = PS3D_SVORTS: the stored
 * tendency, copied the same way.  Horizontal stafft coefficients scale by
 * sqrt(n_dst/n_src) per axis, interior sine coefficients also by
 * sqrt(nz_dst/nz_src); Nyquist slots of either grid are dropped. */
int ps3d_gpu_resample(ps3d_handle *dst, ps3d_handle *src, int which);

/* = vor2vel (src/inversion/inversion.f90:23-226) on the device state. */
int ps3d_gpu_vor2vel(ps3d_handle *h);
/* = source -> vorticity_tendency (src/inversion/inversion.f90:298-388).  With a
 * closure attached, source and step return PS3D_E_UNSUPPORTED: the closure term is
 * formed inside the fused step of ps3d_gpu_advance_n only. */
int ps3d_gpu_source(ps3d_handle *h);
/* = adapt (src/stepper/advance.f90:109-377) incl. bstep%set_diffusion; needs a
 * preceding vor2vel.  Writes dt and a PS3D_NDIAG record (diag may be NULL). */
int ps3d_gpu_adapt(ps3d_handle *h, double t, double *dt, double *diag);
/* = bstep%step (impl_rk4.f90:76-207 or cn2.f90:92-181); needs preceding
 * vor2vel + adapt + source at time *t. Advances *t. */
int ps3d_gpu_step(ps3d_handle *h, double *t, double dt);
/* = advance (src/stepper/advance.f90:77-104), n times, without host round trips.
 * diag_series (may be NULL) receives n records of PS3D_NDIAG doubles. */
int ps3d_gpu_advance_n(ps3d_handle *h, double *t, int nsteps, double *diag_series);
/* = ps3d's time loop `do while (t < time%limit); call advance(bstep, t)`
 * (src/ps3d.f90:114-116) with the loop test on the device: up to nmax steps are
 * queued at once (graph replays) and each runs only while t < min(t_stop, limit);
 * the steps after the test fails return at once.  A host program stops at its
 * output times by passing the next one as t_stop (write_step, src/utils/utils.f90:62-89).
 * *nsteps receives the number of steps taken, diag_series (may be NULL, room for
 * nmax records) their records.  Not with a closure attached (PS3D_E_UNSUPPORTED). */
int ps3d_gpu_advance_until(ps3d_handle *h, double *t, double t_stop, int nmax, double *diag_series, int *nsteps);

/* = pressure + horizontal_divergence (src/fields/fields_derived.f90:67-182), which
 * ps3d's adapt calls (src/stepper/advance.f90:278-282).  Computed from the state of
 * the last ps3d_gpu_vor2vel; call it after vor2vel (and adapt) and before source,
 * where write_step sits in advance (advance.f90:85-91).  Neither field feeds the
 * time step, so ps3d_gpu_advance_n leaves them out; results are read with
 * PS3D_PRES / PS3D_DELTA. */
int ps3d_gpu_derived(ps3d_handle *h);

/* = update_netcdf_field_diagnostics (src/fields/field_diagnostics_netcdf.f90:257-439)
 * together with the values adapt hands to set_netcdf_field_diagnostic
 * (src/stepper/advance.f90:188-193, 315-321, 366): the PS3D_NSTAT record
 * (enum ps3d_stat) of the state of the last ps3d_gpu_vor2vel, written to the host
 * array stats.  diag is the PS3D_NDIAG adapt record of the same time (from
 * ps3d_gpu_adapt, or that step's row of ps3d_gpu_advance_n / _until); NULL leaves
 * the adapt entries 0.  The sums, minima and maxima are reductions over the
 * device's physical velocity and vorticity; usdeltarms uses the horizontal
 * divergence, which this call computes with ps3d_gpu_derived unless that already
 * ran after the vor2vel.  ro_min / ro_max divide by f_cor(3) as ps3d does (inf
 * or nan when it is 0); bfmax, rolling_mean_bfmax and ri_min are 0 (no buoyancy).
 * The partials and the device copy of the record are allocated on first use. */
int ps3d_gpu_field_stats(ps3d_handle *h, const double *diag, double *stats);

/* Domain means of the derived fields of ps3d's nc_reader (plotting/nc_reader.py:
 * 94-120) over the state of the last ps3d_gpu_vor2vel: kinetic energy as
 * get_kinetic_energy (field_diagnostics.f90:85-122), helicity u.w (the adapt
 * record's), the cross-helicity magnitude |u x w| with the lids weighted 1/2, and
 * the largest |u x w|. */
enum ps3d_moment {
    PS3D_MO_KE = 0, PS3D_MO_HELICITY, PS3D_MO_CROSS_HELICITY, PS3D_MO_CROSS_HELICITY_MAX,
    PS3D_NMOMENT
};
int ps3d_gpu_moments(ps3d_handle *h, double *out);

/* = genspec (src/genspec.f90:44-124) on the GPU: the kinetic-energy spectrum of a
 * physical velocity in shells of |K| = sqrt(kx^2 + ky^2 + kz^2).  vel is a host
 * array of 3 volumes (as ps3d_gpu_get's PS3D_VEL), which then replaces the handle's
 * physical velocity (the state needs ps3d_gpu_vor2vel again before it steps), or
 * NULL for the velocity of the last ps3d_gpu_vor2vel.  Shell m = int(nint(|K|)/dk),
 * m = 0..kmax, holds spec[m] = sum of squared coefficients * 4/3 pi dk^3
 * ((m+1)^3 - m^3) / num[m], normalised so that sum(spec) dk = get_kinetic_energy.
 * *nshell = kmax + 1 and *dk are always set; spec == NULL only queries them; cap
 * (the length of spec and num) below *nshell gives PS3D_E_ARG.  num (may be NULL)
 * receives the point count of each shell.  Uses the y-derivative volumes as scratch:
 * ps3d_gpu_derived and ps3d_gpu_source need a new vor2vel afterwards.  Buffers are
 * allocated on first use. */
int ps3d_gpu_spectrum(ps3d_handle *h, const double *vel, int cap, int *nshell, double *dk, double *spec,
                      long long *num);

/* ps3d_gpu_field_stats without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  stats_dev is DEVICE memory of
 * PS3D_NSTAT doubles; the record is written there on the handle's stream (read it
 * after ps3d_gpu_sync).  The per-block partials go to one array the library keeps
 * per GPU (a fixed 1024 blocks x 20 doubles = 160 KiB, held until the process
 * ends); a grid whose partition would need more returns PS3D_E_UNSUPPORTED.  Needs
 * ps3d_gpu_vor2vel and then ps3d_gpu_derived first (PS3D_E_STATE otherwise), since
 * computing the divergence here would allocate.  Every call on one GPU shares
 * that array: calls on one GPU that overlap in time (different handles on
 * different streams) race on it and must be serialised by the caller. */
int ps3d_gpu_field_stats_noalloc(ps3d_handle *h, const double *diag, double *stats_dev);


/* Hyperdiffusion coefficients: vvisc of init_diffusion (inversion_utils.f90:157-184)
 * and the prefactor of the last adapt (get_diffusion_pre_factor, advance.f90:381-410);
 * column (k, l) decays at the rate pref * visc * (k^2 + l^2)^nnu.  Either pointer may
 * be NULL. */
int ps3d_gpu_viscosity(ps3d_handle *h, double *visc, double *pref);

/* Device <-> host copies of a field (enum ps3d_field). */
int ps3d_gpu_get(ps3d_handle *h, int which, double *host);
int ps3d_gpu_set(ps3d_handle *h, int which, const double *host);

/* One operator of enum ps3d_op on host arrays (kernel checks of the reference's
 * unit tests). */
int ps3d_gpu_op(ps3d_handle *h, int op, const double *in, double *out);

/* Batched 1-D transforms of stafft (src/fft/stafft.f90), x(m, n) column-major
 * (m fastest), in place on a host array: kind 0 = forfft (196-287),
 * 1 = revfft (296-403), 2 = dct on x(m, 0:n) (410-483), 3 = dst on x(m, 1:n)
 * (489-550). Runs on the current HIP device. */
int ps3d_gpu_fft1d(int kind, int m, int n, double *x);

/* One integrating-factor RK4 substep (impl_rk4.f90:212-364, stage 1..4) applied
 * from the stored tendency PS3D_SVORTS to PS3D_SVOR with step dt and the per-column
 * dissipation vdiss (ny*nx doubles, ky fastest, as vdiss(ky,kx) of
 * src/fields/fields.f90:33).  The kernel check of unit-tests/test_implicit_rk.f90,
 * which holds the tendency fixed across the four stages. */
int ps3d_gpu_rk_substep(ps3d_handle *h, int stage, double dt, const double *vdiss);

/* Kernel kinds of the fused step, for ps3d_gpu_profile_kernels. */
enum ps3d_kernel_kind {
    PS3D_K_ZINV = 0,      /* z pass of vor2vel                                   */
    PS3D_K_YINV = 1,      /* y inverse, 6 fields                                 */
    PS3D_K_XMAIN = 2,     /* x inverse, u x omega, x forward                     */
    PS3D_K_YFWD = 3,      /* y forward, 3 fields                                 */
    PS3D_K_ZTEND = 4,     /* z pass of the tendency + RK substep                 */
    PS3D_K_YINV_ADAPT = 5,/* y inverse at the adapt stage (+ derivatives, sums)  */
    PS3D_K_XADAPT = 6,    /* strain / CFL / vorticity diagnostics                */
    PS3D_K_REDUCE = 7,    /* one-block reductions                                */
    PS3D_K_MEAN = 8,      /* vorticity-mean restore + time update                */
    PS3D_K_ZTEND_ZINV = 9,/* RK4 stage 1-3 z pass + the next stage's z inverse   */
    PS3D_NKIND = 10
};

/* Runs nsteps of advance with a HIP event pair around every kernel on the
 * handle's stream and reports, per kernel kind, the summed milliseconds, the
 * number of launches and the algorithmic HBM bytes per launch (the bytes a pass
 * must move: every field it reads and writes once).  Arrays hold PS3D_NKIND
 * entries.  Not for the timed region: the events add gaps between kernels. */
int ps3d_gpu_profile_kernels(ps3d_handle *h, int nsteps, double *ms, int *launches, double *bytes);

/* Wait for all work queued on the handle's stream. */
int ps3d_gpu_sync(ps3d_handle *h);

/* Kernel time of the most recent advance_n call, measured with HIP events on the
 * handle's stream: total milliseconds and number of steps. */
int ps3d_gpu_last_timing(ps3d_handle *h, double *ms, int *nsteps);

/* ---- Batched instances and ensembles ------------------------------------------
 * ps3d runs one grid per process: an ensemble there is many runs of the time loop
 * (src/ps3d.f90:114-116), each calling advance (src/stepper/advance.f90:77-104).
 * Here one batched handle holds nbatch instances of one grid and every stepping
 * kernel covers all of them in one launch (the instance is the second grid
 * dimension); an ensemble holds one batched handle per grid size, steps the
 * groups concurrently on forked streams and can replay the whole step as one
 * captured HIP graph.  The plain entry points above refuse batched handles. */

/* nbatch instances of the grid p. */
int ps3d_gpu_create_batch(const ps3d_params *p, int nbatch, ps3d_handle **h);
int ps3d_gpu_batch_size(ps3d_handle *h, int *nbatch);
/* setup_fields of instance b (as ps3d_gpu_set_vorticity). */
int ps3d_gpu_set_vorticity_batch(ps3d_handle *h, int b, const double *vor);
/* svor of instance b (PS3D_SVOR only). */
int ps3d_gpu_get_batch(ps3d_handle *h, int b, int which, double *host);
/* nsteps of advance for every instance; t[b] in/out per instance.  diag_series
 * (may be NULL) receives nbatch x nsteps records of PS3D_NDIAG doubles. */
int ps3d_gpu_advance_n_batch(ps3d_handle *h, double *t, int nsteps, double *diag_series);

typedef struct ps3d_ensemble ps3d_ensemble;
/* ngroups batched handles: group g is the grid p[g] with nbatch[g] instances. */
int ps3d_gpu_ensemble_create(int ngroups, const ps3d_params *p, const int *nbatch, ps3d_ensemble **e);
/* The batched handle of group g (owned by the ensemble; set initial vorticity
 * with ps3d_gpu_set_vorticity_batch, read with ps3d_gpu_get_batch). */
int ps3d_gpu_ensemble_group(ps3d_ensemble *e, int g, ps3d_handle **h);
/* nsteps steps of every instance of every group; use_graph = 1 replays one
 * captured graph per step.  ms: kernel time (HIP events), may be NULL. */
int ps3d_gpu_ensemble_advance(ps3d_ensemble *e, int nsteps, int use_graph, double *ms);
/* Times `steps` steps of each group under each available z launch shape and
 * keeps the fastest, then, one mode at a time, the block size of each
 * register-holding line pass, then the FFT schedule of the transform-only y and x
 * line passes (state restored afterwards).  choice[g] (may be
 * NULL) receives the shape code: 100..102 = one wavefront per column (register
 * budget variant), q < 100 = q quads per block of the quad-per-thread kernels. */
int ps3d_gpu_ensemble_tune(ps3d_ensemble *e, int steps, int *choice);
int ps3d_gpu_ensemble_set_choice(ps3d_ensemble *e, int g, int code);
/* Block sizes (threads) the tuner chose for group g's register-holding line passes:
 * threads[0] x products, [1] adapt-stage y, [2] x adapt (1024, 512 or 256: full,
 * half or quarter tile depth).  No counterpart in the reference. */
int ps3d_gpu_ensemble_line_choice(ps3d_ensemble *e, int g, int *threads);
/* FFT schedules the tuner chose for group g's transform-only line passes:
 * sched[0] x lines, sched[1] y lines; bit 1 = radix-8-first plan (else
 * radix-16-first), bit 2 = the Hermitian pair of the inverse formed in the first
 * radix pass (else a pass of its own), bit 4 = 512 threads per tile (else 256).
 * No counterpart in the reference. */
int ps3d_gpu_ensemble_fft_schedule(ps3d_ensemble *e, int g, int *sched);
int ps3d_gpu_ensemble_destroy(ps3d_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif /* PS3D_GPU_H */
