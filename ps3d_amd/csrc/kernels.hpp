// kernels.hpp -- kernel argument blocks and launchers shared by the passes.
#pragma once
#include <cstdint>

#include "common.hpp"

struct ps3d_cnn;

namespace ps3d {

// Device-resident scalars (one small array per handle).  Kernels read dt and the
// diffusion factor from here, so a whole step can run (and be captured in a HIP
// graph) without host round trips.
enum Scal {
    SC_T = 0,        // time
    SC_DT,           // current step
    SC_DFAC,         // impl-rk4: pref*dt/2 ; cn2: dt or pref*dt (impl_rk4.f90:37-53, cn2.f90:39-80)
    SC_VISC,         // vvisc (inversion_utils.f90:157-184)
    SC_INI_XI,       // ini_vor_mean(1:2) (utils.f90:176)
    SC_INI_ETA,
    SC_VORTRMS,      // sqrt(<|omega|^2>)
    SC_KE, SC_EN, SC_HEL,
    SC_VORMX, SC_VORMY, SC_VORMZ,
    SC_PREF,
    SC_RM_SMA,       // rolling mean state (rolling_mean.f90)
    SC_RM_INEW, SC_RM_IOLD, SC_RM_FILLED,
    SC_STEP,         // step counter (diag series row)
    SC_GATE,         // 1 while the gated step runs (t < SC_TSTOP), else 0
    SC_TSTOP,        // ps3d_gpu_advance_until: steps start only while t < this
    SC_NSCAL = 32
};

// z-kernel modes
enum { ZT_STORE = 0, ZT_RK4 = 1, ZT_CN2 = 2 };
enum { ZOP_DST = 0, ZOP_DCT = 1, ZOP_DIFFZ = 2, ZOP_COMBINE = 3, ZOP_DECOMPOSE = 4, ZOP_POISSON = 5 };
enum { MEAN_CALC = 0, MEAN_ADJUST = 1 };

// Batched instances of one grid (an ensemble group): blockIdx.y selects the
// instance, whose data sit at these element strides from instance 0.
struct Batch {
    int nb;                          // instances (gridDim.y); 0 or 1: a single grid
    long vol;                        // field pool
    long part;                       // per-block partials of a line pass
    long party, partx, partm, diag, roll;   // finishing reductions
};

// 1-D tables on the device
struct ZTabs {
    const double *rkx, *rky;      // wavenumbers (sta3dfft.f90:86-104)
    const double *filx, *fily, *filz;   // filter factors: Hou&Li exponents or 2/3 0/1 (inversion_utils.f90:377-455)
    int hou_li;
    int nnu;
};

struct ZLaunch {
    int zpt, blocks;
    size_t lds;
    int nz;
    int threads;   // TQ * QPB
    int wave;      // 1..3: one wavefront per column (zwave.hip, register variant), 0: quad-per-thread
    int nquads;
};

struct ZInvArgs {
    Grid g;
    ZTabs tb;
    int TQ, QPB, nquads;
    FftPlan pz;
    const double2 *twz;
    double zscale;     // 1/sqrt(2 nz)
    double rkz_sc;     // pi / Lz (deriv1d.f90:20-23)
    const double *svor[3];
    double *svor_out[2];
    double *out[6];
    double *mid[3];
    Batch bt;
};

struct RkArgs {
    double *q[3], *qdi[3], *qdf[3];
    double *scal;
    const double *vdiss;   // optional per-column dissipation (ky fastest); nullptr: dfac*vhdis
};

struct ZTendArgs {
    Grid g;
    ZTabs tb;
    int TQ, QPB, nquads;
    FftPlan pz;
    const double2 *twz;
    double zscale;
    const double *rqp[3];
    double *svorts[3];
    int mode, stage;
    RkArgs rk;
    Batch bt;
    const double *extra[3];   // optional mixed-spectral tendency added to the source (closure)
};

struct ZOpArgs {
    Grid g;
    ZTabs tb;
    int TQ, QPB, nquads;
    FftPlan pz;
    const double2 *twz;
    double zscale, rkz_sc;
    const double *in;
    double *out;
    int op;
};

struct MeanArgs {
    Grid g;
    FftPlan pz;
    const double2 *twz;
    double zscale;
    double *svor[2];
    double *scal;
    int mode, advance_time;
    Batch bt;
};

size_t zpass_lds_bytes(int QPB, int nz);
hipError_t launch_zinv(const ZLaunch &zl, const ZInvArgs &args, hipStream_t stream);
hipError_t launch_ztend(const ZLaunch &zl, const ZTendArgs &args, hipStream_t stream);
hipError_t launch_zrk(const ZLaunch &zl, const ZTendArgs &args, hipStream_t stream);
hipError_t launch_zop(const ZLaunch &zl, const ZOpArgs &args, hipStream_t stream);

// wave-per-column z kernels (zwave.hip), nz in {128, 256, 512}
bool zwave_supported(int nz);
size_t zwave_lds_bytes(int nz);
hipError_t configure_zwave();
// variant 1: tendency pass at three waves per SIMD; 2: register allocation left to
// the compiler in both; 3: both at three waves per SIMD
hipError_t launch_zinv_w(int nz, int variant, int nquads, const ZInvArgs &args, hipStream_t stream);
hipError_t launch_ztend_w(int nz, int variant, int nquads, const ZTendArgs &args, hipStream_t stream);
hipError_t launch_zop_w(int nz, int nquads, const ZOpArgs &args, hipStream_t stream);
// RK4 stage 1..3 of the tendency pass fused with the next stage's vor2vel pass (the
// stage result stays in registers; t.rk.q is not written)
hipError_t launch_ztend_zinv_w(int nz, int variant, int nquads, const ZTendArgs &t, const ZInvArgs &z,
                               hipStream_t stream);
hipError_t launch_mean(const MeanArgs &args, int nz, hipStream_t stream);

// Pointwise kernels of the output-step derived fields (fields_derived.f90:67-182),
// over whole padded volumes.
struct PresArgs {
    long n;                     // doubles per padded volume
    const double *dudx, *dudy, *dvdy, *dwdx, *dwdy, *xi, *eta, *zeta;
    double *pres;
};
hipError_t launch_pres_rhs(const PresArgs &a, hipStream_t stream);

// Spectral resampling of a mixed-representation field between grids of the same
// box: truncation (to a coarser grid) or zero padding (to a finer one).
struct ResampleArgs {
    const double *src[3];
    double *dst[3];
    int snx, sny, snz, szs;
    int dnx, dny, dnz, dzs;
    double hfac, zfac;
};
hipError_t launch_resample(const ResampleArgs &a, hipStream_t stream);
hipError_t launch_add2(long n, const double *a, const double *b, double *out, hipStream_t stream);

// ---------------------------------------------------------------------------
// x / y line passes
// ---------------------------------------------------------------------------
enum LineMode {
    LM_FWD = 0,       // forward Hermitian transform of nf fields
    LM_INV = 1,       // inverse transform of nf fields
    LM_INV_Y1 = 2,    // y inverse at the adapt stage: u,v,w,xi,eta,zeta(,mid) + y-derivatives + Parseval sums
    LM_XMAIN = 3,     // x: inverse 6 fields, u x omega, forward 3
    LM_XADAPT = 4,    // x: strain eigenvalues, maxima, characteristic vorticity
    LM_DIFF = 5       // spectral derivative along the line (diffx / diffy), no transform
};

// Reduction slots written per block by LM_INV_Y1 (sums), LM_XADAPT (maxima / sums)
// and, at the adapt stage, LM_XMAIN (velocity maxima)
enum { RY_UU = 0, RY_WW, RY_UW, RY_MX, RY_MY, RY_MZ, RY_N };
enum { RX_W2MAX = 0, RX_GG, RX_USGG, RX_LSGG, RX_UMAX, RX_VMAX, RX_WMAX, RX_L1, RX_L2, RX_N };
enum { RM_UMAX = 0, RM_VMAX, RM_WMAX, RM_N };

struct LineArgs {
    Grid g;
    int axis;          // 0: lines along x, 1: lines along y
    int n;             // line length
    long estride;      // element stride along the line (doubles)
    long ostride;      // stride of the "other" horizontal index (doubles)
    int nother;        // number of lines in the other horizontal direction
    int ZC, nzt;       // z depth of a tile, number of z tiles
    FftPlan P;
    const double2 *tw; // exp(-2 pi i m/n)
    const double *rk;  // wavenumber table along the line
    int nwl;           // n/2
    int mode, nf;
    int phys_in;       // LM_XADAPT / LM_XMAIN: inputs already physical in x
    const double *in[12];
    double *out[12];
    double *part;      // per-block partial results (LM_XMAIN: velocity maxima when set)
    int velmax_given;  // LM_XADAPT: the velocity maxima come from the LM_XMAIN pass
    const double *scal;
    double fcor[3];
    double ncelli;
    double sqrt_nx;    // for means from the kx = 0 slot
    Batch bt;
    uint16_t *nn_in;   // LM_XMAIN: also write the closure input (bf16 [p][16]) when set
    float nn_scale[6];
    int rthreads;      // register-holding modes: 1024 (standard tile), 512 (half depth) or 256 (quarter)
    int sched;         // transform-only modes: FFT schedule (LS_R8 | LS_FOLD | LS_T512 bits)
};
// FFT schedules of the transform-only line passes: radix-8-first plan instead of
// radix-16-first; Hermitian pair of the inverse formed in the first radix pass; 512
// threads per tile instead of 256.
enum { LS_R8 = 1, LS_FOLD = 2, LS_T512 = 4 };
// the schedule a line length takes unless told otherwise (fold from n = 128 up)
int line_sched_default(int n);
// whether a schedule is built for a line length (others fall back to the default)
bool line_sched_ok(int n, int sched);

size_t line_lds_bytes(int n, int ZC);
// whether the T-thread register build (T = 256, 512, 1024) exists and holds a tile
// of line length n, and the tile depth it takes
bool line_reg_ok(int n, int T);
int line_reg_zc(int n, int T);
hipError_t launch_line(const LineArgs &a, int nthreads, hipStream_t stream);

// Reductions: finishes the Parseval sums (after LM_INV_Y1) and the adapt step
// (after LM_XADAPT): dt, prefactor, rolling mean, diffusion factor, diag row.
struct FinArgs {
    const double *party;  int nby;
    const double *partx;  int nbx;
    const double *partm;  int nbm;   // velocity maxima of LM_XMAIN, or nullptr
    double *scal;
    double *diag;         // series base (row SC_STEP), or nullptr
    int max_rows;
    double ncelli;
    double alpha, limit, dt_fixed;
    double dx, dy, dz;
    int pretype, stepper, nnu;
    int roll_len;
    double *roll_hist;
    Batch bt;
    const double *gate;   // step gate (Grid::gate)
};
hipError_t launch_fin_y(const FinArgs &a, hipStream_t stream);
hipError_t launch_fin_x(const FinArgs &a, hipStream_t stream);
// Opens (1) or closes (0) the step gate of a single-instance handle:
// scal[SC_GATE] = scal[SC_T] < scal[SC_TSTOP].
hipError_t launch_gate(double *scal, hipStream_t stream);

// ---------------------------------------------------------------------------
// output-step reductions (stats.hip)
// ---------------------------------------------------------------------------
// Per-block partial slots of the field_stats reduction: sums, minima, maxima.
enum {
    ST_S0 = 0, ST_S_UV = 0, ST_S_W, ST_S_XE, ST_S_Z, ST_S_ZTOP, ST_S_DTOP,
    ST_MIN0, ST_MIN_X = ST_MIN0, ST_MIN_Y, ST_MIN_Z,
    ST_MAX0, ST_MAX_X = ST_MAX0, ST_MAX_Y, ST_MAX_Z, ST_MAX_HE2, ST_MAX_UOX, ST_MAX_UOY, ST_MAX_UOZ,
    ST_MAX_USUH2, ST_MAX_LOX, ST_MAX_LOY, ST_MAX_LOZ,
    ST_NP,
    ST_NSUM = ST_MIN0 - ST_S0, ST_NMIN = ST_MAX0 - ST_MIN0, ST_NMAX = ST_NP - ST_MAX0
};
constexpr int kStatsMaxBlocks = 1024;   // blocks of the partials pass (any grid)
constexpr int kMaxDevices = 64;

struct StatsArgs {
    int nx, ny, nz, zs;
    const double *vel[3], *vor[3];
    const double *delta;        // physical horizontal divergence, or nullptr (usdeltarms = 0)
    double *part;               // nb * ST_NP partials
    int nb;
};
struct StatsFin {
    const double *part;
    int nb;
    double ncelli, nxnyi, fcor3;
    double ad[32];              // the adapt record of the same time (enum ps3d_diag)
    double *out;                // enum ps3d_stat record
};
int stats_blocks(int nx, int ny);
hipError_t launch_stats(const StatsArgs &a, const StatsFin &f, hipStream_t stream);
// device address of the library's own partials array (kStatsMaxBlocks * ST_NP
// doubles) on device dev
hipError_t stats_scratch(int dev, double **ptr);

// Domain means of the kinetic energy, helicity and cross-helicity (enum ps3d_moment).
enum { MO_S_KE = 0, MO_S_HEL, MO_S_XH, MO_MAX_XH2, MO_NP };
struct MomArgs {
    int nx, ny, nz, zs;
    const double *vel[3], *vor[3];   // vor[0] == nullptr: kinetic energy only
    double *part;                    // nb * MO_NP partials
    int nb;
    double ncelli;
    double *out;                     // PS3D_NMOMENT values
};
hipError_t launch_moments(const MomArgs &a, hipStream_t stream);

// Shell binning of genspec's spectrum: squared coefficients of one z-transformed
// spectral component summed per shell m = int(nint(|K|) / dk), per block.
constexpr int kSpecWaves = 4;
struct SpecArgs {
    const double *f;        // column layout, (x, y)-spectral and z-transformed
    int nx, ny, nz, zs;
    const double *rkx, *rky;
    double rkz1;            // pi / extent(3): rkz(kz) = rkz1 * kz
    double dki;             // 1 / dk
    int nbins;              // shells held per block (largest m + 1)
    double *part;           // nb * nbins sums
    int *cpart;             // nb * nbins point counts, or nullptr
    int nb;
    int accumulate;         // add to part instead of overwriting it
};
size_t spec_lds_bytes(int nbins);
hipError_t launch_spectrum(const SpecArgs &a, hipStream_t stream);
// per-shell totals over the blocks, in block order
hipError_t launch_spectrum_fin(const double *part, const int *cpart, int nb, int nbins, double *spec,
                               long long *cnt, hipStream_t stream);

// closure hooks (closure.hip)
bool cnn_matches(const ps3d_cnn *c, int nx, int ny, int nz);
uint16_t *cnn_input(ps3d_cnn *c);
hipError_t cnn_apply(ps3d_cnn *c, const float *out_scale, int zs, double *o0, double *o1, double *o2, hipStream_t s);

}  // namespace ps3d
