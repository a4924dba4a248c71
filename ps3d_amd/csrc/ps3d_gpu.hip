// ps3d_gpu.hip -- C ABI (include/ps3d_gpu.h) and step orchestration.
//
// One handle = one grid resident in the HBM of the current device, one HIP stream.
// The fused RK4 stage is five passes over HBM (z-inverse, y-inverse, x, y-forward,
// z-forward+update); the adapt stage adds a y-inverse with derivatives, an x pass
// for the strain/CFL/vorticity diagnostics and two one-block reductions.  dt and
// the diffusion factor stay on the device, so advance_n queues whole steps without
// host round trips.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ps3d_closure.h"
#include "../../include/ps3d_gpu.h"
#include "kernels.hpp"

using namespace ps3d;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(PS3D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

namespace ps3d {
hipError_t configure_zpass();
hipError_t configure_linepass();
}

struct LineCfg {
    int n, ZC, nzt, nother;
    long estride, ostride;
};

struct ps3d_handle {
    ps3d_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    // one captured step replayed by advance_n (rebuilt when a captured argument --
    // closure, diagnostics buffer, state set-up -- changes)
    hipGraph_t sgraph = nullptr;
    hipGraphExec_t sexec = nullptr;
    // the same step behind the device-side gate (ps3d_gpu_advance_until)
    hipGraph_t ggraph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // counts the times a captured argument changed (drop_step_graph): graphs that
    // captured this handle elsewhere (an ensemble step) compare it before replaying
    unsigned gen = 0;
    Grid g{};
    size_t nvol = 0;               // doubles per padded volume
    double *pool = nullptr;
    double *svor[3]{}, *qdi[3]{}, *qdf[3]{}, *svel[3]{}, *vel[3]{}, *vor[3]{}, *grad[3]{}, *mid[3]{};
    double *scal = nullptr, *party = nullptr, *partx = nullptr, *partm = nullptr, *roll = nullptr;
    double *pres = nullptr, *delta = nullptr;   // output-step derived fields, allocated on first use
    double *diag = nullptr;
    int diag_cap = 0;
    double *rkx = nullptr, *rky = nullptr, *filx = nullptr, *fily = nullptr, *filz = nullptr;
    double2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
    FftPlan px{}, py{}, pz{};
    ZLaunch zl{};
    int TQ = 0, QPB = 0, nquads = 0;
    bool zfuse = true;             // RK4 tendency pass fused with the next vor2vel pass
    LineCfg lx{}, ly{};
    // tiles of the register-holding line modes (x products pass lxm, x adapt pass lxr,
    // adapt-stage y pass lyr) and their block sizes: 1024 threads on the standard
    // tile, or 512 on half of it (two blocks per CU)
    LineCfg lxm{}, lxr{}, lyr{};
    int rtm = 1024, rtx = 1024, rty = 1024;
    int lsx = 0, lsy = 0;          // FFT schedules of the transform-only x and y line passes
    double visc = 0.0;
    bool have_state = false, have_phys = false, have_adapt = false;
    // batched instances of this grid (ensemble group): instance b's data sit at
    // b * stride from instance 0; launches of the stepping kernels cover all of them
    int nbatch = 1;
    long b_vol = 0, b_party = 0, b_partx = 0, b_partm = 0, b_roll = 0, b_diag = 0;
    std::vector<char> inst_set;    // instances given an initial vorticity
    ps3d_cnn *cnn = nullptr;       // attached closure (owned by the caller)
    float cnn_in[6]{}, cnn_out[3]{};
    bool have_grad = false;   // grad holds the physical y-derivatives of u, v, w (vor2vel, before source)
    bool have_derived = false;
    bool derived_fresh = false;   // pres / delta belong to the state of the last vor2vel
    double *stats_part = nullptr, *stats_out = nullptr;   // ps3d_gpu_field_stats, allocated on first use
    // ps3d_gpu_spectrum / ps3d_gpu_moments buffers, allocated on first use: per-block
    // shell sums and counts (spec_cap shells), then the totals and the moments
    double *spec_part = nullptr, *spec_out = nullptr;
    int *spec_cpart = nullptr;
    long long *spec_cnt = nullptr;
    int spec_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_steps = 0;
    std::vector<double> h_rkx, h_rky;
    // per-kernel profiling (ps3d_gpu_profile_kernels)
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_kind;
    size_t prof_used = 0;
    bool prof_err = false;   // an event could not be created or recorded
};

// Event pair around the next launch when profiling is on.
static void prof_begin(ps3d_handle *h, int kind) {
    if (!h->prof) return;
    if (h->prof_used + 2 > h->prof_ev.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                h->prof_err = true;
                h->prof = false;
                return;
            }
            h->prof_ev.push_back(e);
        }
    }
    if (hipEventRecord(h->prof_ev[h->prof_used], h->stream) != hipSuccess) h->prof_err = true;
    h->prof_kind.push_back(kind);
}
static void prof_end(ps3d_handle *h) {
    if (!h->prof) return;
    if (hipEventRecord(h->prof_ev[h->prof_used + 1], h->stream) != hipSuccess) h->prof_err = true;
    h->prof_used += 2;
}

// ---------------------------------------------------------------------------
// host tables (the reference's formulas, evaluated in the same order)
// ---------------------------------------------------------------------------
static double ipow(double x, int p) {
    double r = 1.0, b = x;
    int e = p;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

// rkx(k) = hrkx(2 min(k, n-k)), hrkx(j) = (pi/L) j  (sta2dfft.f90:70-75, sta3dfft.f90:90-102)
static std::vector<double> wavenumbers(int n, double L) {
    const double pi = 3.141592653589793238462643383279502884;
    const double sc = pi / L;
    std::vector<double> h(n + 1), rk(n, 0.0);
    for (int k = 1; k <= n; ++k) h[k] = sc * (double)k;
    const int nw = n / 2;
    rk[0] = 0.0;
    for (int k = 1; k <= nw - 1; ++k) {
        rk[k] = h[2 * k];
        rk[n - k] = h[2 * k];
    }
    rk[nw] = h[n];
    return rk;
}

static std::vector<double2> twiddles(int n) {
    std::vector<double2> t(n);
    for (int m = 0; m < n; ++m) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        t[m] = make_double2((double)cosl(a), -(double)sinl(a));
    }
    return t;
}

// stafft's factorisen (src/fft/stafft.f90:128-187) takes the primes 2, 3 and 5.
static bool factor235(int n) {
    if (n < 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

static LineCfg line_cfg(int n, int nz, int zs, int nother, long estride, long ostride, int rthreads = 0) {
    LineCfg c;
    c.n = n;
    int zc = 4096 / n;
    if (zc > 16) zc = 16;
    if (zc < 2) zc = 2;
    zc &= ~1;
    if (rthreads) zc = ps3d::line_reg_zc(n, rthreads);
    c.ZC = zc;
    c.nzt = (nz + 1 + zc - 1) / zc;
    c.nother = nother;
    c.estride = estride;
    c.ostride = ostride;
    (void)zs;
    return c;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &v) {
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(1, v.size())));
    HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
}

static ZTabs ztabs(const ps3d_handle *h) {
    ZTabs t;
    t.rkx = h->rkx;
    t.rky = h->rky;
    t.filx = h->filx;
    t.fily = h->fily;
    t.filz = h->filz;
    t.hou_li = h->p.filtering == PS3D_FILTER_HOU_LI;
    t.nnu = h->p.nnu;
    return t;
}

static void drop_step_graph(ps3d_handle *h) {
    ++h->gen;
    if (h->sexec) (void)hipGraphExecDestroy(h->sexec);
    if (h->sgraph) (void)hipGraphDestroy(h->sgraph);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->ggraph) (void)hipGraphDestroy(h->ggraph);
    h->sexec = nullptr;
    h->sgraph = nullptr;
    h->gexec = nullptr;
    h->ggraph = nullptr;
}

static void free_handle(ps3d_handle *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_step_graph(h);
    (void)hipFree(h->pool);
    (void)hipFree(h->pres);
    (void)hipFree(h->delta);
    (void)hipFree(h->stats_part);
    (void)hipFree(h->stats_out);
    (void)hipFree(h->spec_part);
    (void)hipFree(h->spec_out);
    (void)hipFree(h->spec_cpart);
    (void)hipFree(h->spec_cnt);
    (void)hipFree(h->scal);
    (void)hipFree(h->party);
    (void)hipFree(h->partx);
    (void)hipFree(h->partm);
    (void)hipFree(h->roll);
    (void)hipFree(h->diag);
    (void)hipFree(h->rkx);
    (void)hipFree(h->rky);
    (void)hipFree(h->filx);
    (void)hipFree(h->fily);
    (void)hipFree(h->filz);
    (void)hipFree(h->twx);
    (void)hipFree(h->twy);
    (void)hipFree(h->twz);
    for (hipEvent_t e : h->prof_ev) (void)hipEventDestroy(e);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// Builds tables, plans and launch shapes. alloc_fields = false for the 1-D helpers.
static void apply_line_choice(ps3d_handle *h, int tm, int ty, int tx);

static int init_handle(ps3d_handle *h, const ps3d_params *p, bool alloc_fields, bool strict, int nbatch = 1) {
    h->p = *p;
    h->nbatch = nbatch;
    h->inst_set.assign(nbatch, 0);
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    if (strict) {
        if (nx < 2 || ny < 2 || nz < 2 || (nx & 1) || (ny & 1) || (nz & 1))
            return fail(PS3D_E_ARG, "nx, ny, nz must be even and >= 2");
        if (!factor235(nx) || !factor235(ny) || !factor235(nz))
            return fail(PS3D_E_FACTOR, "Factorisation not possible: grid lengths must be 2^a 3^b 5^c");
        if (nx > 4096 || ny > 4096 || nz > 1024) return fail(PS3D_E_UNSUPPORTED, "grid too large (nx,ny <= 4096, nz <= 1024)");
        for (int d = 0; d < 3; ++d)
            if (!(p->extent[d] > 0.0)) return fail(PS3D_E_ARG, "Domain length equal to zero not allowed");
        if (p->nnu < 1) return fail(PS3D_E_ARG, "nnu must be >= 1");
        if (p->filtering != PS3D_FILTER_HOU_LI && p->filtering != PS3D_FILTER_23RULE)
            return fail(PS3D_E_ARG, "filtering must be 'Hou & Li' or '2/3-rule'");
        if (p->stepper != PS3D_STEPPER_IMPL_RK4 && p->stepper != PS3D_STEPPER_CN2)
            return fail(PS3D_E_ARG, "unknown stepper");
        if (p->pretype < 0 || p->pretype > 5) return fail(PS3D_E_ARG, "unknown pretype");
        if (p->length_scale != PS3D_LS_KOLMOGOROV && p->length_scale != PS3D_LS_GEOPHYSICAL)
            return fail(PS3D_E_ARG, "We only support 'Kolmogorov' or 'geophysical'");
        if (p->roll_mean_win_size < 1) return fail(PS3D_E_ARG, "roll_mean_win_size must be >= 1");
    }
    HIPCHK(hipGetDevice(&h->device));
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    Grid &g = h->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.zs = ((nz + 1 + kZAlign - 1) / kZAlign) * kZAlign;
    g.nwx = nx / 2;
    g.nwy = ny / 2;
    g.dz = p->extent[2] / (double)nz;
    g.dzi = 1.0 / g.dz;
    g.hdzi = 0.5 * g.dzi;
    g.lz = p->extent[2];
    g.zlo = p->lower[2];
    h->nvol = (size_t)nx * ny * g.zs;

    h->h_rkx = wavenumbers(nx, p->extent[0]);
    h->h_rky = wavenumbers(ny, p->extent[1]);
    std::vector<double> rkz(nz + 1, 0.0);
    for (int k = 1; k <= nz; ++k) rkz[k] = (3.141592653589793238462643383279502884 / p->extent[2]) * (double)k;
    // filters (inversion_utils.f90:377-455)
    auto maxv = [](const std::vector<double> &v) { double m = 0; for (double x : v) m = std::max(m, x); return m; };
    std::vector<double> fx(nx), fy(ny), fz(nz + 1);
    const double mx = maxv(h->h_rkx), my = maxv(h->h_rky), mz = maxv(rkz);
    if (p->filtering == PS3D_FILTER_HOU_LI) {
        const double ix = 1.0 / mx, iy = 1.0 / my, iz = 1.0 / mz;
        for (int k = 0; k < nx; ++k) fx[k] = -36.0 * ipow(ix * h->h_rkx[k], 36);
        for (int k = 0; k < ny; ++k) fy[k] = -36.0 * ipow(iy * h->h_rky[k], 36);
        for (int k = 0; k <= nz; ++k) fz[k] = -36.0 * ipow(iz * rkz[k], 36);
    } else {
        const double f23 = 2.0 / 3.0;
        for (int k = 0; k < nx; ++k) fx[k] = (h->h_rkx[k] <= f23 * mx) ? 1.0 : 0.0;
        for (int k = 0; k < ny; ++k) fy[k] = (h->h_rky[k] <= f23 * my) ? 1.0 : 0.0;
        for (int k = 0; k <= nz; ++k) fz[k] = (rkz[k] <= f23 * mz) ? 1.0 : 0.0;
    }
    int st;
    if ((st = upload(&h->rkx, h->h_rkx)) || (st = upload(&h->rky, h->h_rky)) || (st = upload(&h->filx, fx)) ||
        (st = upload(&h->fily, fy)) || (st = upload(&h->filz, fz)))
        return st;
    if (!make_plan(nx, h->px) || !make_plan(ny, h->py) || !make_plan(2 * nz, h->pz))
        return fail(PS3D_E_FACTOR, "Factorisation not possible: grid lengths must be 2^a 3^b 5^c");
    if ((st = upload(&h->twx, twiddles(nx))) || (st = upload(&h->twy, twiddles(ny))) ||
        (st = upload(&h->twz, twiddles(2 * nz))))
        return st;

    // z launch shape: TQ threads per quad (a multiple of 64) own ZPT levels each;
    // small columns share a block (QPB quads), long ones take up to 512 threads
    const int nz1 = nz + 1;
    auto up64 = [](int v) { return (v + 63) / 64 * 64; };
    if (nz1 <= 64) { h->TQ = 64; h->zl.zpt = 1; }
    else if (nz1 <= 128) { h->TQ = 128; h->zl.zpt = 1; }
    else {
        int zpt = 1;
        while (up64((nz1 + zpt - 1) / zpt) > 512) ++zpt;
        h->zl.zpt = zpt;
        h->TQ = up64((nz1 + zpt - 1) / zpt);
    }
    if (h->zl.zpt > 3) return fail(PS3D_E_UNSUPPORTED, "nz too large");
    h->QPB = h->TQ >= 256 ? 1 : 256 / h->TQ;
    h->zl.threads = h->TQ * h->QPB;
    h->nquads = (nx / 2 + 1) * (ny / 2 + 1);
    h->zl.blocks = (h->nquads + h->QPB - 1) / h->QPB;
    h->zl.lds = zpass_lds_bytes(h->QPB, nz);
    h->zl.nz = nz;
    if (h->zl.lds > 160 * 1024) return fail(PS3D_E_UNSUPPORTED, "nz too large for the LDS z transform");
    // one wavefront per column where a wave FFT exists for 2 nz (PS3D_ZWAVE=0 keeps
    // the quad-per-thread kernels, for comparison)
    h->zl.nquads = h->nquads;
    {
        const char *ev = std::getenv("PS3D_ZWAVE");
        const bool off = ev && ev[0] == '0';
        // register variant: the two-waves-per-SIMD builds (no spills, the update's
        // operands loaded ahead) measure fastest at nz = 256 and 512
        int var = 2;
        if (ev && ev[0] >= '1' && ev[0] <= '3') var = ev[0] - '0';
        h->zl.wave = (ps3d::zwave_supported(nz) && !off) ? var : 0;
        const char *fz = std::getenv("PS3D_ZFUSE");
        h->zfuse = !(fz && fz[0] == '0');
    }

    h->ly = line_cfg(ny, nz, g.zs, nx, (long)g.zs, (long)ny * g.zs);
    h->lx = line_cfg(nx, nz, g.zs, ny, (long)ny * g.zs, (long)g.zs);
    {
        // Block sizes of the register-holding line modes (x products pass, adapt-stage
        // y pass, x adapt pass): 1024 threads on the standard tile, 512 on half of it,
        // 256 on a quarter.  Measured fastest: 512, 512, 256 at 256^3, 512 for all
        // three at 512^3, and 1024, 1024, 512 at 1024^3 (where a half tile holds two z
        // levels); the ensemble tuner picks 256 for all three at n <= 96.
        // PS3D_LINE_T="m,y,x" overrides.
        int tm = nx <= 512 ? 512 : 1024, ty = ny <= 512 ? 512 : 1024, tx = nx <= 256 ? 256 : 512;
        if (std::max(nx, ny) <= 96) tm = ty = tx = 256;
        if (const char *ev = std::getenv("PS3D_LINE_T")) std::sscanf(ev, "%d,%d,%d", &tm, &ty, &tx);
        apply_line_choice(h, tm, ty, tx);
        // FFT schedules of the transform-only passes; PS3D_LINE_SCHED="x,y" overrides
        int sx = line_sched_default(nx), sy = line_sched_default(ny);
        if (const char *ev = std::getenv("PS3D_LINE_SCHED")) std::sscanf(ev, "%d,%d", &sx, &sy);
        h->lsx = line_sched_ok(nx, sx) ? sx : line_sched_default(nx);
        h->lsy = line_sched_ok(ny, sy) ? sy : line_sched_default(ny);
    }

    HIPCHK(ps3d::configure_zpass());
    HIPCHK(ps3d::configure_zwave());
    HIPCHK(ps3d::configure_linepass());
    {   // device address of the allocation-free partials array (once per device)
        double *scr = nullptr;
        HIPCHK(ps3d::stats_scratch(h->device, &scr));
    }

    const size_t nb = (size_t)nbatch;
    // per-tile partial sums sized for the shallowest tiles (2 levels), so that any
    // block size of the adapt passes fits
    const long nzt_max = (nz + 2) / 2;
    h->b_party = (long)h->lyr.nother * nzt_max * RY_N;
    h->b_partx = (long)h->lxr.nother * nzt_max * RX_N;
    h->b_partm = (long)h->lxm.nother * nzt_max * RM_N;
    h->b_roll = std::max(1, p->roll_mean_win_size);
    HIPCHK(hipMalloc((void **)&h->scal, sizeof(double) * SC_NSCAL * nb));
    // clears go on the handle's stream: the stream is non-blocking, so a clear on
    // the null stream would not be ordered before the handle's first kernels
    HIPCHK(hipMemsetAsync(h->scal, 0, sizeof(double) * SC_NSCAL * nb, h->stream));
    HIPCHK(hipMalloc((void **)&h->party, sizeof(double) * h->b_party * nb));
    HIPCHK(hipMalloc((void **)&h->partx, sizeof(double) * h->b_partx * nb));
    HIPCHK(hipMalloc((void **)&h->partm, sizeof(double) * h->b_partm * nb));
    HIPCHK(hipMalloc((void **)&h->roll, sizeof(double) * h->b_roll * nb));
    if (alloc_fields) {
        const size_t nv = 24;
        h->b_vol = (long)(h->nvol * nv);
        const size_t bytes = sizeof(double) * h->nvol * nv * nb;
        hipError_t e = hipMalloc((void **)&h->pool, bytes);
        if (e != hipSuccess) return fail(PS3D_E_NOMEM, "device allocation of the field pool failed");
        HIPCHK(hipMemsetAsync(h->pool, 0, bytes, h->stream));
        double *q = h->pool;
        auto take3 = [&](double **d) { for (int c = 0; c < 3; ++c) { d[c] = q; q += h->nvol; } };
        take3(h->svor); take3(h->qdi); take3(h->qdf); take3(h->svel);
        take3(h->vel); take3(h->vor); take3(h->grad); take3(h->mid);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static Batch batch(const ps3d_handle *h) {
    Batch b;
    b.nb = h->nbatch;
    b.vol = h->b_vol;
    b.part = 0;
    b.party = h->b_party;
    b.partx = h->b_partx;
    b.partm = h->b_partm;
    b.diag = h->b_diag;
    b.roll = h->b_roll;
    return b;
}

// Block sizes of the register-holding line modes (1024, 512 or 256 threads; a size
// without a build for the line length falls back to 1024) and their tiles.
static void apply_line_choice(ps3d_handle *h, int tm, int ty, int tx) {
    const Grid &g = h->g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    auto pick = [](int n, int T) { return (T != 1024 && ps3d::line_reg_ok(n, T)) ? T : 1024; };
    h->rtm = pick(nx, tm);
    h->rty = pick(ny, ty);
    h->rtx = pick(nx, tx);
    h->lxm = line_cfg(nx, nz, g.zs, ny, (long)ny * g.zs, (long)g.zs, h->rtm);
    h->lyr = line_cfg(ny, nz, g.zs, nx, (long)g.zs, (long)ny * g.zs, h->rty);
    h->lxr = line_cfg(nx, nz, g.zs, ny, (long)ny * g.zs, (long)g.zs, h->rtx);
}

static LineArgs line_args(const ps3d_handle *h, int axis, int mode) {
    LineArgs a;
    std::memset(&a, 0, sizeof a);
    a.bt = batch(h);
    const LineCfg &c = mode == LM_XMAIN ? h->lxm : mode == LM_XADAPT ? h->lxr : mode == LM_INV_Y1 ? h->lyr
                     : (axis == 0 ? h->lx : h->ly);
    a.rthreads = mode == LM_XMAIN ? h->rtm : mode == LM_XADAPT ? h->rtx : mode == LM_INV_Y1 ? h->rty : 256;
    a.sched = axis == 0 ? h->lsx : h->lsy;
    a.g = h->g;
    a.axis = axis;
    a.n = c.n;
    a.estride = c.estride;
    a.ostride = c.ostride;
    a.nother = c.nother;
    a.ZC = c.ZC;
    a.nzt = c.nzt;
    a.P = axis == 0 ? h->px : h->py;
    a.tw = axis == 0 ? h->twx : h->twy;
    a.rk = axis == 0 ? h->rkx : h->rky;
    a.nwl = c.n / 2;
    a.mode = mode;
    a.scal = h->scal;
    for (int d = 0; d < 3; ++d) a.fcor[d] = h->p.f_cor[d];
    a.ncelli = 1.0 / ((double)h->g.nx * h->g.ny * h->g.nz);
    a.sqrt_nx = std::sqrt((double)h->g.nx);
    return a;
}

static int run_line(ps3d_handle *h, const LineArgs &a) {
    HIPCHK(launch_line(a, 0, h->stream));
    return 0;
}

static int line_fields(ps3d_handle *h, int axis, int mode, int nf, double *const *in, double *const *out) {
    LineArgs a = line_args(h, axis, mode);
    a.nf = nf;
    for (int f = 0; f < nf; ++f) {
        a.in[f] = in[f];
        a.out[f] = out[f];
    }
    return run_line(h, a);
}

static FinArgs fin_args(ps3d_handle *h) {
    FinArgs f;
    std::memset(&f, 0, sizeof f);
    f.party = h->party;
    f.nby = h->lyr.nother * h->lyr.nzt;
    f.partx = h->partx;
    f.nbx = h->lxr.nother * h->lxr.nzt;
    f.scal = h->scal;
    f.diag = h->diag;
    f.max_rows = h->diag_cap;
    f.ncelli = 1.0 / ((double)h->g.nx * h->g.ny * h->g.nz);
    f.alpha = h->p.alpha;
    f.limit = h->p.limit;
    f.dt_fixed = h->p.dt_fixed;
    f.dx = h->p.extent[0] / h->g.nx;
    f.dy = h->p.extent[1] / h->g.ny;
    f.dz = h->p.extent[2] / h->g.nz;
    f.pretype = h->p.pretype;
    f.stepper = h->p.stepper;
    f.nnu = h->p.nnu;
    f.roll_len = h->p.roll_mean_win_size;
    f.roll_hist = h->roll;
    f.bt = batch(h);
    f.gate = h->g.gate;
    return f;
}

static ZInvArgs zinv_args(ps3d_handle *h) {
    ZInvArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = h->g;
    a.tb = ztabs(h);
    a.TQ = h->TQ;
    a.QPB = h->QPB;
    a.nquads = h->nquads;
    a.pz = h->pz;
    a.twz = h->twz;
    a.zscale = 1.0 / std::sqrt(2.0 * h->g.nz);
    a.rkz_sc = 3.141592653589793238462643383279502884 / h->p.extent[2];
    a.bt = batch(h);
    for (int c = 0; c < 3; ++c) a.svor[c] = h->svor[c];
    for (int c = 0; c < 3; ++c) {
        a.out[c] = h->svel[c];
        a.out[3 + c] = h->vor[c];
    }
    return a;
}

static ZTendArgs ztend_args(ps3d_handle *h, int mode, int stage, double *const *rqp) {
    ZTendArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = h->g;
    a.tb = ztabs(h);
    a.TQ = h->TQ;
    a.QPB = h->QPB;
    a.nquads = h->nquads;
    a.pz = h->pz;
    a.twz = h->twz;
    a.zscale = 1.0 / std::sqrt(2.0 * h->g.nz);
    for (int c = 0; c < 3; ++c) {
        a.rqp[c] = rqp ? rqp[c] : nullptr;
        a.svorts[c] = h->mid[c];
        a.rk.q[c] = h->svor[c];
        a.rk.qdi[c] = h->qdi[c];
        a.rk.qdf[c] = h->qdf[c];
    }
    a.rk.scal = h->scal;
    a.mode = mode;
    a.stage = stage;
    a.bt = batch(h);
    return a;
}

static ZOpArgs zop_args(ps3d_handle *h, int op, const double *in, double *out) {
    ZOpArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = h->g;
    a.tb = ztabs(h);
    a.TQ = h->TQ;
    a.QPB = h->QPB;
    a.nquads = h->nquads;
    a.pz = h->pz;
    a.twz = h->twz;
    a.zscale = 1.0 / std::sqrt(2.0 * h->g.nz);
    a.rkz_sc = 3.141592653589793238462643383279502884 / h->p.extent[2];
    a.in = in;
    a.out = out;
    a.op = op;
    return a;
}

static int run_mean(ps3d_handle *h, int mode, int advance_time) {
    MeanArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = h->g;
    a.pz = h->pz;
    a.twz = h->twz;
    a.zscale = 1.0 / std::sqrt(2.0 * h->g.nz);
    a.svor[0] = h->svor[0];
    a.svor[1] = h->svor[1];
    a.scal = h->scal;
    a.mode = mode;
    a.advance_time = advance_time;
    a.bt = batch(h);
    prof_begin(h, PS3D_K_MEAN);
    HIPCHK(launch_mean(a, h->g.nz, h->stream));
    prof_end(h);
    return 0;
}

// vor2vel + the adapt-stage y pass: leaves u, v, w, xi, eta, zeta, their y
// derivatives and z midpoints x-spectral / y-physical, svel semi-spectral, the
// projected svor (project = true) and the Parseval sums finished into scal.
static int inverse_adapt_stage(ps3d_handle *h, bool project) {
    ZInvArgs z = zinv_args(h);
    if (project) {
        z.svor_out[0] = h->svor[0];
        z.svor_out[1] = h->svor[1];
    }
    for (int c = 0; c < 3; ++c) z.mid[c] = h->mid[c];
    prof_begin(h, PS3D_K_ZINV);
    HIPCHK(launch_zinv(h->zl, z, h->stream));
    prof_end(h);
    LineArgs y = line_args(h, 1, LM_INV_Y1);
    for (int c = 0; c < 3; ++c) {
        y.in[c] = h->svel[c];
        y.out[c] = h->vel[c];
        y.in[3 + c] = h->vor[c];
        y.out[3 + c] = h->vor[c];
        y.in[6 + c] = h->mid[c];
        y.out[6 + c] = h->mid[c];
        y.out[9 + c] = h->grad[c];
    }
    y.part = h->party;
    y.bt.part = h->b_party;
    prof_begin(h, PS3D_K_YINV_ADAPT);
    int st = run_line(h, y);
    prof_end(h);
    if (st) return st;
    prof_begin(h, PS3D_K_REDUCE);
    HIPCHK(launch_fin_y(fin_args(h), h->stream));
    prof_end(h);
    return 0;
}

// velmax_given: the products pass of this stage has written the velocity maxima
static int x_adapt(ps3d_handle *h, int phys_in, bool velmax_given = false) {
    LineArgs x = line_args(h, 0, LM_XADAPT);
    x.phys_in = phys_in;
    x.velmax_given = velmax_given;
    for (int c = 0; c < 3; ++c) {
        x.in[c] = h->vel[c];
        x.in[3 + c] = h->vor[c];
        x.in[6 + c] = h->grad[c];
        x.in[9 + c] = h->mid[c];
    }
    x.part = h->partx;
    x.bt.part = h->b_partx;
    prof_begin(h, PS3D_K_XADAPT);
    int st = run_line(h, x);
    prof_end(h);
    if (st) return st;
    FinArgs f = fin_args(h);
    if (velmax_given) {
        f.partm = h->partm;
        f.nbm = h->lxm.nother * h->lxm.nzt;
    }
    prof_begin(h, PS3D_K_REDUCE);
    HIPCHK(launch_fin_x(f, h->stream));
    prof_end(h);
    return 0;
}

// x pass (u x omega) + y forward: r, q, p left semi-spectral in dst[]; with
// velmax, the x pass also writes the per-wave velocity maxima into partm
static int products(ps3d_handle *h, int phys_in, double *const *dst, bool nn_export = false, bool velmax = false) {
    LineArgs x = line_args(h, 0, LM_XMAIN);
    x.phys_in = phys_in;
    if (velmax) {
        x.part = h->partm;
        x.bt.part = h->b_partm;
    }
    for (int c = 0; c < 3; ++c) {
        x.in[c] = h->vel[c];
        x.in[3 + c] = h->vor[c];
        x.out[c] = dst[c];
    }
    if (h->cnn && nn_export) {   // the x pass also writes the closure input
        x.nn_in = ps3d::cnn_input(h->cnn);
        for (int k = 0; k < 6; ++k) x.nn_scale[k] = h->cnn_in[k];
    }
    prof_begin(h, PS3D_K_XMAIN);
    int st = run_line(h, x);
    prof_end(h);
    if (st) return st;
    prof_begin(h, PS3D_K_YFWD);
    st = line_fields(h, 1, LM_FWD, 3, dst, dst);
    prof_end(h);
    return st;
}

// Whether RK4 stage `stage`'s tendency pass also runs the vor2vel z pass of the
// next stage (wave kernels; PS3D_ZFUSE=0 keeps them apart, for comparison).
static bool fuse_next(const ps3d_handle *h, int mode, int stage) {
    return mode == ZT_RK4 && stage <= 3 && h->zl.wave != 0 && h->zfuse;
}

// One fused RK stage (impl_rk4) or cn2 iteration.  stage 1 includes adapt.
static int fused_stage(ps3d_handle *h, int mode, int stage) {
    int st;
    double *const *rqp = h->vel;
    if (stage == 1) {
        if ((st = inverse_adapt_stage(h, true))) return st;
        // the products pass goes first and also takes the velocity maxima; its output
        // goes to svel (free after the y pass), as the adapt pass still reads vel
        if ((st = products(h, 0, h->svel, true, true))) return st;
        if ((st = x_adapt(h, 0, true))) return st;
        rqp = h->svel;
    } else {
        if (!fuse_next(h, mode, stage - 1)) {
            ZInvArgs z = zinv_args(h);
            prof_begin(h, PS3D_K_ZINV);
            HIPCHK(launch_zinv(h->zl, z, h->stream));
            prof_end(h);
        }
        double *in[6] = {h->svel[0], h->svel[1], h->svel[2], h->vor[0], h->vor[1], h->vor[2]};
        double *out[6] = {h->vel[0], h->vel[1], h->vel[2], h->vor[0], h->vor[1], h->vor[2]};
        prof_begin(h, PS3D_K_YINV);
        st = line_fields(h, 1, LM_INV, 6, in, out);
        prof_end(h);
        if (st) return st;
        if ((st = products(h, 0, h->vel))) return st;
    }
    ZTendArgs t = ztend_args(h, mode, stage, rqp);
    if (h->cnn) {
        // closure: net on (u, v, w, omega) at the step's first stage -> physical
        // tendency in grad, decomposed like a physical field into mid and added to the
        // source of every stage of the step (held fixed over the step, as the
        // subgrid term is updated once per step; mid is free after the adapt stage)
        if (stage == 1) {
            HIPCHK(ps3d::cnn_apply(h->cnn, h->cnn_out, h->g.zs, h->grad[0], h->grad[1], h->grad[2], h->stream));
            if ((st = line_fields(h, 1, LM_FWD, 3, h->grad, h->grad))) return st;
            if ((st = line_fields(h, 0, LM_FWD, 3, h->grad, h->grad))) return st;
            for (int c = 0; c < 3; ++c)
                HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DECOMPOSE, h->grad[c], h->mid[c]), h->stream));
        }
        for (int c = 0; c < 3; ++c) t.extra[c] = h->mid[c];
    }
    if (fuse_next(h, mode, stage)) {
        prof_begin(h, PS3D_K_ZTEND_ZINV);
        HIPCHK(launch_ztend_zinv_w(h->zl.nz, h->zl.wave, h->zl.nquads, t, zinv_args(h), h->stream));
        prof_end(h);
        return 0;
    }
    prof_begin(h, PS3D_K_ZTEND);
    HIPCHK(launch_ztend(h->zl, t, h->stream));
    prof_end(h);
    return 0;
}

static int one_step(ps3d_handle *h) {
    int st;
    if (h->p.stepper == PS3D_STEPPER_CN2) {
        for (int it = 0; it < 3; ++it) {
            if ((st = fused_stage(h, ZT_CN2, it == 0 ? 1 : 2))) return st;
            if ((st = run_mean(h, MEAN_ADJUST, it == 2 ? 2 : 0))) return st;
        }
        return 0;
    }
    for (int s = 1; s <= 4; ++s)
        if ((st = fused_stage(h, ZT_RK4, s))) return st;
    return run_mean(h, MEAN_ADJUST, 1);
}

// host <-> device copies between ps3d's (0:nz, y, x) layout and the padded columns
// Host <-> device copies between ps3d's (0:nz, y, x) layout and the padded columns.
// A strided 2-D copy from pageable memory moves one short row at a time; instead
// whole padded columns go through one pinned staging buffer of the process in large
// contiguous copies, and the (un)padding is a memcpy per column on the host.
namespace {
constexpr size_t kStageBytes = 16u << 20;
std::mutex g_stage_mu;
double *g_stage = nullptr;
}  // namespace

static int with_stage(double **stage) {
    if (!g_stage && hipHostMalloc((void **)&g_stage, kStageBytes, hipHostMallocDefault) != hipSuccess) {
        g_stage = nullptr;
        return fail(PS3D_E_NOMEM, "pinned staging buffer");
    }
    *stage = g_stage;
    return 0;
}

static int copy_in(ps3d_handle *h, double *dst, const double *src) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    double *stage;
    int st;
    if ((st = with_stage(&stage))) return st;
    const size_t nz1 = (size_t)h->g.nz + 1, zs = h->g.zs, ncol = (size_t)h->g.nx * h->g.ny;
    const size_t cpc = std::max<size_t>(1, kStageBytes / (sizeof(double) * zs));
    for (size_t c0 = 0; c0 < ncol; c0 += cpc) {
        const size_t nc = std::min(cpc, ncol - c0);
        for (size_t c = 0; c < nc; ++c) {
            std::memcpy(stage + c * zs, src + (c0 + c) * nz1, sizeof(double) * nz1);
            std::memset(stage + c * zs + nz1, 0, sizeof(double) * (zs - nz1));
        }
        HIPCHK(hipMemcpyAsync(dst + c0 * zs, stage, sizeof(double) * nc * zs, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));   // the buffer is refilled next
    }
    return 0;
}
static int copy_out(ps3d_handle *h, double *dst, const double *src) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    double *stage;
    int st;
    if ((st = with_stage(&stage))) return st;
    const size_t nz1 = (size_t)h->g.nz + 1, zs = h->g.zs, ncol = (size_t)h->g.nx * h->g.ny;
    const size_t cpc = std::max<size_t>(1, kStageBytes / (sizeof(double) * zs));
    for (size_t c0 = 0; c0 < ncol; c0 += cpc) {
        const size_t nc = std::min(cpc, ncol - c0);
        HIPCHK(hipMemcpyAsync(stage, src + c0 * zs, sizeof(double) * nc * zs, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (size_t c = 0; c < nc; ++c) std::memcpy(dst + (c0 + c) * nz1, stage + c * zs, sizeof(double) * nz1);
    }
    return 0;
}

static int read_scal(ps3d_handle *h, double *out) {
    HIPCHK(hipMemcpyAsync(out, h->scal, sizeof(double) * SC_NSCAL, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}
static int write_scal(ps3d_handle *h, int idx, double v) {
    HIPCHK(hipMemcpyAsync(h->scal + idx, &v, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int ensure_diag(ps3d_handle *h, int rows) {
    if (rows <= h->diag_cap) return 0;
    drop_step_graph(h);
    (void)hipFree(h->diag);
    h->diag = nullptr;
    HIPCHK(hipMalloc((void **)&h->diag, sizeof(double) * PS3D_NDIAG * (size_t)rows * h->nbatch));
    h->diag_cap = rows;
    h->b_diag = (long)PS3D_NDIAG * rows;
    return 0;
}

// A single-instance view of instance b of a batched handle: the same tables,
// stream and launch shapes, data pointers moved to the instance.  Used for the
// per-instance set-up and copies; never freed.
static ps3d_handle make_view(const ps3d_handle *h, int b) {
    ps3d_handle v;
    v.p = h->p;
    v.device = h->device;
    v.stream = h->stream;
    v.g = h->g;
    v.nvol = h->nvol;
    const long o = (long)b * h->b_vol;
    for (int c = 0; c < 3; ++c) {
        v.svor[c] = h->svor[c] + o;
        v.qdi[c] = h->qdi[c] + o;
        v.qdf[c] = h->qdf[c] + o;
        v.svel[c] = h->svel[c] + o;
        v.vel[c] = h->vel[c] + o;
        v.vor[c] = h->vor[c] + o;
        v.grad[c] = h->grad[c] + o;
        v.mid[c] = h->mid[c] + o;
    }
    v.scal = h->scal + (size_t)b * SC_NSCAL;
    v.party = h->party + (size_t)b * h->b_party;
    v.partx = h->partx + (size_t)b * h->b_partx;
    v.partm = h->partm + (size_t)b * h->b_partm;
    v.roll = h->roll + (size_t)b * h->b_roll;
    v.diag = h->diag ? h->diag + (size_t)b * h->b_diag : nullptr;
    v.diag_cap = h->diag_cap;
    v.rkx = h->rkx; v.rky = h->rky; v.filx = h->filx; v.fily = h->fily; v.filz = h->filz;
    v.twx = h->twx; v.twy = h->twy; v.twz = h->twz;
    v.px = h->px; v.py = h->py; v.pz = h->pz;
    v.zl = h->zl;
    v.TQ = h->TQ; v.QPB = h->QPB; v.nquads = h->nquads;
    v.lx = h->lx; v.ly = h->ly;
    v.lxm = h->lxm; v.lxr = h->lxr; v.lyr = h->lyr;
    v.rtm = h->rtm; v.rtx = h->rtx; v.rty = h->rty;
    v.h_rkx = h->h_rkx;
    v.h_rky = h->h_rky;
    v.nbatch = 1;
    v.have_state = true;
    return v;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *ps3d_gpu_last_error(void) { return g_err.c_str(); }
const char *ps3d_gpu_version(void) { return "ps3d-mi355x 0.1 (gfx950, FP64)"; }

void ps3d_params_default(ps3d_params *p) {
    std::memset(p, 0, sizeof *p);
    p->nx = p->ny = p->nz = 32;
    for (int d = 0; d < 3; ++d) p->extent[d] = 1.0;
    p->filtering = PS3D_FILTER_HOU_LI;
    p->stepper = PS3D_STEPPER_IMPL_RK4;
    p->nnu = 3;
    p->prediss = 10.0;
    p->pretype = PS3D_PRE_ROLL_MEAN_MAX_STRAIN;
    p->roll_mean_win_size = 1000;
    p->length_scale = PS3D_LS_KOLMOGOROV;
    p->alpha = 0.1;
    p->limit = 0.0;
    p->dt_fixed = 0.0;
}

int ps3d_gpu_create(const ps3d_params *p, ps3d_handle **out) {
    if (!p || !out) return fail(PS3D_E_ARG, "null argument");
    *out = nullptr;
    ps3d_handle *h = new ps3d_handle();
    int st = init_handle(h, p, true, true);
    if (st) {
        std::string msg = g_err;
        free_handle(h);
        g_err = msg;
        return st;
    }
    *out = h;
    return 0;
}

int ps3d_gpu_destroy(ps3d_handle *h) {
    free_handle(h);
    return 0;
}

int ps3d_gpu_sync(ps3d_handle *h) {
    if (!h) return fail(PS3D_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// setup_fields (utils.f90:136-184)
static int finish_setup(ps3d_handle *h);

int ps3d_gpu_set_vorticity(ps3d_handle *h, const double *vor) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !vor) return fail(PS3D_E_ARG, "null argument");
    drop_step_graph(h);
    const size_t hv = (size_t)h->g.nx * h->g.ny * (h->g.nz + 1);
    int st;
    for (int c = 0; c < 3; ++c)
        if ((st = copy_in(h, h->vel[c], vor + c * hv))) return st;
    // field_decompose_physical: fftxyp2s (y then x) + decompose
    if ((st = line_fields(h, 1, LM_FWD, 3, h->vel, h->vel))) return st;
    if ((st = line_fields(h, 0, LM_FWD, 3, h->vel, h->vel))) return st;
    for (int c = 0; c < 3; ++c) HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DECOMPOSE, h->vel[c], h->svor[c]), h->stream));
    return finish_setup(h);
}

// The part of setup_fields after svor holds the initial vorticity: ini_vor_mean,
// t = 0, rolling-mean state, vor2vel and init_diffusion (utils.f90:160-184).
static int finish_setup(ps3d_handle *h) {
    int st;
    if ((st = run_mean(h, MEAN_CALC, 0))) return st;
    HIPCHK(hipMemsetAsync(h->scal + SC_T, 0, sizeof(double), h->stream));
    double init[4] = {0.0, 1.0, 1.0, 0.0};   // rolling mean: sma, inew, iold, filled
    HIPCHK(hipMemcpyAsync(h->scal + SC_RM_SMA, init, sizeof init, hipMemcpyHostToDevice, h->stream));
    h->have_state = true;
    if ((st = ps3d_gpu_vor2vel(h))) return st;
    double sc[SC_NSCAL];
    if ((st = read_scal(h, sc))) return st;
    // init_diffusion -> get_viscosity (inversion_utils.f90:124-184)
    const double te = sc[SC_KE], en = sc[SC_EN];
    double rkxmax = 0, rkymax = 0;
    for (double v : h->h_rkx) rkxmax = std::max(rkxmax, v);
    for (double v : h->h_rky) rkymax = std::max(rkymax, v);
    const double K2max = std::pow(std::max(rkxmax, rkymax), 2);
    const double rkmsi = 1.0 / K2max;
    if (h->p.length_scale == PS3D_LS_KOLMOGOROV)
        h->visc = h->p.prediss * std::pow(K2max * te / en, 1.0 / 3.0) * ipow(rkmsi, h->p.nnu);
    else
        h->visc = h->p.prediss * ipow(rkmsi, h->p.nnu);
    return write_scal(h, SC_VISC, h->visc);
}

// vor2vel: device state -> physical vel, vor (and y-derivatives / midpoints for adapt)
int ps3d_gpu_vor2vel(ps3d_handle *h) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h) return fail(PS3D_E_ARG, "null handle");
    if (!h->have_state) return fail(PS3D_E_STATE, "vor2vel before set_vorticity");
    int st;
    if ((st = inverse_adapt_stage(h, true))) return st;
    double *f[12];
    for (int c = 0; c < 3; ++c) {
        f[c] = h->vel[c];
        f[3 + c] = h->vor[c];
        f[6 + c] = h->grad[c];
        f[9 + c] = h->mid[c];
    }
    if ((st = line_fields(h, 0, LM_INV, 12, f, f))) return st;
    h->have_phys = true;
    h->have_grad = true;
    h->have_adapt = false;
    h->derived_fresh = false;
    return 0;
}

int ps3d_gpu_adapt(ps3d_handle *h, double t, double *dt, double *diag) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h) return fail(PS3D_E_ARG, "null handle");
    if (!h->have_phys) return fail(PS3D_E_STATE, "adapt needs a preceding vor2vel");
    int st;
    if ((st = ensure_diag(h, 1))) return st;
    double zero = 0.0;
    HIPCHK(hipMemcpyAsync(h->scal + SC_T, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->scal + SC_STEP, &zero, sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((st = x_adapt(h, 1))) return st;
    double row[PS3D_NDIAG];
    HIPCHK(hipMemcpyAsync(row, h->diag, sizeof row, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (dt) *dt = row[PS3D_D_DT];
    if (diag) std::memcpy(diag, row, sizeof row);
    h->have_adapt = true;
    return 0;
}

int ps3d_gpu_viscosity(ps3d_handle *h, double *visc, double *pref) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h) return fail(PS3D_E_ARG, "null handle");
    double sc[SC_NSCAL];
    int st;
    if ((st = read_scal(h, sc))) return st;
    if (visc) *visc = sc[SC_VISC];
    if (pref) *pref = sc[SC_PREF];
    return 0;
}

int ps3d_gpu_source(ps3d_handle *h) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h) return fail(PS3D_E_ARG, "null handle");
    if (!h->have_phys) return fail(PS3D_E_STATE, "source needs a preceding vor2vel");
    // the closure term is formed inside the fused step only (its tendency is held in
    // the buffer this path keeps svorts in)
    if (h->cnn) return fail(PS3D_E_UNSUPPORTED, "source with a closure attached: use ps3d_gpu_advance_n");
    int st;
    // r, q, p into grad (adapt has consumed it), then the z part into svorts (= mid)
    h->have_grad = false;
    if ((st = products(h, 1, h->grad))) return st;
    ZTendArgs t = ztend_args(h, ZT_STORE, 0, h->grad);
    HIPCHK(launch_ztend(h->zl, t, h->stream));
    return 0;
}

int ps3d_gpu_step(ps3d_handle *h, double *t, double dt) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !t) return fail(PS3D_E_ARG, "null argument");
    if (!h->have_adapt) return fail(PS3D_E_STATE, "step needs vor2vel, adapt and source first");
    if (h->cnn) return fail(PS3D_E_UNSUPPORTED, "step with a closure attached: use ps3d_gpu_advance_n");
    int st;
    HIPCHK(hipMemcpyAsync(h->scal + SC_T, t, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->scal + SC_DT, &dt, sizeof(double), hipMemcpyHostToDevice, h->stream));
    const int mode = h->p.stepper == PS3D_STEPPER_CN2 ? ZT_CN2 : ZT_RK4;
    ZTendArgs a = ztend_args(h, mode, 1, nullptr);
    HIPCHK(launch_zrk(h->zl, a, h->stream));
    if (mode == ZT_CN2) {
        if ((st = run_mean(h, MEAN_ADJUST, 0))) return st;
        for (int it = 1; it <= 2; ++it) {
            if ((st = fused_stage(h, ZT_CN2, 2))) return st;
            if ((st = run_mean(h, MEAN_ADJUST, it == 2 ? 2 : 0))) return st;
        }
    } else {
        for (int s = 2; s <= 4; ++s)
            if ((st = fused_stage(h, ZT_RK4, s))) return st;
        if ((st = run_mean(h, MEAN_ADJUST, 1))) return st;
    }
    double sc[SC_NSCAL];
    if ((st = read_scal(h, sc))) return st;
    *t = sc[SC_T];
    h->have_phys = false;
    h->have_adapt = false;
    h->have_grad = false;
    return 0;
}

// Captures one step (with gated: the gate kernel first and every kernel of the
// step behind it) into a graph and instantiates it.
static int capture_step(ps3d_handle *h, bool gated, hipGraph_t *graph, hipGraphExec_t *exec) {
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    int st = 0;
    if (gated) {
        h->g.gate = h->scal + SC_GATE;
        const hipError_t e = launch_gate(h->scal, h->stream);
        if (e != hipSuccess) st = fail(PS3D_E_HIP, std::string("gate: ") + hipGetErrorString(e));
    }
    if (!st) st = one_step(h);
    h->g.gate = nullptr;
    hipGraph_t gr = nullptr;
    const hipError_t ce = hipStreamEndCapture(h->stream, &gr);
    if (st || ce != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return st ? st : fail(PS3D_E_HIP, std::string("step capture: ") + hipGetErrorString(ce));
    }
    const hipError_t ie = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(gr);
        *exec = nullptr;
        return fail(PS3D_E_HIP, std::string("step graph: ") + hipGetErrorString(ie));
    }
    *graph = gr;
    return 0;
}

// After a run of steps: time, kernel timing, the diagnostics rows (rows < 0: as
// many as the device's row counter says were written).
static int finish_run(ps3d_handle *h, double *t, int rows, double *diag_series) {
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    double sc[SC_NSCAL];
    int st;
    if ((st = read_scal(h, sc))) return st;
    *t = sc[SC_T];
    if (rows < 0) rows = (int)sc[SC_STEP];
    h->last_steps = rows;
    if (diag_series && rows > 0)
        HIPCHK(hipMemcpy(diag_series, h->diag, sizeof(double) * PS3D_NDIAG * (size_t)rows, hipMemcpyDeviceToHost));
    h->have_phys = false;
    h->have_adapt = false;
    h->have_grad = false;
    return 0;
}

int ps3d_gpu_advance_n(ps3d_handle *h, double *t, int nsteps, double *diag_series) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !t || nsteps < 0) return fail(PS3D_E_ARG, "invalid argument");
    if (!h->have_state) return fail(PS3D_E_STATE, "advance before set_vorticity");
    int st;
    if ((st = ensure_diag(h, std::max(1, nsteps)))) return st;
    double zero = 0.0;
    HIPCHK(hipMemcpyAsync(h->scal + SC_T, t, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->scal + SC_STEP, &zero, sizeof(double), hipMemcpyHostToDevice, h->stream));
    // from two steps on, the step is captured once and replayed (it has no host
    // work: dt and the diagnostics row counter live on the device); PS3D_GRAPH=0
    // launches every kernel instead
    const char *gv = std::getenv("PS3D_GRAPH");
    const bool graph = nsteps >= 2 && !(gv && gv[0] == '0');
    if (graph && !h->sexec && (st = capture_step(h, false, &h->sgraph, &h->sexec))) return st;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int s = 0; s < nsteps; ++s) {
        if (graph)
            HIPCHK(hipGraphLaunch(h->sexec, h->stream));
        else if ((st = one_step(h)))
            return st;
    }
    return finish_run(h, t, nsteps, diag_series);
}

int ps3d_gpu_advance_until(ps3d_handle *h, double *t, double t_stop, int nmax, double *diag_series, int *nsteps) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !t || nmax < 0) return fail(PS3D_E_ARG, "invalid argument");
    if (!h->have_state) return fail(PS3D_E_STATE, "advance before set_vorticity");
    if (h->cnn) return fail(PS3D_E_UNSUPPORTED, "advance_until with a closure attached: use ps3d_gpu_advance_n");
    int st;
    if ((st = ensure_diag(h, std::max(1, nmax)))) return st;
    // SC_STEP, SC_GATE, SC_TSTOP are consecutive slots
    const double init[3] = {0.0, 0.0, std::min(t_stop, h->p.limit)};
    HIPCHK(hipMemcpyAsync(h->scal + SC_T, t, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->scal + SC_STEP, init, sizeof init, hipMemcpyHostToDevice, h->stream));
    if (!h->gexec && (st = capture_step(h, true, &h->ggraph, &h->gexec))) return st;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int s = 0; s < nmax; ++s) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    if ((st = finish_run(h, t, -1, diag_series))) return st;
    if (nsteps) *nsteps = h->last_steps;
    return 0;
}

int ps3d_gpu_rk_substep(ps3d_handle *h, int stage, double dt, const double *vdiss) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !vdiss || stage < 1 || stage > 4) return fail(PS3D_E_ARG, "invalid argument");
    const size_t n2 = (size_t)h->g.nx * h->g.ny;
    double *dv = nullptr;
    HIPCHK(hipMalloc((void **)&dv, sizeof(double) * n2));
    HIPCHK(hipMemcpy(dv, vdiss, sizeof(double) * n2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpyAsync(h->scal + SC_DT, &dt, sizeof(double), hipMemcpyHostToDevice, h->stream));
    ZTendArgs a = ztend_args(h, ZT_RK4, stage, nullptr);
    a.rk.vdiss = dv;
    hipError_t e = launch_zrk(h->zl, a, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dv);
    if (e != hipSuccess) return fail(PS3D_E_HIP, std::string("rk_substep: ") + hipGetErrorString(e));
    return 0;
}

int ps3d_gpu_profile_kernels(ps3d_handle *h, int nsteps, double *ms, int *launches, double *bytes) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || nsteps < 1) return fail(PS3D_E_ARG, "invalid argument");
    if (!h->have_state) return fail(PS3D_E_STATE, "profile before set_vorticity");
    int st;
    if ((st = ensure_diag(h, nsteps))) return st;
    double zero = 0.0;
    HIPCHK(hipMemcpyAsync(h->scal + SC_STEP, &zero, sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->prof = true;
    h->prof_err = false;
    h->prof_used = 0;
    h->prof_kind.clear();
    for (int s = 0; s < nsteps && !st; ++s) st = one_step(h);
    h->prof = false;
    if (st) return st;
    if (h->prof_err) return fail(PS3D_E_HIP, "profiling events could not be created or recorded");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < PS3D_NKIND; ++k) {
        if (ms) ms[k] = 0.0;
        if (launches) launches[k] = 0;
    }
    for (size_t i = 0; i < h->prof_kind.size(); ++i) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, h->prof_ev[2 * i], h->prof_ev[2 * i + 1]));
        const int k = h->prof_kind[i];
        if (ms) ms[k] += t;
        if (launches) launches[k] += 1;
    }
    if (bytes) {
        // algorithmic bytes per launch: each field a pass reads and writes, once
        const double vol = 8.0 * (double)h->g.nx * h->g.ny * (h->g.nz + 1);
        const bool cn2 = h->p.stepper == PS3D_STEPPER_CN2;
        // z inverse: 3 R + 6 W; at the adapt stage + 2 W (projected svor) + 3 W (midpoints)
        bytes[PS3D_K_ZINV] = (cn2 ? (14.0 + 2 * 9.0) / 3.0 : (14.0 + 3 * 9.0) / 4.0) * vol;
        bytes[PS3D_K_YINV] = 12.0 * vol;        // 6 R + 6 W
        bytes[PS3D_K_XMAIN] = 9.0 * vol;        // 6 R + 3 W
        bytes[PS3D_K_YFWD] = 6.0 * vol;         // 3 R + 3 W
        // z forward + update, per stage: rk4 15, 15, 15, 9 ; cn2 12, 9, 9
        bytes[PS3D_K_ZTEND] = (cn2 ? 10.0 : 13.5) * vol;
        // fused: stage 1..3 tendency + update without the stage result (12) and the
        // next stage's z inverse without its svor read (6 W); stage 4 alone is 9; the
        // z inverse left is the adapt stage's (14)
        bytes[PS3D_K_ZTEND_ZINV] = 18.0 * vol;
        if (!cn2 && h->zl.wave && h->zfuse) {
            bytes[PS3D_K_ZTEND] = 9.0 * vol;
            bytes[PS3D_K_ZINV] = 14.0 * vol;
        }
        bytes[PS3D_K_YINV_ADAPT] = 21.0 * vol;  // 9 R + 12 W
        // x adapt in a step: 11 R (u, w for their x derivatives, the y derivatives,
        // omega, the midpoints); the velocity maxima come from the products pass
        bytes[PS3D_K_XADAPT] = 11.0 * vol;
        bytes[PS3D_K_REDUCE] = 0.0;
        bytes[PS3D_K_MEAN] = 0.0;
    }
    return 0;
}

int ps3d_gpu_last_timing(ps3d_handle *h, double *ms, int *nsteps) {
    if (!h) return fail(PS3D_E_ARG, "null handle");
    if (ms) *ms = h->last_ms;
    if (nsteps) *nsteps = h->last_steps;
    return 0;
}

// pressure + horizontal_divergence (fields_derived.f90:67-182) from the state of the
// last vor2vel: svel (semi-spectral), vor and the y-derivatives du/dy, dv/dy, dw/dy
// (physical).  qdi / qdf serve as scratch: they only carry data inside a step.
int ps3d_gpu_derived(ps3d_handle *h) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h) return fail(PS3D_E_ARG, "null handle");
    if (!h->have_phys || !h->have_grad)
        return fail(PS3D_E_STATE, "derived needs a preceding vor2vel (and no source since)");
    const size_t bytes = sizeof(double) * h->nvol;
    if (!h->pres) HIPCHK(hipMalloc((void **)&h->pres, bytes));
    if (!h->delta) HIPCHK(hipMalloc((void **)&h->delta, bytes));
    int st;
    double *dudx = h->qdi[0], *dwdx = h->qdi[1], *dvdy_s = h->qdi[2], *div_s = h->qdf[0];
    {   // du/dx, dw/dx (spectral) and dv/dy (spectral)
        double *in[2] = {h->svel[0], h->svel[2]};
        double *out[2] = {dudx, dwdx};
        if ((st = line_fields(h, 0, LM_DIFF, 2, in, out))) return st;
        double *in1[1] = {h->svel[1]};
        double *out1[1] = {dvdy_s};
        if ((st = line_fields(h, 1, LM_DIFF, 1, in1, out1))) return st;
    }
    HIPCHK(launch_add2((long)h->nvol, dudx, dvdy_s, div_s, h->stream));
    {   // fftxys2p of du/dx, dw/dx and the divergence
        double *in[3] = {dudx, dwdx, div_s};
        double *out[3] = {dudx, dwdx, h->delta};
        if ((st = line_fields(h, 0, LM_INV, 3, in, out))) return st;
        if ((st = line_fields(h, 1, LM_INV, 3, out, out))) return st;
    }
    PresArgs pa;
    pa.n = (long)h->nvol;
    pa.dudx = dudx;
    pa.dudy = h->grad[0];
    pa.dvdy = h->grad[1];
    pa.dwdx = dwdx;
    pa.dwdy = h->grad[2];
    pa.xi = h->vor[0];
    pa.eta = h->vor[1];
    pa.zeta = h->vor[2];
    pa.pres = h->pres;
    HIPCHK(launch_pres_rhs(pa, h->stream));
    double *p1[1] = {h->pres};
    if ((st = line_fields(h, 1, LM_FWD, 1, p1, p1)) || (st = line_fields(h, 0, LM_FWD, 1, p1, p1))) return st;
    HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_POISSON, h->pres, h->pres), h->stream));
    if ((st = line_fields(h, 0, LM_INV, 1, p1, p1)) || (st = line_fields(h, 1, LM_INV, 1, p1, p1))) return st;
    h->have_derived = true;
    h->derived_fresh = true;
    return 0;
}

// field_stats record (enum ps3d_stat) of the physical state of the last vor2vel, on
// the handle's stream: partials in part, the record in out (device memory).
static int field_stats_launch(ps3d_handle *h, const double *diag, double *part, double *out) {
    StatsArgs a;
    std::memset(&a, 0, sizeof a);
    a.nx = h->g.nx;
    a.ny = h->g.ny;
    a.nz = h->g.nz;
    a.zs = h->g.zs;
    for (int c = 0; c < 3; ++c) {
        a.vel[c] = h->vel[c];
        a.vor[c] = h->vor[c];
    }
    a.delta = h->delta;
    a.part = part;
    a.nb = stats_blocks(h->g.nx, h->g.ny);
    if (a.nb > kStatsMaxBlocks) return fail(PS3D_E_UNSUPPORTED, "field_stats: partition beyond the partials array");
    StatsFin f;
    std::memset(&f, 0, sizeof f);
    f.part = part;
    f.nb = a.nb;
    f.ncelli = 1.0 / ((double)h->g.nx * h->g.ny * h->g.nz);
    f.nxnyi = 1.0 / ((double)h->g.nx * h->g.ny);
    f.fcor3 = h->p.f_cor[2];
    static_assert(PS3D_NDIAG <= (int)(sizeof f.ad / sizeof f.ad[0]), "adapt record fits");
    if (diag) std::memcpy(f.ad, diag, sizeof(double) * PS3D_NDIAG);
    f.out = out;
    HIPCHK(launch_stats(a, f, h->stream));
    return 0;
}

int ps3d_gpu_field_stats(ps3d_handle *h, const double *diag, double *stats) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !stats) return fail(PS3D_E_ARG, "null argument");
    if (!h->have_phys) return fail(PS3D_E_STATE, "field_stats needs a preceding vor2vel");
    int st;
    if (!h->derived_fresh && (st = ps3d_gpu_derived(h))) return st;
    if (!h->stats_part) HIPCHK(hipMalloc((void **)&h->stats_part, sizeof(double) * kStatsMaxBlocks * ST_NP));
    if (!h->stats_out) HIPCHK(hipMalloc((void **)&h->stats_out, sizeof(double) * PS3D_NSTAT));
    if ((st = field_stats_launch(h, diag, h->stats_part, h->stats_out))) return st;
    HIPCHK(hipMemcpyAsync(stats, h->stats_out, sizeof(double) * PS3D_NSTAT, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int ps3d_gpu_field_stats_noalloc(ps3d_handle *h, const double *diag, double *stats_dev) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !stats_dev) return fail(PS3D_E_ARG, "null argument");
    if (!h->have_phys || !h->derived_fresh)
        return fail(PS3D_E_STATE, "field_stats_noalloc needs vor2vel and then ps3d_gpu_derived first");
    if ((long)stats_blocks(h->g.nx, h->g.ny) * ST_NP > (long)kStatsMaxBlocks * ST_NP)
        return fail(PS3D_E_UNSUPPORTED, "field_stats_noalloc: grid needs more partials than the library's array holds");
    double *scr = nullptr;
    HIPCHK(stats_scratch(h->device, &scr));
    return field_stats_launch(h, diag, scr, stats_dev);
}

// ---------------------------------------------------------------------------
// genspec's spectrum and nc_reader's derived means
// ---------------------------------------------------------------------------
constexpr int kSpecMaxBlocks = 1024;

static int spec_buffers(ps3d_handle *h, int nbins) {
    if (nbins <= h->spec_cap && h->spec_part) return 0;
    (void)hipFree(h->spec_part);
    (void)hipFree(h->spec_out);
    (void)hipFree(h->spec_cpart);
    (void)hipFree(h->spec_cnt);
    h->spec_part = h->spec_out = nullptr;
    h->spec_cpart = nullptr;
    h->spec_cnt = nullptr;
    h->spec_cap = 0;
    const size_t np = (size_t)kSpecMaxBlocks * nbins;
    if (hipMalloc((void **)&h->spec_part, sizeof(double) * std::max<size_t>(np, (size_t)kStatsMaxBlocks * MO_NP)) != hipSuccess ||
        hipMalloc((void **)&h->spec_cpart, sizeof(int) * np) != hipSuccess ||
        hipMalloc((void **)&h->spec_out, sizeof(double) * (nbins + PS3D_NMOMENT)) != hipSuccess ||
        hipMalloc((void **)&h->spec_cnt, sizeof(long long) * nbins) != hipSuccess)
        return fail(PS3D_E_NOMEM, "spectrum buffers");
    h->spec_cap = nbins;
    return 0;
}

// kinetic energy, helicity and cross-helicity of the handle's physical fields into
// out (device, PS3D_NMOMENT doubles); with_vor = false: kinetic energy only
static int moments_launch(ps3d_handle *h, bool with_vor, double *part, double *out) {
    MomArgs a;
    std::memset(&a, 0, sizeof a);
    a.nx = h->g.nx;
    a.ny = h->g.ny;
    a.nz = h->g.nz;
    a.zs = h->g.zs;
    for (int c = 0; c < 3; ++c) {
        a.vel[c] = h->vel[c];
        a.vor[c] = with_vor ? h->vor[c] : nullptr;
    }
    a.part = part;
    a.nb = stats_blocks(h->g.nx, h->g.ny);
    a.ncelli = 1.0 / ((double)h->g.nx * h->g.ny * h->g.nz);
    a.out = out;
    HIPCHK(launch_moments(a, h->stream));
    return 0;
}

int ps3d_gpu_moments(ps3d_handle *h, double *out) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !out) return fail(PS3D_E_ARG, "null argument");
    if (!h->have_phys) return fail(PS3D_E_STATE, "moments need a preceding vor2vel");
    int st;
    if ((st = spec_buffers(h, std::max(1, h->spec_cap)))) return st;
    double *res = h->spec_out + h->spec_cap;
    if ((st = moments_launch(h, true, h->spec_part, res))) return st;
    HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * PS3D_NMOMENT, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// genspec.f90:62-81: largest shell radius kmax = nint(|K|max) and the shell width dk.
// The largest |K|^2 is the sum of the three largest squared wavenumbers (each term
// and the rounded sum are monotone), so no pass over the grid is needed.
static void spectrum_shells(const ps3d_handle *h, int *kmax, double *dk, double *dki, int *nbins) {
#pragma clang fp contract(off)
    double mx = 0.0, my = 0.0;
    for (double v : h->h_rkx) mx = std::max(mx, v * v);
    for (double v : h->h_rky) my = std::max(my, v * v);
    const double rz = (3.141592653589793238462643383279502884 / h->p.extent[2]) * (double)h->g.nz;
    const double k2 = mx + my + rz * rz;   // rounded as the kernel rounds (no contraction)
    *kmax = (int)std::lround(std::sqrt((double)k2));
    const double nx = h->g.nx, ny = h->g.ny, nz = h->g.nz;
    *dk = (double)*kmax / std::sqrt((0.5 * nx) * (0.5 * nx) + (0.5 * ny) * (0.5 * ny) + nz * nz);
    *dki = 1.0 / *dk;
    // shells the binning keeps: m <= int(kmax * dki) is all it can reach, and shells
    // beyond kmax (reachable when dk < 1) fall outside genspec's spec(0:kmax)
    *nbins = std::min(*kmax, (int)((double)*kmax * *dki)) + 1;
}

// genspec.f90:101-117: shell m scaled by the volume of its spherical shell over its
// point count, then the whole spectrum by ke / sum(spec dk).
static void spectrum_scale(int kmax, double dk, double ke, const double *raw, const long long *cnt, double *spec) {
#pragma clang fp contract(off)
    const double prefactor = 4.0 / 3.0 * 3.141592653589793238462643383279502884 * (dk * dk * dk);
    double tot = 0.0;
    for (int m = 0; m <= kmax; ++m) {
        const long long m1 = m + 1, m0 = m;
        spec[m] = cnt[m] > 0 ? raw[m] * prefactor * (double)(m1 * m1 * m1 - m0 * m0 * m0) / (double)cnt[m] : 0.0;
        tot += spec[m] * dk;
    }
    const double snorm = ke / tot;
    for (int m = 0; m <= kmax; ++m) spec[m] *= snorm;
}

int ps3d_gpu_spectrum(ps3d_handle *h, const double *vel, int cap, int *nshell, double *dk, double *spec,
                      long long *num) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !nshell) return fail(PS3D_E_ARG, "null argument");
    int kmax, nbins;
    double dkv, dki;
    spectrum_shells(h, &kmax, &dkv, &dki, &nbins);
    *nshell = kmax + 1;
    if (dk) *dk = dkv;
    if (!spec) return 0;                    // size query
    if (cap < kmax + 1) return fail(PS3D_E_ARG, "spectrum: cap is smaller than the number of shells");
    if (!(dkv > 0.0)) return fail(PS3D_E_ARG, "spectrum: no nonzero wavenumber");
    if (spec_lds_bytes(nbins) > 160 * 1024) return fail(PS3D_E_UNSUPPORTED, "spectrum: too many shells for the LDS histogram");
    int st;
    if (vel) {
        // a velocity from the caller replaces the handle's physical velocity
        drop_step_graph(h);
        const size_t hv = (size_t)h->g.nx * h->g.ny * (h->g.nz + 1);
        for (int c = 0; c < 3; ++c)
            if ((st = copy_in(h, h->vel[c], vel + c * hv))) return st;
        h->have_phys = h->have_adapt = h->derived_fresh = false;
    } else if (!h->have_phys) {
        return fail(PS3D_E_STATE, "spectrum needs a preceding vor2vel (or a velocity)");
    }
    if ((st = spec_buffers(h, nbins))) return st;
    h->have_grad = false;                   // the y-derivative volumes serve as scratch
    double *S = h->grad[0], *T = h->grad[1];
    double *s1[1] = {S};
    double *res = h->spec_out + h->spec_cap;
    // ke = get_kinetic_energy() of the same velocity (genspec.f90:46)
    if ((st = moments_launch(h, false, h->spec_part, res))) return st;
    HIPCHK(hipMemcpyAsync(spec, res, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double ke = spec[0];
    SpecArgs a;
    std::memset(&a, 0, sizeof a);
    a.nx = h->g.nx;
    a.ny = h->g.ny;
    a.nz = h->g.nz;
    a.zs = h->g.zs;
    a.rkx = h->rkx;
    a.rky = h->rky;
    a.rkz1 = 3.141592653589793238462643383279502884 / h->p.extent[2];
    a.dki = dki;
    a.nbins = nbins;
    a.part = h->spec_part;
    a.nb = (int)std::min<long>(kSpecMaxBlocks, std::max<long>(1, (long)h->g.nx * h->g.ny / 16));
    a.f = T;
    // (1) fftxyp2s of each component, dct of u and v, dst of w (genspec.f90:50-60);
    // (3) squared amplitudes summed per shell, point counts with the first component
    for (int c = 0; c < 3; ++c) {
        double *v1[1] = {h->vel[c]};
        if ((st = line_fields(h, 1, LM_FWD, 1, v1, s1)) || (st = line_fields(h, 0, LM_FWD, 1, s1, s1))) return st;
        HIPCHK(launch_zop(h->zl, zop_args(h, c < 2 ? ZOP_DCT : ZOP_DST, S, T), h->stream));
        a.accumulate = c > 0;
        a.cpart = c == 0 ? h->spec_cpart : nullptr;
        HIPCHK(launch_spectrum(a, h->stream));
    }
    HIPCHK(launch_spectrum_fin(h->spec_part, h->spec_cpart, a.nb, nbins, h->spec_out, h->spec_cnt, h->stream));
    // shells nbins..kmax (dk > 1: beyond the reach of every point) stay empty
    std::vector<double> raw(kmax + 1, 0.0);
    std::vector<long long> cnt(kmax + 1, 0);
    HIPCHK(hipMemcpyAsync(raw.data(), h->spec_out, sizeof(double) * nbins, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), h->spec_cnt, sizeof(long long) * nbins, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // shells beyond kmax (reachable only when dk < 1) are not part of genspec's spectrum
    spectrum_scale(kmax, dkv, ke, raw.data(), cnt.data(), spec);
    if (num)
        for (int m = 0; m <= kmax; ++m) num[m] = cnt[m];
    return 0;
}

int ps3d_gpu_get(ps3d_handle *h, int which, double *host) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !host) return fail(PS3D_E_ARG, "null argument");
    const size_t hv = (size_t)h->g.nx * h->g.ny * (h->g.nz + 1);
    double *const *src = nullptr;
    switch (which) {
        case PS3D_SVOR: src = h->svor; break;
        case PS3D_VOR: src = h->vor; break;
        case PS3D_VEL: src = h->vel; break;
        case PS3D_SVEL: src = h->svel; break;
        case PS3D_SVORTS: src = h->mid; break;
        case PS3D_PRES:
        case PS3D_DELTA:
            if (!h->have_derived) return fail(PS3D_E_STATE, "pressure / divergence need ps3d_gpu_derived first");
            src = which == PS3D_PRES ? &h->pres : &h->delta;
            break;
        default: return fail(PS3D_E_UNSUPPORTED, "field not available");
    }
    const int nc = (which == PS3D_PRES || which == PS3D_DELTA) ? 1 : 3;
    int st;
    for (int c = 0; c < nc; ++c)
        if ((st = copy_out(h, host + c * hv, src[c]))) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int ps3d_gpu_set(ps3d_handle *h, int which, const double *host) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !host) return fail(PS3D_E_ARG, "null argument");
    const size_t hv = (size_t)h->g.nx * h->g.ny * (h->g.nz + 1);
    double *const *dst = nullptr;
    switch (which) {
        case PS3D_SVOR: dst = h->svor; h->have_state = true; break;
        case PS3D_SVORTS: dst = h->mid; break;
        default: return fail(PS3D_E_UNSUPPORTED, "field cannot be set");
    }
    int st;
    for (int c = 0; c < 3; ++c)
        if ((st = copy_in(h, dst[c], host + c * hv))) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int ps3d_gpu_op(ps3d_handle *h, int op, const double *in, double *out) {
    if (h && h->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: use the _batch entry points");
    if (!h || !in || !out) return fail(PS3D_E_ARG, "null argument");
    double *A = h->grad[0], *B = h->grad[1];
    h->have_grad = false;
    int st;
    if ((st = copy_in(h, A, in))) return st;
    double *a1[1] = {A};
    double *b1[1] = {B};
    double *res = B;
    switch (op) {
        case PS3D_OP_FFTXYP2S:
            if ((st = line_fields(h, 1, LM_FWD, 1, a1, b1)) || (st = line_fields(h, 0, LM_FWD, 1, b1, b1))) return st;
            break;
        case PS3D_OP_FFTXYS2P:
            if ((st = line_fields(h, 0, LM_INV, 1, a1, b1)) || (st = line_fields(h, 1, LM_INV, 1, b1, b1))) return st;
            break;
        case PS3D_OP_DIFFX:
            if ((st = line_fields(h, 0, LM_DIFF, 1, a1, b1))) return st;
            break;
        case PS3D_OP_DIFFY:
            if ((st = line_fields(h, 1, LM_DIFF, 1, a1, b1))) return st;
            break;
        case PS3D_OP_FFTSINE: HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DST, A, B), h->stream)); break;
        case PS3D_OP_FFTCOSINE: HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DCT, A, B), h->stream)); break;
        case PS3D_OP_CENTRAL_DIFFZ: HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DIFFZ, A, B), h->stream)); break;
        case PS3D_OP_COMBINE_SEMI: HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_COMBINE, A, B), h->stream)); break;
        case PS3D_OP_DECOMPOSE_SEMI: HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DECOMPOSE, A, B), h->stream)); break;
        case PS3D_OP_COMBINE_PHYS:
            HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_COMBINE, A, B), h->stream));
            if ((st = line_fields(h, 0, LM_INV, 1, b1, b1)) || (st = line_fields(h, 1, LM_INV, 1, b1, b1))) return st;
            break;
        case PS3D_OP_DECOMPOSE_PHYS:
            if ((st = line_fields(h, 1, LM_FWD, 1, a1, a1)) || (st = line_fields(h, 0, LM_FWD, 1, a1, a1))) return st;
            HIPCHK(launch_zop(h->zl, zop_args(h, ZOP_DECOMPOSE, A, B), h->stream));
            break;
        default:
            return fail(PS3D_E_ARG, "unknown op");
    }
    if ((st = copy_out(h, out, res))) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// Batched stafft transforms on x(m, n) (m fastest) through the line / z kernels.
int ps3d_gpu_fft1d(int kind, int m, int n, double *x) {
    if (!x || m < 1 || n < 2 || (n & 1) || !factor235(n)) return fail(PS3D_E_FACTOR, "fft1d: n must be even and 2^a 3^b 5^c");
    if (kind < 0 || kind > 3) return fail(PS3D_E_ARG, "fft1d: unknown kind");
    ps3d_params p;
    ps3d_params_default(&p);
    ps3d_handle *h = new ps3d_handle();
    int st;
    if (kind <= 1) {
        // lines of length n along "y", m lines along "z": x(m, n) is the column
        // layout with one x slot, ny = n and nz + 1 = m
        p.nx = 2; p.ny = n; p.nz = 2;
        st = init_handle(h, &p, false, false);
        if (!st) {
            h->g.nz = m - 1;
            h->g.zs = ((m + kZAlign - 1) / kZAlign) * kZAlign;
            h->ly = line_cfg(n, m - 1, h->g.zs, 1, (long)h->g.zs, (long)n * h->g.zs);
            double *buf = nullptr;
            if (hipMalloc((void **)&buf, sizeof(double) * (size_t)n * h->g.zs) != hipSuccess) st = fail(PS3D_E_NOMEM, "fft1d alloc");
            if (!st) {
                const size_t w = sizeof(double) * m;
                if (hipMemcpy2D(buf, sizeof(double) * h->g.zs, x, w, w, n, hipMemcpyHostToDevice) != hipSuccess)
                    st = fail(PS3D_E_HIP, "fft1d copy in");
                double *a1[1] = {buf};
                if (!st) st = line_fields(h, 1, kind == 0 ? LM_FWD : LM_INV, 1, a1, a1);
                if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = fail(PS3D_E_HIP, "fft1d sync");
                if (!st && hipMemcpy2D(x, w, buf, sizeof(double) * h->g.zs, w, n, hipMemcpyDeviceToHost) != hipSuccess)
                    st = fail(PS3D_E_HIP, "fft1d copy out");
                (void)hipFree(buf);
            }
        }
    } else {
        // columns of the z kernels: nx = 2, ny = ceil(m/2) (even), nz = n; column c
        // of x is column c of the device volume
        const int ncol = m;
        p.nx = 2;
        p.ny = 2;
        while (2 * p.ny < ncol) p.ny *= 2;   // a power of two keeps the unused y plan valid
        p.nz = n;
        if (n > 1024) { delete h; return fail(PS3D_E_UNSUPPORTED, "fft1d: dct/dst length <= 1024"); }
        st = init_handle(h, &p, false, false);
        if (!st) {
            const size_t nv = (size_t)p.nx * p.ny * h->g.zs;
            double *buf = nullptr;
            if (hipMalloc((void **)&buf, sizeof(double) * 2 * nv) != hipSuccess) st = fail(PS3D_E_NOMEM, "fft1d alloc");
            if (!st && hipMemset(buf, 0, sizeof(double) * 2 * nv) != hipSuccess) st = fail(PS3D_E_HIP, "fft1d clear");
            if (!st) {
                std::vector<double> hb((size_t)p.nx * p.ny * h->g.zs, 0.0);
                for (int c = 0; c < ncol; ++c)
                    for (int j = 0; j <= n; ++j) {
                        double v = 0.0;
                        if (kind == 2) v = x[c + (size_t)m * j];
                        else if (j >= 1) v = x[c + (size_t)m * (j - 1)];
                        hb[(size_t)c * h->g.zs + j] = v;
                    }
                if (hipMemcpy(buf, hb.data(), sizeof(double) * nv, hipMemcpyHostToDevice) != hipSuccess)
                    st = fail(PS3D_E_HIP, "fft1d copy in");
                ZOpArgs za = zop_args(h, kind == 2 ? ZOP_DCT : ZOP_DST, buf, buf + nv);
                if (!st && launch_zop(h->zl, za, h->stream) != hipSuccess) st = fail(PS3D_E_HIP, "fft1d z launch");
                if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = fail(PS3D_E_HIP, "fft1d sync");
                if (!st && hipMemcpy(hb.data(), buf + nv, sizeof(double) * nv, hipMemcpyDeviceToHost) != hipSuccess)
                    st = fail(PS3D_E_HIP, "fft1d copy out");
                if (!st) {
                    for (int c = 0; c < ncol; ++c)
                        for (int j = 0; j <= n; ++j) {
                            if (kind == 2) x[c + (size_t)m * j] = hb[(size_t)c * h->g.zs + j];
                            else if (j >= 1) x[c + (size_t)m * (j - 1)] = hb[(size_t)c * h->g.zs + j];
                        }
                }
            }
            (void)hipFree(buf);   // (a null pointer is a no-op)
        }
    }
    std::string msg = g_err;
    free_handle(h);
    g_err = msg;
    return st;
}


int ps3d_gpu_resample(ps3d_handle *dst, ps3d_handle *src, int which) {
    if (!dst || !src) return fail(PS3D_E_ARG, "null handle");
    if (dst->nbatch != 1 || src->nbatch != 1) return fail(PS3D_E_STATE, "batched handle: resample one instance at a time");
    if (which != PS3D_SVOR && which != PS3D_SVORTS) return fail(PS3D_E_UNSUPPORTED, "resample: svor or svorts only");
    if (which == PS3D_SVOR && !src->have_state) return fail(PS3D_E_STATE, "resample of svor before set_vorticity");
    drop_step_graph(dst);
    for (int d = 0; d < 3; ++d)
        if (dst->p.extent[d] != src->p.extent[d] || dst->p.lower[d] != src->p.lower[d])
            return fail(PS3D_E_ARG, "resample needs the same box");
    HIPCHK(hipStreamSynchronize(src->stream));
    ResampleArgs a;
    std::memset(&a, 0, sizeof a);
    double *const *s = which == PS3D_SVOR ? src->svor : src->mid;
    double *const *d = which == PS3D_SVOR ? dst->svor : dst->mid;
    for (int c = 0; c < 3; ++c) {
        a.src[c] = s[c];
        a.dst[c] = d[c];
    }
    a.snx = src->g.nx; a.sny = src->g.ny; a.snz = src->g.nz; a.szs = src->g.zs;
    a.dnx = dst->g.nx; a.dny = dst->g.ny; a.dnz = dst->g.nz; a.dzs = dst->g.zs;
    // stafft coefficients scale as sqrt(n) per axis (1/sqrt(n) forward transform);
    // the orthonormal DST-I interior as sqrt(nz)
    a.hfac = std::sqrt(((double)a.dnx * a.dny) / ((double)a.snx * a.sny));
    a.zfac = std::sqrt((double)a.dnz / (double)a.snz);
    HIPCHK(launch_resample(a, dst->stream));
    if (which == PS3D_SVORTS) {
        HIPCHK(hipStreamSynchronize(dst->stream));
        return 0;
    }
    dst->have_state = true;
    return finish_setup(dst);
}

// ---------------------------------------------------------------------------
// Batched instances of one grid, and ensembles of several grids
// ---------------------------------------------------------------------------
int ps3d_gpu_create_batch(const ps3d_params *p, int nbatch, ps3d_handle **out) {
    if (!p || !out || nbatch < 1) return fail(PS3D_E_ARG, "invalid argument");
    *out = nullptr;
    ps3d_handle *h = new ps3d_handle();
    int st = init_handle(h, p, true, true, nbatch);
    if (st) {
        std::string msg = g_err;
        free_handle(h);
        g_err = msg;
        return st;
    }
    *out = h;
    return 0;
}

int ps3d_gpu_batch_size(ps3d_handle *h, int *nbatch) {
    if (!h || !nbatch) return fail(PS3D_E_ARG, "null argument");
    *nbatch = h->nbatch;
    return 0;
}

int ps3d_gpu_set_vorticity_batch(ps3d_handle *h, int b, const double *vor) {
    if (!h || !vor || b < 0 || b >= h->nbatch) return fail(PS3D_E_ARG, "invalid argument");
    ps3d_handle v = make_view(h, b);
    v.have_state = false;
    int st = ps3d_gpu_set_vorticity(&v, vor);
    if (st) return st;
    h->inst_set[b] = 1;
    h->have_state = std::all_of(h->inst_set.begin(), h->inst_set.end(), [](char c) { return c != 0; });
    h->have_phys = h->have_adapt = h->have_grad = false;
    return 0;
}

int ps3d_gpu_get_batch(ps3d_handle *h, int b, int which, double *host) {
    if (!h || !host || b < 0 || b >= h->nbatch) return fail(PS3D_E_ARG, "invalid argument");
    if (which != PS3D_SVOR) return fail(PS3D_E_UNSUPPORTED, "batched handles hold svor between steps");
    ps3d_handle v = make_view(h, b);
    return ps3d_gpu_get(&v, which, host);
}

// Enqueues nsteps fused steps of every instance (no host synchronisation).
static int enqueue_steps(ps3d_handle *h, int nsteps) {
    int st;
    for (int s = 0; s < nsteps; ++s)
        if ((st = one_step(h))) return st;
    return 0;
}

static int put_times(ps3d_handle *h, const double *t) {
    std::vector<double> sc((size_t)SC_NSCAL * h->nbatch);
    HIPCHK(hipMemcpyAsync(sc.data(), h->scal, sizeof(double) * sc.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int b = 0; b < h->nbatch; ++b) {
        sc[(size_t)b * SC_NSCAL + SC_T] = t[b];
        sc[(size_t)b * SC_NSCAL + SC_STEP] = 0.0;
    }
    HIPCHK(hipMemcpyAsync(h->scal, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int get_times(ps3d_handle *h, double *t) {
    std::vector<double> sc((size_t)SC_NSCAL * h->nbatch);
    HIPCHK(hipMemcpyAsync(sc.data(), h->scal, sizeof(double) * sc.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int b = 0; b < h->nbatch; ++b) t[b] = sc[(size_t)b * SC_NSCAL + SC_T];
    return 0;
}

int ps3d_gpu_advance_n_batch(ps3d_handle *h, double *t, int nsteps, double *diag_series) {
    if (!h || !t || nsteps < 0) return fail(PS3D_E_ARG, "invalid argument");
    if (!h->have_state) return fail(PS3D_E_STATE, "advance before every instance has an initial vorticity");
    int st;
    if (diag_series && (st = ensure_diag(h, std::max(1, nsteps)))) return st;
    if ((st = put_times(h, t))) return st;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    if ((st = enqueue_steps(h, nsteps))) return st;
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->last_steps = nsteps;
    if ((st = get_times(h, t))) return st;
    if (diag_series && nsteps > 0)
        HIPCHK(hipMemcpy2D(diag_series, sizeof(double) * PS3D_NDIAG * nsteps, h->diag, sizeof(double) * h->b_diag,
                           sizeof(double) * PS3D_NDIAG * nsteps, h->nbatch, hipMemcpyDeviceToHost));
    h->have_phys = h->have_adapt = h->have_grad = false;
    return 0;
}

struct ps3d_ensemble {
    std::vector<ps3d_handle *> g;
    hipStream_t s0 = nullptr;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned> gens;   // the groups' ps3d_handle::gen when the graph was captured
    double last_ms = 0.0;
};

static void drop_graph(ps3d_ensemble *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

int ps3d_gpu_ensemble_destroy(ps3d_ensemble *e) {
    if (!e) return 0;
    if (e->s0) (void)hipStreamSynchronize(e->s0);
    drop_graph(e);
    for (ps3d_handle *h : e->g) free_handle(h);
    for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->s0) (void)hipStreamDestroy(e->s0);
    delete e;
    return 0;
}

int ps3d_gpu_ensemble_create(int ngroups, const ps3d_params *p, const int *nbatch, ps3d_ensemble **out) {
    if (ngroups < 1 || !p || !nbatch || !out) return fail(PS3D_E_ARG, "invalid argument");
    *out = nullptr;
    ps3d_ensemble *e = new ps3d_ensemble();
    auto bail = [&](int st) {
        std::string msg = g_err;
        ps3d_gpu_ensemble_destroy(e);
        g_err = msg;
        return st;
    };
    for (int i = 0; i < ngroups; ++i) {
        ps3d_handle *h = nullptr;
        int st = ps3d_gpu_create_batch(&p[i], nbatch[i], &h);
        if (st) return bail(st);
        e->g.push_back(h);
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return bail(fail(PS3D_E_HIP, "hipEventCreate failed"));
        e->join.push_back(ev);
    }
    if (hipStreamCreateWithFlags(&e->s0, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&e->ev0) != hipSuccess ||
        hipEventCreate(&e->ev1) != hipSuccess)
        return bail(fail(PS3D_E_HIP, "stream / event creation failed"));
    *out = e;
    return 0;
}

int ps3d_gpu_ensemble_group(ps3d_ensemble *e, int g, ps3d_handle **h) {
    if (!e || !h || g < 0 || g >= (int)e->g.size()) return fail(PS3D_E_ARG, "invalid argument");
    *h = e->g[g];
    return 0;
}

// One step of every group: the groups' streams fork from s0 and join back, so the
// groups run concurrently.
static int enqueue_ensemble_step(ps3d_ensemble *e) {
    HIPCHK(hipEventRecord(e->fork, e->s0));
    for (size_t i = 0; i < e->g.size(); ++i) {
        ps3d_handle *h = e->g[i];
        HIPCHK(hipStreamWaitEvent(h->stream, e->fork, 0));
        int st = one_step(h);
        if (st) return st;
        HIPCHK(hipEventRecord(e->join[i], h->stream));
        HIPCHK(hipStreamWaitEvent(e->s0, e->join[i], 0));
    }
    return 0;
}

static int build_graph(ps3d_ensemble *e) {
    drop_graph(e);
    HIPCHK(hipStreamBeginCapture(e->s0, hipStreamCaptureModeRelaxed));
    int st = enqueue_ensemble_step(e);
    hipGraph_t gr = nullptr;
    hipError_t ce = hipStreamEndCapture(e->s0, &gr);
    if (st) {
        if (gr) (void)hipGraphDestroy(gr);
        return st;
    }
    if (ce != hipSuccess) return fail(PS3D_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    e->graph = gr;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->gens.clear();
    for (ps3d_handle *h : e->g) e->gens.push_back(h->gen);
    return 0;
}

int ps3d_gpu_ensemble_advance(ps3d_ensemble *e, int nsteps, int use_graph, double *ms) {
    if (!e || nsteps < 0) return fail(PS3D_E_ARG, "invalid argument");
    for (ps3d_handle *h : e->g)
        if (!h->have_state) return fail(PS3D_E_STATE, "advance before every instance has an initial vorticity");
    int st;
    if (use_graph && e->exec) {
        // a group's captured arguments changed (its diagnostics buffer was
        // reallocated by advance_n_batch, say): capture again
        for (size_t i = 0; i < e->g.size(); ++i)
            if (e->gens[i] != e->g[i]->gen) {
                drop_graph(e);
                break;
            }
    }
    if (use_graph && !e->exec && (st = build_graph(e))) return st;
    HIPCHK(hipEventRecord(e->ev0, e->s0));
    for (int s = 0; s < nsteps; ++s) {
        if (use_graph)
            HIPCHK(hipGraphLaunch(e->exec, e->s0));
        else if ((st = enqueue_ensemble_step(e)))
            return st;
    }
    HIPCHK(hipEventRecord(e->ev1, e->s0));
    HIPCHK(hipStreamSynchronize(e->s0));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, e->ev0, e->ev1));
    e->last_ms = t;
    if (ms) *ms = t;
    for (ps3d_handle *h : e->g) h->have_phys = h->have_adapt = h->have_grad = false;
    return 0;
}

// Launch-shape candidates of one grid: the z kernels (one wave per column where
// available, else quad-per-thread with QPB quads per block).  Code = wave * 100 + QPB.
static std::vector<int> z_candidates(const ps3d_handle *h) {
    std::vector<int> c;
    if (ps3d::zwave_supported(h->g.nz)) {
        c.push_back(100);
        c.push_back(101);
        c.push_back(102);
    }
    for (int q = 1; q <= 8; q *= 2) {
        if (h->TQ * q > 512) break;
        if (zpass_lds_bytes(q, h->g.nz) > 160 * 1024) break;
        c.push_back(q);
    }
    return c;
}

static void apply_z_choice(ps3d_handle *h, int code) {
    h->zl.wave = code >= 100 ? 1 + (code - 100) : 0;
    if (!h->zl.wave) {
        h->QPB = code;
        h->zl.threads = h->TQ * h->QPB;
        h->zl.blocks = (h->nquads + h->QPB - 1) / h->QPB;
        h->zl.lds = zpass_lds_bytes(h->QPB, h->g.nz);
    }
}

static int current_z_choice(const ps3d_handle *h) { return h->zl.wave ? 99 + h->zl.wave : h->QPB; }

// Times `steps` fused steps of each group under every candidate shape and keeps the
// fastest; the state (svor, device scalars, rolling-mean history) is restored.
int ps3d_gpu_ensemble_tune(ps3d_ensemble *e, int steps, int *choice) {
    if (!e || steps < 1) return fail(PS3D_E_ARG, "invalid argument");
    int st;
    for (size_t gi = 0; gi < e->g.size(); ++gi) {
        ps3d_handle *h = e->g[gi];
        if (!h->have_state) return fail(PS3D_E_STATE, "tune needs the initial vorticity of every instance");
        const size_t nb = h->nbatch, sv = 3 * h->nvol;
        double *bk = nullptr;
        HIPCHK(hipMalloc((void **)&bk, sizeof(double) * (sv * nb + SC_NSCAL * nb + h->b_roll * nb)));
        double *bsc = bk + sv * nb, *brl = bsc + SC_NSCAL * nb;
        for (size_t b = 0; b < nb; ++b)
            HIPCHK(hipMemcpyAsync(bk + b * sv, h->svor[0] + b * h->b_vol, sizeof(double) * sv, hipMemcpyDeviceToDevice,
                                  h->stream));
        HIPCHK(hipMemcpyAsync(bsc, h->scal, sizeof(double) * SC_NSCAL * nb, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(brl, h->roll, sizeof(double) * h->b_roll * nb, hipMemcpyDeviceToDevice, h->stream));
        const std::vector<int> cand = z_candidates(h);
        int best = current_z_choice(h);
        float best_ms = 1e30f;
        for (int code : cand) {
            apply_z_choice(h, code);
            if ((st = enqueue_steps(h, 1))) break;       // warm-up
            HIPCHK(hipEventRecord(h->ev0, h->stream));
            if ((st = enqueue_steps(h, steps))) break;
            HIPCHK(hipEventRecord(h->ev1, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
            if (ms < best_ms) {
                best_ms = ms;
                best = code;
            }
        }
        apply_z_choice(h, best);
        // line passes: block size (tile depth) of each register-holding mode in turn
        if (!st) {
            int cur[3] = {h->rtm, h->rty, h->rtx};
            for (int mode = 0; mode < 3 && !st; ++mode) {
                int bestT = cur[mode];
                float bt = 1e30f;
                for (int T : {1024, 512, 256}) {
                    int c[3] = {cur[0], cur[1], cur[2]};
                    c[mode] = T;
                    apply_line_choice(h, c[0], c[1], c[2]);
                    if ((mode == 0 ? h->rtm : mode == 1 ? h->rty : h->rtx) != T) continue;   // no build
                    if ((st = enqueue_steps(h, 1))) break;
                    HIPCHK(hipEventRecord(h->ev0, h->stream));
                    if ((st = enqueue_steps(h, steps))) break;
                    HIPCHK(hipEventRecord(h->ev1, h->stream));
                    HIPCHK(hipStreamSynchronize(h->stream));
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                    if (ms < bt) {
                        bt = ms;
                        bestT = T;
                    }
                }
                cur[mode] = bestT;
            }
            apply_line_choice(h, cur[0], cur[1], cur[2]);
        }
        // FFT schedule of the transform-only passes: y lines, then x lines
        for (int axis = 1; axis >= 0 && !st; --axis) {
            const int n = axis == 0 ? h->g.nx : h->g.ny;
            int &cur = axis == 0 ? h->lsx : h->lsy;
            int bests = cur;
            float bt = 1e30f;
            for (int sc = 0; sc < 8; ++sc) {
                if (!line_sched_ok(n, sc)) continue;
                cur = sc;
                if ((st = enqueue_steps(h, 1))) break;
                HIPCHK(hipEventRecord(h->ev0, h->stream));
                if ((st = enqueue_steps(h, steps))) break;
                HIPCHK(hipEventRecord(h->ev1, h->stream));
                HIPCHK(hipStreamSynchronize(h->stream));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                if (ms < bt) {
                    bt = ms;
                    bests = sc;
                }
            }
            cur = bests;
        }
        for (size_t b = 0; b < nb; ++b)
            HIPCHK(hipMemcpyAsync(h->svor[0] + b * h->b_vol, bk + b * sv, sizeof(double) * sv, hipMemcpyDeviceToDevice,
                                  h->stream));
        HIPCHK(hipMemcpyAsync(h->scal, bsc, sizeof(double) * SC_NSCAL * nb, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->roll, brl, sizeof(double) * h->b_roll * nb, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        (void)hipFree(bk);
        if (st) return st;
        if (choice) choice[gi] = best;
    }
    drop_graph(e);
    return 0;
}

int ps3d_gpu_ensemble_fft_schedule(ps3d_ensemble *e, int g, int *sched) {
    if (!e || !sched || g < 0 || g >= (int)e->g.size()) return fail(PS3D_E_ARG, "invalid argument");
    sched[0] = e->g[g]->lsx;
    sched[1] = e->g[g]->lsy;
    return 0;
}

int ps3d_gpu_ensemble_line_choice(ps3d_ensemble *e, int g, int *threads) {
    if (!e || !threads || g < 0 || g >= (int)e->g.size()) return fail(PS3D_E_ARG, "invalid argument");
    const ps3d_handle *h = e->g[g];
    threads[0] = h->rtm;
    threads[1] = h->rty;
    threads[2] = h->rtx;
    return 0;
}

int ps3d_gpu_ensemble_set_choice(ps3d_ensemble *e, int g, int code) {
    if (!e || g < 0 || g >= (int)e->g.size()) return fail(PS3D_E_ARG, "invalid argument");
    ps3d_handle *h = e->g[g];
    const std::vector<int> cand = z_candidates(h);
    if (std::find(cand.begin(), cand.end(), code) == cand.end()) return fail(PS3D_E_ARG, "launch shape not available");
    apply_z_choice(h, code);
    drop_graph(e);
    return 0;
}


int ps3d_gpu_set_closure(ps3d_handle *h, ps3d_cnn *net, const float *in_scale, const float *out_scale) {
    if (!h) return fail(PS3D_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    drop_step_graph(h);
    if (!net) {
        h->cnn = nullptr;
        return 0;
    }
    if (!in_scale || !out_scale) return fail(PS3D_E_ARG, "null scales");
    if (h->nbatch != 1) return fail(PS3D_E_STATE, "closure on batched handles is not supported");
    if (!ps3d::cnn_matches(net, h->g.nx, h->g.ny, h->g.nz)) return fail(PS3D_E_ARG, "closure grid differs from the handle's");
    h->cnn = net;
    for (int k = 0; k < 6; ++k) h->cnn_in[k] = in_scale[k];
    for (int k = 0; k < 3; ++k) h->cnn_out[k] = out_scale[k];
    return 0;
}

}  // extern "C"
