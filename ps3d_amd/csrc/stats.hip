// stats.hip -- output-step reductions over the physical fields: the field_stats
// quantities of ps3d (src/fields/field_diagnostics_netcdf.f90:257-439) and the
// shell spectrum of genspec (src/genspec.f90:50-124).
//
// Both are two-kernel reductions with a fixed partition of the grid columns over
// the blocks and fixed reduction trees, so that a result does not depend on the
// scheduling: per-block partials first, then one block that combines them in
// block order.
#include "kernels.hpp"
#include "ps3d_gpu.h"

#include <mutex>

namespace ps3d {

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr int kStatThreads = 256;

}  // namespace

// ---------------------------------------------------------------------------
// field_stats
// ---------------------------------------------------------------------------
// Per block: a contiguous range of columns; thread t takes the points t, t+256, ...
// of that range (z fastest).  Partial slot ST_* of block b at part[b * ST_NP + slot].
__global__ void __launch_bounds__(kStatThreads) k_stats_part(StatsArgs a) {
    const long ncol = (long)a.nx * a.ny;
    const int nz1 = a.nz + 1;
    const long c0 = ncol * blockIdx.x / gridDim.x, c1 = ncol * (blockIdx.x + 1) / gridDim.x;
    double s[ST_NSUM], mn[ST_NMIN], mx[ST_NMAX];
#pragma unroll
    for (int k = 0; k < ST_NSUM; ++k) s[k] = 0.0;
#pragma unroll
    for (int k = 0; k < ST_NMIN; ++k) mn[k] = HUGE_VAL;
#pragma unroll
    for (int k = 0; k < ST_NMAX; ++k) mx[k] = -HUGE_VAL;
    for (long i = c0 * nz1 + threadIdx.x; i < c1 * nz1; i += kStatThreads) {
        const long col = i / nz1;
        const int z = (int)(i - col * nz1);
        const size_t p = (size_t)col * a.zs + z;
        const double u = a.vel[0][p], v = a.vel[1][p], w = a.vel[2][p];
        const double xi = a.vor[0][p], eta = a.vor[1][p], zeta = a.vor[2][p];
        // trapezoid weights in z: 1/2 inside, 1/4 on the lids
        // (get_kinetic_energy, get_enstrophy and their parts: field_diagnostics.f90:85-262)
        const double wt = (z == 0 || z == a.nz) ? 0.25 : 0.5;
        s[ST_S_UV - ST_S0] += wt * (u * u + v * v);
        s[ST_S_W - ST_S0] += wt * (w * w);
        s[ST_S_XE - ST_S0] += wt * (xi * xi + eta * eta);
        s[ST_S_Z - ST_S0] += wt * (zeta * zeta);
        mn[ST_MIN_X - ST_MIN0] = fmin(mn[ST_MIN_X - ST_MIN0], xi);
        mn[ST_MIN_Y - ST_MIN0] = fmin(mn[ST_MIN_Y - ST_MIN0], eta);
        mn[ST_MIN_Z - ST_MIN0] = fmin(mn[ST_MIN_Z - ST_MIN0], zeta);
        mx[ST_MAX_X - ST_MAX0] = fmax(mx[ST_MAX_X - ST_MAX0], xi);
        mx[ST_MAX_Y - ST_MAX0] = fmax(mx[ST_MAX_Y - ST_MAX0], eta);
        mx[ST_MAX_Z - ST_MAX0] = fmax(mx[ST_MAX_Z - ST_MAX0], zeta);
        mx[ST_MAX_HE2 - ST_MAX0] = fmax(mx[ST_MAX_HE2 - ST_MAX0], xi * xi + eta * eta);
        if (z == a.nz) {
            // upper surface (field_diagnostics_netcdf.f90:300-306, 396-406)
            s[ST_S_ZTOP - ST_S0] += zeta * zeta;
            const double d = a.delta ? a.delta[p] : 0.0;
            s[ST_S_DTOP - ST_S0] += d * d;
            mx[ST_MAX_UOX - ST_MAX0] = fmax(mx[ST_MAX_UOX - ST_MAX0], xi);
            mx[ST_MAX_UOY - ST_MAX0] = fmax(mx[ST_MAX_UOY - ST_MAX0], eta);
            mx[ST_MAX_UOZ - ST_MAX0] = fmax(mx[ST_MAX_UOZ - ST_MAX0], zeta);
            mx[ST_MAX_USUH2 - ST_MAX0] = fmax(mx[ST_MAX_USUH2 - ST_MAX0], u * u + v * v);
        }
        if (z == 0) {
            mx[ST_MAX_LOX - ST_MAX0] = fmax(mx[ST_MAX_LOX - ST_MAX0], xi);
            mx[ST_MAX_LOY - ST_MAX0] = fmax(mx[ST_MAX_LOY - ST_MAX0], eta);
            mx[ST_MAX_LOZ - ST_MAX0] = fmax(mx[ST_MAX_LOZ - ST_MAX0], zeta);
        }
    }
    __shared__ double red[kStatThreads / 64][ST_NP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < ST_NSUM; ++k) {
        const double r = wave_sum(s[k]);
        if (lane == 0) red[wv][ST_S0 + k] = r;
    }
#pragma unroll
    for (int k = 0; k < ST_NMIN; ++k) {
        const double r = wave_min(mn[k]);
        if (lane == 0) red[wv][ST_MIN0 + k] = r;
    }
#pragma unroll
    for (int k = 0; k < ST_NMAX; ++k) {
        const double r = wave_max(mx[k]);
        if (lane == 0) red[wv][ST_MAX0 + k] = r;
    }
    __syncthreads();
    if (threadIdx.x < ST_NP) {
        const int k = threadIdx.x;
        double r = red[0][k];
        for (int j = 1; j < kStatThreads / 64; ++j) {
            const double o = red[j][k];
            r = k < ST_MIN0 ? r + o : (k < ST_MAX0 ? fmin(r, o) : fmax(r, o));
        }
        a.part[(size_t)blockIdx.x * ST_NP + k] = r;
    }
}

// One block: partials combined in block order, then the record in ps3d's order
// (NC_KE .. NC_ROMAX, field_diagnostics_netcdf.f90:36-75).
__global__ void __launch_bounds__(64) k_stats_fin(StatsFin f) {
    __shared__ double tot[ST_NP];
    const int k = threadIdx.x;
    if (k < ST_NP) {
        double r = f.part[k];
        for (int b = 1; b < f.nb; ++b) {
            const double o = f.part[(size_t)b * ST_NP + k];
            r = k < ST_MIN0 ? r + o : (k < ST_MAX0 ? fmin(r, o) : fmax(r, o));
        }
        tot[k] = r;
    }
    __syncthreads();
    if (k != 0) return;
    double *o = f.out;
    const double kexy = tot[ST_S_UV] * f.ncelli, kez = tot[ST_S_W] * f.ncelli;
    const double enxy = tot[ST_S_XE] * f.ncelli, enz = tot[ST_S_Z] * f.ncelli;
    o[PS3D_ST_KE] = (tot[ST_S_UV] + tot[ST_S_W]) * f.ncelli;
    o[PS3D_ST_EN] = (tot[ST_S_XE] + tot[ST_S_Z]) * f.ncelli;
    o[PS3D_ST_OMAX] = f.ad[PS3D_D_VORTMAX];
    o[PS3D_ST_ORMS] = f.ad[PS3D_D_VORTRMS];
    o[PS3D_ST_OCHAR] = f.ad[PS3D_D_VORCH];
    o[PS3D_ST_OXMEAN] = f.ad[PS3D_D_VORMEAN_X];
    o[PS3D_ST_OYMEAN] = f.ad[PS3D_D_VORMEAN_Y];
    o[PS3D_ST_OZMEAN] = f.ad[PS3D_D_VORMEAN_Z];
    o[PS3D_ST_KEXY] = kexy;
    o[PS3D_ST_KEZ] = kez;
    o[PS3D_ST_ENXY] = enxy;
    o[PS3D_ST_ENZ] = enz;
    o[PS3D_ST_OXMIN] = tot[ST_MIN_X];
    o[PS3D_ST_OYMIN] = tot[ST_MIN_Y];
    o[PS3D_ST_OZMIN] = tot[ST_MIN_Z];
    o[PS3D_ST_OXMAX] = tot[ST_MAX_X];
    o[PS3D_ST_OYMAX] = tot[ST_MAX_Y];
    o[PS3D_ST_OZMAX] = tot[ST_MAX_Z];
    o[PS3D_ST_HEMAX] = sqrt(tot[ST_MAX_HE2]);
    o[PS3D_ST_GMAX] = f.ad[PS3D_D_GGMAX];
    o[PS3D_ST_BFMAX] = 0.0;                 // no buoyancy
    o[PS3D_ST_UMAX] = f.ad[PS3D_D_UMAX];
    o[PS3D_ST_VMAX] = f.ad[PS3D_D_VMAX];
    o[PS3D_ST_WMAX] = f.ad[PS3D_D_WMAX];
    o[PS3D_ST_USOXMAX] = tot[ST_MAX_UOX];
    o[PS3D_ST_LSOXMAX] = tot[ST_MAX_LOX];
    o[PS3D_ST_USOYMAX] = tot[ST_MAX_UOY];
    o[PS3D_ST_LSOYMAX] = tot[ST_MAX_LOY];
    o[PS3D_ST_USOZMAX] = tot[ST_MAX_UOZ];
    o[PS3D_ST_LSOZMAX] = tot[ST_MAX_LOZ];
    o[PS3D_ST_USUHMAX] = sqrt(tot[ST_MAX_USUH2]);
    o[PS3D_ST_USGMAX] = f.ad[PS3D_D_USGGMAX];
    o[PS3D_ST_LSGMAX] = f.ad[PS3D_D_LSGGMAX];
    o[PS3D_ST_USZRMS] = sqrt(tot[ST_S_ZTOP] * f.nxnyi);
    o[PS3D_ST_USDELRMS] = sqrt(tot[ST_S_DTOP] * f.nxnyi);
    o[PS3D_ST_RGMAX] = f.ad[PS3D_D_RGMAX];
    o[PS3D_ST_RBFMAX] = 0.0;                // no buoyancy
    o[PS3D_ST_RIMIN] = 0.0;                 // no buoyancy
    // Ro = zeta / f_cor(3) (get_min/max_rossby_number, field_diagnostics.f90:293-320)
    o[PS3D_ST_ROMIN] = tot[ST_MIN_Z] / f.fcor3;
    o[PS3D_ST_ROMAX] = tot[ST_MAX_Z] / f.fcor3;
}

int stats_blocks(int nx, int ny) {
    const long ncol = (long)nx * ny;
    return (int)std::min<long>(kStatsMaxBlocks, std::max<long>(1, ncol / 16));
}

hipError_t launch_stats(const StatsArgs &a, const StatsFin &f, hipStream_t stream) {
    hipLaunchKernelGGL(k_stats_part, dim3(a.nb), dim3(kStatThreads), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stats_fin, dim3(1), dim3(64), 0, stream, f);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free partials: one array per device, found once per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kStatsMaxBlocks * ST_NP];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on device dev (the array's host-side name is not
// a device address; each device holds its own copy).  Looked up once per device,
// with that device current, and kept.
hipError_t stats_scratch(int dev, double **ptr) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace ps3d
