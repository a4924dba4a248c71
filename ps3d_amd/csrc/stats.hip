// stats.hip -- output-step reductions over the physical fields: the field_stats
// quantities of ps3d (src/fields/field_diagnostics_netcdf.f90:257-439), the means
// of the kinetic energy, helicity and cross-helicity, and the shell binning of
// genspec's spectrum (src/genspec.f90:62-99).
//
// All are two-kernel reductions with a fixed partition of the grid columns over
// the blocks and fixed reduction trees, so that a result does not depend on the
// scheduling: per-block partials first, then a pass that combines them in block
// order.
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
// Kinetic energy, helicity and cross-helicity
// ---------------------------------------------------------------------------
// Same column partition as k_stats_part.  Kinetic energy with get_kinetic_energy's
// weights (1/2 inside, 1/4 on the lids, field_diagnostics.f90:85-122); helicity
// u.w and the cross-helicity magnitude |u x w| (plotting/nc_reader.py:94-120) as
// means with the lids weighted 1/2, the helicity of the adapt record.
__global__ void __launch_bounds__(kStatThreads) k_moments_part(MomArgs a) {
    const long ncol = (long)a.nx * a.ny;
    const int nz1 = a.nz + 1;
    const long c0 = ncol * blockIdx.x / gridDim.x, c1 = ncol * (blockIdx.x + 1) / gridDim.x;
    const bool vor = a.vor[0] != nullptr;
    double ske = 0.0, shel = 0.0, sxh = 0.0, mxh = 0.0;
    for (long i = c0 * nz1 + threadIdx.x; i < c1 * nz1; i += kStatThreads) {
        const long col = i / nz1;
        const int z = (int)(i - col * nz1);
        const size_t p = (size_t)col * a.zs + z;
        const double u = a.vel[0][p], v = a.vel[1][p], w = a.vel[2][p];
        const bool lid = z == 0 || z == a.nz;
        ske += (lid ? 0.25 : 0.5) * (u * u + v * v + w * w);
        if (vor) {
            const double xi = a.vor[0][p], eta = a.vor[1][p], zeta = a.vor[2][p];
            const double wt = lid ? 0.5 : 1.0;
            shel += wt * (u * xi + v * eta + w * zeta);
            const double cx = v * zeta - w * eta, cy = w * xi - u * zeta, cz = u * eta - v * xi;
            const double c2 = cx * cx + cy * cy + cz * cz;
            sxh += wt * sqrt(c2);
            mxh = fmax(mxh, c2);
        }
    }
    __shared__ double red[kStatThreads / 64][MO_NP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    ske = wave_sum(ske);
    shel = wave_sum(shel);
    sxh = wave_sum(sxh);
    mxh = wave_max(mxh);
    if (lane == 0) {
        red[wv][MO_S_KE] = ske;
        red[wv][MO_S_HEL] = shel;
        red[wv][MO_S_XH] = sxh;
        red[wv][MO_MAX_XH2] = mxh;
    }
    __syncthreads();
    if (threadIdx.x < MO_NP) {
        const int k = threadIdx.x;
        double r = red[0][k];
        for (int j = 1; j < kStatThreads / 64; ++j) r = k == MO_MAX_XH2 ? fmax(r, red[j][k]) : r + red[j][k];
        a.part[(size_t)blockIdx.x * MO_NP + k] = r;
    }
}

__global__ void __launch_bounds__(64) k_moments_fin(MomArgs a) {
    __shared__ double tot[MO_NP];
    const int k = threadIdx.x;
    if (k < MO_NP) {
        double r = a.part[k];
        for (int b = 1; b < a.nb; ++b) {
            const double o = a.part[(size_t)b * MO_NP + k];
            r = k == MO_MAX_XH2 ? fmax(r, o) : r + o;
        }
        tot[k] = r;
    }
    __syncthreads();
    if (k != 0) return;
    a.out[PS3D_MO_KE] = tot[MO_S_KE] * a.ncelli;
    a.out[PS3D_MO_HELICITY] = tot[MO_S_HEL] * a.ncelli;
    a.out[PS3D_MO_CROSS_HELICITY] = tot[MO_S_XH] * a.ncelli;
    a.out[PS3D_MO_CROSS_HELICITY_MAX] = sqrt(tot[MO_MAX_XH2]);
}

hipError_t launch_moments(const MomArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(k_moments_part, dim3(a.nb), dim3(kStatThreads), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_moments_fin, dim3(1), dim3(64), 0, stream, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// genspec's shell spectrum (src/genspec.f90:62-99)
// ---------------------------------------------------------------------------
// Per block: a contiguous range of (x, y) columns; wave w takes every kSpecWaves-th
// column of it, 64 z levels at a time (lane = kz).  Along a column the shell index
// m(kz) = int(nint(sqrt(rkx^2 + rky^2 + rkz^2)) * dki) never decreases, so the
// lanes of one shell are contiguous: a segmented scan over the wave sums each run
// and its last lane adds the run to the wave's own LDS histogram.  The waves'
// histograms are added in wave order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(64 * kSpecWaves) k_spec_part(SpecArgs a) {
    extern __shared__ double spec_sm[];
    const int nb = a.nbins;
    double *hist = spec_sm;                                   // [kSpecWaves][nbins]
    int *cnt = reinterpret_cast<int *>(spec_sm + (size_t)kSpecWaves * nb);
    for (int i = threadIdx.x; i < kSpecWaves * nb; i += blockDim.x) {
        hist[i] = 0.0;
        cnt[i] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double *hw = hist + (size_t)wv * nb;
    int *cw = cnt + (size_t)wv * nb;
    const long ncol = (long)a.nx * a.ny;
    const long c0 = ncol * blockIdx.x / gridDim.x, c1 = ncol * (blockIdx.x + 1) / gridDim.x;
    for (long col = c0 + wv; col < c1; col += kSpecWaves) {
        const int ix = (int)(col / a.ny), iy = (int)(col - (long)ix * a.ny);
        // rkx^2 + rky^2 + rkz^2 summed left to right, no contraction (genspec.f90:66)
        const double kh2 = __dadd_rn(__dmul_rn(a.rkx[ix], a.rkx[ix]), __dmul_rn(a.rky[iy], a.rky[iy]));
        const double *fc = a.f + (size_t)col * a.zs;
        for (int base = 0; base <= a.nz; base += 64) {
            const int kz = base + lane;
            const bool ok = kz <= a.nz;
            double e = 0.0;
            int m = 0x7fffffff, c = 0;
            if (ok) {
                const double v = fc[kz];
                e = __dmul_rn(v, v);
                const double rz = __dmul_rn(a.rkz1, (double)kz);
                const double kmag = round(sqrt(__dadd_rn(kh2, __dmul_rn(rz, rz))));   // nint
                m = (int)__dmul_rn(kmag, a.dki);
                c = 1;
            }
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const double eu = __shfl_up(e, off, 64);
                const int mu = __shfl_up(m, off, 64), cu = __shfl_up(c, off, 64);
                if (lane >= off && mu == m) {
                    e += eu;
                    c += cu;
                }
            }
            const int mn = __shfl_down(m, 1, 64);
            if (ok && (lane == 63 || mn != m) && m < nb) {
                hw[m] += e;
                cw[m] += c;
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double s = hist[b];
        int n = cnt[b];
#pragma unroll
        for (int j = 1; j < kSpecWaves; ++j) {
            s += hist[(size_t)j * nb + b];
            n += cnt[(size_t)j * nb + b];
        }
        const size_t o = (size_t)blockIdx.x * nb + b;
        a.part[o] = a.accumulate ? a.part[o] + s : s;
        if (a.cpart) a.cpart[o] = n;
    }
}

__global__ void __launch_bounds__(256) k_spec_fin(const double *part, const int *cpart, int nblk, int nbins,
                                                  double *spec, long long *cnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    double s = 0.0;
    long long n = 0;
    for (int j = 0; j < nblk; ++j) {
        s += part[(size_t)j * nbins + b];
        n += cpart[(size_t)j * nbins + b];
    }
    spec[b] = s;
    cnt[b] = n;
}

size_t spec_lds_bytes(int nbins) { return (size_t)kSpecWaves * nbins * (sizeof(double) + sizeof(int)); }

hipError_t launch_spectrum(const SpecArgs &a, hipStream_t stream) {
    const size_t lds = spec_lds_bytes(a.nbins);
    // set on every call: an output-step launch, and the attribute is per device
    hipError_t e = hipFuncSetAttribute((const void *)k_spec_part, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_spec_part, dim3(a.nb), dim3(64 * kSpecWaves), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_spectrum_fin(const double *part, const int *cpart, int nb, int nbins, double *spec,
                               long long *cnt, hipStream_t stream) {
    hipLaunchKernelGGL(k_spec_fin, dim3((nbins + 255) / 256), dim3(256), 0, stream, part, cpart, nb, nbins, spec, cnt);
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
