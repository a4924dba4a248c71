// common.hpp -- shared types, FFT plans and the LDS Stockham FFT engine of the
// MI355X pseudo-spectral step.
//
// Internal volume layout ("column layout"): index (ix, iy, iz) -> (ix*ny + iy)*zs + iz
// with zs = nz+1 rounded up to a multiple of 16 so that every z-column starts on a
// 128-byte boundary.  Spectral and physical fields share it.  The x/y line passes
// read and write 16-deep z runs (128 B) per line element.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#ifndef PS3D_DEVICE_ONLY
#include <string>
#endif

namespace ps3d {

constexpr int kMaxPass = 16;
constexpr int kZAlign = 16;

// Mixed-radix plan: n = prod R[s], radices from {8, 4, 2, 3, 5}.
struct FftPlan {
    int n;
    int npass;
    int R[kMaxPass];
};

#ifndef PS3D_DEVICE_ONLY
// Returns false when n has a prime factor other than 2, 3 and 5 (the primes of
// stafft's factorisen, src/fft/stafft.f90:128-187).
inline bool make_plan(int n, FftPlan &P) {
    P.n = n;
    P.npass = 0;
    int rem = n;
    if (n < 2) return false;
    while (rem % 8 == 0 && P.npass < kMaxPass) { P.R[P.npass++] = 8; rem /= 8; }
    while (rem % 4 == 0 && P.npass < kMaxPass) { P.R[P.npass++] = 4; rem /= 4; }
    while (rem % 2 == 0 && P.npass < kMaxPass) { P.R[P.npass++] = 2; rem /= 2; }
    while (rem % 3 == 0 && P.npass < kMaxPass) { P.R[P.npass++] = 3; rem /= 3; }
    while (rem % 5 == 0 && P.npass < kMaxPass) { P.R[P.npass++] = 5; rem /= 5; }
    return rem == 1;
}
#endif

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (dir < 0, forward) or +i (dir > 0, inverse)
__host__ __device__ inline double2 cmuli(double2 a, int dir) {
    return dir < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// In-register DFT of radix R (2, 3, 4, 5, 8): u[s] <- sum_t u[t] exp(dir 2 pi i s t / R)
__host__ __device__ inline void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3, int dir) {
    double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmuli(csub(a1, a3), dir);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

__host__ __device__ inline void butterfly(int R, double2 *u, int dir) {
    const double r12 = 0.70710678118654752440084436210484903928;   // sqrt(1/2)
    const double s3 = 0.86602540378443864676372317075293618347;    // sin(pi/3)
    switch (R) {
        case 2: {
            double2 a = u[0], b = u[1];
            u[0] = cadd(a, b);
            u[1] = csub(a, b);
            break;
        }
        case 3: {
            double2 t1 = cadd(u[1], u[2]);
            double2 t2 = make_double2(u[0].x - 0.5 * t1.x, u[0].y - 0.5 * t1.y);
            double2 d = csub(u[1], u[2]);
            double2 t3 = cmuli(make_double2(s3 * d.x, s3 * d.y), dir);
            u[0] = cadd(u[0], t1);
            u[1] = cadd(t2, t3);
            u[2] = csub(t2, t3);
            break;
        }
        case 4:
            dft4(u[0], u[1], u[2], u[3], dir);
            break;
        case 5: {
            // pairs t, 5-t share cos(2 pi s t/5); their differences carry the sines
            const double c1 = 0.30901699437494742410229341718281906;    // cos(2 pi/5)
            const double c2 = -0.80901699437494742410229341718281906;   // cos(4 pi/5)
            const double s1 = 0.95105651629515357211643933337938214;    // sin(2 pi/5)
            const double s2 = 0.58778525229247312916870595463907277;    // sin(4 pi/5)
            const double2 t1 = cadd(u[1], u[4]), t2 = cadd(u[2], u[3]);
            const double2 t3 = csub(u[1], u[4]), t4 = csub(u[2], u[3]);
            const double2 a1 = make_double2(u[0].x + c1 * t1.x + c2 * t2.x, u[0].y + c1 * t1.y + c2 * t2.y);
            const double2 a2 = make_double2(u[0].x + c2 * t1.x + c1 * t2.x, u[0].y + c2 * t1.y + c1 * t2.y);
            const double2 b1 = cmuli(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), dir);
            const double2 b2 = cmuli(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), dir);
            u[0] = cadd(u[0], cadd(t1, t2));
            u[1] = cadd(a1, b1);
            u[4] = csub(a1, b1);
            u[2] = cadd(a2, b2);
            u[3] = csub(a2, b2);
            break;
        }
        case 8: {
            double2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
            double2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
            dft4(e0, e1, e2, e3, dir);
            dft4(o0, o1, o2, o3, dir);
            // w8^s = exp(dir i pi s / 4)
            double2 w1 = make_double2(r12, dir < 0 ? -r12 : r12);
            double2 w3 = make_double2(-r12, dir < 0 ? -r12 : r12);
            o1 = cmul(o1, w1);
            o2 = cmuli(o2, dir);
            o3 = cmul(o3, w3);
            u[0] = cadd(e0, o0); u[4] = csub(e0, o0);
            u[1] = cadd(e1, o1); u[5] = csub(e1, o1);
            u[2] = cadd(e2, o2); u[6] = csub(e2, o2);
            u[3] = cadd(e3, o3); u[7] = csub(e3, o3);
            break;
        }
        default:
            break;
    }
}

// One Stockham pass of radix R for every butterfly index e = e0, e0+de, ...
// src/dst lines are `ls` complex apart; tw[m] = exp(-2 pi i m / n).
template <int R>
__host__ __device__ inline void fft_pass_r(const double2 *src, double2 *dst, int nlines, int ls, int n, int p,
                                           const double2 *tw, int dir, int e0, int de) {
    const int m = n / R, tstep = n / (p * R);
    const int tot = nlines * m;
    for (int e = e0; e < tot; e += de) {
        const int line = e / m;
        const int i = e - line * m;
        const int k = i % p;
        const double2 *s = src + (size_t)line * ls;
        double2 *d = dst + (size_t)line * ls;
        double2 u[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            double2 v = s[i + t * m];
            if (t > 0 && k > 0) {
                const double2 w = tw[k * t * tstep];
                v = dir < 0 ? cmul(v, w) : cmulc(v, w);
            }
            u[t] = v;
        }
        butterfly(R, u, dir);
        const int j = (i - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) d[j + t * p] = u[t];
    }
}

__host__ __device__ inline void fft_pass(const double2 *src, double2 *dst, int nlines, int ls, int n, int R, int p,
                                         const double2 *tw, int dir, int e0, int de) {
    switch (R) {
        case 2: fft_pass_r<2>(src, dst, nlines, ls, n, p, tw, dir, e0, de); break;
        case 3: fft_pass_r<3>(src, dst, nlines, ls, n, p, tw, dir, e0, de); break;
        case 4: fft_pass_r<4>(src, dst, nlines, ls, n, p, tw, dir, e0, de); break;
        case 5: fft_pass_r<5>(src, dst, nlines, ls, n, p, tw, dir, e0, de); break;
        default: fft_pass_r<8>(src, dst, nlines, ls, n, p, tw, dir, e0, de); break;
    }
}

#ifdef __HIPCC__
// Batched complex FFT of `nlines` lines of length P.n held in LDS at a (stride ls),
// ping-ponging with b.  dir = -1 forward, +1 inverse; unscaled.  All threads of
// the block must call it.  Returns the buffer holding the result.
__device__ inline double2 *fft_lds(double2 *a, double2 *b, int nlines, int ls, const FftPlan &P,
                                   const double2 *__restrict__ tw, int dir) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    int p = 1;
    for (int s = 0; s < P.npass; ++s) {
        fft_pass(a, b, nlines, ls, P.n, P.R[s], p, tw, dir, tid, nthr);
        __syncthreads();
        double2 *t = a;
        a = b;
        b = t;
        p *= P.R[s];
    }
    return a;
}
#endif

// Per-handle constants every kernel reads (passed by value).
struct Grid {
    int nx, ny, nz, zs;           // zs = padded column stride
    int nwx, nwy;
    double dzi, hdzi;             // 1/dz, 1/(2 dz)
    double lz;                    // extent(3)
    double zlo;                   // lower(3)
    double dz;
    // step gate: when set, every kernel of a step returns at once while gate[0] == 0
    // (written by k_gate at the start of each step of ps3d_gpu_advance_until)
    const double *gate;
};

#ifdef __HIPCC__
__device__ inline bool gated_off(const double *gate) { return gate != nullptr && gate[0] == 0.0; }
#endif

}  // namespace ps3d
