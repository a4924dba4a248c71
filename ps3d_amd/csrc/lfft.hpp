// lfft.hpp -- compile-time FFT plans for the LDS-resident batched FFTs.
//
// For the lengths the grids use (2^a 3^b up to 2048) the radix sequence, strides
// and twiddle steps are template constants, so the Stockham index arithmetic
// folds to shifts/multiplies and every butterfly lives in registers.  Radix 16
// (4 x 4 in registers) halves the number of LDS passes of radix 4.
#pragma once
#include "common.hpp"

namespace ps3d {

// In-register radix-16 DFT (dir = -1 forward): 16 = 4 x 4 with the w16 twiddles.
__host__ __device__ inline void bf16(double2 *u, int dir) {
    const double c1 = 0.92387953251128675613818887496519852;   // cos(pi/8)
    const double s1 = 0.38268343236508977172845998403039887;   // sin(pi/8)
    const double r = 0.70710678118654752440084436210484903928;  // sqrt(1/2)
#pragma unroll
    for (int a = 0; a < 4; ++a) dft4(u[a], u[a + 4], u[a + 8], u[a + 12], dir);
    const double sg = dir < 0 ? -1.0 : 1.0;
    auto tw = [sg](double2 v, double c, double s) {
        return make_double2(v.x * c - v.y * sg * s, v.x * sg * s + v.y * c);
    };
    // u[a + 4b] *= w16^(a b)
    u[5] = tw(u[5], c1, s1);
    u[6] = tw(u[6], r, r);
    u[7] = tw(u[7], s1, c1);
    u[9] = tw(u[9], r, r);
    u[10] = cmuli(u[10], dir);
    u[11] = tw(u[11], -r, r);
    u[13] = tw(u[13], s1, c1);
    u[14] = tw(u[14], -r, r);
    u[15] = tw(u[15], -c1, -s1);
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4(u[4 * b], u[4 * b + 1], u[4 * b + 2], u[4 * b + 3], dir);
    double2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = u[i];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) u[b + 4 * k2] = t[4 * b + k2];
}

template <int R>
__host__ __device__ inline void bfly(double2 *u, int dir) {
    if constexpr (R == 16)
        bf16(u, dir);
    else
        butterfly(R, u, dir);
}

// Radix sequence of a length (0 terminates).
template <int N>
struct CPlan {
    static constexpr bool ok = false;
};
#define PS3D_CPLAN(N_, A, B, C, D)                                 \
    template <>                                                    \
    struct CPlan<N_> {                                             \
        static constexpr bool ok = true;                           \
        static constexpr int r0 = A, r1 = B, r2 = C, r3 = D;       \
    };
PS3D_CPLAN(16, 16, 0, 0, 0)
PS3D_CPLAN(32, 8, 4, 0, 0)
PS3D_CPLAN(48, 16, 3, 0, 0)
PS3D_CPLAN(64, 8, 8, 0, 0)
PS3D_CPLAN(96, 16, 2, 3, 0)
PS3D_CPLAN(128, 16, 8, 0, 0)
PS3D_CPLAN(192, 16, 4, 3, 0)
PS3D_CPLAN(256, 16, 16, 0, 0)
PS3D_CPLAN(384, 16, 8, 3, 0)
PS3D_CPLAN(512, 16, 16, 2, 0)
PS3D_CPLAN(768, 16, 16, 3, 0)
PS3D_CPLAN(1024, 16, 16, 4, 0)
PS3D_CPLAN(1536, 16, 16, 2, 3)
PS3D_CPLAN(2048, 16, 16, 8, 0)
#undef PS3D_CPLAN

// Radix-8 first plans (fewer live registers per butterfly), used by the z kernels
// whose threads also hold several columns of field data.
template <int N>
struct CPlan8 {
    static constexpr bool ok = false;
};
#define PS3D_CPLAN8(N_, A, B, C, D)                                \
    template <>                                                    \
    struct CPlan8<N_> {                                            \
        static constexpr bool ok = true;                           \
        static constexpr int r0 = A, r1 = B, r2 = C, r3 = D;       \
    };
PS3D_CPLAN8(64, 8, 8, 0, 0)
PS3D_CPLAN8(96, 8, 4, 3, 0)
PS3D_CPLAN8(128, 8, 8, 2, 0)
PS3D_CPLAN8(192, 8, 8, 3, 0)
PS3D_CPLAN8(256, 8, 8, 4, 0)
PS3D_CPLAN8(384, 8, 8, 2, 3)
PS3D_CPLAN8(512, 8, 8, 8, 0)
PS3D_CPLAN8(768, 8, 8, 4, 3)
PS3D_CPLAN8(1024, 8, 8, 4, 4)
PS3D_CPLAN8(1536, 8, 8, 8, 3)
PS3D_CPLAN8(2048, 8, 8, 8, 4)
#undef PS3D_CPLAN8

// LDS slot of line element x.  SKEW: one pad slot after every 16 elements, so that
// a pass writing with a stride of R = 4, 8 or 16 elements (the first Stockham pass)
// spreads a wave's 16-byte stores over the LDS banks instead of one bank group.
template <bool SKEW>
__host__ __device__ constexpr int lslot(int x) {
    return SKEW ? x + (x >> 4) : x;
}

// Element x of the complex line whose inverse FFT gives two real lines from their
// Hermitian real spectra (stafft.f90:55-58): slot k < n/2 holds the real parts
// (a(k), b(k)) and slot n-k the imaginary parts; scaled by hs.
template <int N, bool SKEW = false>
__host__ __device__ inline double2 herm_in(const double2 *s, int x, double hs) {
    constexpr int H = N / 2;
    if (x == 0 || x == H) {
        const double2 v = s[lslot<SKEW>(x)];
        return make_double2(hs * v.x, hs * v.y);
    }
    const double2 a = s[lslot<SKEW>(x < H ? x : N - x)], b = s[lslot<SKEW>(x < H ? N - x : x)];   // real, imaginary slots
    return x < H ? make_double2(hs * (a.x - b.y), hs * (b.x + a.y)) : make_double2(hs * (a.x + b.y), hs * (a.y - b.x));
}

// One Stockham pass with compile-time length, radix and span.  HERM (first pass
// only): the input lines are Hermitian real spectra, combined on the fly.
template <int N, int R, int P, bool HERM = false, bool SKEW = false>
__host__ __device__ inline void cpass(const double2 *src, double2 *dst, int nlines, int ls, const double2 *tw, int dir,
                                      int e0, int de, double hs = 1.0) {
    static_assert(!HERM || P == 1, "Hermitian input only in the first pass");
    constexpr int m = N / R, tstep = N / (P * R);
    const int tot = nlines * m;
    for (int e = e0; e < tot; e += de) {
        const int line = e / m;
        const int i = e - line * m;
        const int k = i % P;
        const double2 *s = src + (size_t)line * ls;
        double2 *d = dst + (size_t)line * ls;
        double2 u[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            double2 v = HERM ? herm_in<N, SKEW>(s, i + t * m, hs) : s[lslot<SKEW>(i + t * m)];
            if (P > 1 && t > 0) {
                const double2 w = tw[k * t * tstep];
                v = dir < 0 ? cmul(v, w) : cmulc(v, w);
            }
            u[t] = v;
        }
        bfly<R>(u, dir);
        const int j = (i - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) d[lslot<SKEW>(j + t * P)] = u[t];
    }
}

#ifdef __HIPCC__
// Batched complex FFT of `nlines` lines of length N in LDS (stride ls), ping-pong
// with b; tw[m] = exp(-2 pi i m/N).  All threads of the block call it.
// HERM: the input lines hold Hermitian real spectra (scaled by hs on the way in),
// which saves the separate pass that would form the complex lines.
template <int N, typename C = CPlan<N>, bool HERM = false, bool SKEW = false>
__device__ inline double2 *cfft_lds(double2 *a, double2 *b, int nlines, int ls, const double2 *__restrict__ tw,
                                    int dir, double hs = 1.0) {
    static_assert(C::ok, "no compile-time plan for this length");
    const int tid = threadIdx.x, nthr = blockDim.x;
    cpass<N, C::r0, 1, HERM, SKEW>(a, b, nlines, ls, tw, dir, tid, nthr, hs);
    __syncthreads();
    if constexpr (C::r1 == 0) {
        return b;
    } else {
        cpass<N, C::r1, C::r0, false, SKEW>(b, a, nlines, ls, tw, dir, tid, nthr);
        __syncthreads();
        if constexpr (C::r2 == 0) {
            return a;
        } else {
            cpass<N, C::r2, C::r0 * C::r1, false, SKEW>(a, b, nlines, ls, tw, dir, tid, nthr);
            __syncthreads();
            if constexpr (C::r3 == 0) {
                return b;
            } else {
                cpass<N, C::r3, C::r0 * C::r1 * C::r2, false, SKEW>(b, a, nlines, ls, tw, dir, tid, nthr);
                __syncthreads();
                return a;
            }
        }
    }
}
#endif

}  // namespace ps3d
