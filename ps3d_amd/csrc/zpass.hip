// zpass.hip -- z-column kernels of the MI355X pseudo-spectral step.
//
// Work unit: a "quad" of spectral columns {(kx,ky), (nx-kx,ky), (kx,ny-ky),
// (nx-kx,ny-ky)} for kx in [0,nx/2], ky in [0,ny/2].  ps3d's spectral x/y
// derivative of column k reads column n-k (src/fft/sta3dfft.f90:304-377 with the
// index reversal of src/mpi/mpi_reverse.f90:235-244), so every spectral operator of
// vor2vel and vorticity_tendency is local to a quad.  Members that coincide
// (kx in {0, nx/2} or ky in {0, ny/2}) are computed as duplicates and not stored.
//
// Each thread owns z = tq + s*TQ (s < ZPT) of all four columns of its quad in
// registers; z transforms (DST-I of the interior, DCT-I of the whole column,
// src/fft/stafft.f90:410-550) go through LDS as complex FFTs of length 2nz of the
// odd/even extension, two columns per complex line.
#include <algorithm>

#include "kernels.hpp"
#include "lfft.hpp"
#include "zmath.hpp"

namespace ps3d {

// ---------------------------------------------------------------------------
// per-thread / per-quad context
// ---------------------------------------------------------------------------
template <int ZPT>
struct ZT {
    int tq, qi, TQ, QPB, nz, N;      // N = 2 nz
    int ny;
    int z[ZPT];
    bool has[ZPT];
    bool qvalid;                      // quad exists (last block may be partial)
    bool q00;                         // quad (0,0)
    size_t col[4];                    // column offsets (ix*ny+iy)*zs
    size_t colidx[4];                 // column index ix*ny+iy
    bool store[4];                    // member distinct and quad valid
    double fx[4], fy[4];              // derivative factors
    double k2l2, k2l2i, K;
    int kx, ky;
};

template <int ZPT>
__device__ inline ZT<ZPT> make_zt(const Grid &g, const ZTabs &tb, int TQ, int QPB, int nquads, long ioff = 0) {
    ZT<ZPT> T;
    T.TQ = TQ;
    T.QPB = QPB;
    T.tq = threadIdx.x % TQ;
    T.qi = threadIdx.x / TQ;
    T.nz = g.nz;
    T.ny = g.ny;
    T.N = 2 * g.nz;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        T.z[s] = T.tq + s * TQ;
        T.has[s] = T.z[s] <= g.nz;
    }
    int q = blockIdx.x * QPB + T.qi;
    T.qvalid = q < nquads;
    if (!T.qvalid) q = nquads - 1;
    const int nqy = g.ny / 2 + 1;
    const int kx = q / nqy, ky = q % nqy;
    T.kx = kx;
    T.ky = ky;
    const int kxp = (g.nx - kx) % g.nx, kyp = (g.ny - ky) % g.ny;
    const int mx[4] = {kx, kxp, kx, kxp};
    const int my[4] = {ky, ky, kyp, kyp};
    const bool dx = kxp != kx, dy = kyp != ky;
    const bool dist[4] = {true, dx, dy, dx && dy};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        T.colidx[m] = (size_t)mx[m] * g.ny + my[m];
        T.col[m] = T.colidx[m] * (size_t)g.zs + ioff;   // ioff: batched instance
        T.store[m] = dist[m] && T.qvalid;
        // diffx: ds(k) = si * rk(k) * fs(n-k), si = +1 for k > n/2, -1 otherwise;
        // ds(0) = 0 and ds(n/2) = 0 for even n (sta3dfft.f90:318-335)
        const int ix = mx[m], iy = my[m];
        double fxm = (ix >= g.nwx + 1) ? tb.rkx[ix] : -tb.rkx[ix];
        if (ix == 0 || ix == g.nwx) fxm = 0.0;
        double fym = (iy >= g.nwy + 1) ? tb.rky[iy] : -tb.rky[iy];
        if (iy == 0 || iy == g.nwy) fym = 0.0;
        T.fx[m] = fxm;
        T.fy[m] = fym;
    }
    T.q00 = (kx == 0 && ky == 0);
    // k2l2 = rkx^2 + rky^2; k2l2i = 1/k2l2, 0 at (0,0) (inversion_utils.f90:243-258)
    const double rx = tb.rkx[kx], ry = tb.rky[ky];
    T.k2l2 = rx * rx + ry * ry;
    T.k2l2i = T.q00 ? 0.0 : 1.0 / T.k2l2;
    T.K = sqrt(T.k2l2);
    return T;
}

// ---------------------------------------------------------------------------
// LDS helpers
// ---------------------------------------------------------------------------
struct ZLds {
    double2 *A, *B;    // FFT ping-pong: QPB*lines lines of N complex each
    double *lid;       // QPB * 3 fields * 4 members * 2 (z=0, z=nz)
    double *colbuf;    // QPB * 12 * (nz+1)   (aliases A, B)
    int lines;         // 4: two fields per FFT call; 2 (long columns): one
};

// FFT lines per quad that fit the 160 KB of LDS: 4 (two transforms per call) when
// possible, else 2.
__host__ __device__ inline int zpass_lines(int QPB, int nz) {
    const size_t N = 2 * (size_t)nz;
    const size_t fft4 = 2 * (size_t)QPB * 4 * N * sizeof(double2);
    const size_t col = (size_t)QPB * 12 * (nz + 1) * sizeof(double);
    const size_t need4 = (fft4 > col ? fft4 : col) + (size_t)QPB * 24 * sizeof(double);
    return need4 <= 160 * 1024 ? 4 : 2;
}

__device__ inline ZLds zlds(int QPB, int N, int nz) {
    extern __shared__ double2 zsmem[];
    ZLds L;
    L.lines = zpass_lines(QPB, nz);
    L.A = zsmem;
    L.B = zsmem + (size_t)QPB * L.lines * N;
    L.lid = reinterpret_cast<double *>(L.B + (size_t)QPB * L.lines * N);
    L.colbuf = reinterpret_cast<double *>(L.A);
    return L;
}

// Batched complex FFT of the 2*QPB extension lines (length 2 nz) of a block.
template <int NZ>
__device__ inline double2 *zfft(const ZLds &L, int nlines, int N, const FftPlan &P, const double2 *tw) {
    if constexpr (NZ == 0)
        return fft_lds(L.A, L.B, nlines, N, P, tw, -1);
    else
        return cfft_lds<2 * NZ, CPlan8<2 * NZ>>(L.A, L.B, nlines, 2 * NZ, tw, -1);
}

// Broadcast the lid values (z = 0 and z = nz) of F to every thread of the quad.
template <int ZPT>
__device__ inline void get_lids(const ZT<ZPT> &T, const ZLds &L, const double (&F)[ZPT][4], double (&lo)[4],
                                double (&hi)[4]) {
    double *lid = L.lid + T.qi * 24;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (T.has[s] && T.z[s] == 0)
#pragma unroll
            for (int m = 0; m < 4; ++m) lid[m] = F[s][m];
        if (T.has[s] && T.z[s] == T.nz)
#pragma unroll
            for (int m = 0; m < 4; ++m) lid[4 + m] = F[s][m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        lo[m] = lid[m];
        hi[m] = lid[4 + m];
    }
    __syncthreads();
}

// Interior DST-I (stafft.f90:489-550, orthonormal, self-inverse) of the four
// members of F: z = 1..nz-1 replaced, z = 0 and z = nz untouched.
template <int ZPT, int NZ>
__device__ inline void zdst4(const ZT<ZPT> &T, const ZLds &L, double (&F)[ZPT][4], const FftPlan &P,
                             const double2 *tw, double scale) {
    const int n = T.nz, N = T.N;
    double2 *l0 = L.A + (size_t)(T.qi * 2) * N;
    double2 *l1 = l0 + N;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        if (z == 0 || z == n) {
            l0[z] = make_double2(0.0, 0.0);
            l1[z] = make_double2(0.0, 0.0);
        } else {
            l0[z] = make_double2(F[s][0], F[s][1]);
            l0[N - z] = make_double2(-F[s][0], -F[s][1]);
            l1[z] = make_double2(F[s][2], F[s][3]);
            l1[N - z] = make_double2(-F[s][2], -F[s][3]);
        }
    }
    __syncthreads();
    double2 *R = zfft<NZ>(L, T.QPB * 2, N, P, tw);
    const double2 *r0 = R + (size_t)(T.qi * 2) * N;
    const double2 *r1 = r0 + N;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        if (z == 0 || z == n) continue;
        const double2 a = r0[z], b = r1[z];
        F[s][0] = -a.y * scale;
        F[s][1] = a.x * scale;
        F[s][2] = -b.y * scale;
        F[s][3] = b.x * scale;
    }
    __syncthreads();
}

// Full DCT-I (stafft.f90:410-483) of the four members of F (z = 0..nz).
template <int ZPT, int NZ>
__device__ inline void zdct4(const ZT<ZPT> &T, const ZLds &L, double (&F)[ZPT][4], const FftPlan &P,
                             const double2 *tw, double scale) {
    const int n = T.nz, N = T.N;
    double2 *l0 = L.A + (size_t)(T.qi * 2) * N;
    double2 *l1 = l0 + N;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        l0[z] = make_double2(F[s][0], F[s][1]);
        l1[z] = make_double2(F[s][2], F[s][3]);
        if (z > 0 && z < n) {
            l0[N - z] = l0[z];
            l1[N - z] = l1[z];
        }
    }
    __syncthreads();
    double2 *R = zfft<NZ>(L, T.QPB * 2, N, P, tw);
    const double2 *r0 = R + (size_t)(T.qi * 2) * N;
    const double2 *r1 = r0 + N;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        const double2 a = r0[z], b = r1[z];
        F[s][0] = a.x * scale;
        F[s][1] = a.y * scale;
        F[s][2] = b.x * scale;
        F[s][3] = b.y * scale;
    }
    __syncthreads();
}

// central_diffz (inversion_utils.f90:653-673) of F into D: one-sided at the lids.
template <int ZPT>
__device__ inline void zdiff(const ZT<ZPT> &T, const ZLds &L, const Grid &g, const double (&F)[ZPT][4],
                             double (&D)[ZPT][4]) {
    const int nz1 = T.nz + 1;
    double *cb = L.colbuf + (size_t)T.qi * 4 * nz1;
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
        if (T.has[s])
#pragma unroll
            for (int m = 0; m < 4; ++m) cb[m * nz1 + T.z[s]] = F[s][m];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double *c = cb + m * nz1;
            if (z == 0)
                D[s][m] = g.dzi * (c[1] - c[0]);
            else if (z == T.nz)
                D[s][m] = g.dzi * (c[T.nz] - c[T.nz - 1]);
            else
                D[s][m] = (c[z + 1] - c[z - 1]) * g.hdzi;
        }
    }
    __syncthreads();
}

// field_decompose_semi_spectral (inversion_utils.f90:563-592) in registers.
template <int ZPT, int NZ>
__device__ inline void zdecompose(const ZT<ZPT> &T, const ZLds &L, const Grid &g, double (&F)[ZPT][4],
                                  const FftPlan &P, const double2 *tw, double scale) {
    double lo[4], hi[4];
    get_lids(T, L, F, lo, hi);
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s] || T.z[s] == 0 || T.z[s] == T.nz) continue;
        double pm, pp;
        phi_only(g, T.K, T.q00, T.z[s], pm, pp);
#pragma unroll
        for (int m = 0; m < 4; ++m) F[s][m] = F[s][m] - (lo[m] * pm + hi[m] * pp);
    }
    zdst4<ZPT, NZ>(T, L, F, P, tw, scale);
}

// field_combine_semi_spectral (inversion_utils.f90:617-647) in registers.
template <int ZPT, int NZ>
__device__ inline void zcombine(const ZT<ZPT> &T, const ZLds &L, const Grid &g, double (&F)[ZPT][4],
                                const FftPlan &P, const double2 *tw, double scale) {
    double lo[4], hi[4];
    get_lids(T, L, F, lo, hi);
    zdst4<ZPT, NZ>(T, L, F, P, tw, scale);
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s] || T.z[s] == 0 || T.z[s] == T.nz) continue;
        double pm, pp;
        phi_only(g, T.K, T.q00, T.z[s], pm, pp);
#pragma unroll
        for (int m = 0; m < 4; ++m) F[s][m] = F[s][m] + lo[m] * pm + hi[m] * pp;
    }
}


enum { XK_DST = 0, XK_DCT = 1 };

// Extension of the four members of F into two complex lines (DST-I: odd about
// z = 0 and nz, interior only; DCT-I: even), and the read-back of the transform.
template <int ZPT, int K>
__device__ inline void ext_put(const ZT<ZPT> &T, double2 *l0, double2 *l1, const double (&F)[ZPT][4]) {
    const int n = T.nz, N = T.N;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        if (K == XK_DST) {
            if (z == 0 || z == n) {
                l0[z] = make_double2(0.0, 0.0);
                l1[z] = make_double2(0.0, 0.0);
            } else {
                l0[z] = make_double2(F[s][0], F[s][1]);
                l0[N - z] = make_double2(-F[s][0], -F[s][1]);
                l1[z] = make_double2(F[s][2], F[s][3]);
                l1[N - z] = make_double2(-F[s][2], -F[s][3]);
            }
        } else {
            l0[z] = make_double2(F[s][0], F[s][1]);
            l1[z] = make_double2(F[s][2], F[s][3]);
            if (z > 0 && z < n) {
                l0[N - z] = l0[z];
                l1[N - z] = l1[z];
            }
        }
    }
}
template <int ZPT, int K>
__device__ inline void ext_get(const ZT<ZPT> &T, const double2 *r0, const double2 *r1, double (&F)[ZPT][4],
                               double scale) {
    const int n = T.nz;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
        const double2 a = r0[z], b = r1[z];
        if (K == XK_DST) {
            if (z == 0 || z == n) continue;
            F[s][0] = -a.y * scale;
            F[s][1] = a.x * scale;
            F[s][2] = -b.y * scale;
            F[s][3] = b.x * scale;
        } else {
            F[s][0] = a.x * scale;
            F[s][1] = a.y * scale;
            F[s][2] = b.x * scale;
            F[s][3] = b.y * scale;
        }
    }
}

// Two z transforms of the quad in one batched FFT call (4 lines per quad).
template <int ZPT, int NZ, int K>
__device__ inline void zxform1(const ZT<ZPT> &T, const ZLds &L, double (&F)[ZPT][4], const FftPlan &P,
                               const double2 *tw, double scale) {
    const int N = T.N;
    double2 *b = L.A + (size_t)(T.qi * 2) * N;
    ext_put<ZPT, K>(T, b, b + N, F);
    __syncthreads();
    double2 *R = zfft<NZ>(L, T.QPB * 2, N, P, tw);
    const double2 *r = R + (size_t)(T.qi * 2) * N;
    ext_get<ZPT, K>(T, r, r + N, F, scale);
    __syncthreads();
}

template <int ZPT, int NZ, int K0, int K1>
__device__ inline void zxform2(const ZT<ZPT> &T, const ZLds &L, double (&F0)[ZPT][4], double (&F1)[ZPT][4],
                               const FftPlan &P, const double2 *tw, double scale) {
    if (L.lines != 4) {   // long columns: one transform per call
        zxform1<ZPT, NZ, K0>(T, L, F0, P, tw, scale);
        zxform1<ZPT, NZ, K1>(T, L, F1, P, tw, scale);
        return;
    }
    const int N = T.N;
    double2 *b = L.A + (size_t)(T.qi * 4) * N;
    ext_put<ZPT, K0>(T, b, b + N, F0);
    ext_put<ZPT, K1>(T, b + 2 * N, b + 3 * N, F1);
    __syncthreads();
    double2 *R = zfft<NZ>(L, T.QPB * 4, N, P, tw);
    const double2 *r = R + (size_t)(T.qi * 4) * N;
    ext_get<ZPT, K0>(T, r, r + N, F0, scale);
    ext_get<ZPT, K1>(T, r + 2 * N, r + 3 * N, F1, scale);
    __syncthreads();
}

// Lid values of three fields broadcast with one barrier pair.
template <int ZPT>
__device__ inline void get_lids3(const ZT<ZPT> &T, const ZLds &L, const double (&F0)[ZPT][4],
                                 const double (&F1)[ZPT][4], const double (&F2)[ZPT][4], double (&lo)[3][4],
                                 double (&hi)[3][4]) {
    double *lid = L.lid + T.qi * 24;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int off = T.z[s] == 0 ? 0 : (T.z[s] == T.nz ? 12 : -1);
        if (off < 0) continue;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            lid[off + m] = F0[s][m];
            lid[off + 4 + m] = F1[s][m];
            lid[off + 8 + m] = F2[s][m];
        }
    }
    __syncthreads();
    for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            lo[f][m] = lid[4 * f + m];
            hi[f][m] = lid[12 + 4 * f + m];
        }
    __syncthreads();
}

// central_diffz of one or two fields through one LDS column exchange; also returns
// the lid values of each derivative (known to every thread from the columns).
template <int ZPT>
__device__ inline void fd_from_cols(const ZT<ZPT> &T, const Grid &g, const double *cb, double (&D)[ZPT][4],
                                    double (&lo)[4], double (&hi)[4]) {
    const int nz1 = T.nz + 1;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s]) continue;
        const int z = T.z[s];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double *c = cb + m * nz1;
            if (z == 0)
                D[s][m] = g.dzi * (c[1] - c[0]);
            else if (z == T.nz)
                D[s][m] = g.dzi * (c[T.nz] - c[T.nz - 1]);
            else
                D[s][m] = (c[z + 1] - c[z - 1]) * g.hdzi;
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double *c = cb + m * nz1;
        lo[m] = g.dzi * (c[1] - c[0]);
        hi[m] = g.dzi * (c[T.nz] - c[T.nz - 1]);
    }
}
template <int ZPT>
__device__ inline void put_cols(const ZT<ZPT> &T, double *cb, const double (&F)[ZPT][4]) {
    const int nz1 = T.nz + 1;
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
        if (T.has[s])
#pragma unroll
            for (int m = 0; m < 4; ++m) cb[m * nz1 + T.z[s]] = F[s][m];
}
template <int ZPT>
__device__ inline void zdiff2(const ZT<ZPT> &T, const ZLds &L, const Grid &g, const double (&F0)[ZPT][4],
                              double (&D0)[ZPT][4], double (&lo0)[4], double (&hi0)[4], const double (&F1)[ZPT][4],
                              double (&D1)[ZPT][4], double (&lo1)[4], double (&hi1)[4]) {
    const int nz1 = T.nz + 1;
    double *cb0 = L.colbuf + (size_t)T.qi * 12 * nz1;
    double *cb1 = cb0 + 4 * nz1;
    put_cols(T, cb0, F0);
    put_cols(T, cb1, F1);
    __syncthreads();
    fd_from_cols(T, g, cb0, D0, lo0, hi0);
    fd_from_cols(T, g, cb1, D1, lo1, hi1);
    __syncthreads();
}
template <int ZPT>
__device__ inline void zdiff1(const ZT<ZPT> &T, const ZLds &L, const Grid &g, const double (&F0)[ZPT][4],
                              double (&D0)[ZPT][4], double (&lo0)[4], double (&hi0)[4]) {
    const int nz1 = T.nz + 1;
    double *cb0 = L.colbuf + (size_t)T.qi * 12 * nz1;
    put_cols(T, cb0, F0);
    __syncthreads();
    fd_from_cols(T, g, cb0, D0, lo0, hi0);
    __syncthreads();
}

// subtract (decompose) or add (combine) the harmonic interpolant of the lids
template <int ZPT>
__device__ inline void harmonic(const ZT<ZPT> &T, const Grid &g, double (&F)[ZPT][4], const double (&lo)[4],
                                const double (&hi)[4], double sign) {
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s] || T.z[s] == 0 || T.z[s] == T.nz) continue;
        double pm, pp;
        phi_only(g, T.K, T.q00, T.z[s], pm, pp);
#pragma unroll
        for (int m = 0; m < 4; ++m) F[s][m] = F[s][m] + sign * (lo[m] * pm + hi[m] * pp);
    }
}

template <int ZPT>
__device__ inline void zmid(const ZT<ZPT> &T, const ZLds &L, const double (&F)[ZPT][4], double *__restrict__ dst) {
    const int nz1 = T.nz + 1;
    double *cb = L.colbuf + (size_t)T.qi * 4 * nz1;
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
        if (T.has[s])
#pragma unroll
            for (int m = 0; m < 4; ++m) cb[m * nz1 + T.z[s]] = F[s][m];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        const int z = T.z[s];
        if (!T.has[s]) continue;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (!T.store[m]) continue;
            const double v = (z > 0) ? 0.5 * (cb[m * nz1 + z - 1] + cb[m * nz1 + z]) : 0.0;
            dst[T.col[m] + z] = v;
        }
    }
    __syncthreads();
}

template <int ZPT>
__device__ inline void mid_store(const ZT<ZPT> &T, const double *cb, double *__restrict__ dst) {
    const int nz1 = T.nz + 1;
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        const int z = T.z[s];
        if (!T.has[s]) continue;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (!T.store[m]) continue;
            dst[T.col[m] + z] = (z > 0) ? 0.5 * (cb[m * nz1 + z - 1] + cb[m * nz1 + z]) : 0.0;
        }
    }
}
template <int ZPT>
__device__ inline void zmid3(const ZT<ZPT> &T, const ZLds &L, const double (&F0)[ZPT][4], const double (&F1)[ZPT][4],
                             const double (&F2)[ZPT][4], double *d0, double *d1, double *d2) {
    const int nz1 = T.nz + 1;
    double *cb = L.colbuf + (size_t)T.qi * 12 * nz1;
    put_cols(T, cb, F0);
    put_cols(T, cb + 4 * nz1, F1);
    put_cols(T, cb + 8 * nz1, F2);
    __syncthreads();
    mid_store(T, cb, d0);
    mid_store(T, cb + 4 * nz1, d1);
    mid_store(T, cb + 8 * nz1, d2);
    __syncthreads();
}

template <int ZPT>
__device__ inline void zload(const ZT<ZPT> &T, const double *__restrict__ src, double (&F)[ZPT][4]) {
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) F[s][m] = T.has[s] ? src[T.col[m] + T.z[s]] : 0.0;
}
template <int ZPT>
__device__ inline void zstore(const ZT<ZPT> &T, double *__restrict__ dst, const double (&F)[ZPT][4]) {
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
        if (T.has[s])
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (T.store[m]) dst[T.col[m] + T.z[s]] = F[s][m];
}

// ---------------------------------------------------------------------------
// vor2vel, vertical part (src/inversion/inversion.f90:23-226) for one quad.
// In : svor (mixed).  Out: semi-spectral u, v, w, xi, eta, zeta; optionally the
// solenoidally projected svor(xi, eta) and the z-midpoint vorticity
// (omega(z-1) + omega(z))/2 at index z = 1..nz (for the characteristic vorticity).
// ---------------------------------------------------------------------------
template <int ZPT, int NZ>
__global__ void __launch_bounds__(512) k_zinv(ZInvArgs a) {
    const Grid g = a.g;
    if (gated_off(g.gate)) return;
    ZT<ZPT> T = make_zt<ZPT>(g, a.tb, a.TQ, a.QPB, a.nquads, (long)blockIdx.y * a.bt.vol);
    ZLds L = zlds(a.QPB, T.N, g.nz);
    const FftPlan &P = a.pz;
    const double sc = a.zscale;
    const int nz = g.nz;

    double Xi[ZPT][4], Et[ZPT][4], C[ZPT][4], D[ZPT][4], E[ZPT][4];
    zload(T, a.svor[0], Xi);
    zload(T, a.svor[1], Et);
    zload(T, a.svor[2], C);

    // D = B_x - A_y (inversion.f90:41-45)
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) D[s][m] = T.fx[m] * Et[s][m ^ 1] - T.fy[m] * Xi[s][m ^ 2];
    // C = combine(zeta); E = decompose(central_diffz(C)) (46-48)
    {
        double c0[4], cn[4];
        get_lids(T, L, C, c0, cn);
        zdst4<ZPT, NZ>(T, L, C, P, a.twz, sc);
        harmonic(T, g, C, c0, cn, 1.0);
        double e0[4], en[4];
        zdiff1(T, L, g, C, E, e0, en);
        harmonic(T, g, E, e0, en, -1.0);
        zdst4<ZPT, NZ>(T, L, E, P, a.twz, sc);
    }

    // projection A = k2l2i (E_x + D_y), B = k2l2i (E_y - D_x); (0,0) column kept (56-75)
    double Xp[ZPT][4], Ep[ZPT][4];
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (T.q00) {
                Xp[s][m] = Xi[s][m];
                Ep[s][m] = Et[s][m];
            } else {
                Xp[s][m] = T.k2l2i * (T.fx[m] * E[s][m ^ 1] + T.fy[m] * D[s][m ^ 2]);
                Ep[s][m] = T.k2l2i * (T.fy[m] * E[s][m ^ 2] - T.fx[m] * D[s][m ^ 1]);
            }
        }
    if (a.svor_out[0]) {
        zstore(T, a.svor_out[0], Xp);
        zstore(T, a.svor_out[1], Ep);
    }

    // source of w: ds = A'_y - B'_x (86-90), its lid values for the boundary part
    double Ds[ZPT][4];
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) Ds[s][m] = T.fy[m] * Xp[s][m ^ 2] - T.fx[m] * Ep[s][m ^ 1];
    // lid values of ds (boundary part of w) and of the projected A, B (mean flow
    // of the (0,0) column and the combine below)
    double lo3[3][4], hi3[3][4];
    get_lids3(T, L, Ds, Xp, Ep, lo3, hi3);
    const double(&d0)[4] = lo3[0];
    const double(&dn)[4] = hi3[0];

    // bs (interior) and es (all z): boundary parts (94-104); green and d/dz of the
    // sine part (107-122)
    double Bs[ZPT][4], As[ZPT][4];
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        const int z = T.z[s];
        const Hyp h = hyp_all(g, T.K, T.k2l2i, T.q00, T.has[s] ? z : 0);
        const double rkz = a.rkz_sc * (double)z;
        const double green = -1.0 / (T.k2l2 + rkz * rkz);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            Bs[s][m] = d0[m] * h.thetam + dn[m] * h.thetap;
            E[s][m] = d0[m] * h.dthetam + dn[m] * h.dthetap;     // es
            if (z > 0 && z < nz) {
                Ds[s][m] = green * Ds[s][m];
                As[s][m] = rkz * Ds[s][m];
            } else {
                As[s][m] = 0.0;
            }
        }
        // quad (0,0): the mean flow (150-165) rides in the duplicate members 1, 2
        if (T.q00) {
            const double rkzi = (z > 0 && z < nz) ? 1.0 / rkz : 0.0;
            As[s][1] = (z > 0 && z < nz) ? -rkzi * Ep[s][0] : 0.0;
            As[s][2] = (z > 0 && z < nz) ? rkzi * Xp[s][0] : 0.0;
        }
    }
    zxform2<ZPT, NZ, XK_DCT, XK_DST>(T, L, As, Ds, P, a.twz, sc);

    // complete w and dw/dz (134-141); xi, eta combined for the physical vorticity
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        const int z = T.z[s];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (z == 0 || z == nz)
                Ds[s][m] = 0.0;
            else
                Ds[s][m] = Ds[s][m] + Bs[s][m];
            E[s][m] = E[s][m] + As[s][m];
        }
    }
    double ubar[ZPT], vbar[ZPT];
    if (T.q00) {
        // gamtop = Lz/2 (phip00^2 - 1/3), gambot(iz) = gamtop(nz-iz) (350-369)
#pragma unroll
        for (int s = 0; s < ZPT; ++s) {
            const int z = T.has[s] ? T.z[s] : 0;
            double zm, zp, zm2, zp2;
            zm_zp(g, z, zm, zp);
            zm_zp(g, nz - z, zm2, zp2);
            const double pt = zp / g.lz, pb = zp2 / g.lz;
            const double gamtop = 0.5 * g.lz * (pt * pt - 1.0 / 3.0);
            const double gambot = 0.5 * g.lz * (pb * pb - 1.0 / 3.0);
            ubar[s] = As[s][1] + hi3[2][0] * gamtop - lo3[2][0] * gambot;
            vbar[s] = As[s][2] - hi3[1][0] * gamtop + lo3[1][0] * gambot;
        }
    }
    zxform2<ZPT, NZ, XK_DST, XK_DST>(T, L, Xp, Ep, P, a.twz, sc);
    harmonic(T, g, Xp, lo3[1], hi3[1], 1.0);
    harmonic(T, g, Ep, lo3[2], hi3[2], 1.0);

    // u = k2l2i (es_x + cs_y), v = k2l2i (es_y - cs_x); (0,0): mean flow (169-213)
    double U[ZPT][4], V[ZPT][4];
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (T.q00) {
                U[s][m] = ubar[s];
                V[s][m] = vbar[s];
            } else {
                U[s][m] = T.k2l2i * (T.fx[m] * E[s][m ^ 1] + T.fy[m] * C[s][m ^ 2]);
                V[s][m] = T.k2l2i * (T.fy[m] * E[s][m ^ 2] - T.fx[m] * C[s][m ^ 1]);
            }
        }
    zstore(T, a.out[0], U);
    zstore(T, a.out[1], V);
    zstore(T, a.out[2], Ds);
    zstore(T, a.out[3], Xp);
    zstore(T, a.out[4], Ep);
    zstore(T, a.out[5], C);

    if (a.mid[0]) {
        // (omega(z-1) + omega(z))/2 at z = 1..nz (0 at z = 0): the cell-midpoint
        // vorticity of get_char_vorticity (field_diagnostics.f90:509-521)
        zmid3(T, L, Xp, Ep, C, a.mid[0], a.mid[1], a.mid[2]);
    }
}

// RK4 / cn2 update of one quad from the mixed-spectral tendency S (registers).
template <int ZPT>
__device__ inline void rk_update(const ZT<ZPT> &T, const RkArgs &rk, int mode, int stage, double (&S)[3][ZPT][4],
                                 const ZTabs &tb) {
    const double *sc = rk.scal + (size_t)blockIdx.y * SC_NSCAL;
    const double dt = sc[SC_DT];
    // filt(0,ky,kx) with filt(:,0,0) = 1 (inversion_utils.f90:263-277, 377-455)
    double f0;
    if (T.q00)
        f0 = 1.0;
    else if (tb.hou_li)
        f0 = exp(tb.filx[T.kx] + tb.fily[T.ky]);
    else
        f0 = tb.filx[T.kx] * tb.fily[T.ky];
    // vhdis = visc * (k^2+l^2)^nnu (inversion_utils.f90:188-218)
    double kp = 1.0;
    for (int i = 0; i < tb.nnu; ++i) kp *= T.k2l2;
    const double vhdis = sc[SC_VISC] * kp;
    if (mode == ZT_CN2) {
        // cn2: vdiss = 1/(1 + dfac vhdis); svor = filt (vortsm + dt/2 svorts) then x vdiss
        const double vd = 1.0 / (1.0 + sc[SC_DFAC] * vhdis);
        const double dt2 = 0.5 * dt;
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int s = 0; s < ZPT; ++s) {
                if (!T.has[s]) continue;
                const int z = T.z[s];
                double fz = f0;
                if (z > 0 && z < T.nz && !T.q00) fz = tb.hou_li ? f0 * exp(tb.filz[z]) : f0 * tb.filz[z];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (!T.store[m]) continue;
                    const size_t i = T.col[m] + z;
                    double vs;
                    if (stage == 1) {
                        vs = rk.q[c][i] + dt2 * S[c][s][m];     // vortsm = svor + dt/2 svorts
                        rk.qdf[c][i] = vs;
                    } else {
                        vs = rk.qdf[c][i];
                    }
                    rk.q[c][i] = vd * (fz * (vs + dt2 * S[c][s][m]));
                }
            }
        }
        return;
    }
    // impl-diff-rk4: vdiss = dfac * vhdis, dfac = vorch*dt/2 (impl_rk4.f90:37-53);
    // epq = exp(vdiss) filt0, emq = 1/exp(vdiss) (87-89)
    const double dt2 = 0.5 * dt, dt3 = dt / 3.0, dt6 = dt / 6.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (!T.store[m]) continue;
        const double D = rk.vdiss ? rk.vdiss[T.colidx[m]] : sc[SC_DFAC] * vhdis;
        const double ep = exp(D);
        const double em = 1.0 / ep;
        const double epf = ep * f0;
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int s = 0; s < ZPT; ++s) {
                if (!T.has[s]) continue;
                const size_t i = T.col[m] + T.z[s];
                const double t = S[c][s][m];
                if (stage == 1) {
                    const double sq = f0 * t;                 // filtered source (227-229)
                    const double q0 = rk.q[c][i];
                    rk.qdi[c][i] = q0;
                    rk.q[c][i] = (q0 + dt2 * sq) * em;
                    rk.qdf[c][i] = q0 + dt6 * sq;
                } else if (stage == 2) {
                    const double sq = epf * t;
                    rk.q[c][i] = em * (rk.qdi[c][i] + dt2 * sq);
                    rk.qdf[c][i] = rk.qdf[c][i] + dt3 * sq;
                } else if (stage == 3) {
                    const double sq = epf * t;
                    rk.q[c][i] = (em * em) * (rk.qdi[c][i] + dt * sq);
                    rk.qdf[c][i] = rk.qdf[c][i] + dt3 * sq;
                } else {
                    const double sq = (epf * epf) * t;
                    rk.q[c][i] = (em * em) * (rk.qdf[c][i] + dt6 * sq);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// vorticity_tendency, vertical part (inversion.f90:298-371), fused with one
// integrating-factor RK4 substep (impl_rk4.f90:212-364) or the cn2 update
// (cn2.f90:92-181).  In: semi-spectral r, q, p.  The substep's
// combine -> multiply by a (kx,ky) factor -> decompose sandwiches are applied as the
// plain multiply they equal (both transforms are linear and act per column).
// ---------------------------------------------------------------------------
template <int ZPT, int NZ>
__global__ void __launch_bounds__(512) k_ztend(ZTendArgs a) {
    const Grid g = a.g;
    if (gated_off(g.gate)) return;
    ZT<ZPT> T = make_zt<ZPT>(g, a.tb, a.TQ, a.QPB, a.nquads, (long)blockIdx.y * a.bt.vol);
    ZLds L = zlds(a.QPB, T.N, g.nz);
    const FftPlan &P = a.pz;
    const double sc = a.zscale;

    double R[ZPT][4], Q[ZPT][4], Pp[ZPT][4], FQ[ZPT][4], FP[ZPT][4];
    zload(T, a.rqp[0], R);
    zload(T, a.rqp[1], Q);
    zload(T, a.rqp[2], Pp);
    // central_diffz of q and p (342, 356) and the lids of all five fields
    double fq0[4], fqn[4], fp0[4], fpn[4];
    zdiff2(T, L, g, Q, FQ, fq0, fqn, Pp, FP, fp0, fpn);
    double lo3[3][4], hi3[3][4];
    get_lids3(T, L, R, Q, Pp, lo3, hi3);
    // field_decompose of r, q, p, dq/dz, dp/dz: subtract the harmonic part, then DST
#pragma unroll
    for (int s = 0; s < ZPT; ++s) {
        if (!T.has[s] || T.z[s] == 0 || T.z[s] == T.nz) continue;
        double pm, pp;
        phi_only(g, T.K, T.q00, T.z[s], pm, pp);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            R[s][m] = R[s][m] - (lo3[0][m] * pm + hi3[0][m] * pp);
            Q[s][m] = Q[s][m] - (lo3[1][m] * pm + hi3[1][m] * pp);
            Pp[s][m] = Pp[s][m] - (lo3[2][m] * pm + hi3[2][m] * pp);
            FQ[s][m] = FQ[s][m] - (fq0[m] * pm + fqn[m] * pp);
            FP[s][m] = FP[s][m] - (fp0[m] * pm + fpn[m] * pp);
        }
    }
    zxform2<ZPT, NZ, XK_DST, XK_DST>(T, L, R, FQ, P, a.twz, sc);
    zxform2<ZPT, NZ, XK_DST, XK_DST>(T, L, FP, Q, P, a.twz, sc);
    zdst4<ZPT, NZ>(T, L, Pp, P, a.twz, sc);
    // dxi/dt = r_y - q_z ; deta/dt = p_z - r_x ; dzeta/dt = q_x - p_y (341-367)
    double S[3][ZPT][4];
#pragma unroll
    for (int s = 0; s < ZPT; ++s)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            S[0][s][m] = T.fy[m] * R[s][m ^ 2] - FQ[s][m];
            S[1][s][m] = FP[s][m] - T.fx[m] * R[s][m ^ 1];
            S[2][s][m] = T.fx[m] * Q[s][m ^ 1] - T.fy[m] * Pp[s][m ^ 2];
        }

    if (a.extra[0]) {           // closure tendency
#pragma unroll
        for (int s = 0; s < ZPT; ++s)
            if (T.has[s])
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    for (int c = 0; c < 3; ++c) S[c][s][m] += a.extra[c][T.col[m] + T.z[s]];
    }
    if (a.mode == ZT_STORE) {   // plain source: svorts
        for (int c = 0; c < 3; ++c) zstore(T, a.svorts[c], S[c]);
        return;
    }
    rk_update(T, a.rk, a.mode, a.stage, S, a.tb);
}

// RK substep from a stored tendency (the API path: source then step).
template <int ZPT>
__global__ void __launch_bounds__(512) k_zrk(ZTendArgs a) {
    const Grid g = a.g;
    ZT<ZPT> T = make_zt<ZPT>(g, a.tb, a.TQ, a.QPB, a.nquads, (long)blockIdx.y * a.bt.vol);
    double S[3][ZPT][4];
    for (int c = 0; c < 3; ++c) zload(T, a.svorts[c], S[c]);
    rk_update(T, a.rk, a.mode, a.stage, S, a.tb);
}

// Single z operators on one volume (kernel checks): fftsine, fftcosine,
// central_diffz, combine / decompose semi-spectral.
template <int ZPT, int NZ>
__global__ void __launch_bounds__(512) k_zop(ZOpArgs a) {
    const Grid g = a.g;
    ZT<ZPT> T = make_zt<ZPT>(g, a.tb, a.TQ, a.QPB, a.nquads);
    ZLds L = zlds(a.QPB, T.N, g.nz);
    double F[ZPT][4], O[ZPT][4];
    zload(T, a.in, F);
    switch (a.op) {
        case ZOP_DST: {
            zdst4<ZPT, NZ>(T, L, F, a.pz, a.twz, a.zscale);
            // stafft dst sets x(n) = 0 (stafft.f90:546-549); x(0) is not part of it
#pragma unroll
            for (int s = 0; s < ZPT; ++s)
                if (T.has[s] && T.z[s] == T.nz)
#pragma unroll
                    for (int m = 0; m < 4; ++m) F[s][m] = 0.0;
            zstore(T, a.out, F);
            break;
        }
        case ZOP_DCT:
            zdct4<ZPT, NZ>(T, L, F, a.pz, a.twz, a.zscale);
            zstore(T, a.out, F);
            break;
        case ZOP_DIFFZ:
            zdiff(T, L, g, F, O);
            zstore(T, a.out, O);
            break;
        case ZOP_COMBINE:
            zcombine<ZPT, NZ>(T, L, g, F, a.pz, a.twz, a.zscale);
            zstore(T, a.out, F);
            break;
        case ZOP_DECOMPOSE:
            zdecompose<ZPT, NZ>(T, L, g, F, a.pz, a.twz, a.zscale);
            zstore(T, a.out, F);
            break;
        case ZOP_POISSON: {
            // pressure's z part (fields_derived.f90:117-137): dct, times green, dct.
            // green = -1/(k2l2 + rkz^2) for kz >= 1 and -k2l2i at kz = 0
            // (inversion_utils.f90:240-258); k2l2 is the same for the 4 quad members.
            zdct4<ZPT, NZ>(T, L, F, a.pz, a.twz, a.zscale);
#pragma unroll
            for (int s = 0; s < ZPT; ++s) {
                const int z = T.z[s];
                const double rkz = a.rkz_sc * (double)z;
                const double green = z == 0 ? -T.k2l2i : -1.0 / (T.k2l2 + rkz * rkz);
#pragma unroll
                for (int m = 0; m < 4; ++m) F[s][m] = green * F[s][m];
            }
            zdct4<ZPT, NZ>(T, L, F, a.pz, a.twz, a.zscale);
            zstore(T, a.out, F);
            break;
        }
        default:
            break;
    }
}

// Right-hand side of the pressure Poisson equation, 2 (J_xy(u,v) + J_yz(v,w) +
// J_zx(w,u)) with dv/dx = zeta + du/dy, du/dz = eta + dw/dx, dv/dz = dw/dy - xi,
// dw/dz = -(du/dx + dv/dy) (fields_derived.f90:89-104), same operation order.
__global__ void __launch_bounds__(256) k_pres_rhs(PresArgs a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
        const double dudx = a.dudx[i], dudy = a.dudy[i], dvdy = a.dvdy[i], dwdx = a.dwdx[i], dwdy = a.dwdy[i];
        const double dwdz = -(dudx + dvdy);
        a.pres[i] = 2.0 * (dudx * dvdy - dudy * (a.zeta[i] + dudy) + dvdy * dwdz - dwdy * (dwdy - a.xi[i]) +
                           dwdz * dudx - dwdx * (dwdx + a.eta[i]));
    }
}

__global__ void __launch_bounds__(256) k_add2(long n, const double *__restrict__ x, const double *__restrict__ y,
                                               double *__restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = x[i] + y[i];
}

// Horizontal slot map of the stafft Hermitian form (cos slot k <= n/2, sine slot
// n-k of wavenumber k) from a destination index to the source index of the same
// wavenumber, or -1.  The Nyquist slot of either grid is left out (zero).
__device__ inline int slot_map(int i, int nd, int ns) {
    if (i == 0) return 0;
    const int hd = nd / 2, hs = ns / 2;
    if (i < hd) return i < hs ? i : -1;
    if (i == hd) return -1;
    const int m = nd - i;
    return m < hs ? ns - m : -1;
}

// One destination column per block: lids scaled by hfac, interior sine
// coefficients kz < min(nz) by hfac * zfac, the rest zero.
__global__ void __launch_bounds__(256) k_resample(ResampleArgs a) {
    const int col = blockIdx.x;
    const int ix = col / a.dny, iy = col % a.dny;
    const int sx = slot_map(ix, a.dnx, a.snx), sy = slot_map(iy, a.dny, a.sny);
    const bool have = sx >= 0 && sy >= 0;
    const size_t dcol = (size_t)col * a.dzs;
    const size_t scol = have ? ((size_t)sx * a.sny + sy) * a.szs : 0;
    for (int z = threadIdx.x; z <= a.dnz; z += blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
            if (have) {
                if (z == 0)
                    v = a.src[c][scol] * a.hfac;
                else if (z == a.dnz)
                    v = a.src[c][scol + a.snz] * a.hfac;
                else if (z < a.snz)
                    v = a.src[c][scol + z] * (a.hfac * a.zfac);
            }
            a.dst[c][dcol + z] = v;
        }
    }
}

hipError_t launch_resample(const ResampleArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(k_resample, dim3(a.dnx * a.dny), dim3(256), 0, stream, a);
    return hipGetLastError();
}

static unsigned pointwise_blocks(long n) {
    const long b = (n + 255) / 256;
    return (unsigned)std::min<long>(b, 256L * 64L);
}

hipError_t launch_pres_rhs(const PresArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(k_pres_rhs, dim3(pointwise_blocks(a.n)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_add2(long n, const double *a, const double *b, double *out, hipStream_t stream) {
    hipLaunchKernelGGL(k_add2, dim3(pointwise_blocks(n)), dim3(256), 0, stream, n, a, b, out);
    return hipGetLastError();
}

// calc_vorticity_mean / adjust_vorticity_mean (field_diagnostics.f90:584-619) on
// the (0,0) column, plus the end-of-step time update t += dt.  One block.
__global__ void __launch_bounds__(256) k_mean(MeanArgs a) {
    extern __shared__ double2 msmem[];
    if (gated_off(a.g.gate)) return;
    const int n = a.g.nz, N = 2 * n, tid = threadIdx.x;
    double2 *A = msmem, *B = msmem + N;
    double *red = reinterpret_cast<double *>(B + N);
    // column (0,0) starts at offset 0 of each instance
    double *xi = a.svor[0] + (size_t)blockIdx.y * a.bt.vol, *eta = a.svor[1] + (size_t)blockIdx.y * a.bt.vol;
    double *scal = a.scal + (size_t)blockIdx.y * SC_NSCAL;
    for (int z = tid; z <= n; z += blockDim.x) {
        if (z == 0 || z == n) {
            A[z] = make_double2(0.0, 0.0);
        } else {
            A[z] = make_double2(xi[z], eta[z]);
            A[N - z] = make_double2(-xi[z], -eta[z]);
        }
    }
    __syncthreads();
    double2 *R = fft_lds(A, B, 1, N, a.pz, a.twz, -1);
    // interior of the DST: xi -> -Im, eta -> Re (scaled)
    double s0 = 0.0, s1 = 0.0;
    for (int z = tid; z < n; z += blockDim.x)
        if (z > 0) {
            s0 += -R[z].y * a.zscale;
            s1 += R[z].x * a.zscale;
        }
    // deterministic block reduction
    red[tid] = s0;
    red[blockDim.x + tid] = s1;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[tid] += red[tid + w];
            red[blockDim.x + tid] += red[blockDim.x + tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double fnzi = 1.0 / (double)n;
        const double savg0 = 0.5 * (xi[0] + xi[n]) + fnzi * red[0];
        const double savg1 = 0.5 * (eta[0] + eta[n]) + fnzi * red[blockDim.x];
        if (a.mode == MEAN_CALC) {
            scal[SC_INI_XI] = savg0;
            scal[SC_INI_ETA] = savg1;
        } else {
            const double d0 = scal[SC_INI_XI] - savg0, d1 = scal[SC_INI_ETA] - savg1;
            xi[0] = xi[0] + d0;
            xi[n] = xi[n] + d0;
            eta[0] = eta[0] + d1;
            eta[n] = eta[n] + d1;
        }
        // impl_rk4 advances t by dt/2 twice (impl_rk4.f90:123,149), cn2 by dt (cn2.f90:179)
        if (a.advance_time == 1) {
            const double dt2 = 0.5 * scal[SC_DT];
            scal[SC_T] = (scal[SC_T] + dt2) + dt2;
        } else if (a.advance_time == 2) {
            scal[SC_T] = scal[SC_T] + scal[SC_DT];
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
size_t zpass_lds_bytes(int QPB, int nz) {
    const size_t N = 2 * (size_t)nz;
    size_t fft = 2 * (size_t)QPB * zpass_lines(QPB, nz) * N * sizeof(double2);
    size_t col = (size_t)QPB * 12 * (nz + 1) * sizeof(double);
    size_t need = fft > col ? fft : col;
    return need + (size_t)QPB * 24 * sizeof(double);
}

static inline int zbatch(const ZInvArgs &a) { return a.bt.nb > 1 ? a.bt.nb : 1; }
static inline int zbatch(const ZTendArgs &a) { return a.bt.nb > 1 ? a.bt.nb : 1; }
static inline int zbatch(const ZOpArgs &) { return 1; }

// (ZPT, nz) pairs with a compile-time FFT plan; other nz use the runtime plan.
#define PS3D_Z_SIZES(X) X(1, 32) X(1, 48) X(1, 64) X(1, 96) X(1, 128) X(1, 192) X(1, 256) X(1, 384) \
    X(2, 512) X(2, 768) X(3, 1024)

#define ZLAUNCH(KERNEL, Z_, N_, ARGS, LAUNCH) \
    hipLaunchKernelGGL((KERNEL<Z_, N_>), dim3(LAUNCH.blocks, zbatch(ARGS)), dim3(LAUNCH.threads), LAUNCH.lds, stream, ARGS)

#define ZDISPATCH_FFT(KERNEL, ARGS, LAUNCH)                                              \
    do {                                                                                 \
        bool done_ = false;                                                              \
        _Pragma("clang diagnostic push")                                                 \
        if (false) {}                                                                    \
        PS3D_Z_SIZES(ZTRY_##KERNEL)                                                      \
        _Pragma("clang diagnostic pop")                                                  \
        if (!done_) {                                                                    \
            switch (LAUNCH.zpt) {                                                        \
                case 1: ZLAUNCH(KERNEL, 1, 0, ARGS, LAUNCH); break;                      \
                case 2: ZLAUNCH(KERNEL, 2, 0, ARGS, LAUNCH); break;                      \
                case 3: ZLAUNCH(KERNEL, 3, 0, ARGS, LAUNCH); break;                      \
                default: return hipErrorInvalidValue;                                    \
            }                                                                            \
        }                                                                                \
    } while (0)

#define ZTRY_k_zinv(Z_, N_) else if (zl.zpt == Z_ && zl.nz == N_) { ZLAUNCH(k_zinv, Z_, N_, args, zl); done_ = true; }
#define ZTRY_k_ztend(Z_, N_) else if (zl.zpt == Z_ && zl.nz == N_) { ZLAUNCH(k_ztend, Z_, N_, args, zl); done_ = true; }
#define ZTRY_k_zop(Z_, N_) else if (zl.zpt == Z_ && zl.nz == N_) { ZLAUNCH(k_zop, Z_, N_, args, zl); done_ = true; }

hipError_t configure_zpass() {
    const int lds = 160 * 1024;
    hipError_t e = hipSuccess;
#define SETA(K) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)(K), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
#define ZCFG3(Z_, N_) SETA((k_zinv<Z_, N_>)) SETA((k_ztend<Z_, N_>)) SETA((k_zop<Z_, N_>))
    PS3D_Z_SIZES(ZCFG3)
    ZCFG3(1, 0) ZCFG3(2, 0) ZCFG3(3, 0)
    SETA(k_zrk<1>) SETA(k_zrk<2>) SETA(k_zrk<3>)
    SETA(k_mean)
#undef ZCFG3
#undef SETA
    return e;
}

hipError_t launch_zinv(const ZLaunch &zl, const ZInvArgs &args, hipStream_t stream) {
    if (zl.wave) return launch_zinv_w(zl.nz, zl.wave, zl.nquads, args, stream);
    ZDISPATCH_FFT(k_zinv, args, zl);
    return hipGetLastError();
}
hipError_t launch_ztend(const ZLaunch &zl, const ZTendArgs &args, hipStream_t stream) {
    if (zl.wave) return launch_ztend_w(zl.nz, zl.wave, zl.nquads, args, stream);
    ZDISPATCH_FFT(k_ztend, args, zl);
    return hipGetLastError();
}
hipError_t launch_zrk(const ZLaunch &zl, const ZTendArgs &args, hipStream_t stream) {
    switch (zl.zpt) {
        case 1: hipLaunchKernelGGL(k_zrk<1>, dim3(zl.blocks, zbatch(args)), dim3(zl.threads), zl.lds, stream, args); break;
        case 2: hipLaunchKernelGGL(k_zrk<2>, dim3(zl.blocks, zbatch(args)), dim3(zl.threads), zl.lds, stream, args); break;
        case 3: hipLaunchKernelGGL(k_zrk<3>, dim3(zl.blocks, zbatch(args)), dim3(zl.threads), zl.lds, stream, args); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_zop(const ZLaunch &zl, const ZOpArgs &args, hipStream_t stream) {
    if (zl.wave) return launch_zop_w(zl.nz, zl.nquads, args, stream);
    ZDISPATCH_FFT(k_zop, args, zl);
    return hipGetLastError();
}
hipError_t launch_mean(const MeanArgs &args, int nz, hipStream_t stream) {
    size_t lds = 2 * 2 * (size_t)nz * sizeof(double2) + 2 * 256 * sizeof(double);
    hipLaunchKernelGGL(k_mean, dim3(1, args.bt.nb > 1 ? args.bt.nb : 1), dim3(256), lds, stream, args);
    return hipGetLastError();
}

}  // namespace ps3d
