// linepass.hip -- horizontal (x or y) line passes of the MI355X pseudo-spectral step.
//
// A block owns a tile of n line elements x ZC z levels at one fixed index of the
// other horizontal direction.  In the column layout the tile is read as n runs of
// ZC consecutive doubles (128 B for ZC = 16).  Levels 2p and 2p+1 share one
// complex line (re, im), so one complex FFT of length n gives the real transforms
// of two levels; the Hermitian real form of stafft (src/fft/stafft.f90:55-58,
// 196-403) keeps the same packing: slot j of line p holds (level 2p, level 2p+1).
#include "kernels.hpp"
#include "lfft.hpp"

namespace ps3d {

constexpr int kEpt = 2;        // complex elements per thread held in registers
constexpr int kLineThreads = 1024;

struct Tile {
    int n, ZC, nl, LS, zce, z0, other;
    long base;
    const double2 *twl;   // LDS copy of the twiddles (compile-time lengths)
    bool skew;            // LDS lines in the skewed slot layout (lslot, compile-time lengths)
    // LDS slot of element x of a line
    __device__ int sl(int x) const { return skew ? x + (x >> 4) : x; }
};

// LDS line stride (complex elements) of the skewed layout: n slots, n/16 pads, one
// more so that consecutive lines start four banks apart
__host__ __device__ constexpr int line_ls(int n) { return n + (n >> 4) + 1; }

// z depth of a tile for a line length: n*ZC <= 4096 doubles, 2 <= ZC <= 16, even
__host__ __device__ constexpr int zc_for(int n) {
    return ((4096 / n > 16 ? 16 : (4096 / n < 2 ? 2 : 4096 / n)) & ~1);
}
// z depth of the register-holding modes at T threads per block: the 1024-thread
// build takes the standard tile, T = 512 half of it and T = 256 a quarter (at least 2)
__host__ __device__ constexpr int zc_reg(int n, int T) {
    return T >= kLineThreads ? zc_for(n) : ((zc_for(n) * T / kLineThreads) < 2 ? 2 : ((zc_for(n) * T / kLineThreads) & ~1));
}
// the registers of T threads (kEpt complex values each) must hold the tile
__host__ __device__ constexpr bool reg_fits(int n, int T) { return n * zc_reg(n, T) / 2 <= kEpt * T; }

// Register-holding modes run 1024 threads per tile and use the radix-8 plans to
// stay within 128 VGPRs; transform-only modes use the radix-16 plans unless their
// schedule asks for the radix-8 ones (R8).
template <int N, bool REG = false, bool R8 = false>
__device__ inline double2 *lfft(const Tile &t, const LineArgs &a, double2 *A, double2 *B, int dir) {
    if constexpr (N == 0)
        return fft_lds(A, B, t.nl, t.LS, a.P, a.tw, dir);
    else if constexpr ((REG || R8) && CPlan8<N>::ok)
        return cfft_lds<N, CPlan8<N>, false, true>(A, B, t.nl, t.LS, t.twl, dir);
    else
        return cfft_lds<N, CPlan<N>, false, true>(A, B, t.nl, t.LS, t.twl, dir);
}

// Tile geometry fixed at compile time for a built line length: the index arithmetic
// of the loads, stores and transforms then divides by constants.
__device__ inline void fix_tile(Tile &t, int n, int zc) {
    t.n = n;
    t.ZC = zc;
    t.nl = zc / 2;
    t.LS = line_ls(n);
    t.skew = true;
    const int rem = t.zce;
    t.zce = rem < zc ? rem : zc;
}

// Tile handled by this block.  Workgroups go to the 8 XCDs round-robin (XCD =
// blockIdx.x mod 8); the remap hands each XCD a contiguous range of tiles, so the z
// tiles of one line set -- which share cache lines (a tile of ZC < 16 levels covers
// part of a 128-byte line) and pages -- meet in one L2.
__device__ inline int tile_index() {
    const int B = gridDim.x, i = blockIdx.x;
    const int q = B >> 3, r = B & 7, x = i & 7, k = i >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ inline Tile make_tile(const LineArgs &a) {
    Tile t;
    t.n = a.n;
    t.ZC = a.ZC;
    t.nl = a.ZC / 2;
    t.LS = a.n + 1;
    t.skew = false;
    const int ti = tile_index();
    const int zt = ti % a.nzt;
    t.other = ti / a.nzt;
    t.z0 = zt * a.ZC;
    const int rem = a.g.nz + 1 - t.z0;
    t.zce = rem < a.ZC ? rem : a.ZC;
    t.base = (long)blockIdx.y * a.bt.vol + (long)t.other * a.ostride + t.z0;
    t.twl = nullptr;
    return t;
}

// Global accesses move both z levels of a line slot at once (16 bytes): every
// tile start, column stride and volume offset is even, so the pair is aligned.
// A tile whose last level is odd reads one level past it (inside the column's
// padding) and zeroes it.
__device__ inline double2 gpair(const Tile &t, const LineArgs &a, const double *__restrict__ src, int j, int p) {
    const int z = 2 * p;
    if (z >= t.zce) return make_double2(0.0, 0.0);
    const double2 v = *reinterpret_cast<const double2 *>(src + t.base + (long)j * a.estride + z);
    return z + 1 < t.zce ? v : make_double2(v.x, 0.0);
}
__device__ inline void spair(const Tile &t, const LineArgs &a, double *__restrict__ dst, int j, int p, double2 v) {
    const int z = 2 * p;
    double *o = dst + t.base + (long)j * a.estride + z;
    if (z + 1 < t.zce)
        *reinterpret_cast<double2 *>(o) = v;
    else if (z < t.zce)
        o[0] = v.x;
}

__device__ inline void tile_load(const Tile &t, const LineArgs &a, const double *__restrict__ src, double2 *buf) {
    const int tot = t.n * t.nl;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int p = e % t.nl, j = e / t.nl;
        buf[(size_t)p * t.LS + t.sl(j)] = gpair(t, a, src, j, p);
    }
}

// Deep tiles (16 levels: 128-byte runs) store level pairs; shallower tiles store
// single levels, so that a wave's stores span fewer line slots.
__device__ inline void tile_store(const Tile &t, const LineArgs &a, double *__restrict__ dst, const double2 *buf,
                                  double scale) {
    if (t.ZC >= 16) {
        const int tot = t.n * t.nl;
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int p = e % t.nl, j = e / t.nl;
            const double2 v = buf[(size_t)p * t.LS + t.sl(j)];
            spair(t, a, dst, j, p, make_double2(v.x * scale, v.y * scale));
        }
        return;
    }
    const double *d = reinterpret_cast<const double *>(buf);
    const int tot = t.n * t.ZC;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int zz = e % t.ZC, j = e / t.ZC;
        if (zz < t.zce) dst[t.base + (long)j * a.estride + zz] = d[((size_t)(zz >> 1) * t.LS + t.sl(j)) * 2 + (zz & 1)] * scale;
    }
}

// Hermitian pair -> complex spectrum, then inverse FFT.  Returns result buffer,
// values to be scaled by 1/n.  FOLD (a compile-time plan only): the pair is formed
// inside the first FFT pass instead of in a pass of its own.
template <int N, bool REG = false, bool FOLD = (N >= 128), bool R8 = false>
__device__ inline double2 *tile_inverse(const Tile &t, const LineArgs &a, double2 *A, double2 *B) {
    const int n = t.n, h = n / 2;
    const double sn = sqrt((double)n);
    // (measured on single grids: the fold pays from n = 128 up, which is the default;
    // the transform-only passes take it as part of their schedule)
    if constexpr (FOLD && N > 0) {
        if constexpr ((REG || R8) && CPlan8<N>::ok)
            return cfft_lds<N, CPlan8<N>, true, true>(A, B, t.nl, t.LS, t.twl, +1, sn);
        else
            return cfft_lds<N, CPlan<N>, true, true>(A, B, t.nl, t.LS, t.twl, +1, sn);
    }
    const int tot = t.nl * (h + 1);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int line = e / (h + 1), k = e % (h + 1);
        double2 *L = A + (size_t)line * t.LS;
        const int sk_ = t.sl(k), snk = t.sl(n - k);
        if (k == 0 || k == h) {
            const double2 v = L[sk_];
            L[sk_] = make_double2(sn * v.x, sn * v.y);
        } else {
            const double2 ck = L[sk_], sk = L[snk];   // (a(k), b(k)), (a(n-k), b(n-k))
            L[sk_] = make_double2(sn * (ck.x - sk.y), sn * (sk.x + ck.y));
            L[snk] = make_double2(sn * (ck.x + sk.y), sn * (-sk.x + ck.y));
        }
    }
    __syncthreads();
    return lfft<N, REG, R8>(t, a, A, B, +1);
}

// Forward FFT then split into the Hermitian pair, in the returned buffer, scaled.
template <int N, bool REG = false, bool R8 = false>
__device__ inline double2 *tile_forward(const Tile &t, const LineArgs &a, double2 *A, double2 *B) {
    double2 *R = lfft<N, REG, R8>(t, a, A, B, -1);
    const int n = t.n, h = n / 2;
    const double sc = 1.0 / sqrt((double)n);
    const int tot = t.nl * (h + 1);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int line = e / (h + 1), k = e % (h + 1);
        double2 *L = R + (size_t)line * t.LS;
        const int nk = (n - k) % n;
        const double2 zk = L[t.sl(k)], znk = L[t.sl(nk)];
        const double Ar = 0.5 * (zk.x + znk.x), Ai = 0.5 * (zk.y - znk.y);
        const double Br = 0.5 * (zk.y + znk.y), Bi = -0.5 * (zk.x - znk.x);
        L[t.sl(k)] = make_double2(Ar * sc, Br * sc);
        if (k != 0 && k != h) L[t.sl(nk)] = make_double2(Ai * sc, Bi * sc);
    }
    __syncthreads();
    return R;
}

// Store of a forward-transformed tile (complex FFT output R, unscaled) as the
// Hermitian real spectra of its two z levels per line, split on the fly: slot
// k <= n/2 takes (Re A_k, Re B_k), slot n-k (Im A_k, Im B_k), scaled by 1/sqrt(n)
// (the split of tile_forward without its LDS pass).  Each thread writes both
// levels of one slot (16 contiguous bytes).
template <int N>
__device__ inline void tile_store_fwd(const Tile &t, const LineArgs &a, double *__restrict__ dst, const double2 *R) {
    constexpr int H = N / 2;
    const double sc = 0.5 / sqrt((double)N);
    const int tot = t.nl * N;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int p = e % t.nl, j = e / t.nl;
        const double2 *L = R + (size_t)p * t.LS;
        const double2 zj = L[t.sl(j)], zm = L[t.sl(j == 0 ? 0 : N - j)];
        double2 v;
        if (j <= H)
            v = make_double2(sc * (zj.x + zm.x), sc * (zj.y + zm.y));
        else
            v = make_double2(sc * (zm.y - zj.y), sc * (zj.x - zm.x));
        spair(t, a, dst, j, p, v);
    }
}

// Spectral derivative along the line (diffx / diffy, sta3dfft.f90:304-377):
// dst[j] = si(j) rk(j) src[n-j], zero at j = 0 and j = n/2.
__device__ inline void tile_diff(const Tile &t, const LineArgs &a, const double2 *src, double2 *dst) {
    const int n = t.n, h = n / 2;
    const int tot = t.nl * n;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int line = e / n, j = e % n;
        double f = (j >= h + 1) ? a.rk[j] : -a.rk[j];
        if (j == 0 || j == h) f = 0.0;
        const double2 v = src[(size_t)line * t.LS + t.sl((n - j) % n)];
        dst[(size_t)line * t.LS + t.sl(j)] = make_double2(f * v.x, f * v.y);
    }
    __syncthreads();
}

// registers <-> tile
__device__ inline void to_regs(const Tile &t, const double2 *buf, double2 (&r)[kEpt], double scale) {
    const int tot = t.nl * t.n;
#pragma unroll
    for (int k = 0; k < kEpt; ++k) {
        const int e = threadIdx.x + k * blockDim.x;
        if (e < tot) {
            const int line = e / t.n, j = e % t.n;
            const double2 v = buf[(size_t)line * t.LS + t.sl(j)];
            r[k] = make_double2(v.x * scale, v.y * scale);
        } else {
            r[k] = make_double2(0.0, 0.0);
        }
    }
}
__device__ inline void from_regs(const Tile &t, double2 *buf, const double2 (&r)[kEpt]) {
    const int tot = t.nl * t.n;
#pragma unroll
    for (int k = 0; k < kEpt; ++k) {
        const int e = threadIdx.x + k * blockDim.x;
        if (e < tot) {
            const int line = e / t.n, j = e % t.n;
            buf[(size_t)line * t.LS + t.sl(j)] = r[k];
        }
    }
}
// z level (absolute) of component c (0 = re, 1 = im) of register element k
__device__ inline int reg_z(const Tile &t, int k, int c) {
    const int e = threadIdx.x + k * blockDim.x;
    const int line = e / t.n;
    return t.z0 + 2 * line + c;
}
__device__ inline bool reg_valid(const Tile &t, int k, int c, int nz) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e >= t.nl * t.n) return false;
    const int zz = 2 * (e / t.n) + c;
    return zz < t.zce && t.z0 + zz <= nz;
}
__device__ inline int reg_j(const Tile &t, int k) { return (threadIdx.x + k * blockDim.x) % t.n; }

// Load one field, transform it to physical space along the line (unless the input
// is already physical), and keep it in registers.
template <int N>
__device__ inline void field_phys(const Tile &t, const LineArgs &a, const double *src, double2 *A, double2 *B,
                                  double2 (&r)[kEpt], bool diff) {
    tile_load(t, a, src, A);
    __syncthreads();
    double2 *S = A, *O = B;
    if (diff) {
        if (a.phys_in) {
            S = tile_forward<N, true>(t, a, A, B);
            O = (S == A) ? B : A;
        }
        tile_diff(t, a, S, O);
        double2 *tmp = S;
        S = O;
        O = tmp;
    } else if (a.phys_in) {
        to_regs(t, A, r, 1.0);
        __syncthreads();
        return;
    }
    double2 *R = tile_inverse<N, true>(t, a, S, O);
    to_regs(t, R, r, 1.0 / (double)t.n);
    __syncthreads();
}

// registers -> forward transform -> global
template <int N>
__device__ inline void fwd_store(const Tile &t, const LineArgs &a, double2 *A, double2 *B, const double2 (&r)[kEpt],
                                 double *dst) {
    from_regs(t, A, r);
    __syncthreads();
    if constexpr (N > 0) {
        double2 *R = lfft<N, true>(t, a, A, B, -1);
        tile_store_fwd<N>(t, a, dst, R);
    } else {
        double2 *R = tile_forward<N, true>(t, a, A, B);
        tile_store(t, a, dst, R, 1.0);
    }
    __syncthreads();
}

// closure input: channel ch of the bf16 tensor [p][16], p = (ix*ny + iy)*(nz+1) + iz,
// from a physical field in registers (lines along x)
__device__ inline void nn_export(const Tile &t, const LineArgs &a, const double2 (&r)[kEpt], int ch) {
    if (!a.nn_in) return;
    const float s = a.nn_scale[ch];
#pragma unroll
    for (int k = 0; k < kEpt; ++k)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (!reg_valid(t, k, c, a.g.nz)) continue;
            const long p = ((long)reg_j(t, k) * a.g.ny + t.other) * (a.g.nz + 1) + reg_z(t, k, c);
            const float v = (float)(c ? r[k].y : r[k].x) * s;
            uint32_t u = __builtin_bit_cast(uint32_t, v);
            u += 0x7fffu + ((u >> 16) & 1u);
            a.nn_in[p * 16 + ch] = (uint16_t)(u >> 16);
        }
}

// Register prefetch of a tile for the register-holding kernels (T threads): the
// global loads of the next field are issued before the transform of the current one.
// (8 bytes per thread: a wave's load then spans half as many line slots as with
// the 16-byte pairs, which the x passes -- slots a whole y-z plane apart -- prefer)
template <int N, int T = kLineThreads>
struct Pf {
    static constexpr int ZC = N > 0 ? zc_reg(N, T) : 2;
    static constexpr int PER = N > 0 ? (N * ZC + T - 1) / T : 1;
    double pre[PER];
};
template <int N, int T>
__device__ inline void pf_load(Pf<N, T> &p, const Tile &t, const LineArgs &a, const double *__restrict__ src) {
    constexpr int ZC = Pf<N, T>::ZC;
#pragma unroll
    for (int k = 0; k < Pf<N, T>::PER; ++k) {
        const int e = threadIdx.x + k * T, zz = e % ZC, j = e / ZC;
        p.pre[k] = (j < t.n && zz < t.zce) ? src[t.base + (long)j * a.estride + zz] : 0.0;
    }
}
template <int N, int T>
__device__ inline void pf_commit(const Pf<N, T> &p, const Tile &t, double2 *A) {
    constexpr int ZC = Pf<N, T>::ZC;
    double *d = reinterpret_cast<double *>(A);
#pragma unroll
    for (int k = 0; k < Pf<N, T>::PER; ++k) {
        const int e = threadIdx.x + k * T, zz = e % ZC, j = e / ZC;
        if (j < t.n) d[((size_t)(zz >> 1) * t.LS + t.sl(j)) * 2 + (zz & 1)] = p.pre[k];
    }
}
// field_phys on the prefetched tile (spectral input; diff: x/y derivative first),
// starting the loads of `next` (nullptr: none).  N = 0 falls back to field_phys.
template <int N, int T>
__device__ inline void phys_pf(Pf<N, T> &pf, const Tile &t, const LineArgs &a, double2 *A, double2 *B,
                               double2 (&r)[kEpt], bool diff, const double *cur, const double *next) {
    if constexpr (N == 0) {
        field_phys<N>(t, a, cur, A, B, r, diff);
    } else {
        pf_commit(pf, t, A);
        __syncthreads();
        if (next) pf_load(pf, t, a, next);
        double2 *S = A, *O = B;
        if (diff) {
            tile_diff(t, a, A, B);
            S = B;
            O = A;
        }
        double2 *R = tile_inverse<N, true>(t, a, S, O);
        to_regs(t, R, r, 1.0 / (double)t.n);
        __syncthreads();
    }
}

// block-wide reductions (deterministic tree in LDS)
__device__ inline double block_sum(double v, double *scratch) {
    const int tid = threadIdx.x;
    scratch[tid] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) scratch[tid] += scratch[tid + w];
        __syncthreads();
    }
    const double r = scratch[0];
    __syncthreads();
    return r;
}
__device__ inline double block_max(double v, double *scratch) {
    const int tid = threadIdx.x;
    scratch[tid] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) scratch[tid] = fmax(scratch[tid], scratch[tid + w]);
        __syncthreads();
    }
    const double r = scratch[0];
    __syncthreads();
    return r;
}

// Cyclic Jacobi eigenvalues of a symmetric 3x3 matrix (Rutishauser 1966, as in
// src/utils/jacobi.f90:20-305): returns max |lambda|.  Bounded number of sweeps.
__device__ inline void jac_rot(double (&A)[3][3], double (&D)[3], double (&Z)[3], int i, int j) {
    const double eps = 2.220446049250313e-16;
    const double aij = A[i][j];
    double g = 100.0 * fabs(aij);
    double h = D[j] - D[i];
    double t;
    if (fabs(h) + g == fabs(h)) {
        t = aij / (h + copysign(eps, h));
    } else {
        const double theta = 0.5 * h / (aij + copysign(eps, aij));
        t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
        if (theta < 0.0) t = -t;
    }
    const double c = 1.0 / sqrt(1.0 + t * t);
    const double s = t * c;
    const double tau = s / (1.0 + c);
    h = t * aij;
    Z[i] -= h;
    Z[j] += h;
    D[i] -= h;
    D[j] += h;
    A[i][j] = 0.0;
    for (int k = 0; k < i; ++k) {
        g = A[k][i]; h = A[k][j];
        A[k][i] = g - s * (h + g * tau);
        A[k][j] = h + s * (g - h * tau);
    }
    for (int k = i + 1; k < j; ++k) {
        g = A[i][k]; h = A[k][j];
        A[i][k] = g - s * (h + g * tau);
        A[k][j] = h + s * (g - h * tau);
    }
    for (int k = j + 1; k < 3; ++k) {
        g = A[i][k]; h = A[j][k];
        A[i][k] = g - s * (h + g * tau);
        A[j][k] = h + s * (g - h * tau);
    }
}
__device__ inline double jacobi_absmax(double s11, double s12, double s13, double s22, double s23, double s33) {
    double A[3][3] = {{s11, s12, s13}, {0.0, s22, s23}, {0.0, 0.0, s33}};
    double D[3] = {s11, s22, s33}, B[3] = {s11, s22, s33}, Z[3] = {0.0, 0.0, 0.0};
    double sm = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    for (int sweep = 0; sweep < 64 && sm > 1.0e-15; ++sweep) {
        jac_rot(A, D, Z, 0, 1);
        jac_rot(A, D, Z, 0, 2);
        jac_rot(A, D, Z, 1, 2);
        for (int k = 0; k < 3; ++k) {
            B[k] += Z[k];
            D[k] = B[k];
            Z[k] = 0.0;
        }
        sm = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    }
    return fmax(fmax(fabs(D[0]), fabs(D[1])), fabs(D[2]));
}

// Largest |eigenvalue| of a symmetric 3x3 matrix from the trigonometric solution of
// its characteristic cubic (Smith, CACM 4(4), 1961).  Within a few 1e-8 of the exact
// value relative to the matrix norm (acos loses half the digits next to a double
// root), so it only selects the points whose Jacobi value can be a block maximum.
__device__ inline double eig_absmax_est(double s11, double s12, double s13, double s22, double s23, double s33) {
    const double p1 = s12 * s12 + s13 * s13 + s23 * s23;
    const double q = (s11 + s22 + s33) * (1.0 / 3.0);
    const double d1 = s11 - q, d2 = s22 - q, d3 = s33 - q;
    const double p2 = d1 * d1 + d2 * d2 + d3 * d3 + 2.0 * p1;
    if (!(p2 > 0.0)) return fabs(q);
    const double p = sqrt(p2 * (1.0 / 6.0));
    const double det = d1 * (d2 * d3 - s23 * s23) - s12 * (s12 * d3 - s23 * s13) + s13 * (s12 * s23 - d2 * s13);
    const double ip = 1.0 / p;
    const double r = fmin(1.0, fmax(-1.0, 0.5 * det * ip * ip * ip));
    const double c = cos(acos(r) * (1.0 / 3.0));            // angle in [0, pi/3]
    const double sn = sqrt(fmax(0.0, 1.0 - c * c));
    const double l1 = q + 2.0 * p * c;                        // largest eigenvalue
    const double l3 = q - p * (c + 1.7320508075688772 * sn);  // smallest eigenvalue
    return fmax(fabs(l1), fabs(l3));
}

// Transform-only modes (forward, inverse, derivative) of nf fields; the global
// loads of field f+1 are issued before the FFT of field f.  SCHED: the FFT
// schedule, bit LS_R8 = radix-8-first plan instead of radix-16-first, bit LS_FOLD
// = Hermitian pair of the inverse formed in the first radix pass, bit LS_T512 = 512
// threads per tile instead of 256 (twice the waves on the same LDS).
template <int N, int SCHED>
__global__ void __launch_bounds__((SCHED & LS_T512) ? 512 : 256) k_line_simple(LineArgs a) {
    constexpr bool R8 = (SCHED & LS_R8) != 0, FOLD = (SCHED & LS_FOLD) != 0;
    constexpr int T = (SCHED & LS_T512) ? 512 : 256;
    extern __shared__ double2 lsmem[];
    if (gated_off(a.g.gate)) return;
    Tile t = make_tile(a);
    if constexpr (N > 0) fix_tile(t, N, zc_for(N));
    const size_t tb = (size_t)t.nl * t.LS;
    double2 *A = lsmem, *B = lsmem + tb;
    if constexpr (N > 0) {
        double2 *twl = B + tb;
        for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = a.tw[i];
        t.twl = twl;
    }
    if (a.mode == LM_DIFF || N == 0) {
        for (int f = 0; f < a.nf; ++f) {
            tile_load(t, a, a.in[f], A);
            __syncthreads();
            if (a.mode == LM_FWD) {
                double2 *R = tile_forward<N, false, R8>(t, a, A, B);
                tile_store(t, a, a.out[f], R, 1.0);
            } else if (a.mode == LM_INV) {
                double2 *R = tile_inverse<N, false, FOLD, R8>(t, a, A, B);
                tile_store(t, a, a.out[f], R, 1.0 / (double)t.n);
            } else {
                tile_diff(t, a, A, B);
                tile_store(t, a, a.out[f], B, 1.0);
            }
            __syncthreads();
        }
        return;
    }
    if constexpr (N > 0) {
        // deep tiles load level pairs (16 bytes), shallower ones single levels
        constexpr int ZC = zc_for(N), NL = ZC / 2;
        constexpr bool PAIR = ZC >= 16;
        constexpr int PER = PAIR ? (N * NL + T - 1) / T : (N * ZC + T - 1) / T;
        double2 pre[PAIR ? PER : 1];
        double pre1[PAIR ? 1 : PER];
        auto gload = [&](const double *__restrict__ src) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int e = threadIdx.x + k * T;
                if constexpr (PAIR) {
                    const int q = e % NL, j = e / NL;
                    pre[k] = j < N ? gpair(t, a, src, j, q) : make_double2(0.0, 0.0);
                } else {
                    const int zz = e % ZC, j = e / ZC;
                    pre1[k] = (j < N && zz < t.zce) ? src[t.base + (long)j * a.estride + zz] : 0.0;
                }
            }
        };
        gload(a.in[0]);
        for (int f = 0; f < a.nf; ++f) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int e = threadIdx.x + k * T;
                if constexpr (PAIR) {
                    const int q = e % NL, j = e / NL;
                    if (j < N) A[(size_t)q * t.LS + t.sl(j)] = pre[k];
                } else {
                    const int zz = e % ZC, j = e / ZC;
                    if (j < N) reinterpret_cast<double *>(A)[((size_t)(zz >> 1) * t.LS + t.sl(j)) * 2 + (zz & 1)] = pre1[k];
                }
            }
            __syncthreads();
            if (f + 1 < a.nf) gload(a.in[f + 1]);
            if (a.mode == LM_FWD) {
                double2 *R = lfft<N, false, R8>(t, a, A, B, -1);
                tile_store_fwd<N>(t, a, a.out[f], R);
            } else {
                double2 *R = tile_inverse<N, false, FOLD, R8>(t, a, A, B);
                tile_store(t, a, a.out[f], R, 1.0 / (double)t.n);
            }
            __syncthreads();
        }
    }
}

// Modes that keep physical values in registers (products, adapt diagnostics); one
// instantiation per mode M so that each gets its own register allocation.
// T = 1024: one block per CU; T = 512 (half-depth tiles, at most 128 VGPRs): two
// blocks per CU, so that one block's barriers overlap the other's loads; T = 256:
// four blocks per CU on quarter-depth tiles.
template <int N, int M, int T>
__global__ void __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(T >= kLineThreads ? 1 : 4)))
k_line_reg(LineArgs a) {
    extern __shared__ double2 lsmem[];
    if (gated_off(a.g.gate)) return;
    Tile t = make_tile(a);
    if constexpr (N > 0) fix_tile(t, N, zc_reg(N, T));
    const size_t tb = (size_t)t.nl * t.LS;
    // the products pass transforms field pairs: A and B hold two tiles each
    constexpr bool PAIRS = ((M == LM_XMAIN || M == LM_XADAPT) && N > 0);
    double2 *A = lsmem, *B = lsmem + (PAIRS ? 2 * tb : tb);
    if constexpr (N > 0) {
        double2 *twl = B + (PAIRS ? 2 * tb : tb);
        for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = a.tw[i];
        t.twl = twl;
    }
    double *scratch = reinterpret_cast<double *>(lsmem);
    const int nz = a.g.nz;

    if constexpr (M == LM_INV_Y1) {
        // u, v, w with their y derivatives; xi, eta, zeta; optional midpoints.
        // Parseval in x (this tile is one x slot): sum_x f^2 = sum_k w(k) f_k^2,
        // w = 1 at k = 0, nx/2 and 2 elsewhere.
        const int kx = t.other;
        const double wk = (kx == 0 || kx == a.g.nx / 2) ? 1.0 : 2.0;
        double2 U[3][kEpt];
        double suu = 0.0, sww = 0.0, suw = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
        // every field is loaded through the register prefetch: the loads of the
        // next source are in flight during the current transform
        Pf<N, T> pf;
        auto take = [&](const double *cur, const double *next) {
            if constexpr (N > 0) {
                pf_commit(pf, t, A);
                __syncthreads();
                if (next) pf_load(pf, t, a, next);
            } else {
                tile_load(t, a, cur, A);
                __syncthreads();
            }
        };
        if constexpr (N > 0) pf_load(pf, t, a, a.in[0]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            take(a.in[c], a.in[c]);
            tile_diff(t, a, A, B);            // y derivative in B
            double2 *R = tile_inverse<N, true>(t, a, B, A);
            tile_store(t, a, a.out[9 + c], R, 1.0 / (double)t.n);
            __syncthreads();
            take(a.in[c], a.in[c + 1]);       // in[3] (xi) follows w
            R = tile_inverse<N, true>(t, a, A, B);
            tile_store(t, a, a.out[c], R, 1.0 / (double)t.n);
            to_regs(t, R, U[c], 1.0 / (double)t.n);
            __syncthreads();
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            take(a.in[3 + c], (c < 2 || a.in[6]) ? a.in[4 + c] : nullptr);   // in[6]: midpoints
            double2 *R = tile_inverse<N, true>(t, a, A, B);
            tile_store(t, a, a.out[3 + c], R, 1.0 / (double)t.n);
            double2 W[kEpt];
            to_regs(t, R, W, 1.0 / (double)t.n);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kEpt; ++k)
                for (int h = 0; h < 2; ++h) {
                    if (!reg_valid(t, k, h, nz)) continue;
                    const int z = reg_z(t, k, h);
                    const double wz = (z == 0 || z == nz) ? 0.5 : 1.0;
                    const double w = h ? W[k].y : W[k].x;
                    const double u = h ? U[c][k].y : U[c][k].x;
                    suu += wz * wk * u * u;
                    sww += wz * wk * w * w;
                    suw += wz * wk * u * w;
                    if (kx == 0) {
                        if (c == 0) m0 += wz * w;
                        if (c == 1) m1 += wz * w;
                        if (c == 2) m2 += wz * w;
                    }
                }
        }
        if (a.in[6]) {
            for (int c = 0; c < 3; ++c) {
                take(a.in[6 + c], c < 2 ? a.in[7 + c] : nullptr);
                double2 *R = tile_inverse<N, true>(t, a, A, B);
                tile_store(t, a, a.out[6 + c], R, 1.0 / (double)t.n);
                __syncthreads();
            }
        }
        suu = block_sum(suu, scratch);
        sww = block_sum(sww, scratch);
        suw = block_sum(suw, scratch);
        m0 = block_sum(m0, scratch);
        m1 = block_sum(m1, scratch);
        m2 = block_sum(m2, scratch);
        if (threadIdx.x == 0) {
            double *p = a.part + (size_t)blockIdx.y * a.bt.part + (size_t)tile_index() * RY_N;
            p[RY_UU] = suu;
            p[RY_WW] = sww;
            p[RY_UW] = suw;
            p[RY_MX] = m0 * a.sqrt_nx;
            p[RY_MY] = m1 * a.sqrt_nx;
            p[RY_MZ] = m2 * a.sqrt_nx;
        }
        return;
    }

    if constexpr (M == LM_XMAIN) {
        // u x omega in flux form (inversion.f90:316-352) with omega + f_cor.  Fields
        // go through the transforms in pairs (16 lines per FFT call: twice the
        // butterflies per pass, half the barriers); the global loads of the next
        // pair are in flight during the current pair's transform.
        double2 u[kEpt], v[kEpt], w[kEpt], r[kEpt], q[kEpt], p[kEpt], o[kEpt], o2[kEpt];
        if constexpr (N > 0) {
            Tile t2 = t;
            t2.nl = 2 * t.nl;
            Pf<N, T> p0, p1;
            auto phys2 = [&](int f, double2(&r0)[kEpt], double2(&r1)[kEpt]) {
                pf_commit(p0, t, A);
                pf_commit(p1, t, A + tb);
                __syncthreads();
                if (f + 2 < 6) {
                    pf_load(p0, t, a, a.in[f + 2]);
                    pf_load(p1, t, a, a.in[f + 3]);
                }
                if (a.phys_in) {
                    to_regs(t, A, r0, 1.0);
                    to_regs(t, A + tb, r1, 1.0);
                } else {
                    double2 *R = tile_inverse<N, true>(t2, a, A, B);
                    to_regs(t, R, r0, 1.0 / (double)t.n);
                    to_regs(t, R + tb, r1, 1.0 / (double)t.n);
                }
                __syncthreads();
            };
            pf_load(p0, t, a, a.in[0]);
            pf_load(p1, t, a, a.in[1]);
            phys2(0, u, v);
            phys2(2, w, o);     // o = xi
            nn_export(t, a, u, 0);
            nn_export(t, a, v, 1);
            nn_export(t, a, w, 2);
            nn_export(t, a, o, 3);
#pragma unroll
            for (int k = 0; k < kEpt; ++k) {
                const double xr = o[k].x + a.fcor[0], xi = o[k].y + a.fcor[0];
                r[k] = make_double2(-v[k].x * xr, -v[k].y * xi);
                q[k] = make_double2(w[k].x * xr, w[k].y * xi);
            }
            phys2(4, o, o2);    // o = eta, o2 = zeta
        } else {
            field_phys<N>(t, a, a.in[0], A, B, u, false);
            field_phys<N>(t, a, a.in[1], A, B, v, false);
            field_phys<N>(t, a, a.in[2], A, B, w, false);
            field_phys<N>(t, a, a.in[3], A, B, o, false);
            nn_export(t, a, u, 0);
            nn_export(t, a, v, 1);
            nn_export(t, a, w, 2);
            nn_export(t, a, o, 3);
#pragma unroll
            for (int k = 0; k < kEpt; ++k) {
                const double xr = o[k].x + a.fcor[0], xi = o[k].y + a.fcor[0];
                r[k] = make_double2(-v[k].x * xr, -v[k].y * xi);
                q[k] = make_double2(w[k].x * xr, w[k].y * xi);
            }
            field_phys<N>(t, a, a.in[4], A, B, o, false);
            field_phys<N>(t, a, a.in[5], A, B, o2, false);
        }
        nn_export(t, a, o, 4);
        nn_export(t, a, o2, 5);
        if (a.part) {
            // adapt stage: signed maxima of u, v, w (advance.f90:319-321), one partial
            // per block, so that the x adapt pass need not load the velocity again
            // (the tile buffers are free here: the last transform ended on a barrier)
            double um = -INFINITY, vm = -INFINITY, wm = -INFINITY;
#pragma unroll
            for (int k = 0; k < kEpt; ++k)
                for (int h = 0; h < 2; ++h)
                    if (reg_valid(t, k, h, nz)) {
                        um = fmax(um, h ? u[k].y : u[k].x);
                        vm = fmax(vm, h ? v[k].y : v[k].x);
                        wm = fmax(wm, h ? w[k].y : w[k].x);
                    }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) {
                um = fmax(um, __shfl_xor(um, m));
                vm = fmax(vm, __shfl_xor(vm, m));
                wm = fmax(wm, __shfl_xor(wm, m));
            }
            const int wv = threadIdx.x / 64;
            if ((threadIdx.x & 63) == 0) {
                scratch[RM_N * wv + RM_UMAX] = um;
                scratch[RM_N * wv + RM_VMAX] = vm;
                scratch[RM_N * wv + RM_WMAX] = wm;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int i = 1; i < T / 64; ++i) {
                    um = fmax(um, scratch[RM_N * i + RM_UMAX]);
                    vm = fmax(vm, scratch[RM_N * i + RM_VMAX]);
                    wm = fmax(wm, scratch[RM_N * i + RM_WMAX]);
                }
                double *pm = a.part + (size_t)blockIdx.y * a.bt.part + (size_t)tile_index() * RM_N;
                pm[RM_UMAX] = um;
                pm[RM_VMAX] = vm;
                pm[RM_WMAX] = wm;
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kEpt; ++k) {
            const double er = o[k].x + a.fcor[1], ei = o[k].y + a.fcor[1];
            r[k] = make_double2(u[k].x * er + r[k].x, u[k].y * ei + r[k].y);
            p[k] = make_double2(-w[k].x * er, -w[k].y * ei);
        }
#pragma unroll
        for (int k = 0; k < kEpt; ++k) {
            const double zr = o2[k].x + a.fcor[2], zi = o2[k].y + a.fcor[2];
            q[k] = make_double2(q[k].x - u[k].x * zr, q[k].y - u[k].y * zi);
            p[k] = make_double2(v[k].x * zr + p[k].x, v[k].y * zi + p[k].y);
        }
        // r = u eta - v xi ; q = w xi - u zeta ; p = v zeta - w eta  -> forward in x
        if constexpr (N > 0) {
            Tile t2 = t;
            t2.nl = 2 * t.nl;
            from_regs(t, A, r);
            from_regs(t, A + tb, q);
            __syncthreads();
            double2 *R = lfft<N, true>(t2, a, A, B, -1);
            tile_store_fwd<N>(t, a, a.out[0], R);
            tile_store_fwd<N>(t, a, a.out[1], R + tb);
            __syncthreads();
        } else {
            fwd_store<N>(t, a, A, B, r, a.out[0]);
            fwd_store<N>(t, a, A, B, q, a.out[1]);
        }
        fwd_store<N>(t, a, A, B, p, a.out[2]);
        return;
    }

    if constexpr (M == LM_XADAPT) {
        // adapt (advance.f90:171-289): max |omega|, signed max of u, v, w, largest
        // |eigenvalue| of the symmetrised strain (Jacobi per point), lid maxima, and
        // the sums of the characteristic vorticity.
        const double vortrms = a.scal[(size_t)blockIdx.y * SC_NSCAL + SC_VORTRMS];
        double umax = -INFINITY, vmax = -INFINITY, wmax = -INFINITY, w2max = 0.0;
        double gg = 2.220446049250313e-16, usgg = 0.0, lsgg = 0.0, l1 = 0.0, l2 = 0.0;
        double2 ux[kEpt], uy[kEpt], vy[kEpt], wx[kEpt], wy[kEpt], xi[kEpt], et[kEpt], ze[kEpt];
        // Fields in pairs (16 lines per FFT call): (u, v) (w, xi) (u_x, w_x) (u_y, v_y)
        // (w_y, eta) (zeta) then the midpoint vorticity (mxi, meta) (mzeta); spectral
        // inputs prefetch the next pair during the current transform.
        Pf<N, T> p0, p1;
        Tile t2 = t;
        t2.nl = 2 * t.nl;
        auto get2 = [&](int fa, int fb, bool diff, double2(&ra)[kEpt], double2(&rb)[kEpt], int na, int nb) {
            if (N == 0 || a.phys_in) {
                field_phys<N>(t, a, a.in[fa], A, B, ra, diff);
                if (fb >= 0) field_phys<N>(t, a, a.in[fb], A, B, rb, diff);
                return;
            }
            pf_commit(p0, t, A);
            if (fb >= 0) pf_commit(p1, t, A + tb);
            __syncthreads();
            if (na >= 0) pf_load(p0, t, a, a.in[na]);
            if (nb >= 0) pf_load(p1, t, a, a.in[nb]);
            const Tile &tt = fb >= 0 ? t2 : t;
            double2 *S = A, *O = B;
            if (diff) {
                tile_diff(tt, a, A, B);
                S = B;
                O = A;
            }
            double2 *R = tile_inverse<N, true>(tt, a, S, O);
            to_regs(t, R, ra, 1.0 / (double)t.n);
            if (fb >= 0) to_regs(t, R + tb, rb, 1.0 / (double)t.n);
            __syncthreads();
        };
        if constexpr (N > 0) {
            if (!a.phys_in) {
                pf_load(p0, t, a, a.in[0]);
                pf_load(p1, t, a, a.in[a.velmax_given ? 2 : 1]);
            }
        }
        if (a.velmax_given) {
            // the products pass has taken the velocity maxima: (u_x, w_x) (u_y, v_y)
            // (w_y, eta) (xi, zeta), eleven tile loads in all instead of fourteen
            get2(0, 2, true, ux, wx, 6, 7);    // du/dx, dw/dx
            get2(6, 7, false, uy, vy, 8, 4);   // du/dy, dv/dy
            get2(8, 4, false, wy, et, 3, 5);   // dw/dy, eta
            get2(3, 5, false, xi, ze, 9, 10);  // xi, zeta
        } else {
            get2(0, 1, false, ux, vy, 2, 3);   // u, v (maxima)
#pragma unroll
            for (int k = 0; k < kEpt; ++k)
                for (int h = 0; h < 2; ++h)
                    if (reg_valid(t, k, h, nz)) {
                        umax = fmax(umax, h ? ux[k].y : ux[k].x);
                        vmax = fmax(vmax, h ? vy[k].y : vy[k].x);
                    }
            get2(2, 3, false, wx, xi, 0, 2);   // w (maximum), xi
#pragma unroll
            for (int k = 0; k < kEpt; ++k)
                for (int h = 0; h < 2; ++h)
                    if (reg_valid(t, k, h, nz)) wmax = fmax(wmax, h ? wx[k].y : wx[k].x);
            get2(0, 2, true, ux, wx, 6, 7);    // du/dx, dw/dx
            get2(6, 7, false, uy, vy, 8, 4);   // du/dy, dv/dy
            get2(8, 4, false, wy, et, 5, -1);  // dw/dy, eta
            get2(5, -1, false, ze, ze, 9, 10); // zeta
        }
        // S11 = u_x, S12 = u_y + zeta/2, S13 = w_x + eta/2, S22 = v_y,
        // S23 = w_y - xi/2, S33 = -(u_x + v_y)  (advance.f90:225-257).  The block's
        // maximum of the Jacobi values is found by running Jacobi only where the
        // closed-form estimate comes within 1e-6 of the block's largest estimate (the
        // estimate's error is far below that margin, so the maximising point is always
        // among them); lid points, which have maxima of their own, always run it.
#define C_(V) (h ? (V).y : (V).x)
#define STRAIN_ARGS C_(ux[k]), C_(uy[k]) + 0.5 * C_(ze[k]), C_(wx[k]) + 0.5 * C_(et[k]), C_(vy[k]), \
                    C_(wy[k]) - 0.5 * C_(xi[k]), -(C_(ux[k]) + C_(vy[k]))
        float est[kEpt][2];   // float keeps the registers down; its rounding is far below the margin
        float emax = 0.0f;
#pragma unroll
        for (int k = 0; k < kEpt; ++k)
            for (int h = 0; h < 2; ++h) {
                est[k][h] = -1.0f;
                if (!reg_valid(t, k, h, nz)) continue;
                const double o1 = C_(xi[k]), o2 = C_(et[k]), o3 = C_(ze[k]);
                w2max = fmax(w2max, o1 * o1 + o2 * o2 + o3 * o3);
                est[k][h] = (float)eig_absmax_est(STRAIN_ARGS);
                emax = fmaxf(emax, est[k][h]);
            }
        const float ecut = (float)(block_max((double)emax, scratch) * (1.0 - 1.0e-6));
#pragma unroll
        for (int k = 0; k < kEpt; ++k)
            for (int h = 0; h < 2; ++h) {
                if (!reg_valid(t, k, h, nz)) continue;
                const int z = reg_z(t, k, h);
                const bool lid = z == 0 || z == nz;
                if (!lid && !(est[k][h] >= ecut)) continue;
                const double lmax = jacobi_absmax(STRAIN_ARGS);
                gg = fmax(gg, lmax);
                if (z == nz) usgg = fmax(usgg, lmax);
                if (z == 0) lsgg = fmax(lsgg, lmax);
            }
#undef STRAIN_ARGS
        // characteristic vorticity <|w|^2>/<|w|> over cells with |w| > w_rms
        // (field_diagnostics.f90:501-545), from the z-midpoint vorticity
        get2(9, 10, false, xi, et, 11, -1);
        get2(11, -1, false, ze, ze, -1, -1);
#pragma unroll
        for (int k = 0; k < kEpt; ++k)
            for (int h = 0; h < 2; ++h) {
                if (!reg_valid(t, k, h, nz)) continue;
                if (reg_z(t, k, h) < 1) continue;
                const double t1 = fabs(C_(xi[k])), t2 = fabs(C_(et[k])), t3 = fabs(C_(ze[k]));
                if (t1 + t2 + t3 > vortrms) {
                    l1 += t1 + t2 + t3;
                    l2 += t1 * t1 + t2 * t2 + t3 * t3;
                }
            }
#undef C_
        w2max = block_max(w2max, scratch);
        gg = block_max(gg, scratch);
        usgg = block_max(usgg, scratch);
        lsgg = block_max(lsgg, scratch);
        umax = block_max(umax, scratch);
        vmax = block_max(vmax, scratch);
        wmax = block_max(wmax, scratch);
        l1 = block_sum(l1, scratch);
        l2 = block_sum(l2, scratch);
        if (threadIdx.x == 0) {
            double *p = a.part + (size_t)blockIdx.y * a.bt.part + (size_t)tile_index() * RX_N;
            p[RX_W2MAX] = w2max;
            p[RX_GG] = gg;
            p[RX_USGG] = usgg;
            p[RX_LSGG] = lsgg;
            p[RX_UMAX] = umax;
            p[RX_VMAX] = vmax;
            p[RX_WMAX] = wmax;
            p[RX_L1] = l1;
            p[RX_L2] = l2;
        }
        return;
    }
}

size_t line_lds_bytes(int n, int ZC) {
    const size_t tb = (size_t)(ZC / 2) * line_ls(n) * sizeof(double2);
    size_t need = 2 * tb + (size_t)n * sizeof(double2);
    const size_t red = kLineThreads * sizeof(double);
    return need > red ? need : red;
}

#define PS3D_LINE_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
// lengths with a 256-thread (quarter-depth) register build
#define PS3D_LINE_SIZES_Q(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

hipError_t configure_linepass() {
    const int lds = 160 * 1024;
    hipError_t e = hipSuccess;
#define CFGT(N_, T_)                                                                                         \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)k_line_reg<N_, LM_INV_Y1, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)k_line_reg<N_, LM_XMAIN, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)k_line_reg<N_, LM_XADAPT, T_>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
#define CFGS(N_, S_)                                                                                         \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)k_line_simple<N_, S_>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
#define CFG(N_) CFGS(N_, 0) CFGS(N_, 1) CFGS(N_, 2) CFGS(N_, 3) CFGS(N_, 4) CFGS(N_, 5) CFGS(N_, 6) CFGS(N_, 7) \
    CFGT(N_, kLineThreads)
#define CFG2(N_) CFG(N_) CFGT(N_, 512)
#define CFG4(N_) CFGT(N_, 256)
    CFGS(0, 0)
    CFGT(0, kLineThreads)
    PS3D_LINE_SIZES(CFG2)
    PS3D_LINE_SIZES_Q(CFG4)
#undef CFG4
#undef CFG2
#undef CFG
#undef CFGS
#undef CFGT
    return e;
}

template <int N, int T>
static void launch_reg_t(const LineArgs &a, hipStream_t stream, int blocks, int nb, size_t lds) {
    if (a.mode == LM_INV_Y1)
        hipLaunchKernelGGL((k_line_reg<N, LM_INV_Y1, T>), dim3(blocks, nb), dim3(T), lds, stream, a);
    else if (a.mode == LM_XMAIN)
        hipLaunchKernelGGL((k_line_reg<N, LM_XMAIN, T>), dim3(blocks, nb), dim3(T), lds, stream, a);
    else
        hipLaunchKernelGGL((k_line_reg<N, LM_XADAPT, T>), dim3(blocks, nb), dim3(T), lds, stream, a);
}

template <int N>
static void launch_line_t(const LineArgs &a, hipStream_t stream) {
    const int blocks = a.nother * a.nzt;
    const int nb = a.bt.nb > 1 ? a.bt.nb : 1;
    size_t lds = line_lds_bytes(a.n, a.ZC);
    if ((a.mode == LM_XMAIN || a.mode == LM_XADAPT) && N > 0)
        lds += (size_t)2 * (a.ZC / 2) * line_ls(a.n) * sizeof(double2);   // field pairs
    if (a.mode == LM_FWD || a.mode == LM_INV || a.mode == LM_DIFF) {
        if constexpr (N == 0) {
            hipLaunchKernelGGL((k_line_simple<0, 0>), dim3(blocks, nb), dim3(256), lds, stream, a);
        } else {
            switch (a.sched & 7) {
#define SCH(S_) \
    case S_: hipLaunchKernelGGL((k_line_simple<N, S_>), dim3(blocks, nb), dim3((S_ & LS_T512) ? 512 : 256), lds, stream, a); break;
                SCH(0) SCH(1) SCH(2) SCH(3) SCH(4) SCH(5) SCH(6) SCH(7)
#undef SCH
            }
        }
    }
    else if (a.rthreads == 512 && N > 0)
        launch_reg_t<N, 512>(a, stream, blocks, nb, lds);
    else if constexpr (N > 0 && N <= 512) {
        if (a.rthreads == 256)
            launch_reg_t<N, 256>(a, stream, blocks, nb, lds);
        else
            launch_reg_t<N, kLineThreads>(a, stream, blocks, nb, lds);
    } else
        launch_reg_t<N, kLineThreads>(a, stream, blocks, nb, lds);
}

bool line_reg_ok(int n, int T) {
    if (T == kLineThreads) return true;
    bool built = false;
    switch (n) {
#define CASE(N_) case N_: built = T == 512 || (T == 256 && N_ <= 512); break;
        PS3D_LINE_SIZES(CASE)
#undef CASE
        default: break;
    }
    return built && reg_fits(n, T);
}

int line_reg_zc(int n, int T) { return zc_reg(n, T); }

// Measured on the plain workload (steps/s under each schedule of the y passes):
// 128^3 and 512^3 run fastest with the fold at 256 threads, 256^3 with 512 threads
// and the pair pass (116.5 against 115.6 steps/s); the small grids of the ensemble
// go to the tuner.
int line_sched_default(int n) { return n == 256 ? LS_T512 : (n >= 128 ? LS_FOLD : 0); }

bool line_sched_ok(int n, int sched) {
    if (sched < 0 || sched > 7) return false;
    bool built = false, r8 = false;
    switch (n) {
#define CASE(N_) case N_: built = true; r8 = CPlan8<N_>::ok; break;
        PS3D_LINE_SIZES(CASE)
#undef CASE
        default: break;
    }
    if (!built) return sched == 0;
    return !(sched & LS_R8) || r8;
}

hipError_t launch_line(const LineArgs &a, int nthreads, hipStream_t stream) {
    (void)nthreads;
    // compile-time plans need the standard tile depth of their build
    const bool reg = a.mode == LM_INV_Y1 || a.mode == LM_XMAIN || a.mode == LM_XADAPT;
    const int rt = (a.rthreads == 512 || a.rthreads == 256) ? a.rthreads : kLineThreads;
    const bool std_zc = a.ZC == (reg ? zc_reg(a.n, rt) : zc_for(a.n));
    if (reg && rt != kLineThreads && (!std_zc || !line_reg_ok(a.n, rt))) return hipErrorInvalidValue;
    switch (std_zc ? a.n : 0) {
#define CASE(N_) case N_: launch_line_t<N_>(a, stream); break;
        PS3D_LINE_SIZES(CASE)
#undef CASE
        default: launch_line_t<0>(a, stream); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// finishing reductions (one block)
// ---------------------------------------------------------------------------
// one block per instance; 1024 threads read the per-tile partials
constexpr int kFinThreads = 1024;

// NV values per thread reduced over the block (MAX: fmax, else sum) for thread 0:
// shuffles inside each wave, then one pass over the per-wave results in LDS (two
// barriers in all, where one block_sum/block_max per value takes ten each).
template <int NV, bool MAX>
__device__ inline void fin_reduce(double (&v)[NV], double *red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double o = __shfl_xor(v[q], m);
            v[q] = MAX ? fmax(v[q], o) : v[q] + o;
        }
    if (lane == 0)
        for (int q = 0; q < NV; ++q) red[wv * NV + q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < nw; ++i)
            for (int q = 0; q < NV; ++q) v[q] = MAX ? fmax(v[q], red[i * NV + q]) : v[q] + red[i * NV + q];
    __syncthreads();
}

__global__ void __launch_bounds__(kFinThreads) k_fin_y(FinArgs a) {
    __shared__ double red[kFinThreads];
    if (gated_off(a.gate)) return;
    const double *party = a.party + (size_t)blockIdx.y * a.bt.party;
    double *scal = a.scal + (size_t)blockIdx.y * SC_NSCAL;
    double s[RY_N];
    for (int q = 0; q < RY_N; ++q) s[q] = 0.0;
    for (int b = threadIdx.x; b < a.nby; b += blockDim.x)
        for (int q = 0; q < RY_N; ++q) s[q] += party[(size_t)b * RY_N + q];
    fin_reduce<RY_N, false>(s, red);
    const double(&tot)[RY_N] = s;
    if (threadIdx.x == 0) {
        // get_kinetic_energy / get_enstrophy: (1/2 interior + 1/4 lids) / ncell
        // (field_diagnostics.f90:85-206); vortrms = sqrt(get_mean(|w|^2)) (advance.f90:177)
        scal[SC_KE] = 0.5 * tot[RY_UU] * a.ncelli;
        scal[SC_EN] = 0.5 * tot[RY_WW] * a.ncelli;
        scal[SC_HEL] = tot[RY_UW] * a.ncelli;
        scal[SC_VORTRMS] = sqrt(tot[RY_WW] * a.ncelli);
        scal[SC_VORMX] = tot[RY_MX] * a.ncelli;
        scal[SC_VORMY] = tot[RY_MY] * a.ncelli;
        scal[SC_VORMZ] = tot[RY_MZ] * a.ncelli;
    }
}

__global__ void __launch_bounds__(kFinThreads) k_fin_x(FinArgs a) {
    __shared__ double red[kFinThreads];
    if (gated_off(a.gate)) return;
    double w2 = 0.0, gg = 2.220446049250313e-16, us = 0.0, ls = 0.0;
    double um = -INFINITY, vm = -INFINITY, wm = -INFINITY, l1 = 0.0, l2 = 0.0;
    const double *partx = a.partx + (size_t)blockIdx.y * a.bt.partx;
    double *roll_hist = a.roll_hist + (size_t)blockIdx.y * a.bt.roll;
    double *diag = a.diag ? a.diag + (size_t)blockIdx.y * a.bt.diag : nullptr;
    for (int b = threadIdx.x; b < a.nbx; b += blockDim.x) {
        const double *p = partx + (size_t)b * RX_N;
        w2 = fmax(w2, p[RX_W2MAX]);
        gg = fmax(gg, p[RX_GG]);
        us = fmax(us, p[RX_USGG]);
        ls = fmax(ls, p[RX_LSGG]);
        um = fmax(um, p[RX_UMAX]);
        vm = fmax(vm, p[RX_VMAX]);
        wm = fmax(wm, p[RX_WMAX]);
        l1 += p[RX_L1];
        l2 += p[RX_L2];
    }
    if (a.partm) {
        const double *partm = a.partm + (size_t)blockIdx.y * a.bt.partm;
        for (int b = threadIdx.x; b < a.nbm; b += blockDim.x) {
            const double *p = partm + (size_t)b * RM_N;
            um = fmax(um, p[RM_UMAX]);
            vm = fmax(vm, p[RM_VMAX]);
            wm = fmax(wm, p[RM_WMAX]);
        }
    }
    double mx[7] = {w2, gg, us, ls, um, vm, wm}, sm[2] = {l1, l2};
    fin_reduce<7, true>(mx, red);
    fin_reduce<2, false>(sm, red);
    if (threadIdx.x != 0) return;
    w2 = mx[0];
    gg = mx[1];
    us = mx[2];
    ls = mx[3];
    um = mx[4];
    vm = mx[5];
    wm = mx[6];
    l1 = sm[0];
    l2 = sm[1];
    double *sc = a.scal + (size_t)blockIdx.y * SC_NSCAL;
    const double small = 1.0e-12, cflmax = 0.8;
    const double t = sc[SC_T];
    const double vortmax = sqrt(w2);
    const double vorch = l2 / (small + l1);            // vorl1 starts at small
    const double bfmax = 0.0;
    // CFL and strain limits (advance.f90:325-333)
    const double dtcfl = cflmax * fmin(fmin(a.dx / (um + small), a.dy / (vm + small)), a.dz / (wm + small));
    double dt = fmin(fmin(a.alpha / (gg + small), a.alpha / (bfmax + small)), fmin(dtcfl, a.limit - t));
    if (a.dt_fixed > 0.0) dt = fmin(a.dt_fixed, a.limit - t);
    // rolling mean (rolling_mean.f90:36-69); indices are 1-based as there
    double sma = sc[SC_RM_SMA];
    int inew = (int)sc[SC_RM_INEW], iold = (int)sc[SC_RM_IOLD], filled = (int)sc[SC_RM_FILLED];
    const int len = a.roll_len;
    if (filled) {
        const double vold = roll_hist[iold - 1];
        iold = (iold % len) + 1;
        sma = sma + (gg - vold) / (double)len;
        roll_hist[inew - 1] = gg;
        inew = (inew % len) + 1;
    } else {
        roll_hist[inew - 1] = gg;
        double s = 0.0;
        for (int i = 0; i < inew; ++i) s += roll_hist[i];
        sma = s / (double)inew;
        filled = (len == inew);
        inew = (inew % len) + 1;
    }
    sc[SC_RM_SMA] = sma;
    sc[SC_RM_INEW] = inew;
    sc[SC_RM_IOLD] = iold;
    sc[SC_RM_FILLED] = filled;
    // get_diffusion_pre_factor (advance.f90:381-410)
    double pref;
    switch (a.pretype) {
        case 0: pref = 1.0; break;
        case 1: pref = vorch; break;
        case 2: pref = bfmax; break;
        case 3: pref = sma; break;
        case 4: pref = gg; break;
        default: pref = us; break;
    }
    sc[SC_PREF] = pref;
    sc[SC_DT] = dt;
    if (a.stepper == 1)
        sc[SC_DFAC] = (a.nnu == 1) ? dt : pref * dt;    // cn2.f90:52-58
    else
        sc[SC_DFAC] = 0.5 * pref * dt;                   // impl_rk4.f90:43
    if (diag) {
        const int row = (int)sc[SC_STEP];
        if (row < a.max_rows) {
            double *d = diag + (size_t)row * 19;
            d[0] = t; d[1] = dt; d[2] = sc[SC_KE]; d[3] = sc[SC_EN]; d[4] = sc[SC_HEL];
            d[5] = vortmax; d[6] = sc[SC_VORTRMS]; d[7] = vorch;
            d[8] = sc[SC_VORMX]; d[9] = sc[SC_VORMY]; d[10] = sc[SC_VORMZ];
            d[11] = gg; d[12] = um; d[13] = vm; d[14] = wm; d[15] = us; d[16] = ls; d[17] = sma; d[18] = pref;
        }
        sc[SC_STEP] = row + 1;
    }
}

hipError_t launch_fin_y(const FinArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(k_fin_y, dim3(1, a.bt.nb > 1 ? a.bt.nb : 1), dim3(kFinThreads), 0, stream, a);
    return hipGetLastError();
}
hipError_t launch_fin_x(const FinArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(k_fin_x, dim3(1, a.bt.nb > 1 ? a.bt.nb : 1), dim3(kFinThreads), 0, stream, a);
    return hipGetLastError();
}

__global__ void k_gate(double *scal) {
    if (threadIdx.x == 0) scal[SC_GATE] = (scal[SC_T] < scal[SC_TSTOP]) ? 1.0 : 0.0;
}

hipError_t launch_gate(double *scal, hipStream_t stream) {
    hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, stream, scal);
    return hipGetLastError();
}

}  // namespace ps3d
