// zwave.hip -- z-column kernels with one wavefront per spectral column.
//
// A block is one quad of partner columns {(kx,ky), (nx-kx,ky), (kx,ny-ky),
// (nx-kx,ny-ky)} (the partners of ps3d's spectral x/y derivatives,
// src/fft/sta3dfft.f90:304-377); wave m of the block owns member m.  Lane l holds
// z = l + 64 s (s < nz/64) of each field in registers and the top lid z = nz as a
// wave-uniform value.  Everything along z stays inside the wave:
//   * lid values by readlane, central_diffz by lane shuffles;
//   * the DST-I / DCT-I (src/fft/stafft.f90:410-550) as a complex FFT of length
//     2 nz = 64 P of the odd / even extension, two fields per complex line, in
//     registers: a radix-P pass over the registers, then the 64-point DFT across
//     lanes as 8 x 8 with two transposes through the wave's own LDS slice (no
//     block barrier);
// and only the x/y derivative products exchange data between the four waves
// (one LDS exchange with two block barriers).
//
// Sizes: nz in {64, 128, 256, 512} (P = 2, 4, 8, 16).  Other nz use the
// quad-per-thread kernels of zpass.hip.
#include "kernels.hpp"
#include "lfft.hpp"
#include "zmath.hpp"

namespace ps3d {

namespace {

enum { WK_DST = 0, WK_DCT = 1, WK_NONE = 2 };

__device__ inline double rdlane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

__device__ inline double2 shfl2(double2 v, int src) {
    return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// LDS ordering inside one wavefront (the wave's lanes execute together; this
// orders the LDS accesses and stops the compiler from moving them across).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NZ>
struct WDim {
    static constexpr int S = NZ / 64;      // register slots per lane (z < nz)
    static constexpr int P = NZ / 32;      // complex points per lane of a length-2nz line
    static constexpr int IPL = (8 * P + 63) / 64;   // 8-point DFT items per lane across lanes
    static constexpr int WS = 72 * P;      // double2 of LDS per wave
    // nz <= 256: the exchanges of vor2vel's first and third phase go to a second LDS
    // region (two fields per member) beside the wave slices, so that no exchange
    // needs a barrier after its reads (the next write to a region comes after the
    // next exchange's barrier); nz = 512 keeps one region (the LDS of two blocks
    // per CU would not hold a second one)
    static constexpr bool DB = NZ <= 256;
};

// One field of one column: z = lane + 64 s in v[s], z = nz in top (same in all lanes).
template <int NZ>
struct WF {
    double v[WDim<NZ>::S];
    double top;
};

// Per-wave context: the member column this wave owns.
struct WC {
    int lane, m;
    bool store, q00;
    size_t col, colidx;
    double fx, fy, k2l2, k2l2i, K;
    int kx, ky;
};

__device__ inline WC make_wc(const Grid &g, const ZTabs &tb, long ioff) {
    WC c;
    c.lane = threadIdx.x & 63;
    c.m = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int q = blockIdx.x;
    const int nqy = g.ny / 2 + 1;
    const int kx = q / nqy, ky = q % nqy;
    c.kx = kx;
    c.ky = ky;
    const int kxp = (g.nx - kx) % g.nx, kyp = (g.ny - ky) % g.ny;
    const int ix = (c.m & 1) ? kxp : kx;
    const int iy = (c.m & 2) ? kyp : ky;
    const bool dx = kxp != kx, dy = kyp != ky;
    // a member that coincides with member 0 / 1 / 2 is computed but not stored
    c.store = c.m == 0 || (c.m == 1 && dx) || (c.m == 2 && dy) || (c.m == 3 && dx && dy);
    c.colidx = (size_t)ix * g.ny + iy;
    c.col = c.colidx * (size_t)g.zs + ioff;   // ioff: batched instance
    // diffx: ds(k) = si * rk(k) * fs(n-k), si = +1 for k > n/2, -1 otherwise;
    // ds(0) = 0 and ds(n/2) = 0 for even n (sta3dfft.f90:318-335)
    double fxm = (ix >= g.nwx + 1) ? tb.rkx[ix] : -tb.rkx[ix];
    if (ix == 0 || ix == g.nwx) fxm = 0.0;
    double fym = (iy >= g.nwy + 1) ? tb.rky[iy] : -tb.rky[iy];
    if (iy == 0 || iy == g.nwy) fym = 0.0;
    c.fx = fxm;
    c.fy = fym;
    c.q00 = (kx == 0 && ky == 0);
    // k2l2 = rkx^2 + rky^2; k2l2i = 1/k2l2, 0 at (0,0) (inversion_utils.f90:243-258)
    const double rx = tb.rkx[kx], ry = tb.rky[ky];
    c.k2l2 = rx * rx + ry * ry;
    c.k2l2i = c.q00 ? 0.0 : 1.0 / c.k2l2;
    c.K = sqrt(c.k2l2);
    return c;
}

template <int NZ>
__device__ inline void wload(const WC &c, const double *__restrict__ src, WF<NZ> &F) {
    const double *p = src + c.col;
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) F.v[s] = p[c.lane + 64 * s];
    F.top = p[NZ];
}

template <int NZ>
__device__ inline void wstore(const WC &c, double *__restrict__ dst, const WF<NZ> &F) {
    if (!c.store) return;
    double *p = dst + c.col;
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) p[c.lane + 64 * s] = F.v[s];
    if (c.lane == 0) p[NZ] = F.top;
}

template <int NZ>
__device__ inline double wlo(const WF<NZ> &F) {
    return rdlane(F.v[0], 0);
}

// central_diffz (inversion_utils.f90:653-673): 2nd-order interior, one-sided lids.
template <int NZ>
__device__ inline void wdiff(const WC &c, const Grid &g, const WF<NZ> &F, WF<NZ> &D) {
    constexpr int S = WDim<NZ>::S;
    const int l = c.lane;
    double f1 = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double up = __shfl(F.v[s], (l + 1) & 63, 64);
        const double nxt = (s + 1 < S) ? rdlane(F.v[(s + 1) % S], 0) : F.top;
        if (l == 63) up = nxt;
        double dn = __shfl(F.v[s], (l + 63) & 63, 64);
        const double prv = (s > 0) ? rdlane(F.v[(s + S - 1) % S], 63) : 0.0;
        if (l == 0) dn = prv;
        if (s == 0) f1 = up;
        D.v[s] = (up - dn) * g.hdzi;
    }
    if (l == 0) D.v[0] = g.dzi * (f1 - F.v[0]);
    D.top = g.dzi * (F.top - rdlane(F.v[S - 1], 63));
}

// (F(z-1) + F(z)) / 2 at z >= 1, 0 at z = 0: the cell-midpoint values of
// get_char_vorticity (field_diagnostics.f90:509-521).
template <int NZ>
__device__ inline void wmid(const WC &c, const WF<NZ> &F, WF<NZ> &M) {
    constexpr int S = WDim<NZ>::S;
    const int l = c.lane;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double dn = __shfl(F.v[s], (l + 63) & 63, 64);
        const double prv = (s > 0) ? rdlane(F.v[(s + S - 1) % S], 63) : 0.0;
        if (l == 0) dn = prv;
        M.v[s] = 0.5 * (dn + F.v[s]);
    }
    if (l == 0) M.v[0] = 0.0;
    M.top = 0.5 * (rdlane(F.v[S - 1], 63) + F.top);
}

// exp(-K zp) and exp(-K zm) at the lane's levels z = lane + 64 s (and z = nz in
// slot S), by recurrence along z: two exponentials per lane and one per column
// instead of two per level.  The products add a few ulps to each value.
template <int NZ>
struct WExp {
    double ep[WDim<NZ>::S + 1], em[WDim<NZ>::S + 1];
    double ef, div;
};
// The column's exponentials and division, formed once per kernel (a kernel that
// runs both the tendency and the vor2vel pass expands them three times).
struct WSeed {
    double e, f, e64, ef, div;
};
__device__ inline WSeed make_wseed(const WC &c, const Grid &g) {
    WSeed x;
    const double kd = c.K * g.dz;
    x.e64 = exp(-64.0 * kd);
    x.ef = exp(-c.K * g.lz);
    x.div = 1.0 / (1.0 - x.ef * x.ef);
    x.e = exp(-kd * (double)c.lane);
    x.f = exp(-kd * (double)(64 - c.lane));   // z = lane + 64 (S - 1): zm = dz (64 - lane)
    return x;
}
template <int NZ>
__device__ inline WExp<NZ> make_wexp(const WSeed &sd) {
    constexpr int S = WDim<NZ>::S;
    static_assert(NZ == 64 * S, "wave layout covers nz = 64 S");
    WExp<NZ> x;
    x.ef = sd.ef;
    x.div = sd.div;
    double e = sd.e;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        x.ep[s] = e;
        e *= sd.e64;
    }
    double f = sd.f;
#pragma unroll
    for (int s = S - 1; s >= 0; --s) {
        x.em[s] = f;
        f *= sd.e64;
    }
    x.ep[S] = x.ef;
    x.em[S] = 1.0;
    return x;
}

// phi-(z), phi+(z) of the lane's interior levels (inversion_utils.f90:484-520)
template <int NZ>
struct WPhi {
    double pm[WDim<NZ>::S], pp[WDim<NZ>::S];
};
template <int NZ>
__device__ inline WPhi<NZ> make_phi(const WC &c, const Grid &g, const WSeed &sd) {
    WPhi<NZ> h;
    const WExp<NZ> x = make_wexp<NZ>(sd);
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) {
        if (c.q00) {
            double zm, zp;
            zm_zp(g, c.lane + 64 * s, zm, zp);
            h.pm[s] = zm / g.lz;
            h.pp[s] = zp / g.lz;
        } else {
            h.pm[s] = x.div * (x.ep[s] - x.ef * x.em[s]);
            h.pp[s] = x.div * (x.em[s] - x.ef * x.ep[s]);
        }
    }
    return h;
}

// Add (sign +1: combine) or subtract (sign -1: decompose) the harmonic
// interpolant of the lid values, interior levels only.
template <int NZ>
__device__ inline void wharm(const WC &c, const WPhi<NZ> &h, WF<NZ> &F, double lo, double hi, double sign) {
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s)
        if (c.lane + 64 * s > 0) F.v[s] = F.v[s] + sign * (lo * h.pm[s] + hi * h.pp[s]);
}

// Twiddles of the wave FFT of length N = 64 P.  The lane's roots are read from the
// global table tw[m] = exp(-2 pi i m / N) once per kernel (those loads overlap the
// column loads); each transform forms the powers it needs from them by binary
// splitting.  Reading the table inside every transform made each transform's first
// use wait for all of the wave's global accesses in flight (the vector memory
// counter is in order), stores of the previous phase included.
//   r1 = w_N^lane (register pass);
//   P = 8, 16: a = w_64^(lane mod 8) (8 x 8 lane stage);
//   P = 2, 4:  a = w_64^(lane mod 16), b = w_16^(lane mod 4) (radix-4 lane stages).
template <int P>
struct WTw {
    double2 r1, a, b;
};
template <int P>
__device__ inline WTw<P> make_tw(const double2 *__restrict__ tw, int lane) {
    WTw<P> t;
    t.r1 = tw[lane];
    if constexpr (P >= 8) {
        t.a = tw[P * (lane & 7)];
        t.b = t.a;
    } else {
        t.a = tw[P * (lane & 15)];
        t.b = tw[4 * P * (lane & 3)];
    }
    return t;
}
// An opaque copy of a root: keeps the compiler from sharing the powers between
// transforms (they would stay live across the kernel).
__device__ inline double2 wroot(double2 r) {
    asm volatile("" : "+v"(r.x), "+v"(r.y));
    return r;
}

// Forward complex FFT of length 64 P (P = 8, 16).  In and out: lane l, register j
// holds element l + 64 j.  ws: the wave's LDS slice (72 P double2).
template <int P>
__device__ inline void wfft(double2 (&x)[P], double2 *__restrict__ ws, const WTw<P> &tw, int lane) {
    constexpr int IPL = (8 * P + 63) / 64;
    constexpr int NIT = 8 * P;
    // powers of the lane's roots by binary splitting (product depth <= log2 k;
    // measured: reading every power from the table instead made the fused
    // tendency + vor2vel kernel 1.7x slower).  Item lane + 64 r of the lane stage
    // has l1 = lane mod 8 for every r (8 P is a multiple of 64).
    double2 ta[P], tb[IPL][8];
    ta[1] = wroot(tw.r1);
#pragma unroll
    for (int k = 2; k < P; ++k) ta[k] = cmul(ta[k / 2], ta[k - k / 2]);
#pragma unroll
    for (int r = 0; r < IPL; ++r) {
        tb[r][1] = wroot(tw.a);
#pragma unroll
        for (int k = 2; k < 8; ++k) tb[r][k] = cmul(tb[r][k / 2], tb[r][k - k / 2]);
    }
    // length-P DFT over the registers, then the twiddle w_N^(lane k1)
    bfly<P>(x, -1);
#pragma unroll
    for (int k = 1; k < P; ++k) x[k] = cmul(x[k], ta[k]);
    // 64-point DFTs over the lanes (one per k1): rows k1 of 64
#pragma unroll
    for (int k = 0; k < P; ++k) ws[k * 64 + lane] = x[k];
    wave_sync();
    double2 u[IPL][8];
#pragma unroll
    for (int r = 0; r < IPL; ++r) {
        const int it = (lane + 64 * r) % NIT;
        const int k1 = it >> 3, l1 = it & 7;
#pragma unroll
        for (int j = 0; j < 8; ++j) u[r][j] = ws[k1 * 64 + l1 + 8 * j];
    }
    wave_sync();
    // 8-point DFT over l2, twiddle w_64^(l1 k2a); rows k1 + P k2a of 9 (padded)
#pragma unroll
    for (int r = 0; r < IPL; ++r) {
        const int it = (lane + 64 * r) % NIT;
        const int k1 = it >> 3, l1 = it & 7;
        bfly<8>(u[r], -1);
#pragma unroll
        for (int k = 1; k < 8; ++k) u[r][k] = cmul(u[r][k], tb[r][k]);
#pragma unroll
        for (int k = 0; k < 8; ++k) ws[(k1 + P * k) * 9 + l1] = u[r][k];
    }
    wave_sync();
    // 8-point DFT over l1: item it = k1 + P k2a yields X[it + 8 P k2b]
#pragma unroll
    for (int r = 0; r < IPL; ++r) {
        const int it = (lane + 64 * r) % NIT;
#pragma unroll
        for (int j = 0; j < 8; ++j) u[r][j] = ws[it * 9 + j];
        bfly<8>(u[r], -1);
    }
    wave_sync();   // the slice is free for the caller again
    if constexpr (P == 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = u[0][j];
    } else {
        static_assert(P == 16, "8 x 8 lane stages for P = 8, 16 (wfft4 takes P = 2, 4)");
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            x[2 * j] = u[0][j];
            x[2 * j + 1] = u[1][j];
        }
    }
}

// Forward complex FFT of length 64 P for P = 2, 4, same layout as wfft: the 64-point
// DFTs across the lanes as 4 x 4 x 4, three radix-4 stages through the wave's LDS
// slice.  The 8 x 8 form has 8 P items per stage, so at P = 4 half the lanes and at
// P = 2 three quarters repeat the others' work; here there are 16 P items.
//   stage 1: item (k1, m), m < 16: DFT over l3 of row k1's elements m + 16 l3 -> k2a,
//            twiddle w_64^(m k2a);
//   stage 2: item (k1, k2a, l1): DFT over l2 of stage 1's (k1, k2a, l1 + 4 l2) -> qa,
//            twiddle w_16^(l1 qa);
//   stage 3: item k1 + P k2a + 4 P qa: DFT over l1 -> X[item + 16 P qb].
template <int P>
__device__ inline void wfft4(double2 (&x)[P], double2 *__restrict__ ws, const WTw<P> &tw, int lane) {
    static_assert(P == 2 || P == 4, "radix-4 lane stages for P = 2, 4");
    constexpr int NIT = 16 * P;
    double2 ta[P];
    ta[1] = wroot(tw.r1);
#pragma unroll
    for (int k = 2; k < P; ++k) ta[k] = cmul(ta[k / 2], ta[k - k / 2]);
    // length-P DFT over the registers, then the twiddle w_N^(lane k1)
    bfly<P>(x, -1);
#pragma unroll
    for (int k = 1; k < P; ++k) x[k] = cmul(x[k], ta[k]);
#pragma unroll
    for (int k = 0; k < P; ++k) ws[k * 64 + lane] = x[k];
    wave_sync();
    // (at P = 2 the upper lanes repeat the items of the lower ones: same values to
    // the same addresses)
    const int it = lane % NIT;
    double2 u[4];
    {
        const int k1 = it >> 4, m = it & 15;
#pragma unroll
        for (int l3 = 0; l3 < 4; ++l3) u[l3] = ws[k1 * 64 + m + 16 * l3];
        const double2 w1 = wroot(tw.a);   // w_64^m, m = lane mod 16
        const double2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        wave_sync();
        bfly<4>(u, -1);
        u[1] = cmul(u[1], w1);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
#pragma unroll
        for (int k2a = 0; k2a < 4; ++k2a) ws[k1 * 64 + k2a * 16 + m] = u[k2a];
    }
    wave_sync();
    {
        const int k1 = it >> 4, k2a = (it >> 2) & 3, l1 = it & 3;
#pragma unroll
        for (int l2 = 0; l2 < 4; ++l2) u[l2] = ws[k1 * 64 + k2a * 16 + l1 + 4 * l2];
        const double2 w1 = wroot(tw.b);   // w_16^l1, l1 = lane mod 4
        const double2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        wave_sync();
        bfly<4>(u, -1);
        u[1] = cmul(u[1], w1);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
#pragma unroll
        for (int qa = 0; qa < 4; ++qa) ws[((qa * 4 + k2a) * 4 + l1) * P + k1] = u[qa];
    }
    wave_sync();
    {
        const int k1 = it % P, k2a = (it / P) & 3, qa = it / (4 * P);
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1) u[l1] = ws[((qa * 4 + k2a) * 4 + l1) * P + k1];
        bfly<4>(u, -1);
    }
    wave_sync();   // the slice is free for the caller again
    if constexpr (P == 4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = u[j];
    } else {
        // item lane mod 32 holds X[item + 32 qb]; element lane + 64 j is
        // X[(lane mod 32) + 32 ((lane >> 5) + 2 j)]
        const bool hi = lane >= 32;
        x[0] = hi ? u[1] : u[0];
        x[1] = hi ? u[3] : u[2];
    }
}

// Values of the extension of F (K = WK_DST: odd about z = 0 and nz, interior only;
// WK_DCT: even; WK_NONE: zero) at the line positions lane + 64 j, j < P.
template <int NZ, int K>
__device__ inline void wext(const WC &c, const WF<NZ> &F, double (&x)[WDim<NZ>::P]) {
    constexpr int S = WDim<NZ>::S, P = WDim<NZ>::P;
    if constexpr (K == WK_NONE) {
#pragma unroll
        for (int j = 0; j < P; ++j) x[j] = 0.0;
        return;
    } else {
        const int l = c.lane;
        const int rl = (64 - l) & 63;
#pragma unroll
        for (int j = 0; j < S; ++j) x[j] = F.v[j];
#pragma unroll
        for (int j = S; j < P; ++j) {
            // position l + 64 j mirrors z = 2 nz - l - 64 j: lane 64 - l, slot P-1-j
            // (l >= 1), or lane 0's own slot P - j (slot S is the top lid)
            const double r = __shfl(F.v[P - 1 - j], rl, 64);
            const double own = (P - j == S) ? F.top : F.v[(P - j) % S];
            const double val = l == 0 ? own : r;
            x[j] = (K == WK_DST) ? -val : val;
        }
        if (K == WK_DST && l == 0) {
            x[0] = 0.0;
            x[S] = 0.0;
        }
    }
}

// Two z transforms of this wave's column in one FFT: F0 in the real part, F1 in
// the imaginary part.  DST-I replaces z = 1..nz-1 and keeps the lids; DCT-I
// replaces z = 0..nz.  Same kinds separate directly; a DCT paired with a DST is
// separated through the mirror X(2nz - k).  scale = 1/sqrt(2 nz).
template <int NZ, int K0, int K1>
__device__ inline void wxform(const WC &c, WF<NZ> &F0, WF<NZ> &F1, double2 *ws, const WTw<WDim<NZ>::P> &tw,
                              double scale) {
    constexpr int S = WDim<NZ>::S, P = WDim<NZ>::P;
    double a[P], b[P];
    wext<NZ, K0>(c, F0, a);
    wext<NZ, K1>(c, F1, b);
    double2 x[P];
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = make_double2(a[j], b[j]);
    if constexpr (P <= 4)
        wfft4<P>(x, ws, tw, c.lane);
    else
        wfft<P>(x, ws, tw, c.lane);
    const int l = c.lane;
    if constexpr (K0 == K1 || K1 == WK_NONE) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (K0 == WK_DST) {
                if (l + 64 * s > 0) {
                    F0.v[s] = -x[s].y * scale;
                    if (K1 == WK_DST) F1.v[s] = x[s].x * scale;
                }
            } else {
                F0.v[s] = x[s].x * scale;
                if (K1 == WK_DCT) F1.v[s] = x[s].y * scale;
            }
        }
        if (K0 == WK_DCT) {
            F0.top = rdlane(x[S].x, 0) * scale;
            if (K1 == WK_DCT) F1.top = rdlane(x[S].y, 0) * scale;
        }
    } else {
        static_assert(K0 == WK_DCT && K1 == WK_DST, "mixed pair: DCT in the real part, DST in the imaginary");
        // even a: A(k) real, even; odd b: i B(k) real, odd.  X(k) = A(k) + i B(k)
        // A(k) = (X(k) + X(2nz-k)) / 2, i B(k) = (X(k) - X(2nz-k)) / 2
        const int rl = (64 - l) & 63;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double r = __shfl(x[P - 1 - s].x, rl, 64);
            const double own = x[(P - s) % P].x;
            const double xm = l == 0 ? own : r;
            F0.v[s] = 0.5 * (x[s].x + xm) * scale;
            if (l + 64 * s > 0) F1.v[s] = 0.5 * (x[s].x - xm) * scale;
        }
        F0.top = rdlane(x[S].x, 0) * scale;
    }
}

// LDS exchange between the four member waves: each wave publishes up to three
// fields in its own slice (BUF = 0) or two in the second region (BUF = 1); readers
// take member m' after the barrier.
template <int NZ, int BUF>
__device__ inline double *xslot(double2 *wsm, int member, int field) {
    if constexpr (BUF == 0)
        return reinterpret_cast<double *>(wsm + (size_t)member * WDim<NZ>::WS) + field * (NZ + 1);
    else
        return reinterpret_cast<double *>(wsm + (size_t)4 * WDim<NZ>::WS) + (member * 2 + field) * (NZ + 1);
}
template <int NZ, int BUF = 0>
__device__ inline void xput(const WC &c, double2 *wsm, int field, const WF<NZ> &F) {
    double *p = xslot<NZ, BUF>(wsm, c.m, field);
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) p[c.lane + 64 * s] = F.v[s];
    if (c.lane == 0) p[NZ] = F.top;
}
template <int NZ, int BUF = 0>
__device__ inline void xget(const WC &c, double2 *wsm, int member, int field, WF<NZ> &F) {
    const double *p = xslot<NZ, BUF>(wsm, member, field);
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) F.v[s] = p[c.lane + 64 * s];
    F.top = p[NZ];
}

// RK4 / cn2 update of this wave's column from the tendency S (same formulas as the
// quad kernels: impl_rk4.f90:212-364, cn2.f90:92-181).  The stored operands of one
// component are loaded ahead of its update (wrk_load), so that their latency overlaps
// the partner exchange and the previous component's update.
template <int NZ>
struct WOps {
    WF<NZ> a, b;
};
// KEEP: the next stage's input stays in registers (fused with the next vor2vel
// pass), so every member -- also one that repeats another member's column -- forms it
template <int NZ, bool KEEP = false>
__device__ __attribute__((always_inline)) inline void wrk_load(const WC &c, const RkArgs &rk, int mode, int stage, int cc, WOps<NZ> &o) {
    if (!KEEP && !c.store) return;
    // fixed destinations (a: q, qdi or qdf; b: qdf at RK4 stages 2 and 3)
    const bool two = mode != ZT_CN2 && (stage == 2 || stage == 3);
    const double *pa = (stage == 1) ? rk.q[cc] : (two ? rk.qdi[cc] : rk.qdf[cc]);
    wload(c, pa, o.a);
    if (two) wload(c, rk.qdf[cc], o.b);
}

// The column's RK coefficients.  wrk_setup reads their inputs (the step's scalars,
// the filter and dissipation of the column: wave-uniform) where they are used;
// the standalone tendency kernel reads them at its top instead (wrk_inputs), so that
// those loads overlap the column loads rather than wait, in the middle of the pass,
// behind the operands in flight (measured: 0.396 -> 0.380 ms per step at 256^3; in
// the fused kernel the early reads measured slower, 2.37 -> 2.42 ms).
struct WRk {
    double dt, f0, vd, em, epf;
};
struct WRkIn {
    double dt, visc, dfac, fx, fy, vdiss;
};
__device__ inline WRkIn wrk_inputs(const WC &c, const RkArgs &rk, const ZTabs &tb) {
    WRkIn in;
    const double *sc = rk.scal + (size_t)blockIdx.y * SC_NSCAL;
    in.dt = sc[SC_DT];
    in.visc = sc[SC_VISC];
    in.dfac = sc[SC_DFAC];
    in.fx = tb.filx[c.kx];
    in.fy = tb.fily[c.ky];
    in.vdiss = rk.vdiss ? rk.vdiss[c.colidx] : 0.0;
    return in;
}
template <int NZ>
__device__ inline WRk wrk_setup_from(const WC &c, const RkArgs &rk, int mode, const ZTabs &tb, const WRkIn &in) {
    WRk r;
    r.dt = in.dt;
    if (c.q00)
        r.f0 = 1.0;
    else if (tb.hou_li)
        r.f0 = exp(in.fx + in.fy);
    else
        r.f0 = in.fx * in.fy;
    double kp = 1.0;
    for (int i = 0; i < tb.nnu; ++i) kp *= c.k2l2;
    const double vhdis = in.visc * kp;
    if (mode == ZT_CN2) {
        r.vd = 1.0 / (1.0 + in.dfac * vhdis);
        r.em = r.epf = 0.0;
    } else {
        const double D = rk.vdiss ? in.vdiss : in.dfac * vhdis;
        const double ep = exp(D);
        r.em = 1.0 / ep;
        r.epf = ep * r.f0;
        r.vd = 0.0;
    }
    return r;
}
template <int NZ>
__device__ inline WRk wrk_setup(const WC &c, const RkArgs &rk, int mode, const ZTabs &tb) {
    WRk r;
    const double *sc = rk.scal + (size_t)blockIdx.y * SC_NSCAL;
    r.dt = sc[SC_DT];
    if (c.q00)
        r.f0 = 1.0;
    else if (tb.hou_li)
        r.f0 = exp(tb.filx[c.kx] + tb.fily[c.ky]);
    else
        r.f0 = tb.filx[c.kx] * tb.fily[c.ky];
    double kp = 1.0;
    for (int i = 0; i < tb.nnu; ++i) kp *= c.k2l2;
    const double vhdis = sc[SC_VISC] * kp;
    if (mode == ZT_CN2) {
        r.vd = 1.0 / (1.0 + sc[SC_DFAC] * vhdis);
        r.em = r.epf = 0.0;
    } else {
        const double D = rk.vdiss ? rk.vdiss[c.colidx] : sc[SC_DFAC] * vhdis;
        const double ep = exp(D);
        r.em = 1.0 / ep;
        r.epf = ep * r.f0;
        r.vd = 0.0;
    }
    return r;
}

// KEEP (RK4 stages 1-3): the stage's result q -- the next stage's svor -- goes to
// keep (every lane holds the top value) instead of memory; qdi / qdf are stored.
template <int NZ, bool KEEP = false>
__device__ __attribute__((always_inline)) inline void wrk_apply(const WC &c, const RkArgs &rk, int mode, int stage, int cc, const WF<NZ> &S,
                                 const WOps<NZ> &o, const WRk &r, const ZTabs &tb, WF<NZ> *keep = nullptr) {
    if (!KEEP && !c.store) return;
    constexpr int NS = WDim<NZ>::S;
    const int l = c.lane;
    const double dt = r.dt;
    if (mode == ZT_CN2) {
        const double dt2 = 0.5 * dt;
#pragma unroll
        for (int s = 0; s <= NS; ++s) {
            const int z = s < NS ? l + 64 * s : NZ;
            if (s == NS && l != 0) continue;
            const double t = s < NS ? S.v[s % NS] : S.top;
            const double a0 = s < NS ? o.a.v[s % NS] : o.a.top;
            double fz = r.f0;
            if (z > 0 && z < NZ && !c.q00) fz = tb.hou_li ? r.f0 * exp(tb.filz[z]) : r.f0 * tb.filz[z];
            const size_t i = c.col + z;
            double vs;
            if (stage == 1) {
                vs = a0 + dt2 * t;
                rk.qdf[cc][i] = vs;
            } else {
                vs = a0;
            }
            rk.q[cc][i] = r.vd * (fz * (vs + dt2 * t));
        }
        return;
    }
    const double dt2 = 0.5 * dt, dt3 = dt / 3.0, dt6 = dt / 6.0;
    const double em = r.em, epf = r.epf;
    // with KEEP every lane forms the (wave-uniform) top value; lane 0 stores it
    const bool st = c.store;
#pragma unroll
    for (int s = 0; s <= NS; ++s) {
        if (!KEEP && s == NS && l != 0) continue;
        const bool mine = s < NS || l == 0;
        const int z = s < NS ? l + 64 * s : NZ;
        const double t = s < NS ? S.v[s % NS] : S.top;
        const double a0 = s < NS ? o.a.v[s % NS] : o.a.top;
        const size_t i = c.col + z;
        double q;
        if (stage == 2 || stage == 3) {
            const double b0 = s < NS ? o.b.v[s % NS] : o.b.top;
            const double sq = epf * t;
            q = (stage == 2 ? em : em * em) * (a0 + (stage == 2 ? dt2 : dt) * sq);
            if (st && mine) rk.qdf[cc][i] = b0 + dt3 * sq;
        } else if (stage == 1) {
            const double sq = r.f0 * t;
            q = (a0 + dt2 * sq) * em;
            if (st && mine) {
                rk.qdi[cc][i] = a0;
                rk.qdf[cc][i] = a0 + dt6 * sq;
            }
        } else {
            const double sq = (epf * epf) * t;
            q = (em * em) * (a0 + dt6 * sq);
        }
        if constexpr (KEEP) {
            if (s < NS)
                keep->v[s % NS] = q;
            else
                keep->top = q;
        } else {
            rk.q[cc][i] = q;
        }
    }
}

template <int NZ>
__device__ inline void wzero(WF<NZ> &F) {
#pragma unroll
    for (int s = 0; s < WDim<NZ>::S; ++s) F.v[s] = 0.0;
    F.top = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// vor2vel, vertical part (src/inversion/inversion.f90:23-226) for one quad, one
// wave per member.  Same inputs and outputs as the quad-per-thread k_zinv.
// ---------------------------------------------------------------------------
// OCC > 1 asks the compiler for OCC waves per SIMD (fewer registers, possibly
// spills); OCC = 1 leaves it free, except at nz <= 128, where it asks for four
// (12-36 bytes per lane of spills, measured faster than three waves without).  Both
// variants are built; the launch shape (and the ensemble tuner) choose.
// The body: svor from memory, or (reg != nullptr) the three components already in
// this wave's registers (the fused tendency + next-stage pass below).
template <int NZ>
__device__ __attribute__((always_inline)) inline void zinv_body(const ZInvArgs &a, const WC &c, double2 *wsm,
                                                                const WF<NZ> *reg0, const WF<NZ> *reg1,
                                                                const WF<NZ> *reg2, const WSeed &sd,
                                                                const WTw<WDim<NZ>::P> &tw) {
    constexpr int S = WDim<NZ>::S;
    const Grid g = a.g;
    double2 *ws = wsm + (size_t)c.m * WDim<NZ>::WS;
    const double sc = a.zscale;
    const int l = c.lane;
    const int m1 = c.m ^ 1, m2 = c.m ^ 2;

    WF<NZ> Xi, Et, C, D, E;
    if (reg0) {
        Xi = *reg0;
        Et = *reg1;
        C = *reg2;
    } else {
        wload(c, a.svor[0], Xi);
        wload(c, a.svor[1], Et);
        wload(c, a.svor[2], C);
    }

    // D = B_x - A_y (inversion.f90:41-45)
    constexpr bool DB = WDim<NZ>::DB;
    constexpr int XB = DB ? 1 : 0;
    xput<NZ, XB>(c, wsm, 0, Et);
    xput<NZ, XB>(c, wsm, 1, Xi);
    __syncthreads();
    {
        WF<NZ> t1, t2;
        xget<NZ, XB>(c, wsm, m1, 0, t1);
        xget<NZ, XB>(c, wsm, m2, 1, t2);
#pragma unroll
        for (int s = 0; s < S; ++s) D.v[s] = c.fx * t1.v[s] - c.fy * t2.v[s];
        D.top = c.fx * t1.top - c.fy * t2.top;
    }
    if constexpr (!DB) __syncthreads();   // the slices are the FFTs' next

    const WPhi<NZ> ph = make_phi<NZ>(c, g, sd);
    // C = combine(zeta); E = decompose(central_diffz(C)) (46-48)
    {
        const double c0 = wlo(C), cn = C.top;
        WF<NZ> none;
        wxform<NZ, WK_DST, WK_NONE>(c, C, none, ws, tw, sc);
        wharm(c, ph, C, c0, cn, 1.0);
        wdiff(c, g, C, E);
        const double e0 = wlo(E), en = E.top;
        wharm(c, ph, E, e0, en, -1.0);
        wxform<NZ, WK_DST, WK_NONE>(c, E, none, ws, tw, sc);
    }

    // projection A = k2l2i (E_x + D_y), B = k2l2i (E_y - D_x); (0,0) column kept (56-75)
    WF<NZ> Xp, Ep;
    xput(c, wsm, 0, E);
    xput(c, wsm, 1, D);
    __syncthreads();
    {
        WF<NZ> e1, e2, d1, d2;
        xget(c, wsm, m1, 0, e1);
        xget(c, wsm, m2, 0, e2);
        xget(c, wsm, m1, 1, d1);
        xget(c, wsm, m2, 1, d2);
        if (c.q00) {
            Xp = Xi;
            Ep = Et;
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                Xp.v[s] = c.k2l2i * (c.fx * e1.v[s] + c.fy * d2.v[s]);
                Ep.v[s] = c.k2l2i * (c.fy * e2.v[s] - c.fx * d1.v[s]);
            }
            Xp.top = c.k2l2i * (c.fx * e1.top + c.fy * d2.top);
            Ep.top = c.k2l2i * (c.fy * e2.top - c.fx * d1.top);
        }
    }
    if constexpr (!DB) __syncthreads();   // the next exchange writes the same slots
    if (a.svor_out[0]) {
        wstore(c, a.svor_out[0], Xp);
        wstore(c, a.svor_out[1], Ep);
    }

    // source of w: ds = A'_y - B'_x (86-90)
    WF<NZ> Ds;
    xput<NZ, XB>(c, wsm, 0, Xp);
    xput<NZ, XB>(c, wsm, 1, Ep);
    __syncthreads();
    {
        WF<NZ> x2, e1;
        xget<NZ, XB>(c, wsm, m2, 0, x2);
        xget<NZ, XB>(c, wsm, m1, 1, e1);
#pragma unroll
        for (int s = 0; s < S; ++s) Ds.v[s] = c.fy * x2.v[s] - c.fx * e1.v[s];
        Ds.top = c.fy * x2.top - c.fx * e1.top;
    }
    if constexpr (!DB) __syncthreads();   // the slices are the FFTs' next
    const double d0 = wlo(Ds), dn = Ds.top;
    const double xp0 = wlo(Xp), xpn = Xp.top, ep0 = wlo(Ep), epn = Ep.top;

    // bs (interior) and es (all z): boundary parts (94-104); green and d/dz of the
    // sine part (107-122); quad (0,0): the mean flow (150-165) in members 1, 2
    WF<NZ> Bs, As;
    const WExp<NZ> wx = make_wexp<NZ>(sd);
#pragma unroll
    for (int s = 0; s <= S; ++s) {
        const int z = s < S ? l + 64 * s : NZ;
        const Hyp h = hyp_from(g, c.K, c.k2l2i, c.q00, z, wx.ep[s], wx.em[s], wx.ef, wx.div);
        const double rkz = a.rkz_sc * (double)z;
        const double green = -1.0 / (c.k2l2 + rkz * rkz);
        const double bs = d0 * h.thetam + dn * h.thetap;
        const double es = d0 * h.dthetam + dn * h.dthetap;
        const bool interior = z > 0 && z < NZ;
        const double dsv = s < S ? Ds.v[s % S] : Ds.top;
        const double dg = interior ? green * dsv : dsv;
        double as = interior ? rkz * dg : 0.0;
        if (c.q00 && (c.m == 1 || c.m == 2)) {
            const double rkzi = interior ? 1.0 / rkz : 0.0;
            const double src = c.m == 1 ? -(s < S ? Ep.v[s % S] : Ep.top) : (s < S ? Xp.v[s % S] : Xp.top);
            as = interior ? rkzi * src : 0.0;
        }
        if (s < S) {
            Bs.v[s] = bs;
            E.v[s] = es;
            Ds.v[s] = dg;
            As.v[s] = as;
        } else {
            Bs.top = bs;
            E.top = es;
            Ds.top = dg;
            As.top = as;
        }
    }
    wxform<NZ, WK_DCT, WK_DST>(c, As, Ds, ws, tw, sc);

    // complete w and dw/dz (134-141)
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int z = l + 64 * s;
        Ds.v[s] = z == 0 ? 0.0 : Ds.v[s] + Bs.v[s];
        E.v[s] = E.v[s] + As.v[s];
    }
    Ds.top = 0.0;
    E.top = E.top + As.top;

    // gamtop = Lz/2 (phip00^2 - 1/3), gambot(iz) = gamtop(nz-iz) (350-369)
    WF<NZ> Bar;
    if (c.q00) {
#pragma unroll
        for (int s = 0; s <= S; ++s) {
            const int z = s < S ? l + 64 * s : NZ;
            double zm, zp, zm2, zp2;
            zm_zp(g, z, zm, zp);
            zm_zp(g, NZ - z, zm2, zp2);
            const double pt = zp / g.lz, pb = zp2 / g.lz;
            const double gamtop = 0.5 * g.lz * (pt * pt - 1.0 / 3.0);
            const double gambot = 0.5 * g.lz * (pb * pb - 1.0 / 3.0);
            const double as = s < S ? As.v[s % S] : As.top;
            const double v = c.m == 1 ? as + epn * gamtop - ep0 * gambot : as - xpn * gamtop + xp0 * gambot;
            if (s < S)
                Bar.v[s] = v;
            else
                Bar.top = v;
        }
    }
    wxform<NZ, WK_DST, WK_DST>(c, Xp, Ep, ws, tw, sc);
    wharm(c, ph, Xp, xp0, xpn, 1.0);
    wharm(c, ph, Ep, ep0, epn, 1.0);

    // u = k2l2i (es_x + cs_y), v = k2l2i (es_y - cs_x); (0,0): ubar, vbar (169-213)
    WF<NZ> U, V;
    xput(c, wsm, 0, (c.q00 && c.m == 1) ? Bar : E);
    xput(c, wsm, 1, (c.q00 && c.m == 2) ? Bar : C);
    __syncthreads();
    {
        WF<NZ> e1, e2, c1, c2;
        xget(c, wsm, m1, 0, e1);
        xget(c, wsm, m2, 0, e2);
        xget(c, wsm, m1, 1, c1);
        xget(c, wsm, m2, 1, c2);
        if (c.q00) {
            U = e1;     // member 0 stores: ubar from member 1, vbar from member 2
            V = c2;
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                U.v[s] = c.k2l2i * (c.fx * e1.v[s] + c.fy * c2.v[s]);
                V.v[s] = c.k2l2i * (c.fy * e2.v[s] - c.fx * c1.v[s]);
            }
            U.top = c.k2l2i * (c.fx * e1.top + c.fy * c2.top);
            V.top = c.k2l2i * (c.fy * e2.top - c.fx * c1.top);
        }
    }
    wstore(c, a.out[0], U);
    wstore(c, a.out[1], V);
    wstore(c, a.out[2], Ds);
    wstore(c, a.out[3], Xp);
    wstore(c, a.out[4], Ep);
    wstore(c, a.out[5], C);
    if (a.mid[0]) {
        WF<NZ> M;
        wmid(c, Xp, M);
        wstore(c, a.mid[0], M);
        wmid(c, Ep, M);
        wstore(c, a.mid[1], M);
        wmid(c, C, M);
        wstore(c, a.mid[2], M);
    }
}

template <int NZ, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NZ <= 128 && OCC == 1) ? 4 : OCC, 8))) k_zinv_w(ZInvArgs a) {
    extern __shared__ double2 wsm[];
    if (gated_off(a.g.gate)) return;
    const WC c = make_wc(a.g, a.tb, (long)blockIdx.y * a.bt.vol);
    zinv_body<NZ>(a, c, wsm, nullptr, nullptr, nullptr, make_wseed(c, a.g), make_tw<WDim<NZ>::P>(a.twz, c.lane));
}

// ---------------------------------------------------------------------------
// vorticity_tendency, vertical part (inversion.f90:298-371), fused with one RK4
// substep or cn2 update; one wave per member.  Same contract as k_ztend.
// ---------------------------------------------------------------------------
// The body; KEEP (RK4 stages 1-3): the stage's result stays in keep[3] for the
// next stage's vor2vel pass instead of going to memory.
template <int NZ, bool KEEP>
__device__ __attribute__((always_inline)) inline void ztend_body(const ZTendArgs &a, const WC &c, double2 *wsm,
                                                                 WF<NZ> *keep0, WF<NZ> *keep1, WF<NZ> *keep2,
                                                                 const WSeed &sd,
                                                                 const WTw<WDim<NZ>::P> &tw, const WRkIn *rin) {
    constexpr int S = WDim<NZ>::S;
    const Grid g = a.g;
    double2 *ws = wsm + (size_t)c.m * WDim<NZ>::WS;
    const double sc = a.zscale;

    WF<NZ> R, Q, Pp, FQ, FP;
    wload(c, a.rqp[0], R);
    wload(c, a.rqp[1], Q);
    wload(c, a.rqp[2], Pp);
    // central_diffz of q and p (342, 356)
    wdiff(c, g, Q, FQ);
    wdiff(c, g, Pp, FP);
    // field_decompose of r, q, p, dq/dz, dp/dz: subtract the harmonic part, then DST
    {
        const WPhi<NZ> ph = make_phi<NZ>(c, g, sd);
        const double r0 = wlo(R), rn = R.top, q0 = wlo(Q), qn = Q.top, p0 = wlo(Pp), pn = Pp.top;
        const double fq0 = wlo(FQ), fqn = FQ.top, fp0 = wlo(FP), fpn = FP.top;
        wharm(c, ph, R, r0, rn, -1.0);
        wharm(c, ph, Q, q0, qn, -1.0);
        wharm(c, ph, Pp, p0, pn, -1.0);
        wharm(c, ph, FQ, fq0, fqn, -1.0);
        wharm(c, ph, FP, fp0, fpn, -1.0);
    }
    wxform<NZ, WK_DST, WK_DST>(c, R, FQ, ws, tw, sc);
    wxform<NZ, WK_DST, WK_DST>(c, FP, Q, ws, tw, sc);
    {
        WF<NZ> none;
        wxform<NZ, WK_DST, WK_NONE>(c, Pp, none, ws, tw, sc);
    }
    // operands of the first component's update: in flight across the exchange
    const bool upd = KEEP || a.mode != ZT_STORE;
    WOps<NZ> ops;
    if (upd) wrk_load<NZ, KEEP>(c, a.rk, a.mode, a.stage, 0, ops);
    // dxi/dt = r_y - q_z ; deta/dt = p_z - r_x ; dzeta/dt = q_x - p_y (341-367)
    xput(c, wsm, 0, R);
    xput(c, wsm, 1, Q);
    xput(c, wsm, 2, Pp);
    __syncthreads();
    WF<NZ> St[3];
    {
        WF<NZ> r1, r2, q1, p2;
        xget(c, wsm, c.m ^ 1, 0, r1);
        xget(c, wsm, c.m ^ 2, 0, r2);
        xget(c, wsm, c.m ^ 1, 1, q1);
        xget(c, wsm, c.m ^ 2, 2, p2);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            St[0].v[s] = c.fy * r2.v[s] - FQ.v[s];
            St[1].v[s] = FP.v[s] - c.fx * r1.v[s];
            St[2].v[s] = c.fx * q1.v[s] - c.fy * p2.v[s];
        }
        St[0].top = c.fy * r2.top - FQ.top;
        St[1].top = FP.top - c.fx * r1.top;
        St[2].top = c.fx * q1.top - c.fy * p2.top;
    }
    if (a.extra[0]) {   // closure tendency
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            WF<NZ> ex;
            wload(c, a.extra[cc], ex);
#pragma unroll
            for (int s = 0; s < S; ++s) St[cc].v[s] += ex.v[s];
            St[cc].top += ex.top;
        }
    }
    if (!upd) {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) wstore(c, a.svorts[cc], St[cc]);
        return;
    }
    const WRk rr = rin ? wrk_setup_from<NZ>(c, a.rk, a.mode, a.tb, *rin) : wrk_setup<NZ>(c, a.rk, a.mode, a.tb);
    if constexpr (NZ >= 512) {
        // nz = 512: the three components written out in sequence, the kept results in
        // three separate variables (as one array indexed in the loop below they were
        // kept in scratch memory, 224 bytes per lane): fused kernel 30.6 -> 28.2-29.8
        // ms per 512^3 step (two builds with identical code for it measured both ends:
        // its 96 KB of code exceed the instruction cache, and where the library places
        // it matters).  At nz <= 256 the loop and the array measured as fast or faster.
        WOps<NZ> o1, o2;
        wrk_load<NZ, KEEP>(c, a.rk, a.mode, a.stage, 1, o1);
        wrk_apply<NZ, KEEP>(c, a.rk, a.mode, a.stage, 0, St[0], ops, rr, a.tb, KEEP ? keep0 : nullptr);
        wrk_load<NZ, KEEP>(c, a.rk, a.mode, a.stage, 2, o2);
        wrk_apply<NZ, KEEP>(c, a.rk, a.mode, a.stage, 1, St[1], o1, rr, a.tb, KEEP ? keep1 : nullptr);
        wrk_apply<NZ, KEEP>(c, a.rk, a.mode, a.stage, 2, St[2], o2, rr, a.tb, KEEP ? keep2 : nullptr);
    } else {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            WOps<NZ> nxt;
            if (cc < 2) wrk_load<NZ, KEEP>(c, a.rk, a.mode, a.stage, cc + 1, nxt);
            // (below nz = 512 the kept results are one array: keep1 = keep0 + 1, ...)
            wrk_apply<NZ, KEEP>(c, a.rk, a.mode, a.stage, cc, St[cc], ops, rr, a.tb, KEEP ? keep0 + cc : nullptr);
            if (cc < 2) ops = nxt;
        }
    }
}

template <int NZ, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NZ <= 128 && OCC == 1) ? 4 : OCC, 8))) k_ztend_w(ZTendArgs a) {
    extern __shared__ double2 wsm[];
    if (gated_off(a.g.gate)) return;
    const WC c = make_wc(a.g, a.tb, (long)blockIdx.y * a.bt.vol);
    const WRkIn rin = wrk_inputs(c, a.rk, a.tb);
    ztend_body<NZ, false>(a, c, wsm, nullptr, nullptr, nullptr, make_wseed(c, a.g), make_tw<WDim<NZ>::P>(a.twz, c.lane),
                          &rin);
}

// RK4 stage s (1..3) of the tendency pass and the vor2vel z pass of stage s+1 in
// one kernel: the new svor of each column never leaves the registers of the wave
// that formed it (saves writing and re-reading three volumes per stage boundary).
// At nz = 512 one barrier separates the two partner exchanges' uses of the block's
// LDS.
template <int NZ, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NZ <= 128 && OCC == 1) ? 4 : OCC, 8)))
k_ztend_zinv_w(ZTendArgs a, ZInvArgs b) {
    extern __shared__ double2 wsm[];
    if (gated_off(a.g.gate)) return;
    const WC c = make_wc(a.g, a.tb, (long)blockIdx.y * a.bt.vol);
    const WSeed sd = make_wseed(c, a.g);
    const WTw<WDim<NZ>::P> tw = make_tw<WDim<NZ>::P>(a.twz, c.lane);   // one set of roots for both passes
    // the stage results: three separate variables at nz = 512, one array below
    if constexpr (NZ >= 512) {
        WF<NZ> q0, q1, q2;
        ztend_body<NZ, true>(a, c, wsm, &q0, &q1, &q2, sd, tw, nullptr);
        __syncthreads();   // one LDS region: the exchanges of both passes use the slices
        zinv_body<NZ>(b, c, wsm, &q0, &q1, &q2, sd, tw);
    } else {
        static_assert(WDim<NZ>::DB, "second exchange region below nz = 512");
        // (with the second region the vor2vel pass's first exchange does not touch the
        // slices the tendency pass's exchange used)
        WF<NZ> q[3];
        ztend_body<NZ, true>(a, c, wsm, q, q + 1, q + 2, sd, tw, nullptr);
        zinv_body<NZ>(b, c, wsm, q, q + 1, q + 2, sd, tw);
    }
}

// Single z operators on one volume (kernel checks of the wave path).
template <int NZ>
__global__ void __launch_bounds__(256) k_zop_w(ZOpArgs a) {
    constexpr int S = WDim<NZ>::S, P = WDim<NZ>::P;
    extern __shared__ double2 wsm[];
    const Grid g = a.g;
    const WC c = make_wc(g, a.tb, 0);
    double2 *ws = wsm + (size_t)c.m * WDim<NZ>::WS;
    const WTw<P> tw = make_tw<P>(a.twz, c.lane);
    const double sc = a.zscale;
    WF<NZ> F, O, none;
    wload(c, a.in, F);
    switch (a.op) {
        case ZOP_DST:
            wxform<NZ, WK_DST, WK_NONE>(c, F, none, ws, tw, sc);
            F.top = 0.0;   // stafft dst sets x(n) = 0 (stafft.f90:546-549)
            wstore(c, a.out, F);
            break;
        case ZOP_DCT:
            wxform<NZ, WK_DCT, WK_NONE>(c, F, none, ws, tw, sc);
            wstore(c, a.out, F);
            break;
        case ZOP_DIFFZ:
            wdiff(c, g, F, O);
            wstore(c, a.out, O);
            break;
        case ZOP_COMBINE: {
            const WPhi<NZ> ph = make_phi<NZ>(c, g, make_wseed(c, g));
            const double lo = wlo(F), hi = F.top;
            wxform<NZ, WK_DST, WK_NONE>(c, F, none, ws, tw, sc);
            wharm(c, ph, F, lo, hi, 1.0);
            wstore(c, a.out, F);
            break;
        }
        case ZOP_DECOMPOSE: {
            const WPhi<NZ> ph = make_phi<NZ>(c, g, make_wseed(c, g));
            wharm(c, ph, F, wlo(F), F.top, -1.0);
            wxform<NZ, WK_DST, WK_NONE>(c, F, none, ws, tw, sc);
            wstore(c, a.out, F);
            break;
        }
        case ZOP_POISSON: {
            wxform<NZ, WK_DCT, WK_NONE>(c, F, none, ws, tw, sc);
#pragma unroll
            for (int s = 0; s <= S; ++s) {
                const int z = s < S ? c.lane + 64 * s : NZ;
                const double rkz = a.rkz_sc * (double)z;
                const double green = z == 0 ? -c.k2l2i : -1.0 / (c.k2l2 + rkz * rkz);
                if (s < S)
                    F.v[s % S] = green * F.v[s % S];
                else
                    F.top = green * F.top;
            }
            wxform<NZ, WK_DCT, WK_NONE>(c, F, none, ws, tw, sc);
            wstore(c, a.out, F);
            break;
        }
        default:
            break;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
bool zwave_supported(int nz) { return nz == 64 || nz == 128 || nz == 256 || nz == 512; }

size_t zwave_lds_bytes(int nz) {
    // four wave slices, and for nz <= 256 the second exchange region
    return 4 * (size_t)72 * (size_t)(nz / 32) * sizeof(double2) + (nz <= 256 ? (size_t)8 * (nz + 1) * sizeof(double) : 0);
}

#define PS3D_ZW_SIZES(X) X(64) X(128) X(256) X(512)

hipError_t configure_zwave() {
    hipError_t e = hipSuccess;
#define SETW(NZ_)                                                                                            \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_zinv_w<NZ_, 1>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_ztend_w<NZ_, 1>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_zinv_w<NZ_, 3>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_ztend_w<NZ_, 3>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_ztend_zinv_w<NZ_, 1>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_ztend_zinv_w<NZ_, 3>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)zwave_lds_bytes(NZ_));                                                  \
    if (e == hipSuccess)                                                                                     \
        e = hipFuncSetAttribute((const void *)(k_zop_w<NZ_>), hipFuncAttributeMaxDynamicSharedMemorySize,    \
                                (int)zwave_lds_bytes(NZ_));
    PS3D_ZW_SIZES(SETW)
#undef SETW
    return e;
}

#define ZW_DISPATCH(KERNEL, NQ, NB, ARGS)                                                                   \
    switch (nz) {                                                                                       \
        case 64: hipLaunchKernelGGL(KERNEL<64>, dim3(NQ, NB), dim3(256), zwave_lds_bytes(64), stream, ARGS); break; \
        case 128: hipLaunchKernelGGL(KERNEL<128>, dim3(NQ, NB), dim3(256), zwave_lds_bytes(128), stream, ARGS); break; \
        case 256: hipLaunchKernelGGL(KERNEL<256>, dim3(NQ, NB), dim3(256), zwave_lds_bytes(256), stream, ARGS); break; \
        case 512: hipLaunchKernelGGL(KERNEL<512>, dim3(NQ, NB), dim3(256), zwave_lds_bytes(512), stream, ARGS); break; \
        default: return hipErrorInvalidValue;                                                           \
    }

// variant 1: vor2vel pass free, tendency pass at three waves per SIMD (measured best
// at nz = 256); 2: both free; 3: both at three waves per SIMD
#define ZW_DISPATCH2(KERNEL, NQ, NB, ARGS, OCC3)                                                        \
    switch (nz * 10 + ((OCC3) ? 3 : 1)) {                                                          \
        case 641: hipLaunchKernelGGL((KERNEL<64, 1>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(64), stream, ARGS); break; \
        case 1281: hipLaunchKernelGGL((KERNEL<128, 1>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(128), stream, ARGS); break; \
        case 2561: hipLaunchKernelGGL((KERNEL<256, 1>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(256), stream, ARGS); break; \
        case 5121: hipLaunchKernelGGL((KERNEL<512, 1>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(512), stream, ARGS); break; \
        case 643: hipLaunchKernelGGL((KERNEL<64, 3>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(64), stream, ARGS); break; \
        case 1283: hipLaunchKernelGGL((KERNEL<128, 3>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(128), stream, ARGS); break; \
        case 2563: hipLaunchKernelGGL((KERNEL<256, 3>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(256), stream, ARGS); break; \
        case 5123: hipLaunchKernelGGL((KERNEL<512, 3>), dim3(NQ, NB), dim3(256), zwave_lds_bytes(512), stream, ARGS); break; \
        default: return hipErrorInvalidValue;                                                           \
    }

hipError_t launch_zinv_w(int nz, int variant, int nquads, const ZInvArgs &args, hipStream_t stream) {
    ZW_DISPATCH2(k_zinv_w, nquads, (args.bt.nb > 1 ? args.bt.nb : 1), args, variant == 3);
    return hipGetLastError();
}
hipError_t launch_ztend_w(int nz, int variant, int nquads, const ZTendArgs &args, hipStream_t stream) {
    ZW_DISPATCH2(k_ztend_w, nquads, (args.bt.nb > 1 ? args.bt.nb : 1), args, variant == 1 || variant == 3);
    return hipGetLastError();
}
hipError_t launch_ztend_zinv_w(int nz, int variant, int nquads, const ZTendArgs &t, const ZInvArgs &z,
                               hipStream_t stream) {
    const int nb = t.bt.nb > 1 ? t.bt.nb : 1;
    const bool occ3 = variant == 3;
    switch (nz * 10 + (occ3 ? 3 : 1)) {
#define ZTZ(NZ_, O_)                                                                                         \
    case NZ_ * 10 + O_:                                                                                      \
        hipLaunchKernelGGL((k_ztend_zinv_w<NZ_, O_>), dim3(nquads, nb), dim3(256), zwave_lds_bytes(NZ_), stream, t, z); \
        break;
        ZTZ(64, 1) ZTZ(128, 1) ZTZ(256, 1) ZTZ(512, 1) ZTZ(64, 3) ZTZ(128, 3) ZTZ(256, 3) ZTZ(512, 3)
#undef ZTZ
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_zop_w(int nz, int nquads, const ZOpArgs &args, hipStream_t stream) {
    ZW_DISPATCH(k_zop_w, nquads, 1, args);
    return hipGetLastError();
}

}  // namespace ps3d
