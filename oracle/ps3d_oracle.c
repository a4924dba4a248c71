/*
 * ps3d_oracle.c -- CPU FP64 restatement of ps3d's pseudo-spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY (see ps3d_oracle.h).  It is the parity oracle for the
 * HIP path and, timed on the host cores, the CPU baseline.  It restates, from the
 * mathematics, the operation sequence of:
 *   src/fft/stafft.f90       forfft/revfft (Hermitian real form), dct, dst
 *   src/fft/sta3dfft.f90     wavenumbers, fftxyp2s/fftxys2p, diffx/diffy, fftsine/fftcosine
 *   src/fft/sta2dfft.f90     init2dfft (half wavenumbers hrkx)
 *   src/fft/deriv1d.f90      init_deriv (rkz)
 *   src/inversion/inversion_utils.f90   tables, filters, viscosity, combine/decompose, central_diffz
 *   src/inversion/inversion.f90         vor2vel, vorticity_tendency, source
 *   src/stepper/impl_rk4.f90            integrating-factor RK4 (unfused, as written)
 *   src/stepper/cn2.f90                 iterated Crank-Nicolson
 *   src/stepper/advance.f90             advance, adapt, get_diffusion_pre_factor
 *   src/fields/field_diagnostics.f90    KE, enstrophy, means, vorch, vorticity-mean restore
 *   src/fields/fields_derived.f90       pressure, horizontal divergence
 *   src/utils/jacobi.f90, rolling_mean.f90, constants.f90, parameters.f90
 * The 1-D FFT engine underneath is an independent mixed-radix Stockham FFT (not
 * stafft's radix kernels); DCT-I/DST-I use the symmetric 2n extension.  Both are
 * pinned against brute-force sums and against the reference's own stafft.f90
 * compiled by oracle/Makefile into oracle/_ref.
 */
#include "ps3d_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *orc_last_error(void) { return g_err; }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* constants.f90:8-49 */
static const double PI = 3.141592653589793238462643383279502884;
static const double SMALL = 1.0e-12;   /* small  */
static const double CFLMAX = 0.8;      /* cflmax */

/* ------------------------------------------------------------------------- */
/* Complex FFT engine (Stockham autosort, radices 4,2,3,5).                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, nf, fac[64];
    double *twr, *twi;   /* exp(-2 pi i m / n), m = 0..n-1 */
} cplan;

static int cplan_init(cplan *P, int n) {
    memset(P, 0, sizeof *P);
    if (n < 1) return PS3D_E_ARG;
    P->n = n;
    int rem = n, nf = 0;
    const int rad[4] = {4, 2, 3, 5};
    for (int r = 0; r < 4; ++r)
        while (rem % rad[r] == 0) { P->fac[nf++] = rad[r]; rem /= rad[r]; }
    if (rem != 1) return PS3D_E_FACTOR;   /* stafft.f90:86-96: only factors 2..6 */
    P->nf = nf;
    P->twr = malloc(sizeof(double) * n);
    P->twi = malloc(sizeof(double) * n);
    for (int m = 0; m < n; ++m) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        P->twr[m] = (double)cosl(a);
        P->twi[m] = -(double)sinl(a);
    }
    return 0;
}
static void cplan_free(cplan *P) { free(P->twr); free(P->twi); memset(P, 0, sizeof *P); }

/* In-place transform of (re, im) using (wr, wi) as work; sign = -1 forward
 * (exp(-2 pi i jk/n)), +1 backward; no scaling. */
static void cfft(const cplan *P, double *re, double *im, double *wr, double *wi, int sign) {
    const int n = P->n;
    double *sr = re, *si = im, *dr = wr, *di = wi;
    int p = 1;
    double ur[5], ui[5], vr[5], vi[5];
    for (int f = 0; f < P->nf; ++f) {
        const int R = P->fac[f], m = n / R, tstep = n / (p * R), rstep = n / R;
        for (int i = 0; i < m; ++i) {
            const int k = i % p;
            for (int t = 0; t < R; ++t) {
                double a = sr[i + t * m], b = si[i + t * m];
                if (t > 0 && k > 0) {
                    int idx = (k * t * tstep) % n;
                    double c = P->twr[idx], s = (sign < 0) ? P->twi[idx] : -P->twi[idx];
                    double na = a * c - b * s;
                    double nb = a * s + b * c;
                    a = na; b = nb;
                }
                ur[t] = a; ui[t] = b;
            }
            for (int s = 0; s < R; ++s) {
                double ar = 0.0, ai = 0.0;
                for (int t = 0; t < R; ++t) {
                    int idx = (s * t * rstep) % n;
                    double c = P->twr[idx], w = (sign < 0) ? P->twi[idx] : -P->twi[idx];
                    ar += ur[t] * c - ui[t] * w;
                    ai += ur[t] * w + ui[t] * c;
                }
                vr[s] = ar; vi[s] = ai;
            }
            const int j = (i - k) * R + k;
            for (int t = 0; t < R; ++t) { dr[j + t * p] = vr[t]; di[j + t * p] = vi[t]; }
        }
        double *tr = sr, *ti = si; sr = dr; si = di; dr = tr; di = ti;
        p *= R;
    }
    if (sr != re) { memcpy(re, sr, sizeof(double) * n); memcpy(im, si, sizeof(double) * n); }
}

/* ---- real transforms in stafft's Hermitian form ------------------------- */
/* Forward of two real lines a, b (length n) -> Hermitian forms in place
 * (stafft.f90:55-58, 196-287): xh(0)=X0/sqrt(n); xh(k)=Re Xk/sqrt(n), 1<=k<=n/2;
 * xh(n-k)=Im Xk/sqrt(n), 1<=k<(n+1)/2, with Xk = sum_j x_j exp(-2 pi i jk/n). */
static void rfft_fwd_pair(const cplan *P, double *a, double *b, double *w) {
    const int n = P->n;
    double *zr = w, *zi = w + n, *wr = w + 2 * n, *wi = w + 3 * n;
    for (int j = 0; j < n; ++j) { zr[j] = a[j]; zi[j] = b ? b[j] : 0.0; }
    cfft(P, zr, zi, wr, wi, -1);
    const double sc = 1.0 / sqrt((double)n);
    for (int k = 0; k <= n / 2; ++k) {
        int nk = (n - k) % n;
        /* A_k = (Z_k + conj Z_{n-k})/2 ; B_k = (Z_k - conj Z_{n-k})/(2i) */
        double Ar = 0.5 * (zr[k] + zr[nk]), Ai = 0.5 * (zi[k] - zi[nk]);
        double Br = 0.5 * (zi[k] + zi[nk]), Bi = -0.5 * (zr[k] - zr[nk]);
        a[k] = Ar * sc;
        if (b) b[k] = Br * sc;
        if (k > 0 && nk != k) {
            a[nk] = Ai * sc;
            if (b) b[nk] = Bi * sc;
        }
    }
}

/* Inverse of rfft_fwd_pair (stafft.f90:296-403). */
static void rfft_inv_pair(const cplan *P, double *a, double *b, double *w) {
    const int n = P->n;
    double *zr = w, *zi = w + n, *wr = w + 2 * n, *wi = w + 3 * n;
    const double sn = sqrt((double)n);
    for (int k = 0; k <= n / 2; ++k) {
        int nk = (n - k) % n;
        double Ar = a[k], Ai = 0.0, Br = b ? b[k] : 0.0, Bi = 0.0;
        if (k > 0 && nk != k) { Ai = a[nk]; Bi = b ? b[nk] : 0.0; }
        Ar *= sn; Ai *= sn; Br *= sn; Bi *= sn;
        /* Z_k = A_k + i B_k ; Z_{n-k} = conj(A_k) + i conj(B_k) */
        zr[k] = Ar - Bi; zi[k] = Ai + Br;
        if (k > 0 && nk != k) { zr[nk] = Ar + Bi; zi[nk] = -Ai + Br; }
    }
    cfft(P, zr, zi, wr, wi, +1);
    const double sc = 1.0 / (double)n;
    for (int j = 0; j < n; ++j) { a[j] = zr[j] * sc; if (b) b[j] = zi[j] * sc; }
}

/* DCT-I (stafft.f90:410-483) of x[0..n] and DST-I (489-550) of x[1..n-1]
 * (x[n] set to 0) for two columns at once through the 2n-point symmetric
 * extension. P is the plan of length 2n.  Either column pointer may be NULL. */
static void dct_pair(const cplan *P2, int n, double *a, double *b, double *w) {
    const int N = 2 * n;
    double *zr = w, *zi = w + N, *wr = w + 2 * N, *wi = w + 3 * N;
    for (int j = 0; j <= n; ++j) { zr[j] = a ? a[j] : 0.0; zi[j] = b ? b[j] : 0.0; }
    for (int j = n + 1; j < N; ++j) { zr[j] = zr[N - j]; zi[j] = zi[N - j]; }
    cfft(P2, zr, zi, wr, wi, -1);
    const double sc = 1.0 / sqrt((double)N);
    for (int k = 0; k <= n; ++k) {
        if (a) a[k] = zr[k] * sc;
        if (b) b[k] = zi[k] * sc;
    }
}
static void dst_pair(const cplan *P2, int n, double *a, double *b, double *w) {
    const int N = 2 * n;
    double *zr = w, *zi = w + N, *wr = w + 2 * N, *wi = w + 3 * N;
    zr[0] = zi[0] = 0.0; zr[n] = zi[n] = 0.0;
    for (int j = 1; j < n; ++j) {
        zr[j] = a ? a[j] : 0.0; zi[j] = b ? b[j] : 0.0;
        zr[N - j] = -zr[j]; zi[N - j] = -zi[j];
    }
    cfft(P2, zr, zi, wr, wi, -1);
    const double sc = 1.0 / sqrt((double)N);
    for (int k = 1; k < n; ++k) {
        if (a) a[k] = -zi[k] * sc;
        if (b) b[k] = zr[k] * sc;
    }
    if (a) a[n] = 0.0;
    if (b) b[n] = 0.0;
}

/* ---- stafft-style batched entry points (x(m, n), m fastest) ------------- */
static int batch_lines(int m, int n, double *x, int kind) {
    cplan P;
    int plen = (kind >= 2) ? 2 * n : n;
    int st = cplan_init(&P, plen);
    if (st) { cplan_free(&P); set_err("Factorisation not possible. Only factors from 2-6 allowed."); return st; }
    int len = (kind == 2) ? n + 1 : n;  /* dct uses 0..n; dst uses 1..n */
    int off = (kind == 3) ? 0 : 0;
    (void)off;
    #pragma omp parallel
    {
        double *buf = malloc(sizeof(double) * (2 * (len + 1) + 8 * plen));
        double *a = buf, *b = buf + len + 1, *w = buf + 2 * (len + 1);
        #pragma omp for schedule(static)
        for (int i = 0; i < m; i += 2) {
            int has_b = (i + 1 < m);
            if (kind == 3) {
                /* dst: x(i, 1:n) -> column c[1..n] */
                a[0] = 0.0; b[0] = 0.0;
                for (int j = 1; j <= n; ++j) { a[j] = x[i + (size_t)m * (j - 1)]; if (has_b) b[j] = x[i + 1 + (size_t)m * (j - 1)]; }
                dst_pair(&P, n, a, has_b ? b : NULL, w);
                for (int j = 1; j <= n; ++j) { x[i + (size_t)m * (j - 1)] = a[j]; if (has_b) x[i + 1 + (size_t)m * (j - 1)] = b[j]; }
            } else {
                for (int j = 0; j < len; ++j) { a[j] = x[i + (size_t)m * j]; if (has_b) b[j] = x[i + 1 + (size_t)m * j]; }
                if (kind == 0) rfft_fwd_pair(&P, a, has_b ? b : NULL, w);
                else if (kind == 1) rfft_inv_pair(&P, a, has_b ? b : NULL, w);
                else dct_pair(&P, n, a, has_b ? b : NULL, w);
                for (int j = 0; j < len; ++j) { x[i + (size_t)m * j] = a[j]; if (has_b) x[i + 1 + (size_t)m * j] = b[j]; }
            }
        }
        free(buf);
    }
    cplan_free(&P);
    return 0;
}
int orc_forfft(int m, int n, double *x) { return batch_lines(m, n, x, 0); }
int orc_revfft(int m, int n, double *x) { return batch_lines(m, n, x, 1); }
int orc_dct(int m, int n, double *x) { return batch_lines(m, n, x, 2); }
int orc_dst(int m, int n, double *x) { return batch_lines(m, n, x, 3); }

/* ------------------------------------------------------------------------- */
/* Jacobi eigenvalues, jacobi.f90:20-49 (givens), 241-305 (apply_rotation,   */
/* jacobi_eigenvalues), sort 309-329.  a is column-major 3x3.                 */
/* ------------------------------------------------------------------------- */
#define A_(i, j) a[((j) - 1) * 3 + ((i) - 1)]
static void jac_rotate(double *a, double *D, double *Z, int i, int j) {
    const double eps = 2.220446049250313e-16;  /* epsilon(1d0) */
    double aij = A_(i, j);
    double g = 100.0 * fabs(aij);
    double h = D[j - 1] - D[i - 1];
    double t, theta;
    if (fabs(h) + g == fabs(h)) {
        t = aij / (h + copysign(eps, h));
    } else {
        theta = 0.5 * h / (aij + copysign(eps, aij));
        t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
        if (theta < 0.0) t = -t;
    }
    double c = 1.0 / sqrt(1.0 + t * t);
    double s = t * c;
    double tau = s / (1.0 + c);
    h = t * aij;
    Z[i - 1] -= h; Z[j - 1] += h;
    D[i - 1] -= h; D[j - 1] += h;
    A_(i, j) = 0.0;
    for (int k = 1; k <= i - 1; ++k) {
        g = A_(k, i); h = A_(k, j);
        A_(k, i) = g - s * (h + g * tau);
        A_(k, j) = h + s * (g - h * tau);
    }
    for (int k = i + 1; k <= j - 1; ++k) {
        g = A_(i, k); h = A_(k, j);
        A_(i, k) = g - s * (h + g * tau);
        A_(k, j) = h + s * (g - h * tau);
    }
    for (int k = j + 1; k <= 3; ++k) {
        g = A_(i, k); h = A_(j, k);
        A_(i, k) = g - s * (h + g * tau);
        A_(j, k) = h + s * (g - h * tau);
    }
}
void orc_jacobi_eigenvalues(double *a, double *D) {
    const double atol = 1.0e-15;
    double B[3], Z[3];
    for (int j = 1; j <= 3; ++j) { D[j - 1] = A_(j, j); B[j - 1] = D[j - 1]; Z[j - 1] = 0.0; }
    double sm = fabs(A_(1, 2)) + fabs(A_(1, 3)) + fabs(A_(2, 3));
    int sweeps = 0;
    while (sm > atol && sweeps < 100) {
        for (int i = 1; i <= 2; ++i)
            for (int j = i + 1; j <= 3; ++j) jac_rotate(a, D, Z, i, j);
        for (int k = 0; k < 3; ++k) { B[k] += Z[k]; D[k] = B[k]; Z[k] = 0.0; }
        sm = fabs(A_(1, 2)) + fabs(A_(1, 3)) + fabs(A_(2, 3));
        ++sweeps;
    }
    double tv;
    if (D[1] > D[0]) { tv = D[0]; D[0] = D[1]; D[1] = tv; }
    if (D[2] > D[1]) { tv = D[1]; D[1] = D[2]; D[2] = tv; }
    if (D[1] > D[0]) { tv = D[0]; D[0] = D[1]; D[1] = tv; }
}
#undef A_

/* ------------------------------------------------------------------------- */
/* Solver state                                                              */
/* ------------------------------------------------------------------------- */
struct orc {
    ps3d_params p;
    int nx, ny, nz, nz1, nwx, nwy;
    size_t n2, npts;            /* ny*nx columns, nz1*ny*nx points */
    double dx[3], dxi[3], upper[3], ncelli, fnzi, dzi, hdzi;
    long ncell;
    cplan px, py, pz2;
    double *hrkx, *hrky, *rkx, *rky, *rkz, *rkzi;
    double *k2l2, *k2l2i, *vhdis, *vdiss, *filt, *green;
    double *phim, *phip, *thetam, *thetap, *dthetam, *dthetap, *gamtop, *gambot;
    double vvisc;
    double *svor, *vor, *vel, *svel, *svorts, *pres, *delta;
    double ini_vor_mean[2];
    double *svorf, *svori, *epq, *emq, *vortsm;
    /* rolling mean (rolling_mean.f90) */
    double *hist; int r_inew, r_iold, r_filled, r_len; double r_sma;
    int have_vor, have_diff;
    double bfmax, vorch, ggmax, usggmax, lsggmax, rmv;
};

#define IDX(o, iz, iy, ix) ((size_t)(iz) + (size_t)(o)->nz1 * ((size_t)(iy) + (size_t)(o)->ny * (size_t)(ix)))
#define C2(o, iy, ix) ((size_t)(iy) + (size_t)(o)->ny * (size_t)(ix))

static double *vol(orc_t *o, int k) { return calloc((size_t)k * o->npts, sizeof(double)); }

/* integer power as gfortran evaluates x**p for integer p (binary method) */
static double ipow(double x, int p) {
    double r = 1.0, b = x;
    int e = p < 0 ? -p : p;
    while (e) { if (e & 1) r *= b; b *= b; e >>= 1; }
    return p < 0 ? 1.0 / r : r;
}

static void set_hyperbolic(orc_t *o, int kx, int ky, const double *zm, const double *zp) {
    /* inversion_utils.f90:484-542 (release build: no NDEBUG clamps) */
    const int nz = o->nz;
    size_t c = C2(o, ky, kx);
    double kl = sqrt(o->k2l2[c]);
    double fac = kl * o->p.extent[2];
    double ef = exp(-fac);
    double div = 1.0 / (1.0 - ef * ef);
    double k2ifac = 0.5 * o->k2l2i[c];
    for (int iz = 0; iz <= nz; ++iz) {
        double Lm = kl * zm[iz], Lp = kl * zp[iz];
        double ep = exp(-Lp), em = exp(-Lm);
        double pm = div * (ep - ef * em);
        double pp = div * (em - ef * ep);
        double dpm = -kl * div * (ep + ef * em);
        double dpp = kl * div * (em + ef * ep);
        double Q = div * (1.0 + ef * ef);
        double R = div * 2.0 * ef;
        size_t i = IDX(o, iz, ky, kx);
        o->phim[i] = pm;
        o->phip[i] = pp;
        o->thetam[i] = k2ifac * (R * Lm * pp - Q * Lp * pm);
        o->thetap[i] = k2ifac * (R * Lp * pm - Q * Lm * pp);
        o->dthetam[i] = -k2ifac * ((Q * Lp - 1.0) * dpm - R * Lm * dpp);
        o->dthetap[i] = -k2ifac * ((Q * Lm - 1.0) * dpp - R * Lp * dpm);
    }
}

orc_t *orc_create(const ps3d_params *p) {
    if (!p || p->nx < 2 || p->ny < 2 || p->nz < 2) { set_err("invalid grid"); return NULL; }
    orc_t *o = calloc(1, sizeof *o);
    o->p = *p;
    o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->nz1 = p->nz + 1;
    o->nwx = o->nx / 2; o->nwy = o->ny / 2;
    o->n2 = (size_t)o->nx * o->ny;
    o->npts = o->n2 * o->nz1;
    /* parameters.f90:61-87 */
    int n3[3] = {o->nx, o->ny, o->nz};
    for (int d = 0; d < 3; ++d) {
        o->upper[d] = p->lower[d] + p->extent[d];
        o->dx[d] = p->extent[d] / (double)n3[d];
        o->dxi[d] = 1.0 / o->dx[d];
    }
    o->ncell = (long)o->nx * o->ny * o->nz;
    o->ncelli = 1.0 / (double)o->ncell;
    o->fnzi = 1.0 / (double)o->nz;
    o->dzi = o->dxi[2];
    o->hdzi = 0.5 * o->dxi[2];
    if (cplan_init(&o->px, o->nx) || cplan_init(&o->py, o->ny) || cplan_init(&o->pz2, 2 * o->nz)) {
        set_err("Factorisation not possible. Only factors from 2-6 allowed.");
        orc_destroy(o);
        return NULL;
    }
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    /* sta2dfft.f90:44-90 + sta3dfft.f90:86-109 + deriv1d.f90:10-33 */
    o->hrkx = calloc(nx + 1, sizeof(double));
    o->hrky = calloc(ny + 1, sizeof(double));
    o->rkx = calloc(nx, sizeof(double));
    o->rky = calloc(ny, sizeof(double));
    o->rkz = calloc(nz + 1, sizeof(double));
    o->rkzi = calloc(nz + 1, sizeof(double));
    double scx = PI / p->extent[0], scy = PI / p->extent[1], scz = PI / p->extent[2];
    for (int k = 1; k <= nx; ++k) o->hrkx[k] = scx * (double)k;
    for (int k = 1; k <= ny; ++k) o->hrky[k] = scy * (double)k;
    o->rkx[0] = 0.0;
    for (int kx = 1; kx <= o->nwx - 1; ++kx) { o->rkx[kx] = o->hrkx[2 * kx]; o->rkx[nx - kx] = o->hrkx[2 * kx]; }
    o->rkx[o->nwx] = o->hrkx[nx];
    o->rky[0] = 0.0;
    for (int ky = 1; ky <= o->nwy - 1; ++ky) { o->rky[ky] = o->hrky[2 * ky]; o->rky[ny - ky] = o->hrky[2 * ky]; }
    o->rky[o->nwy] = o->hrky[ny];
    o->rkz[0] = 0.0;
    for (int k = 1; k <= nz; ++k) o->rkz[k] = scz * (double)k;
    for (int k = 1; k <= nz - 1; ++k) o->rkzi[k] = 1.0 / o->rkz[k];

    /* init_inversion: inversion_utils.f90:222-372 */
    o->k2l2 = calloc(o->n2, sizeof(double));
    o->k2l2i = calloc(o->n2, sizeof(double));
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) o->k2l2[C2(o, ky, kx)] = o->rkx[kx] * o->rkx[kx] + o->rky[ky] * o->rky[ky];
    o->k2l2[0] = 1.0;
    for (size_t c = 0; c < o->n2; ++c) o->k2l2i[c] = 1.0 / o->k2l2[c];
    o->k2l2[0] = 0.0; o->k2l2i[0] = 0.0;

    o->filt = vol(o, 1);
    {
        double rkxmax = 0, rkymax = 0, rkzmax = 0;
        for (int k = 0; k < nx; ++k) rkxmax = fmax(rkxmax, o->rkx[k]);
        for (int k = 0; k < ny; ++k) rkymax = fmax(rkymax, o->rky[k]);
        for (int k = 0; k <= nz; ++k) rkzmax = fmax(rkzmax, o->rkz[k]);
        double *skx = calloc(nx, sizeof(double)), *sky = calloc(ny, sizeof(double)), *skz = calloc(nz + 1, sizeof(double));
        if (p->filtering == PS3D_FILTER_23RULE) {
            /* init_23rd_rule_filter: 407-455 */
            const double f23 = 2.0 / 3.0;
            for (int k = 0; k < nx; ++k) skx[k] = (o->rkx[k] <= f23 * rkxmax) ? 1.0 : 0.0;
            for (int k = 0; k < ny; ++k) sky[k] = (o->rky[k] <= f23 * rkymax) ? 1.0 : 0.0;
            for (int k = 0; k <= nz; ++k) skz[k] = (o->rkz[k] <= f23 * rkzmax) ? 1.0 : 0.0;
            for (int kx = 0; kx < nx; ++kx)
                for (int ky = 0; ky < ny; ++ky) {
                    double f0 = skx[kx] * sky[ky];
                    o->filt[IDX(o, 0, ky, kx)] = f0;
                    o->filt[IDX(o, nz, ky, kx)] = f0;
                    for (int kz = 1; kz <= nz - 1; ++kz) o->filt[IDX(o, kz, ky, kx)] = f0 * skz[kz];
                }
        } else {
            /* init_hou_and_li_filter: 377-403 */
            double kxmaxi = 1.0 / rkxmax, kymaxi = 1.0 / rkymax, kzmaxi = 1.0 / rkzmax;
            for (int k = 0; k < nx; ++k) skx[k] = -36.0 * pow(kxmaxi * o->rkx[k], 36);
            for (int k = 0; k < ny; ++k) sky[k] = -36.0 * pow(kymaxi * o->rky[k], 36);
            for (int k = 0; k <= nz; ++k) skz[k] = -36.0 * pow(kzmaxi * o->rkz[k], 36);
            for (int kx = 0; kx < nx; ++kx)
                for (int ky = 0; ky < ny; ++ky) {
                    double f0 = exp(skx[kx] + sky[ky]);
                    o->filt[IDX(o, 0, ky, kx)] = f0;
                    o->filt[IDX(o, nz, ky, kx)] = f0;
                    for (int kz = 1; kz <= nz - 1; ++kz) o->filt[IDX(o, kz, ky, kx)] = f0 * exp(skz[kz]);
                }
        }
        for (int kz = 0; kz <= nz; ++kz) o->filt[IDX(o, kz, 0, 0)] = 1.0;
        free(skx); free(sky); free(skz);
    }
    o->green = vol(o, 1);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            size_t c = C2(o, ky, kx);
            for (int kz = 1; kz <= nz; ++kz) o->green[IDX(o, kz, ky, kx)] = -1.0 / (o->k2l2[c] + o->rkz[kz] * o->rkz[kz]);
            o->green[IDX(o, 0, ky, kx)] = -o->k2l2i[c];
        }
    double *zm = calloc(nz + 1, sizeof(double)), *zp = calloc(nz + 1, sizeof(double));
    for (int iz = 0; iz <= nz; ++iz) {
        double z = p->lower[2] + o->dx[2] * (double)iz;
        zm[iz] = o->upper[2] - z;
        zp[iz] = z - p->lower[2];
    }
    o->phim = vol(o, 1); o->phip = vol(o, 1);
    o->thetam = vol(o, 1); o->thetap = vol(o, 1);
    o->dthetam = vol(o, 1); o->dthetap = vol(o, 1);
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky)
            if (kx || ky) set_hyperbolic(o, kx, ky, zm, zp);
    o->gamtop = calloc(nz + 1, sizeof(double));
    o->gambot = calloc(nz + 1, sizeof(double));
    for (int iz = 0; iz <= nz; ++iz) {
        o->phim[IDX(o, iz, 0, 0)] = zm[iz] / p->extent[2];
        o->phip[IDX(o, iz, 0, 0)] = zp[iz] / p->extent[2];
    }
    for (int iz = 0; iz <= nz; ++iz) {
        double ph = o->phip[IDX(o, iz, 0, 0)];
        o->gamtop[iz] = 0.5 * p->extent[2] * (ph * ph - 1.0 / 3.0);
    }
    for (int iz = 0; iz <= nz; ++iz) o->gambot[iz] = o->gamtop[nz - iz];
    free(zm); free(zp);

    o->svor = vol(o, 3); o->vor = vol(o, 3); o->vel = vol(o, 3); o->svel = vol(o, 3);
    o->svorts = vol(o, 3); o->pres = vol(o, 1); o->delta = vol(o, 1);
    o->vhdis = calloc(o->n2, sizeof(double));
    o->vdiss = calloc(o->n2, sizeof(double));
    o->epq = calloc(o->n2, sizeof(double));
    o->emq = calloc(o->n2, sizeof(double));
    o->svorf = vol(o, 3); o->svori = vol(o, 3); o->vortsm = vol(o, 3);
    o->r_len = p->roll_mean_win_size > 0 ? p->roll_mean_win_size : 1;
    o->hist = calloc(o->r_len, sizeof(double));
    o->r_inew = 1; o->r_iold = 1;
    return o;
}

void orc_destroy(orc_t *o) {
    if (!o) return;
    cplan_free(&o->px); cplan_free(&o->py); cplan_free(&o->pz2);
    double *ptrs[] = {o->hrkx, o->hrky, o->rkx, o->rky, o->rkz, o->rkzi, o->k2l2, o->k2l2i, o->vhdis, o->vdiss,
                      o->filt, o->green, o->phim, o->phip, o->thetam, o->thetap, o->dthetam, o->dthetap,
                      o->gamtop, o->gambot, o->svor, o->vor, o->vel, o->svel, o->svorts, o->pres, o->delta,
                      o->svorf, o->svori, o->epq, o->emq, o->vortsm, o->hist};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(o);
}

double orc_viscosity(orc_t *o) { return o->vvisc; }

/* ------------------------------------------------------------------------- */
/* sta3dfft restatement                                                       */
/* ------------------------------------------------------------------------- */
/* Transform every line along axis (0 = x, 1 = y) of a volume, kind 0 forward,
 * 1 inverse (stafft forfft / revfft, m = 1 per line as in sta3dfft.f90:161-177). */
static void xy_lines(orc_t *o, double *f, int axis, int inverse) {
    const int n = axis == 0 ? o->nx : o->ny;
    const cplan *P = axis == 0 ? &o->px : &o->py;
    const size_t stride = axis == 0 ? (size_t)o->nz1 * o->ny : (size_t)o->nz1;
    const int nother = axis == 0 ? o->ny : o->nx;
    const size_t ostride = axis == 0 ? (size_t)o->nz1 : (size_t)o->nz1 * o->ny;
    const long nlines = (long)nother * o->nz1;
    #pragma omp parallel
    {
        double *a = malloc(sizeof(double) * (2 * n + 4 * n));
        double *b = a + n, *w = a + 2 * n;
        #pragma omp for schedule(static)
        for (long L = 0; L < nlines; L += 2) {
            int has_b = (L + 1 < nlines);
            size_t baseA = (size_t)(L / o->nz1) * ostride + (size_t)(L % o->nz1);
            size_t baseB = has_b ? (size_t)((L + 1) / o->nz1) * ostride + (size_t)((L + 1) % o->nz1) : 0;
            for (int j = 0; j < n; ++j) { a[j] = f[baseA + stride * j]; if (has_b) b[j] = f[baseB + stride * j]; }
            if (inverse) rfft_inv_pair(P, a, has_b ? b : NULL, w);
            else rfft_fwd_pair(P, a, has_b ? b : NULL, w);
            for (int j = 0; j < n; ++j) { f[baseA + stride * j] = a[j]; if (has_b) f[baseB + stride * j] = b[j]; }
        }
        free(a);
    }
}

/* fftxyp2s: sta3dfft.f90:136-194 (y transform, then x transform) */
static void fftxyp2s(orc_t *o, const double *fp, double *fs) {
    if (fs != fp) memcpy(fs, fp, sizeof(double) * o->npts);
    xy_lines(o, fs, 1, 0);
    xy_lines(o, fs, 0, 0);
}
/* fftxys2p: sta3dfft.f90:202-260 (x back-transform, then y) */
static void fftxys2p(orc_t *o, const double *fs, double *fp) {
    if (fp != fs) memcpy(fp, fs, sizeof(double) * o->npts);
    xy_lines(o, fp, 0, 1);
    xy_lines(o, fp, 1, 1);
}

/* diffx: sta3dfft.f90:304-337 with reverse_x (mpi_reverse.f90:235-244):
 * ds(kx) = si * hrkx(dkx) * fs(nx-kx). */
static void diffx(orc_t *o, const double *fs, double *ds) {
    const int nx = o->nx, nz1 = o->nz1;
    const size_t plane = (size_t)nz1 * o->ny;
    #pragma omp parallel for schedule(static)
    for (int kx = 0; kx < nx; ++kx) {
        double *d = ds + plane * kx;
        if (kx == 0 || ((nx % 2) == 0 && kx == o->nwx)) { memset(d, 0, sizeof(double) * plane); continue; }
        int dkx = (2 * kx < 2 * (nx - kx)) ? 2 * kx : 2 * (nx - kx);
        double si = (kx >= o->nwx + 1) ? 1.0 : -1.0;
        double f = si * o->hrkx[dkx];
        const double *s = fs + plane * (nx - kx);
        for (size_t i = 0; i < plane; ++i) d[i] = f * s[i];
    }
}
/* diffy: sta3dfft.f90:345-377 */
static void diffy(orc_t *o, const double *fs, double *ds) {
    const int ny = o->ny, nz1 = o->nz1;
    #pragma omp parallel for schedule(static)
    for (int kx = 0; kx < o->nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            double *d = ds + IDX(o, 0, ky, kx);
            if (ky == 0 || ((ny % 2) == 0 && ky == o->nwy)) { memset(d, 0, sizeof(double) * nz1); continue; }
            int dky = (2 * ky < 2 * (ny - ky)) ? 2 * ky : 2 * (ny - ky);
            double si = (ky >= o->nwy + 1) ? 1.0 : -1.0;
            double f = si * o->hrky[dky];
            const double *s = fs + IDX(o, 0, ny - ky, kx);
            for (int iz = 0; iz < nz1; ++iz) d[iz] = f * s[iz];
        }
}

/* column transforms: dst on (1:nz) (fftsine 264-278), dct on (0:nz) (fftcosine 282-296) */
static void columns(orc_t *o, double *f, int kind /*0 dst, 1 dct*/) {
    const int nz = o->nz;
    const long ncol = (long)o->n2;
    #pragma omp parallel
    {
        double *w = malloc(sizeof(double) * 8 * 2 * nz);
        #pragma omp for schedule(static)
        for (long c = 0; c < ncol; c += 2) {
            double *a = f + (size_t)c * o->nz1;
            double *b = (c + 1 < ncol) ? a + o->nz1 : NULL;
            if (kind == 0) dst_pair(&o->pz2, nz, a, b, w);
            else dct_pair(&o->pz2, nz, a, b, w);
        }
        free(w);
    }
}

/* ------------------------------------------------------------------------- */
/* inversion_utils restatement                                                */
/* ------------------------------------------------------------------------- */
/* field_decompose_semi_spectral: inversion_utils.f90:563-592 */
static void decompose_semi(orc_t *o, double *sfc) {
    const int nz = o->nz;
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)o->n2; ++c) {
        double *col = sfc + (size_t)c * o->nz1;
        const double *pm = o->phim + (size_t)c * o->nz1, *pp = o->phip + (size_t)c * o->nz1;
        for (int iz = 1; iz <= nz - 1; ++iz) col[iz] = col[iz] - (col[0] * pm[iz] + col[nz] * pp[iz]);
    }
    double *top = malloc(sizeof(double) * o->n2);
    for (size_t c = 0; c < o->n2; ++c) top[c] = sfc[c * o->nz1 + nz];
    columns(o, sfc, 0);
    for (size_t c = 0; c < o->n2; ++c) sfc[c * o->nz1 + nz] = top[c];
    free(top);
}
/* field_combine_semi_spectral: inversion_utils.f90:617-647 */
static void combine_semi(orc_t *o, double *sf) {
    const int nz = o->nz;
    double *top = malloc(sizeof(double) * o->n2);
    for (size_t c = 0; c < o->n2; ++c) { top[c] = sf[c * o->nz1 + nz]; sf[c * o->nz1 + nz] = 0.0; }
    columns(o, sf, 0);
    for (size_t c = 0; c < o->n2; ++c) sf[c * o->nz1 + nz] = top[c];
    free(top);
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)o->n2; ++c) {
        double *col = sf + (size_t)c * o->nz1;
        const double *pm = o->phim + (size_t)c * o->nz1, *pp = o->phip + (size_t)c * o->nz1;
        for (int iz = 1; iz <= nz - 1; ++iz) col[iz] = col[iz] + col[0] * pm[iz] + col[nz] * pp[iz];
    }
}
/* field_decompose_physical: 549-557 (fc is destroyed, like fftxyp2s's input) */
static void decompose_phys(orc_t *o, const double *fc, double *sf) {
    fftxyp2s(o, fc, sf);
    decompose_semi(o, sf);
}
/* field_combine_physical: 599-611 */
static void combine_phys(orc_t *o, const double *sf, double *fc) {
    double *sfc = malloc(sizeof(double) * o->npts);
    memcpy(sfc, sf, sizeof(double) * o->npts);
    combine_semi(o, sfc);
    fftxys2p(o, sfc, fc);
    free(sfc);
}
/* central_diffz: 653-673 */
static void central_diffz(orc_t *o, const double *fs, double *ds) {
    const int nz = o->nz;
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)o->n2; ++c) {
        const double *f = fs + (size_t)c * o->nz1;
        double *d = ds + (size_t)c * o->nz1;
        d[0] = o->dzi * (f[1] - f[0]);
        d[nz] = o->dzi * (f[nz] - f[nz - 1]);
        for (int iz = 1; iz <= nz - 1; ++iz) d[iz] = (f[iz + 1] - f[iz - 1]) * o->hdzi;
    }
}

/* get_viscosity: 157-184;  init_dissipation: 188-218 */
static int init_diffusion(orc_t *o, double te, double en) {
    double rkxmax = 0, rkymax = 0;
    for (int k = 0; k < o->nx; ++k) rkxmax = fmax(rkxmax, o->rkx[k]);
    for (int k = 0; k < o->ny; ++k) rkymax = fmax(rkymax, o->rky[k]);
    double K2max = pow(fmax(rkxmax, rkymax), 2);
    double rkmsi = 1.0 / K2max;
    const int p = o->p.nnu;
    if (o->p.length_scale == PS3D_LS_KOLMOGOROV)
        o->vvisc = o->p.prediss * pow(K2max * te / en, 1.0 / 3.0) * ipow(rkmsi, p);
    else if (o->p.length_scale == PS3D_LS_GEOPHYSICAL)
        o->vvisc = o->p.prediss * ipow(rkmsi, p);
    else { set_err("We only support 'Kolmogorov' or 'geophysical'"); return PS3D_E_ARG; }
    for (size_t c = 0; c < o->n2; ++c) o->vhdis[c] = (p == 1) ? o->vvisc * o->k2l2[c] : o->vvisc * ipow(o->k2l2[c], p);
    o->have_diff = 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* field_diagnostics restatement                                              */
/* ------------------------------------------------------------------------- */
static double get_kinetic_energy(orc_t *o, const double *vv) {   /* 85-122 */
    const int nz = o->nz;
    double s1 = 0, s0 = 0, sn = 0;
    for (int c = 0; c < 3; ++c) {
        const double *v = vv + (size_t)c * o->npts;
        for (size_t col = 0; col < o->n2; ++col) {
            const double *f = v + col * o->nz1;
            for (int iz = 1; iz <= nz - 1; ++iz) s1 += f[iz] * f[iz];
            s0 += f[0] * f[0];
            sn += f[nz] * f[nz];
        }
    }
    return (0.5 * s1 + 0.25 * s0 + 0.25 * sn) * o->ncelli;
}
static double get_mean(orc_t *o, const double *ff) {           /* 405-432 */
    const int nz = o->nz;
    double sl = 0, si = 0;
    for (size_t col = 0; col < o->n2; ++col) {
        const double *f = ff + col * o->nz1;
        sl += f[0] + f[nz];
        for (int iz = 1; iz <= nz - 1; ++iz) si += f[iz];
    }
    return (0.5 * sl + si) / (double)o->ncell;
}
static double get_abs_max(orc_t *o, const double *ff) {        /* 468-495 */
    double m = 0;
    for (size_t i = 0; i < o->npts; ++i) m = fmax(m, fabs(ff[i]));
    return m;
}
static double get_char_vorticity(orc_t *o, double vortrms) {   /* 501-545 */
    const int nz = o->nz;
    double l1 = SMALL, l2 = 0;
    const double *v1 = o->vor, *v2 = o->vor + o->npts, *v3 = o->vor + 2 * o->npts;
    for (size_t col = 0; col < o->n2; ++col)
        for (int iz = 1; iz <= nz; ++iz) {
            size_t a = col * o->nz1 + iz - 1, b = a + 1;
            double t1 = 0.5 * fabs(v1[a] + v1[b]);
            double t2 = 0.5 * fabs(v2[a] + v2[b]);
            double t3 = 0.5 * fabs(v3[a] + v3[b]);
            if (t1 + t2 + t3 > vortrms) { l1 += t1 + t2 + t3; l2 += t1 * t1 + t2 * t2 + t3 * t3; }
        }
    return l2 / l1;
}
static void get_mean_vorticity(orc_t *o, double *vm) {         /* 549-579 */
    const int nz = o->nz;
    for (int c = 0; c < 3; ++c) {
        const double *v = o->vor + (size_t)c * o->npts;
        double si = 0, s0 = 0, sn = 0;
        for (size_t col = 0; col < o->n2; ++col) {
            const double *f = v + col * o->nz1;
            for (int iz = 1; iz <= nz - 1; ++iz) si += f[iz];
            s0 += f[0]; sn += f[nz];
        }
        vm[c] = (si + 0.5 * s0 + 0.5 * sn) * o->ncelli;
    }
}
static double get_helicity(orc_t *o) {
    /* <u.w> with get_mean's trapezoid weights; helicity density as in
     * plotting/nc_reader.py:94-101 */
    const int nz = o->nz;
    double sl = 0, si = 0;
    for (size_t col = 0; col < o->n2; ++col)
        for (int iz = 0; iz <= nz; ++iz) {
            size_t i = col * o->nz1 + iz;
            double h = 0;
            for (int c = 0; c < 3; ++c) h += o->vel[c * o->npts + i] * o->vor[c * o->npts + i];
            if (iz == 0 || iz == nz) sl += h; else si += h;
        }
    return (0.5 * sl + si) / (double)o->ncell;
}
/* Trapezoid sums (1/2 inside, 1/4 on the lids) of f_c^2 over components c0..c1
 * of a 3-volume field: get_horizontal/vertical_kinetic_energy and
 * get_horizontal/vertical_enstrophy (field_diagnostics.f90:128-262). */
static double trapezoid_sq(orc_t *o, const double *vv, int c0, int c1) {
    const int nz = o->nz;
    double t1 = 0, t0 = 0, tn = 0;
    for (size_t col = 0; col < o->n2; ++col) {
        double a1 = 0, a0 = 0, an = 0;
        for (int c = c0; c <= c1; ++c) {
            const double *f = vv + (size_t)c * o->npts + col * o->nz1;
            for (int iz = 1; iz <= nz - 1; ++iz) a1 += f[iz] * f[iz];
            a0 += f[0] * f[0];
            an += f[nz] * f[nz];
        }
        t1 += a1; t0 += a0; tn += an;
    }
    return (0.5 * t1 + 0.25 * t0 + 0.25 * tn) * o->ncelli;
}

/* update_netcdf_field_diagnostics (field_diagnostics_netcdf.f90:257-439) plus the
 * values adapt hands over (advance.f90:188-193, 315-321, 366): the 40 field_stats
 * variables in the order NC_KE .. NC_ROMAX (36-75).  Needs orc_vor2vel and
 * orc_derived of the same state; diag is the adapt record (may be NULL). */
int orc_field_stats(orc_t *o, const double *diag, double *st) {
    const int nz = o->nz;
    const size_t N = o->npts;
    const double *xi = o->vor, *eta = o->vor + N, *zeta = o->vor + 2 * N;
    const double *u = o->vel, *v = o->vel + N;
    for (int k = 0; k < 40; ++k) st[k] = 0.0;
    st[0] = get_kinetic_energy(o, o->vel);                /* NC_KE  */
    st[1] = get_kinetic_energy(o, o->vor);                /* NC_EN: get_enstrophy has the same weights */
    if (diag) {
        st[2] = diag[PS3D_D_VORTMAX];
        st[3] = diag[PS3D_D_VORTRMS];
        st[4] = diag[PS3D_D_VORCH];
        st[5] = diag[PS3D_D_VORMEAN_X];
        st[6] = diag[PS3D_D_VORMEAN_Y];
        st[7] = diag[PS3D_D_VORMEAN_Z];
        st[19] = diag[PS3D_D_GGMAX];
        st[21] = diag[PS3D_D_UMAX];
        st[22] = diag[PS3D_D_VMAX];
        st[23] = diag[PS3D_D_WMAX];
        st[31] = diag[PS3D_D_USGGMAX];
        st[32] = diag[PS3D_D_LSGGMAX];
        st[35] = diag[PS3D_D_RGMAX];
    }
    st[8] = trapezoid_sq(o, o->vel, 0, 1);                /* kexy */
    st[9] = trapezoid_sq(o, o->vel, 2, 2);                /* kez  */
    st[10] = trapezoid_sq(o, o->vor, 0, 1);               /* enxy */
    st[11] = trapezoid_sq(o, o->vor, 2, 2);               /* enz  */
    double mn[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, mx[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL}, hemax = -HUGE_VAL;
    for (size_t i = 0; i < N; ++i) {
        const double w3[3] = {xi[i], eta[i], zeta[i]};
        for (int c = 0; c < 3; ++c) {
            if (w3[c] < mn[c]) mn[c] = w3[c];
            if (w3[c] > mx[c]) mx[c] = w3[c];
        }
        const double he = sqrt(xi[i] * xi[i] + eta[i] * eta[i]);   /* get_max_horizontal_enstrophy */
        if (he > hemax) hemax = he;
    }
    for (int c = 0; c < 3; ++c) { st[12 + c] = mn[c]; st[15 + c] = mx[c]; }
    st[18] = hemax;
    /* lids: maxval(vor(nz,:,:,c)), maxval(vor(0,:,:,c)), sqrt(maxval(u^2 + v^2)) at nz */
    double lid[6] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL}, uh = -HUGE_VAL;
    double sz = 0, sd = 0;
    for (size_t col = 0; col < o->n2; ++col) {
        const size_t t = col * o->nz1 + nz, b = col * o->nz1;
        const double *w3[3] = {xi, eta, zeta};
        for (int c = 0; c < 3; ++c) {
            lid[2 * c] = fmax(lid[2 * c], w3[c][t]);
            lid[2 * c + 1] = fmax(lid[2 * c + 1], w3[c][b]);
        }
        uh = fmax(uh, u[t] * u[t] + v[t] * v[t]);
        sz += zeta[t] * zeta[t];
        sd += o->delta[t] * o->delta[t];
    }
    for (int k = 0; k < 6; ++k) st[24 + k] = lid[k];      /* uoxmax loxmax uoymax loymax uozmax lozmax */
    st[30] = sqrt(uh);                                    /* usuhmax */
    st[33] = sqrt(sz / (double)(o->nx * o->ny));          /* uszrms */
    st[34] = sqrt(sd / (double)(o->nx * o->ny));          /* usdeltarms */
    st[38] = mn[2] / o->p.f_cor[2];                       /* ro_min (field_diagnostics.f90:293-304) */
    st[39] = mx[2] / o->p.f_cor[2];                       /* ro_max (309-320) */
    return 0;
}
/* calc_vorticity_mean: 584-599 */
static void calc_vorticity_mean(orc_t *o, double *savg) {
    const int nz = o->nz;
    double *wk = calloc(2 * (nz + 1), sizeof(double)), *w = malloc(sizeof(double) * 16 * nz);
    for (int nc = 0; nc < 2; ++nc) {
        const double *s = o->svor + (size_t)nc * o->npts;   /* column (0,0) */
        for (int iz = 1; iz <= nz - 1; ++iz) wk[iz] = s[iz];
        wk[nz] = 0.0;
        dst_pair(&o->pz2, nz, wk, NULL, w);
        double sum = 0;
        for (int iz = 1; iz <= nz - 1; ++iz) sum += wk[iz];
        savg[nc] = 0.5 * (s[0] + s[nz]) + o->fnzi * sum;
    }
    free(wk); free(w);
}
/* adjust_vorticity_mean: 604-619 */
static void adjust_vorticity_mean(orc_t *o) {
    double savg[2];
    calc_vorticity_mean(o, savg);
    for (int nc = 0; nc < 2; ++nc) {
        double *s = o->svor + (size_t)nc * o->npts;
        s[0] = s[0] + o->ini_vor_mean[nc] - savg[nc];
        s[o->nz] = s[o->nz] + o->ini_vor_mean[nc] - savg[nc];
    }
}

/* ------------------------------------------------------------------------- */
/* inversion restatement                                                      */
/* ------------------------------------------------------------------------- */
/* vor2vel: inversion.f90:23-226 */
int orc_vor2vel(orc_t *o) {
    const int nz = o->nz, nz1 = o->nz1;
    const size_t N = o->npts;
    double *as = vol(o, 1), *bs = vol(o, 1), *ds = vol(o, 1), *es = vol(o, 1), *cs = vol(o, 1);
    double *ubar = calloc(nz1, sizeof(double)), *vbar = calloc(nz1, sizeof(double));
    double *sv1 = o->svor, *sv2 = o->svor + N, *sv3 = o->svor + 2 * N;
    /* solenoidal projection (35-76) */
    diffx(o, sv2, as);
    diffy(o, sv1, bs);
    for (size_t i = 0; i < N; ++i) { ds[i] = as[i] - bs[i]; cs[i] = sv3[i]; }
    combine_semi(o, cs);
    central_diffz(o, cs, es);
    decompose_semi(o, es);
    memcpy(ubar, sv1, sizeof(double) * nz1);
    memcpy(vbar, sv2, sizeof(double) * nz1);
    diffx(o, es, sv1);
    diffy(o, ds, cs);
    for (size_t c = 0; c < o->n2; ++c)
        for (int iz = 0; iz <= nz; ++iz) { size_t i = c * nz1 + iz; sv1[i] = o->k2l2i[c] * (sv1[i] + cs[i]); }
    diffy(o, es, sv2);
    diffx(o, ds, cs);
    for (size_t c = 0; c < o->n2; ++c)
        for (int iz = 0; iz <= nz; ++iz) { size_t i = c * nz1 + iz; sv2[i] = o->k2l2i[c] * (sv2[i] - cs[i]); }
    memcpy(sv1, ubar, sizeof(double) * nz1);
    memcpy(sv2, vbar, sizeof(double) * nz1);
    /* physical vorticity (80-82) */
    for (int nc = 0; nc < 3; ++nc) combine_phys(o, o->svor + nc * N, o->vor + nc * N);
    /* source for w (86-90) */
    diffy(o, sv1, ds);
    diffx(o, sv2, es);
    for (size_t i = 0; i < N; ++i) ds[i] = ds[i] - es[i];
    /* boundary contributions (94-104) */
    for (size_t c = 0; c < o->n2; ++c) {
        const size_t b0 = c * nz1;
        double d0 = ds[b0], dn = ds[b0 + nz];
        for (int iz = 1; iz <= nz - 1; ++iz) bs[b0 + iz] = d0 * o->thetam[b0 + iz] + dn * o->thetap[b0 + iz];
        for (int iz = 0; iz <= nz; ++iz) es[b0 + iz] = d0 * o->dthetam[b0 + iz] + dn * o->dthetap[b0 + iz];
    }
    /* invert Laplacian for the sine part (107-109) and d/dz (112-122) */
    for (size_t c = 0; c < o->n2; ++c) {
        const size_t b0 = c * nz1;
        for (int iz = 1; iz <= nz - 1; ++iz) ds[b0 + iz] = o->green[b0 + iz] * ds[b0 + iz];
        as[b0] = 0.0;
        for (int kz = 1; kz <= nz - 1; ++kz) as[b0 + kz] = o->rkz[kz] * ds[b0 + kz];
        as[b0 + nz] = 0.0;
    }
    /* 125-131: dct(as) and dst(ds(1:nz)) per column */
    columns(o, as, 1);
    columns(o, ds, 0);
    /* 134-143 */
    for (size_t c = 0; c < o->n2; ++c) {
        const size_t b0 = c * nz1;
        ds[b0] = 0.0;
        for (int iz = 1; iz <= nz - 1; ++iz) ds[b0 + iz] = ds[b0 + iz] + bs[b0 + iz];
        ds[b0 + nz] = 0.0;
    }
    for (size_t i = 0; i < N; ++i) { es[i] = es[i] + as[i]; cs[i] = sv3[i]; }
    combine_semi(o, cs);
    /* horizontally averaged flow (150-165) */
    {
        double *w = malloc(sizeof(double) * 16 * nz);
        ubar[0] = 0.0; vbar[0] = 0.0;
        for (int iz = 1; iz <= nz - 1; ++iz) { ubar[iz] = -o->rkzi[iz] * sv2[iz]; vbar[iz] = o->rkzi[iz] * sv1[iz]; }
        ubar[nz] = 0.0; vbar[nz] = 0.0;
        dct_pair(&o->pz2, nz, ubar, vbar, w);
        for (int iz = 0; iz <= nz; ++iz) {
            ubar[iz] = ubar[iz] + sv2[nz] * o->gamtop[iz] - sv2[0] * o->gambot[iz];
            vbar[iz] = vbar[iz] - sv1[nz] * o->gamtop[iz] + sv1[0] * o->gambot[iz];
        }
        free(w);
    }
    /* u (169-189) */
    diffx(o, es, as);
    diffy(o, cs, bs);
    for (size_t c = 0; c < o->n2; ++c)
        for (int iz = 0; iz <= nz; ++iz) { size_t i = c * nz1 + iz; as[i] = o->k2l2i[c] * (as[i] + bs[i]); }
    memcpy(as, ubar, sizeof(double) * nz1);
    memcpy(o->svel, as, sizeof(double) * N);
    fftxys2p(o, as, o->vel);
    /* v (193-213) */
    diffy(o, es, as);
    diffx(o, cs, bs);
    for (size_t c = 0; c < o->n2; ++c)
        for (int iz = 0; iz <= nz; ++iz) { size_t i = c * nz1 + iz; as[i] = o->k2l2i[c] * (as[i] - bs[i]); }
    memcpy(as, vbar, sizeof(double) * nz1);
    memcpy(o->svel + N, as, sizeof(double) * N);
    fftxys2p(o, as, o->vel + N);
    /* w (218-222) */
    memcpy(o->svel + 2 * N, ds, sizeof(double) * N);
    fftxys2p(o, ds, o->vel + 2 * N);
    free(as); free(bs); free(ds); free(es); free(cs); free(ubar); free(vbar);
    o->have_vor = 1;
    return 0;
}

/* vorticity_tendency: inversion.f90:298-371 (f_cor added to vor in place, 310-314) */
static void vorticity_tendency(orc_t *o) {
    const size_t N = o->npts;
    double *fp = vol(o, 1), *gp = vol(o, 1), *p = vol(o, 1), *q = vol(o, 1), *r = vol(o, 1);
    double *u = o->vel, *v = o->vel + N, *w = o->vel + 2 * N;
    double *xi = o->vor, *eta = o->vor + N, *zeta = o->vor + 2 * N;
    double *t1 = o->svorts, *t2 = o->svorts + N, *t3 = o->svorts + 2 * N;
    for (int nc = 0; nc < 3; ++nc)
        if (o->p.f_cor[nc] != 0.0)
            for (size_t i = 0; i < N; ++i) o->vor[nc * N + i] += o->p.f_cor[nc];
    for (size_t i = 0; i < N; ++i) fp[i] = u[i] * eta[i] - v[i] * xi[i];
    decompose_phys(o, fp, r);
    for (size_t i = 0; i < N; ++i) fp[i] = w[i] * xi[i] - u[i] * zeta[i];
    decompose_phys(o, fp, q);
    diffy(o, r, t1);
    central_diffz(o, fp, gp);
    decompose_phys(o, gp, p);
    for (size_t i = 0; i < N; ++i) t1[i] = t1[i] - p[i];
    for (size_t i = 0; i < N; ++i) fp[i] = v[i] * zeta[i] - w[i] * eta[i];
    decompose_phys(o, fp, p);
    diffx(o, r, t2);
    central_diffz(o, fp, gp);
    decompose_phys(o, gp, r);
    for (size_t i = 0; i < N; ++i) t2[i] = r[i] - t2[i];
    diffx(o, q, t3);
    diffy(o, p, r);
    for (size_t i = 0; i < N; ++i) t3[i] = t3[i] - r[i];
    free(fp); free(gp); free(p); free(q); free(r);
}
int orc_source(orc_t *o) {   /* 378-388 */
    if (!o->have_vor) { set_err("source needs vor2vel"); return PS3D_E_STATE; }
    vorticity_tendency(o);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* fields_derived restatement                                                 */
/* ------------------------------------------------------------------------- */
static void pressure(orc_t *o, const double *dudx, const double *dudy, const double *dvdy,
                     const double *dwdx, const double *dwdy) {  /* 67-157 */
    const size_t N = o->npts;
    double *rs = vol(o, 1);
    const double *xi = o->vor, *eta = o->vor + N, *zeta = o->vor + 2 * N;
    for (size_t i = 0; i < N; ++i) {
        double dwdz = -(dudx[i] + dvdy[i]);
        o->pres[i] = 2.0 * (dudx[i] * dvdy[i] - dudy[i] * (zeta[i] + dudy[i]) +
                            dvdy[i] * dwdz - dwdy[i] * (dwdy[i] - xi[i]) +
                            dwdz * dudx[i] - dwdx[i] * (dwdx[i] + eta[i]));
    }
    fftxyp2s(o, o->pres, rs);
    columns(o, rs, 1);
    for (size_t i = 0; i < N; ++i) rs[i] = o->green[i] * rs[i];
    columns(o, rs, 1);
    fftxys2p(o, rs, o->pres);
    free(rs);
}
static void horizontal_divergence(orc_t *o) {                /* 161-182 */
    const size_t N = o->npts;
    double *uxs = vol(o, 1), *vys = vol(o, 1);
    diffx(o, o->svel, uxs);
    diffy(o, o->svel + N, vys);
    for (size_t i = 0; i < N; ++i) uxs[i] = uxs[i] + vys[i];
    fftxys2p(o, uxs, o->delta);
    free(uxs); free(vys);
}

/* ------------------------------------------------------------------------- */
/* advance / adapt restatement (advance.f90:77-410)                           */
/* ------------------------------------------------------------------------- */
static double rolling_next(orc_t *o, double vnew) {  /* rolling_mean.f90:36-69 */
    if (o->r_filled) {
        double vold = o->hist[o->r_iold - 1];
        o->r_iold = (o->r_iold % o->r_len) + 1;
        o->r_sma = o->r_sma + (vnew - vold) / (double)o->r_len;
        o->hist[o->r_inew - 1] = vnew;
        o->r_inew = (o->r_inew % o->r_len) + 1;
    } else {
        o->hist[o->r_inew - 1] = vnew;
        double s = 0;
        for (int i = 0; i < o->r_inew; ++i) s += o->hist[i];
        o->r_sma = s / (double)o->r_inew;
        o->r_filled = (o->r_len == o->r_inew);
        o->r_inew = (o->r_inew % o->r_len) + 1;
    }
    return o->r_sma;
}

static void set_diffusion(orc_t *o, double dt, double vval) {
    if (o->p.stepper == PS3D_STEPPER_CN2) {          /* cn2.f90:39-80 */
        double dfac = (o->p.nnu == 1) ? dt : vval * dt;
        for (size_t c = 0; c < o->n2; ++c) o->vdiss[c] = 1.0 / (1.0 + dfac * o->vhdis[c]);
    } else {                                          /* impl_rk4.f90:37-53 */
        double dfac = 0.5 * vval * dt;
        for (size_t c = 0; c < o->n2; ++c) o->vdiss[c] = dfac * o->vhdis[c];
    }
}

int orc_adapt(orc_t *o, double t, double *dtout, double *diag) {
    if (!o->have_vor || !o->have_diff) { set_err("adapt needs set_vorticity/vor2vel"); return PS3D_E_STATE; }
    const size_t N = o->npts;
    const int nz = o->nz;
    double *xs = vol(o, 1), *xp = vol(o, 1);
    double *dudx = vol(o, 1), *dudy = vol(o, 1), *dvdy = vol(o, 1), *dwdx = vol(o, 1), *dwdy = vol(o, 1);
    o->bfmax = 0.0;
    for (size_t i = 0; i < N; ++i) xp[i] = o->vor[i] * o->vor[i] + o->vor[N + i] * o->vor[N + i] + o->vor[2 * N + i] * o->vor[2 * N + i];
    double vortmax = sqrt(get_abs_max(o, xp));
    double vortrms = sqrt(get_mean(o, xp));
    o->vorch = get_char_vorticity(o, vortrms);
    double vormean[3];
    get_mean_vorticity(o, vormean);
    diffx(o, o->svel, xs);          fftxys2p(o, xs, dudx);
    diffy(o, o->svel, xs);          fftxys2p(o, xs, dudy);
    diffx(o, o->svel + 2 * N, xs);  fftxys2p(o, xs, dwdx);
    diffy(o, o->svel + N, xs);      fftxys2p(o, xs, dvdy);
    diffy(o, o->svel + 2 * N, xs);  fftxys2p(o, xs, dwdy);
    double gg = 2.220446049250313e-16, us = 0.0, ls = 0.0;   /* ggmax = epsilon(ggmax) */
    #pragma omp parallel for reduction(max:gg, us, ls) schedule(static)
    for (long col = 0; col < (long)o->n2; ++col)
        for (int iz = 0; iz <= nz; ++iz) {
            size_t i = (size_t)col * o->nz1 + iz;
            double a[9] = {0};
            a[0] = dudx[i];                                  /* (1,1) */
            a[3] = dudy[i] + 0.5 * o->vor[2 * N + i];        /* (1,2) */
            a[6] = dwdx[i] + 0.5 * o->vor[N + i];            /* (1,3) */
            a[4] = dvdy[i];                                  /* (2,2) */
            a[7] = dwdy[i] - 0.5 * o->vor[i];                /* (2,3) */
            a[8] = -(dudx[i] + dvdy[i]);                     /* (3,3) */
            double e[3];
            orc_jacobi_eigenvalues(a, e);
            double lmax = fmax(fmax(fabs(e[0]), fabs(e[1])), fabs(e[2]));
            gg = fmax(gg, lmax);
            if (iz == nz) us = fmax(us, lmax);
            if (iz == 0) ls = fmax(ls, lmax);
        }
    o->ggmax = gg; o->usggmax = us; o->lsggmax = ls;
    pressure(o, dudx, dudy, dvdy, dwdx, dwdy);
    horizontal_divergence(o);
    double umax = -INFINITY, vmax = -INFINITY, wmax = -INFINITY;
    for (size_t i = 0; i < N; ++i) {
        umax = fmax(umax, o->vel[i]); vmax = fmax(vmax, o->vel[N + i]); wmax = fmax(wmax, o->vel[2 * N + i]);
    }
    double dtcfl = CFLMAX * fmin(fmin(o->dx[0] / (umax + SMALL), o->dx[1] / (vmax + SMALL)), o->dx[2] / (wmax + SMALL));
    double dt = fmin(fmin(o->p.alpha / (o->ggmax + SMALL), o->p.alpha / (o->bfmax + SMALL)), fmin(dtcfl, o->p.limit - t));
    if (o->p.dt_fixed > 0.0) dt = fmin(o->p.dt_fixed, o->p.limit - t);   /* fixed-dt extension */
    o->rmv = rolling_next(o, o->ggmax);
    double vval;
    switch (o->p.pretype) {                           /* get_diffusion_pre_factor 381-410 */
        case PS3D_PRE_CONSTANT: vval = 1.0; break;
        case PS3D_PRE_VORCH: vval = o->vorch; break;
        case PS3D_PRE_BFMAX: vval = o->bfmax; break;
        case PS3D_PRE_ROLL_MEAN_MAX_STRAIN: vval = o->rmv; break;
        case PS3D_PRE_MAX_STRAIN: vval = o->ggmax; break;
        case PS3D_PRE_US_MAX_STRAIN: vval = o->usggmax; break;
        default: set_err("unknown pretype"); return PS3D_E_ARG;
    }
    set_diffusion(o, dt, vval);
    if (diag) {
        diag[PS3D_D_T] = t; diag[PS3D_D_DT] = dt;
        diag[PS3D_D_KE] = get_kinetic_energy(o, o->vel);
        diag[PS3D_D_EN] = get_kinetic_energy(o, o->vor);   /* get_enstrophy has the same weights */
        diag[PS3D_D_HEL] = get_helicity(o);
        diag[PS3D_D_VORTMAX] = vortmax; diag[PS3D_D_VORTRMS] = vortrms; diag[PS3D_D_VORCH] = o->vorch;
        diag[PS3D_D_VORMEAN_X] = vormean[0]; diag[PS3D_D_VORMEAN_Y] = vormean[1]; diag[PS3D_D_VORMEAN_Z] = vormean[2];
        diag[PS3D_D_GGMAX] = o->ggmax; diag[PS3D_D_UMAX] = umax; diag[PS3D_D_VMAX] = vmax; diag[PS3D_D_WMAX] = wmax;
        diag[PS3D_D_USGGMAX] = o->usggmax; diag[PS3D_D_LSGGMAX] = o->lsggmax; diag[PS3D_D_RGMAX] = o->rmv;
        diag[PS3D_D_PREFACTOR] = vval;
    }
    *dtout = dt;
    free(xs); free(xp); free(dudx); free(dudy); free(dvdy); free(dwdx); free(dwdy);
    return 0;
}

/* ---- impl_rk4 (impl_rk4.f90:76-364), literally with combine/decompose ---- */
static void mul_cols(orc_t *o, double *q, const double *fac) {
    for (size_t c = 0; c < o->n2; ++c)
        for (int iz = 0; iz <= o->nz; ++iz) q[c * o->nz1 + iz] *= fac[c];
}
static void sandwich(orc_t *o, double *q, const double *fac) {
    combine_semi(o, q);
    mul_cols(o, q, fac);
    decompose_semi(o, q);
}
/* fixed_src != NULL replaces "vor2vel; source" by svorts = fixed_src, which is
 * how unit-tests/test_implicit_rk.f90:146-226 drives the scheme. */
static void rk4_tendency(orc_t *o, const double *fixed_src) {
    if (fixed_src) memcpy(o->svorts, fixed_src, sizeof(double) * 3 * o->npts);
    else { orc_vor2vel(o); vorticity_tendency(o); }
}
static void rk4_step(orc_t *o, double *t, double dt, const double *fixed_src) {
    const size_t N = o->npts;
    const double dt2 = 0.5 * dt, dt3 = dt / 3.0, dt6 = dt / 6.0;
    for (size_t c = 0; c < o->n2; ++c) {
        o->epq[c] = exp(o->vdiss[c]);
        o->emq[c] = 1.0 / o->epq[c];
        o->epq[c] = o->epq[c] * o->filt[c * o->nz1];   /* filt(0,:,:) */
    }
    for (int nc = 0; nc < 3; ++nc) {   /* substep one 212-242 */
        double *q = o->svor + nc * N, *sqs = o->svorts + nc * N, *qdi = o->svori + nc * N, *qdf = o->svorf + nc * N;
        for (size_t c = 0; c < o->n2; ++c)
            for (int iz = 0; iz <= o->nz; ++iz) sqs[c * o->nz1 + iz] *= o->filt[c * o->nz1];
        for (size_t i = 0; i < N; ++i) { qdi[i] = q[i]; q[i] = qdi[i] + dt2 * sqs[i]; }
        sandwich(o, q, o->emq);
        for (size_t i = 0; i < N; ++i) qdf[i] = qdi[i] + dt6 * sqs[i];
    }
    rk4_tendency(o, fixed_src);
    *t += dt2;
    for (int nc = 0; nc < 3; ++nc) {   /* substep two 247-284 */
        double *q = o->svor + nc * N, *sqs = o->svorts + nc * N, *qdi = o->svori + nc * N, *qdf = o->svorf + nc * N;
        sandwich(o, sqs, o->epq);
        for (size_t i = 0; i < N; ++i) q[i] = qdi[i] + dt2 * sqs[i];
        sandwich(o, q, o->emq);
        for (size_t i = 0; i < N; ++i) qdf[i] = qdf[i] + dt3 * sqs[i];
    }
    rk4_tendency(o, fixed_src);
    *t += dt2;
    for (size_t c = 0; c < o->n2; ++c) o->emq[c] = o->emq[c] * o->emq[c];
    for (int nc = 0; nc < 3; ++nc) {   /* substep three 289-326 */
        double *q = o->svor + nc * N, *sqs = o->svorts + nc * N, *qdi = o->svori + nc * N, *qdf = o->svorf + nc * N;
        sandwich(o, sqs, o->epq);
        for (size_t i = 0; i < N; ++i) q[i] = qdi[i] + dt * sqs[i];
        sandwich(o, q, o->emq);
        for (size_t i = 0; i < N; ++i) qdf[i] = qdf[i] + dt3 * sqs[i];
    }
    rk4_tendency(o, fixed_src);
    for (size_t c = 0; c < o->n2; ++c) o->epq[c] = o->epq[c] * o->epq[c];
    for (int nc = 0; nc < 3; ++nc) {   /* substep four 331-364 */
        double *q = o->svor + nc * N, *sqs = o->svorts + nc * N, *qdf = o->svorf + nc * N;
        sandwich(o, sqs, o->epq);
        for (size_t i = 0; i < N; ++i) q[i] = qdf[i] + dt6 * sqs[i];
        sandwich(o, q, o->emq);
    }
    if (!fixed_src) adjust_vorticity_mean(o);
}

/* ---- cn2 (cn2.f90:92-181) ----
 * fixed_src != NULL holds the tendency at fixed_src through the iterations and
 * leaves the vorticity mean alone, as rk4_step does for test_implicit_rk. */
static void cn2_step(orc_t *o, double *t, double dt, const double *fixed_src) {
    const size_t N = o->npts;
    const double dt2 = 0.5 * dt;
    if (fixed_src) memcpy(o->svorts, fixed_src, sizeof(double) * 3 * N);
    for (size_t i = 0; i < 3 * N; ++i) o->vortsm[i] = o->svor[i] + dt2 * o->svorts[i];
    for (int iter = 0; iter <= 2; ++iter) {
        if (iter > 0 && !fixed_src) { orc_vor2vel(o); vorticity_tendency(o); }
        for (int nc = 0; nc < 3; ++nc) {
            double *q = o->svor + nc * N, *sm = o->vortsm + nc * N, *ts = o->svorts + nc * N;
            for (size_t i = 0; i < N; ++i) q[i] = o->filt[i] * (sm[i] + dt2 * ts[i]);
            sandwich(o, q, o->vdiss);
        }
        if (!fixed_src) adjust_vorticity_mean(o);
    }
    *t += dt;
}

int orc_step(orc_t *o, double *t, double dt) {
    if (o->p.stepper == PS3D_STEPPER_CN2) cn2_step(o, t, dt, NULL);
    else rk4_step(o, t, dt, NULL);
    return 0;
}

int orc_rk4_check(orc_t *o, double dt, const double *vdiss, const double *src, double *svor) {
    double t = 0.0;
    memcpy(o->vdiss, vdiss, sizeof(double) * o->n2);
    memcpy(o->svor, svor, sizeof(double) * 3 * o->npts);
    memcpy(o->svorts, src, sizeof(double) * 3 * o->npts);
    rk4_step(o, &t, dt, src);
    memcpy(svor, o->svor, sizeof(double) * 3 * o->npts);
    return 0;
}

int orc_cn2_check(orc_t *o, double dt, const double *vdiss, const double *src, double *svor) {
    double t = 0.0;
    memcpy(o->vdiss, vdiss, sizeof(double) * o->n2);
    memcpy(o->svor, svor, sizeof(double) * 3 * o->npts);
    cn2_step(o, &t, dt, src);
    memcpy(svor, o->svor, sizeof(double) * 3 * o->npts);
    return 0;
}

/* advance: advance.f90:77-104 (write_step is host I/O and not part of the oracle) */
int orc_advance_n(orc_t *o, double *t, int nsteps, double *diag_series) {
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        int st = orc_vor2vel(o);
        if (!st) st = orc_adapt(o, *t, &dt, diag_series ? diag_series + (size_t)s * PS3D_NDIAG : NULL);
        if (!st) st = orc_source(o);
        if (!st) st = orc_step(o, t, dt);
        if (st) return st;
    }
    return 0;
}

/* setup_fields: utils.f90:136-184 (without buoyancy) */
int orc_set_vorticity(orc_t *o, const double *vor) {
    const size_t N = o->npts;
    double *tmp = vol(o, 1);
    for (int nc = 0; nc < 3; ++nc) {
        memcpy(tmp, vor + nc * N, sizeof(double) * N);
        decompose_phys(o, tmp, o->svor + nc * N);
    }
    free(tmp);
    calc_vorticity_mean(o, o->ini_vor_mean);
    orc_vor2vel(o);
    double ke = get_kinetic_energy(o, o->vel);
    double en = get_kinetic_energy(o, o->vor);
    return init_diffusion(o, ke, en);
}

int orc_derived(orc_t *o) {
    const size_t N = o->npts;
    double *xs = vol(o, 1);
    double *g[5];
    for (int i = 0; i < 5; ++i) g[i] = vol(o, 1);
    diffx(o, o->svel, xs);          fftxys2p(o, xs, g[0]);
    diffy(o, o->svel, xs);          fftxys2p(o, xs, g[1]);
    diffx(o, o->svel + 2 * N, xs);  fftxys2p(o, xs, g[3]);
    diffy(o, o->svel + N, xs);      fftxys2p(o, xs, g[2]);
    diffy(o, o->svel + 2 * N, xs);  fftxys2p(o, xs, g[4]);
    pressure(o, g[0], g[1], g[2], g[3], g[4]);
    horizontal_divergence(o);
    for (int i = 0; i < 5; ++i) free(g[i]);
    free(xs);
    return 0;
}

static double *field_ptr(orc_t *o, int which, int *nvol) {
    *nvol = 3;
    switch (which) {
        case PS3D_SVOR: return o->svor;
        case PS3D_VOR: return o->vor;
        case PS3D_VEL: return o->vel;
        case PS3D_SVEL: return o->svel;
        case PS3D_SVORTS: return o->svorts;
        case PS3D_PRES: *nvol = 1; return o->pres;
        case PS3D_DELTA: *nvol = 1; return o->delta;
    }
    return NULL;
}
int orc_get(orc_t *o, int which, double *out) {
    int nv; double *f = field_ptr(o, which, &nv);
    if (!f) { set_err("unknown field"); return PS3D_E_ARG; }
    memcpy(out, f, sizeof(double) * nv * o->npts);
    return 0;
}
int orc_set(orc_t *o, int which, const double *in) {
    int nv; double *f = field_ptr(o, which, &nv);
    if (!f) { set_err("unknown field"); return PS3D_E_ARG; }
    memcpy(f, in, sizeof(double) * nv * o->npts);
    if (which == PS3D_VOR || which == PS3D_VEL) o->have_vor = 1;
    return 0;
}

int orc_op(orc_t *o, int op, const double *in, double *out) {
    const size_t N = o->npts;
    double *tmp = vol(o, 1);
    memcpy(tmp, in, sizeof(double) * N);
    int st = 0;
    switch (op) {
        case PS3D_OP_FFTXYP2S: fftxyp2s(o, tmp, out); break;
        case PS3D_OP_FFTXYS2P: fftxys2p(o, tmp, out); break;
        case PS3D_OP_DIFFX: diffx(o, tmp, out); break;
        case PS3D_OP_DIFFY: diffy(o, tmp, out); break;
        case PS3D_OP_FFTSINE: columns(o, tmp, 0); memcpy(out, tmp, sizeof(double) * N); break;
        case PS3D_OP_FFTCOSINE: columns(o, tmp, 1); memcpy(out, tmp, sizeof(double) * N); break;
        case PS3D_OP_CENTRAL_DIFFZ: central_diffz(o, tmp, out); break;
        case PS3D_OP_COMBINE_SEMI: combine_semi(o, tmp); memcpy(out, tmp, sizeof(double) * N); break;
        case PS3D_OP_DECOMPOSE_SEMI: decompose_semi(o, tmp); memcpy(out, tmp, sizeof(double) * N); break;
        case PS3D_OP_COMBINE_PHYS: combine_phys(o, tmp, out); break;
        case PS3D_OP_DECOMPOSE_PHYS: decompose_phys(o, tmp, out); break;
        default: set_err("unknown op"); st = PS3D_E_ARG;
    }
    free(tmp);
    return st;
}

long orc_table(orc_t *o, int which, double *out) {
    const double *src = NULL; long n = 0;
    switch (which) {
        case 0: src = o->k2l2; n = o->n2; break;
        case 1: src = o->k2l2i; n = o->n2; break;
        case 2: src = o->vhdis; n = o->n2; break;
        case 3: src = o->filt; n = o->npts; break;
        case 4: src = o->green; n = o->npts; break;
        case 5: src = o->phim; n = o->npts; break;
        case 6: src = o->phip; n = o->npts; break;
        case 7: src = o->thetam; n = o->npts; break;
        case 8: src = o->thetap; n = o->npts; break;
        case 9: src = o->dthetam; n = o->npts; break;
        case 10: src = o->dthetap; n = o->npts; break;
        case 11: src = o->gamtop; n = o->nz1; break;
        case 12: src = o->gambot; n = o->nz1; break;
        case 13: src = o->rkx; n = o->nx; break;
        case 14: src = o->rky; n = o->ny; break;
        case 15: src = o->rkz; n = o->nz1; break;
        default: return -1;
    }
    if (out) memcpy(out, src, sizeof(double) * n);
    return n;
}
