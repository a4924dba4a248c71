// zmath.hpp -- per-column z functions shared by the z kernels: the harmonic
// (phi), theta and green functions of the mixed z representation.
#pragma once
#include "common.hpp"

namespace ps3d {

// Hyperbolic functions of a quad at level iz (inversion_utils.f90:484-542 and the
// (0,0) column at 326-346).  All four members share K = sqrt(k^2+l^2).
struct Hyp {
    double phim, phip, thetam, thetap, dthetam, dthetap;
};

__device__ inline void zm_zp(const Grid &g, int iz, double &zm, double &zp) {
    // z = lower + dz*iz ; zm = upper - z ; zp = z - lower (inversion_utils.f90:291-297)
    const double z = g.zlo + g.dz * (double)iz;
    zm = (g.zlo + g.lz) - z;
    zp = z - g.zlo;
}

__device__ inline void phi_only(const Grid &g, double K, bool q00, int iz, double &pm, double &pp) {
    double zm, zp;
    zm_zp(g, iz, zm, zp);
    if (q00) {
        pm = zm / g.lz;
        pp = zp / g.lz;
        return;
    }
    const double ef = exp(-K * g.lz);
    const double div = 1.0 / (1.0 - ef * ef);
    const double ep = exp(-K * zp), em = exp(-K * zm);
    pm = div * (ep - ef * em);
    pp = div * (em - ef * ep);
}

// The same functions from ep = exp(-K zp), em = exp(-K zm), ef = exp(-K lz) and
// div = 1 / (1 - ef^2) supplied by the caller (the wave kernels build ep and em by
// recurrence along z).
__device__ inline Hyp hyp_from(const Grid &g, double K, double k2l2i, bool q00, int iz, double ep, double em,
                               double ef, double div) {
    Hyp h;
    double zm, zp;
    zm_zp(g, iz, zm, zp);
    if (q00) {
        h.phim = zm / g.lz;
        h.phip = zp / g.lz;
        h.thetam = h.thetap = h.dthetam = h.dthetap = 0.0;
        return h;
    }
    const double kl = K;
    const double k2ifac = 0.5 * k2l2i;
    const double Lm = kl * zm, Lp = kl * zp;
    h.phim = div * (ep - ef * em);
    h.phip = div * (em - ef * ep);
    const double dphim = -kl * div * (ep + ef * em);
    const double dphip = kl * div * (em + ef * ep);
    const double Q = div * (1.0 + ef * ef);
    const double R = div * 2.0 * ef;
    h.thetam = k2ifac * (R * Lm * h.phip - Q * Lp * h.phim);
    h.thetap = k2ifac * (R * Lp * h.phim - Q * Lm * h.phip);
    h.dthetam = -k2ifac * ((Q * Lp - 1.0) * dphim - R * Lm * dphip);
    h.dthetap = -k2ifac * ((Q * Lm - 1.0) * dphip - R * Lp * dphim);
    return h;
}

__device__ inline Hyp hyp_all(const Grid &g, double K, double k2l2i, bool q00, int iz) {
    if (q00) return hyp_from(g, K, k2l2i, true, iz, 0.0, 0.0, 0.0, 0.0);
    double zm, zp;
    zm_zp(g, iz, zm, zp);
    const double ef = exp(-K * g.lz);
    return hyp_from(g, K, k2l2i, false, iz, exp(-K * zp), exp(-K * zm), ef, 1.0 / (1.0 - ef * ef));
}

}  // namespace ps3d
