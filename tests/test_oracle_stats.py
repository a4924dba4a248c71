"""Pins of the oracle's field_stats record (the 40 variables of
src/fields/field_diagnostics_netcdf.f90:36-75, computed at :257-439) against
closed forms.

* One Beltrami mode (k, l, m) = (2, 1, 1), omega = alpha u, on a 32^3 grid whose
  points reach sin(theta) = +-1 and sin(mz) = 1.  The trapezoid rule in z is exact
  for these trigonometric polynomials and the horizontal means of sin^2, cos^2 are
  1/2, so: kexy = (m^2 + alpha^2) / (8 (k^2+l^2)), kez = 1/8, enxy = alpha^2 kexy,
  enz = alpha^2 / 8; the extrema follow from the z factors of u, v, w
  (ps3d_amd/ics.beltrami_mode) and the lid values; the horizontal divergence is
  -dw/dz = -m cos(mz) cos(theta), so usdeltarms = m / sqrt 2.
* A horizontal vorticity xi = l cos(theta) g(z), eta = -k cos(theta) g(z), g = 1 + z/pi:
  enxy from the trapezoid rule's known error for g^2, the xi, eta extrema and the
  two lids' maxima, which differ.
* A z-independent 2-D vortex psi = cos(kx) cos(ly) (u = -psi_y, v = psi_x, w = 0,
  zeta = (k^2+l^2) psi): uszrms = (k^2+l^2)/2, kexy = (k^2+l^2)/8,
  enz = (k^2+l^2)^2/8 and ro_min, ro_max = min(zeta), max(zeta) / f_cor(3).
* The entries adapt hands over (advance.f90:188-193, 315-321, 366) land in ps3d's
  slots.
"""
import math

import numpy as np
import pytest

from ps3d_amd import ics

PI = math.pi
N = 32
BOX = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
STAT = ["ke", "en", "vortmax", "vortrms", "vorch", "x_vormean", "y_vormean", "z_vormean",
        "kexy", "kez", "enxy", "enz", "x_vormin", "y_vormin", "z_vormin",
        "x_vormax", "y_vormax", "z_vormax", "enxy_max", "gmax", "bfmax", "umax", "vmax", "wmax",
        "uoxmax", "loxmax", "uoymax", "loymax", "uozmax", "lozmax", "usuhmax", "usgmax", "lsgmax",
        "uszrms", "usdeltarms", "rolling_mean_gmax", "rolling_mean_bfmax", "ri_min", "ro_min", "ro_max"]
S = {n: i for i, n in enumerate(STAT)}


def _oracle(O, vor, fcor3=0.5):
    o = O.Oracle(O.make_params(N, N, N, **BOX, filtering="2/3-rule", nnu=3, prediss=30.0, pretype="vorch",
                               f_cor=(0.0, 0.0, fcor3)))
    o.set_vorticity(vor)
    o.vor2vel()
    o.derived()
    return o


def check_beltrami_stats(O):
    k, l, m = 2, 1, 1
    a = math.sqrt(k * k + l * l + m * m)
    f = 1.0 / (k * k + l * l)
    x, y, z = ics.grid(N, N, N, BOX["lower"], BOX["extent"])
    _, vor = ics.beltrami_mode(x, y, z, k, l, m)
    o = _oracle(O, vor, fcor3=0.5)
    st = o.field_stats()
    o.close()
    # velocity-based entries and zeta (vor2vel keeps u, v, w and zeta of the mode to
    # roundoff; its projection moves xi, eta by O(dz^2), so their entries are pinned
    # by the horizontal-vorticity flow below)
    exp = {
        "kexy": f * (m * m + a * a) / 8.0, "kez": 1.0 / 8.0,
        "ke": f * (m * m + a * a) / 8.0 + 1.0 / 8.0,
        "enz": a * a / 8.0, "z_vormax": a, "z_vormin": -a,
        "uozmax": 0.0, "lozmax": 0.0,
        "usuhmax": m * math.sqrt(f),
        "uszrms": 0.0, "usdeltarms": m / math.sqrt(2.0),
        "ro_min": -a / 0.5, "ro_max": a / 0.5,
    }
    for name, v in exp.items():
        assert abs(st[S[name]] - v) <= 1e-12 * max(1.0, abs(v)), (name, st[S[name]], v)
    # nothing from adapt was passed: those slots stay 0, as do the buoyancy ones
    for name in ("vortmax", "vortrms", "vorch", "gmax", "umax", "rolling_mean_gmax", "bfmax",
                 "rolling_mean_bfmax", "ri_min"):
        assert st[S[name]] == 0.0, name


def check_horizontal_vorticity_stats(O):
    """xi = l cos(theta) g(z), eta = -k cos(theta) g(z), zeta = 0 with g = 1 + z/pi
    (divergence-free for any g, so vor2vel keeps it): the lids differ (g = 1 below,
    2 above), and the trapezoid mean of g^2 over n intervals is 7/3 + 1/(6 n^2), so
    enxy = (k^2+l^2)/4 (7/3 + 1/(6 n^2))."""
    k, l = 2, 1
    x, y, z = ics.grid(N, N, N, BOX["lower"], BOX["extent"])
    X, Y = np.meshgrid(x, y, indexing="ij")
    c = np.cos(k * X + l * Y)
    g = 1.0 + z / PI
    vor = np.zeros((3, N, N, N + 1))
    vor[0] = (l * c)[:, :, None] * g
    vor[1] = (-k * c)[:, :, None] * g
    o = _oracle(O, vor, fcor3=1.0)
    st = o.field_stats()
    o.close()
    kl = k * k + l * l
    g2 = 7.0 / 3.0 + 1.0 / (6.0 * N * N)
    exp = {"enxy": kl / 4.0 * g2, "enz": 0.0, "en": kl / 4.0 * g2, "x_vormax": 2 * l, "x_vormin": -2 * l,
           "y_vormax": 2 * k, "y_vormin": -2 * k, "enxy_max": 2 * math.sqrt(kl),
           "uoxmax": 2 * l, "loxmax": l, "uoymax": 2 * k, "loymax": k,
           "uozmax": 0.0, "lozmax": 0.0, "z_vormax": 0.0, "z_vormin": 0.0}
    for name, v in exp.items():
        assert abs(st[S[name]] - v) <= 1e-12 * max(1.0, abs(v)), (name, st[S[name]], v)


def check_vortex_stats(O):
    k, l = 2, 1
    x, y, z = ics.grid(N, N, N, BOX["lower"], BOX["extent"])
    X, Y = np.meshgrid(x, y, indexing="ij")
    zeta = (k * k + l * l) * np.cos(k * X) * np.cos(l * Y)
    vor = np.zeros((3, N, N, N + 1))
    vor[2] = zeta[:, :, None]
    o = _oracle(O, vor, fcor3=-2.0)
    st = o.field_stats()
    o.close()
    kl = k * k + l * l
    exp = {"uszrms": kl / 2.0, "kexy": kl / 8.0, "kez": 0.0, "ke": kl / 8.0, "enz": kl * kl / 8.0,
           "enxy": 0.0, "en": kl * kl / 8.0, "z_vormax": kl, "z_vormin": -kl, "uozmax": kl, "lozmax": kl,
           "usuhmax": max(k, l), "enxy_max": 0.0, "usdeltarms": 0.0,
           # ps3d divides the extrema of zeta by f_cor(3): with f_cor(3) < 0 the
           # "minimum" Rossby number is the larger one
           "ro_min": -kl / -2.0, "ro_max": kl / -2.0}
    for name, v in exp.items():
        assert abs(st[S[name]] - v) <= 1e-12 * max(1.0, abs(v)), (name, st[S[name]], v)


def check_adapt_slots(O):
    x, y, z = ics.grid(N, N, N, BOX["lower"], BOX["extent"])
    _, vor = ics.beltrami_mode(x, y, z, 1, 1, 1)
    o = _oracle(O, vor)
    diag = 100.0 + np.arange(O.NDIAG, dtype=float)
    st = o.field_stats(diag)
    o.close()
    # ps3d_diag order: t dt ke en hel vortmax vortrms vorch vormean(3) ggmax umax vmax wmax usgg lsgg rgmax pref
    want = {"vortmax": 5, "vortrms": 6, "vorch": 7, "x_vormean": 8, "y_vormean": 9, "z_vormean": 10,
            "gmax": 11, "umax": 12, "vmax": 13, "wmax": 14, "usgmax": 15, "lsgmax": 16, "rolling_mean_gmax": 17}
    for name, j in want.items():
        assert st[S[name]] == 100.0 + j, name


def test_beltrami_mode_stats(oracle_mod):
    check_beltrami_stats(oracle_mod)


def test_horizontal_vorticity_stats(oracle_mod):
    check_horizontal_vorticity_stats(oracle_mod)


def test_vortex_stats(oracle_mod):
    check_vortex_stats(oracle_mod)


def test_adapt_values_in_their_slots(oracle_mod):
    check_adapt_slots(oracle_mod)


def test_rossby_number_without_planetary_vorticity(oracle_mod):
    """f_cor(3) = 0: ps3d divides by zero (ro = +-inf), as the record does."""
    x, y, z = ics.grid(N, N, N, BOX["lower"], BOX["extent"])
    _, vor = ics.beltrami_mode(x, y, z, 2, 1, 1)
    o = _oracle(oracle_mod, vor, fcor3=0.0)
    with np.errstate(all="ignore"):
        st = o.field_stats()
    o.close()
    assert st[S["ro_min"]] == -np.inf and st[S["ro_max"]] == np.inf
