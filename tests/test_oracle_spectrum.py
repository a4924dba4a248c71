"""Pins of the oracle's genspec spectrum (src/genspec.f90:44-124) and of the
kinetic-energy / helicity / cross-helicity means (field_diagnostics.f90:85-122,
plotting/nc_reader.py:94-120), from hand-computed shells and closed forms."""
import numpy as np
import pytest

from ps3d_amd import ics

PI = np.pi
EXT = (2 * PI, 2 * PI, PI)


def make(oracle_mod, nx, ny, nz, extent=EXT):
    return oracle_mod.Oracle(oracle_mod.make_params(nx, ny, nz, lower=(0.0, 0.0, 0.0), extent=extent))


def mesh(nx, ny, nz, extent=EXT):
    x, y, z = ics.grid(nx, ny, nz, (0.0, 0.0, 0.0), extent)
    return np.meshgrid(x, y, z, indexing="ij")


def test_kinetic_energy_weights_the_lids_by_half(oracle_mod):
    """A uniform |u| = 1 gives 1/2; energy only on the lids counts a quarter per
    lid level."""
    vel = np.zeros((3, 4, 6, 9))
    vel[0] = 1.0
    assert abs(oracle_mod.kinetic_energy(vel) - 0.5) < 1e-15
    vel[:] = 0.0
    vel[2, :, :, 0] = 2.0
    assert abs(oracle_mod.kinetic_energy(vel) - 0.25 * 4.0 / 8) < 1e-15


def test_fortran_nint_rounds_halves_up(oracle_mod):
    x = np.array([0.5, 1.5, 2.5, 2.4999999999999996, 3.0, 0.49999999999999994])
    assert oracle_mod._nint(x).tolist() == [1.0, 2.0, 3.0, 2.0, 3.0, 0.0]


def test_single_cosine_mode_lands_in_its_shell(oracle_mod):
    """16^3 on (2 pi, 2 pi, pi): kmax = nint(sqrt(8^2 + 8^2 + 16^2)) = nint(19.596) =
    20, dk = 20 / 19.596 = 1.0206.  u = cos 3x cos 4z has |K| = 5, so its energy
    sits in shell int(5 / 1.0206) = 4 (not 5: the shells are dk wide)."""
    o = make(oracle_mod, 16, 16, 16)
    X, Y, Z = mesh(16, 16, 16)
    vel = np.zeros((3, 16, 16, 17))
    vel[0] = np.cos(3 * X) * np.cos(4 * Z)
    k, spec, num = oracle_mod.spectrum(o, vel)
    dk = k[1]
    assert len(spec) == 21
    assert abs(dk - 20.0 / np.sqrt(384.0)) < 1e-15
    assert int(np.argmax(spec)) == 4
    assert np.abs(np.delete(spec, 4)).max() < 1e-13 * spec[4]
    # the spectrum integrates to the kinetic energy, 1/2 <cos^2 3x> <cos^2 4z> = 1/8
    assert abs(np.sum(spec) * dk - 0.125) < 1e-15
    o.close()


def test_sine_mode_of_w_uses_the_sine_transform(oracle_mod):
    """w = cos 2x sin 3z: |K| = sqrt(13) = 3.61, nint 4, shell int(4 / 1.0206) = 3.
    A cosine transform of w would spread it over many kz."""
    o = make(oracle_mod, 16, 16, 16)
    X, Y, Z = mesh(16, 16, 16)
    vel = np.zeros((3, 16, 16, 17))
    vel[2] = np.cos(2 * X) * np.sin(3 * Z)
    k, spec, num = oracle_mod.spectrum(o, vel)
    assert int(np.argmax(spec)) == 3
    assert np.abs(np.delete(spec, 3)).max() < 1e-13 * spec[3]
    o.close()


def test_two_modes_weigh_by_shell_volume_over_count(oracle_mod):
    """Energies a^2/4 and b^2/4 (per unit volume) in shells m1, m2: spec(m) =
    e(m) * 4/3 pi dk^3 ((m+1)^3 - m^3) / num(m), up to the common normalisation."""
    n = 16
    o = make(oracle_mod, n, n, n)
    X, Y, Z = mesh(n, n, n)
    vel = np.zeros((3, n, n, n + 1))
    vel[1] = 2.0 * np.cos(1 * X) * np.cos(2 * Z) + 0.5 * np.cos(6 * Y) * np.cos(8 * Z)
    k, spec, num = oracle_mod.spectrum(o, vel)
    dk = k[1]
    m1, m2 = int(np.rint(np.sqrt(5.0)) / dk), int(np.rint(10.0) / dk)       # int(2 / dk), int(10 / dk)
    assert (m1, m2) == (1, 9)
    nz = np.flatnonzero(spec > 1e-12 * spec.max())
    assert nz.tolist() == [m1, m2]
    shell = lambda m: ((m + 1) ** 3 - m ** 3) / num[m]
    assert abs(spec[m1] / spec[m2] - (4.0 * shell(m1)) / (0.25 * shell(m2))) < 1e-12 * spec[m1] / spec[m2]
    o.close()


def test_point_counts_cover_the_grid_once(oracle_mod):
    """num(m) over all shells counts every (kx, ky, kz) slot once; brute force
    over the slots of a small grid with unequal extents (dk = 2.24 here)."""
    nx, ny, nz = 8, 6, 10
    ext = (2.0, 3.0, 1.5)
    o = make(oracle_mod, nx, ny, nz, ext)
    vel = np.random.default_rng(1).standard_normal((3, nx, ny, nz + 1))
    k, spec, num = oracle_mod.spectrum(o, vel)
    assert num.sum() == nx * ny * (nz + 1)
    kxs = [2 * PI / ext[0] * min(i, nx - i) for i in range(nx)]
    kys = [2 * PI / ext[1] * min(j, ny - j) for j in range(ny)]
    kzs = [PI / ext[2] * q for q in range(nz + 1)]
    mags = [int(np.floor(np.sqrt(a * a + b * b + c * c) + 0.5)) for a in kxs for b in kys for c in kzs]
    kmax = max(mags)
    dk = kmax / np.sqrt((nx / 2) ** 2 + (ny / 2) ** 2 + nz ** 2)
    brute = np.zeros(kmax + 1, dtype=int)
    for g in mags:
        if int(g / dk) <= kmax:
            brute[int(g / dk)] += 1
    assert len(num) == kmax + 1
    assert np.array_equal(num, brute)
    assert abs(np.sum(spec) * k[1] - oracle_mod.kinetic_energy(vel)) < 1e-14
    o.close()


def test_moments_closed_forms(oracle_mod):
    """Beltrami mode (w = alpha u): helicity = 2 alpha ke, |u x w| = 0; uniform
    u = (1, 0, 0) with w = (0, 0, c): |u x w| = c everywhere."""
    nx = ny = nz = 16
    x, y, z = ics.grid(nx, ny, nz, (0.0, 0.0, 0.0), EXT)
    vel, vor = ics.beltrami_mode(x, y, z, 2, 1, 3)
    alpha = np.sqrt(4 + 1 + 9)
    m = oracle_mod.moments(vel, vor)
    assert abs(m["helicity"] - 2 * alpha * m["ke"]) < 1e-14 * abs(m["helicity"])
    assert m["cross_helicity"] < 1e-14 and m["cross_helicity_max"] < 1e-13
    u = np.zeros((3, nx, ny, nz + 1))
    u[0] = 1.0
    w = np.zeros_like(u)
    w[2] = 0.75
    m = oracle_mod.moments(u, w)
    assert abs(m["cross_helicity"] - 0.75) < 1e-15 and m["cross_helicity_max"] == 0.75
    assert m["helicity"] == 0.0 and abs(m["ke"] - 0.5) < 1e-15
