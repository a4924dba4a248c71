"""Inputs, closed-form answers and tolerances of ps3d's own unit tests.

Each builder restates one program under unit-tests/ of the reference (its grid,
box, closed-form field and tolerance) so that the oracle (CPU) and the HIP path
(GPU) are both checked against the same numbers.  Arrays have numpy shape
(nx, ny, nz+1) = Fortran (0:nz, 0:ny-1, 0:nx-1); vector fields (3, nx, ny, nz+1).
"""
import numpy as np

PI = np.pi


def grid(nx, ny, nz, lower, extent):
    dx = np.array(extent, dtype=float) / np.array([nx, ny, nz], dtype=float)
    x = lower[0] + np.arange(nx) * dx[0]
    y = lower[1] + np.arange(ny) * dx[1]
    z = lower[2] + np.arange(nz + 1) * dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return X, Y, Z


def diffx_case():
    """unit-tests/test_diffx.f90:26-75: cos 4x -> fftxyp2s -> diffx -> fftxys2p = -4 sin 4x."""
    g = dict(nx=64, ny=128, nz=64, lower=(-PI, -PI, -PI), extent=(2 * PI,) * 3)
    X, Y, Z = grid(g["nx"], g["ny"], g["nz"], g["lower"], g["extent"])
    return g, np.cos(4 * X), -4 * np.sin(4 * X), 1.0e-12


def diffy_case():
    """unit-tests/test_diffy.f90: nx=128, ny=64, nz=64; cos 4y -> -4 sin 4y."""
    g = dict(nx=128, ny=64, nz=64, lower=(-PI, -PI, -PI), extent=(2 * PI,) * 3)
    X, Y, Z = grid(g["nx"], g["ny"], g["nz"], g["lower"], g["extent"])
    return g, np.cos(4 * Y), -4 * np.sin(4 * Y), 1.0e-12


def diffz_cases():
    """unit-tests/test_diffz_{1,2,3,4}.f90 (central_diffz), tolerances at
    test_diffz_1.f90:60, _2:71, _3:71, _4:59."""
    out = []
    g = dict(nx=32, ny=32, nz=32, lower=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0))
    X, Y, Z = grid(32, 32, 32, g["lower"], g["extent"])
    out.append(("diffz_1", g, Z.copy(), np.ones_like(Z), 1.0e-14))
    g2 = dict(nx=16, ny=32, nz=32, lower=(0.0, -0.5 * PI, 0.0), extent=(PI, 2 * PI, 2 * PI))
    X, Y, Z = grid(16, 32, 32, g2["lower"], g2["extent"])
    out.append(("diffz_2", g2, Z * np.cos(2 * X) * np.sin(Y), np.cos(2 * X) * np.sin(Y), 2.0e-14))
    out.append(("diffz_3", g2, 6 * Z + 3 * np.cos(2 * X) * np.sin(Y) * Z, 6 + 3 * np.cos(2 * X) * np.sin(Y), 2.0e-13))
    g4 = dict(nx=32, ny=32, nz=32, lower=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0))
    X, Y, Z = grid(32, 32, 32, g4["lower"], g4["extent"])
    out.append(("diffz_4", g4, 1 - 3 * Z ** 2, -6 * Z, 1.0e-1))
    return out


def beltrami_mode(X, Y, Z, k, l, m):
    """Single Beltrami mode of unit-tests/test_vor2vel_1.f90:55-77 (and
    src/beltrami.f90:162-181): returns (vel, vor) with vor = alpha * vel."""
    alpha = np.sqrt(k * k + l * l + m * m)
    fk2l2 = 1.0 / (k * k + l * l)
    cosmz, sinmz = np.cos(m * Z), np.sin(m * Z)
    s, c = np.sin(k * X + l * Y), np.cos(k * X + l * Y)
    u = fk2l2 * (k * m * sinmz - l * alpha * cosmz) * s
    v = fk2l2 * (l * m * sinmz + k * alpha * cosmz) * s
    w = cosmz * c
    vel = np.stack([u, v, w])
    return vel, alpha * vel


def vor2vel_cases():
    """unit-tests/test_vor2vel_{1..5}.f90 with their tolerances
    (test_vor2vel_1.f90:99, _2:104, _3, _4, _5)."""
    out = []
    g1 = dict(nx=32, ny=32, nz=32, lower=(-0.5 * PI,) * 3, extent=(PI,) * 3)
    X, Y, Z = grid(32, 32, 32, g1["lower"], g1["extent"])
    vel, vor = beltrami_mode(X, Y, Z, 2.0, 2.0, 1.0)
    out.append(("vor2vel_1", g1, vor, vel, 1.0e-14))

    g2 = dict(nx=32, ny=32, nz=32, lower=(-0.5, -0.5, 0.0), extent=(1.0, 1.0, 1.0))
    X, Y, Z = grid(32, 32, 32, g2["lower"], g2["extent"])
    l = 2 * PI
    k = 2 * l
    klsq = k * k - l * l
    f = 2 * Z - Z ** 2 - Z ** 3
    dfdz = 2 - 2 * Z - 3 * Z ** 2
    d2 = -2 - 6 * Z
    vel = np.stack([k * dfdz * np.cos(k * X) * np.sin(l * Y),
                    -l * dfdz * np.sin(k * X) * np.cos(l * Y),
                    klsq * f * np.sin(k * X) * np.sin(l * Y)])
    vor = np.stack([l * (klsq * f + d2) * np.sin(k * X) * np.cos(l * Y),
                    k * (d2 - klsq * f) * np.cos(k * X) * np.sin(l * Y),
                    -2 * k * l * dfdz * np.cos(k * X) * np.cos(l * Y)])
    out.append(("vor2vel_2", g2, vor, vel, 1.2e-2))

    g3 = dict(nx=32, ny=32, nz=32, lower=(-0.5, -0.5, 0.0), extent=(1.0, 1.0, 1.0))
    X, Y, Z = grid(32, 32, 32, g3["lower"], g3["extent"])
    zero = np.zeros_like(Z)
    upper = g3["lower"][2] + g3["extent"][2]
    out.append(("vor2vel_3", g3, np.stack([zero, np.ones_like(Z), zero]),
                np.stack([Z - 0.5 * (g3["lower"][2] + upper), zero, zero]), 1.0e-15))
    out.append(("vor2vel_4", g3, np.stack([zero, 2 * Z, zero]),
                np.stack([Z ** 2 - 1.0 / 3.0, zero, zero]), 1.0e-15))
    out.append(("vor2vel_5", g3, np.stack([zero, 3 * Z ** 2, zero]),
                np.stack([Z ** 3 - 0.25, zero, zero]), 4.0e-6))
    return out


def implicit_rk_steps():
    """unit-tests/test_implicit_rk.f90:47-94: 32^3 on the unit box, svor = 2,
    source = 1, dissipation rate 1, dt = 0.1 halved 4 times; the error reported
    (and held to 1e-12) is the one of the last, smallest step."""
    g = dict(nx=32, ny=32, nz=32, lower=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0))
    dts = [0.1 / 2 ** i for i in range(4)]
    return g, dts, 1.0e-12
