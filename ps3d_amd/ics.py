"""Initial-condition generators (host side).

ps3d ships one generator, src/beltrami.f90 (one Beltrami mode written to a
NetCDF field file, get_flow_vorticity at :162-181).  The workloads this package
is measured on need two more, written here:

* beltrami_multimode: a sum of Beltrami modes with seeded random phases plus
  seeded Gaussian noise.  Each mode is the reference mode translated in z so
  that w vanishes on both free-slip lids for any integer m.
* random_phase_isotropic: a solenoidal velocity with energy spectrum
  E(k) ~ k^4 exp(-2 (k/k0)^2), w = 0 on the lids, and its vorticity.

Fields come back in ps3d's layout: numpy (3, nx, ny, nz+1) = Fortran
(0:nz, 0:ny-1, 0:nx-1, 3) (src/fields/fields.f90:59-65).  The PRNG is numpy's
PCG64 (numpy.random.default_rng) with the given seed.
"""
import numpy as np

PI = np.pi

BELTRAMI_MODES = ((1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 2))


def grid(nx, ny, nz, lower, extent):
    dx = np.asarray(extent, dtype=float) / np.array([nx, ny, nz], dtype=float)
    x = lower[0] + np.arange(nx) * dx[0]
    y = lower[1] + np.arange(ny) * dx[1]
    z = lower[2] + np.arange(nz + 1) * dx[2]
    return x, y, z


def beltrami_mode(x, y, z, k, l, m, phase=0.0, zmin=0.0):
    """Velocity and vorticity of one Beltrami mode (omega = alpha u,
    alpha = sqrt(k^2+l^2+m^2)).  It is the mode of src/beltrami.f90:162-181
    with z -> z - zmin - pi/(2m), so that w = sin(m (z - zmin)) cos(kx+ly+phase)
    is zero on z = zmin and z = zmin + pi for integer m."""
    alpha = np.sqrt(k * k + l * l + m * m)
    fk2l2 = 1.0 / (k * k + l * l)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    zeta = m * (Z - zmin)
    sm, cm = np.sin(zeta), np.cos(zeta)
    th = k * X + l * Y + phase
    s, c = np.sin(th), np.cos(th)
    u = fk2l2 * (-k * m * cm - l * alpha * sm) * s
    v = fk2l2 * (-l * m * cm + k * alpha * sm) * s
    w = sm * c
    vel = np.stack([u, v, w])
    return vel, alpha * vel


def beltrami_multimode(nx, ny, nz, seed=1, noise=1.0e-3, modes=BELTRAMI_MODES,
                       lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI)):
    """Superposition of Beltrami modes with unit amplitudes and phases drawn
    uniformly in [0, 2 pi), plus `noise` x N(0,1) added to every vorticity
    component at every grid point (same PCG64 stream).  Returns vorticity
    (3, nx, ny, nz+1)."""
    rng = np.random.default_rng(seed)
    phases = rng.uniform(0.0, 2 * PI, size=len(modes))
    x, y, z = grid(nx, ny, nz, lower, extent)
    vor = np.zeros((3, nx, ny, nz + 1))
    for (k, l, m), ph in zip(modes, phases):
        # wavenumbers in units of the box: 2 pi / Lx per k, pi / Lz per m
        kk = k * 2 * PI / extent[0]
        ll = l * 2 * PI / extent[1]
        mm = m * PI / extent[2]
        _, w = beltrami_mode(x, y, z, kk, ll, mm, ph, zmin=lower[2])
        vor += w
    if noise:
        vor += noise * rng.standard_normal(vor.shape)
    return vor


def random_phase_isotropic(nx, ny, nz, k0=4.0, seed=1, lower=(0.0, 0.0, 0.0),
                           extent=(2 * PI, 2 * PI, PI), urms=1.0):
    """Random-phase solenoidal flow with E(k) ~ k^4 exp(-2 (k/k0)^2); returns
    the vorticity (3, nx, ny, nz+1) of random_phase_fields."""
    return random_phase_fields(nx, ny, nz, k0, seed, lower, extent, urms)[1]


def random_phase_fields(nx, ny, nz, k0=4.0, seed=1, lower=(0.0, 0.0, 0.0),
                        extent=(2 * PI, 2 * PI, PI), urms=1.0):
    """Random-phase solenoidal flow with E(k) ~ k^4 exp(-2 (k/k0)^2).

    The velocity is the curl of a vector potential A whose x,y components are
    sine series and z component a cosine series in z, so u,v are cosine series,
    w a sine series vanishing on both lids, and div u = 0 exactly.  Each Fourier
    mode (kx, ky, m) gets complex Gaussian coefficients of rms |A| =
    exp(-(K/k0)^2): then |u| ~ K |A| per mode, the number of modes in a shell
    of radius k grows as k^2, and the shell energy is
    E(k) ~ k^2 * k^2 exp(-2 (k/k0)^2), peaking at k = k0.  Returns
    (velocity, vorticity), each (3, nx, ny, nz+1), scaled to rms speed `urms`."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = extent
    kx = 2 * PI * np.fft.fftfreq(nx, d=1.0 / nx) / Lx
    ky = 2 * PI * np.fft.fftfreq(ny, d=1.0 / ny) / Ly
    mz = PI * np.arange(nz + 1) / Lz
    KX, KY, MZ = np.meshgrid(kx, ky, mz, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + MZ ** 2)
    # per mode |u|^2 ~ K^2 |A|^2 = E(K) / K^2 (shell area)  ->  |A| = exp(-(K/k0)^2)
    amp = np.where(K > 0, np.exp(-((K / k0) ** 2)), 0.0)
    # keep the 2/3 of the resolved band so that the field is smooth
    amp[np.abs(KX) > (2.0 / 3.0) * np.abs(kx).max()] = 0.0
    amp[np.abs(KY) > (2.0 / 3.0) * np.abs(ky).max()] = 0.0
    amp[MZ > (2.0 / 3.0) * mz.max()] = 0.0
    A = [amp * (rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)) for _ in range(3)]
    x, y, z = grid(nx, ny, nz, lower, extent)
    zz = z - lower[2]
    S = np.sin(np.outer(mz, zz))       # (m, iz)
    Cc = np.cos(np.outer(mz, zz))

    def synth(coef, zbasis):
        # sum over m of coef(kx,ky,m) * zbasis(m, z), then inverse 2-D DFT in x,y
        g = np.tensordot(coef, zbasis, axes=([2], [0]))          # (nx, ny, nz+1) complex
        return np.real(np.fft.ifft2(g, axes=(0, 1))) * nx * ny

    ikx, iky = 1j * KX, 1j * KY
    # u = dA3/dy - dA2/dz ; v = dA1/dz - dA3/dx ; w = dA2/dx - dA1/dy
    # A1, A2 ~ sin(m z): d/dz -> m cos ; A3 ~ cos(m z): d/dz -> -m sin
    u = synth(iky * A[2], Cc) - synth(MZ * A[1], Cc)
    v = synth(MZ * A[0], Cc) - synth(ikx * A[2], Cc)
    w = synth(ikx * A[1] - iky * A[0], S)
    # vorticity: xi = dw/dy - dv/dz ; eta = du/dz - dw/dx ; zeta = dv/dx - du/dy
    U = iky * A[2] - MZ * A[1]                  # cos coefficients of u
    V = MZ * A[0] - ikx * A[2]                  # cos coefficients of v
    W = ikx * A[1] - iky * A[0]                 # sin coefficients of w
    xi = synth(iky * W, S) - synth(-MZ * V, S)
    eta = synth(-MZ * U, S) - synth(ikx * W, S)
    zeta = synth(ikx * V - iky * U, Cc)
    vel = np.stack([u, v, w])
    vor = np.stack([xi, eta, zeta])
    rms = np.sqrt(np.mean(vel ** 2))
    if rms > 0:
        vel *= urms / rms
        vor *= urms / rms
    return vel, vor
