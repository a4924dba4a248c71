"""Initial-condition generators (ps3d_amd/ics.py), checked on the CPU.

The random-phase field must have the shell spectrum E(k) ~ k^4 exp(-2 (k/k0)^2)
that its workloads name.  The spectrum is measured here from the physical
velocity alone (numpy FFT in x, y; scipy's DCT-I / DST-I for the cosine /
sine series in z), binned into shells of unit width around integer |k| as
genspec does (genspec.f90:50-124), and the exponent of k is fitted after
dividing out the Gaussian factor.
"""
import numpy as np
import pytest
import scipy.fft as sf

from ps3d_amd import ics

PI = np.pi


def shell_spectrum(vel):
    """Shell energy E(k), k = 0, 1, ..., of a velocity on the 2 pi x 2 pi x pi
    box (integer wavenumbers in x, y and z) whose u, v are cosine series and w
    a sine series in z.  sum_k E(k) is the mean of |u|^2 / 2 over the box."""
    _, nx, ny, nz1 = vel.shape
    nz = nz1 - 1
    kx = np.fft.fftfreq(nx, d=1.0 / nx)
    ky = np.fft.fftfreq(ny, d=1.0 / ny)
    m = np.arange(nz + 1)
    KX, KY, M = np.meshgrid(kx, ky, m, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + M ** 2)
    e = np.zeros(K.shape)
    for c in range(3):
        h = np.fft.fft2(vel[c], axes=(0, 1)) / (nx * ny)      # f = sum c_k e^{i k.x}
        if c < 2:
            a = sf.dct(h.real, type=1, axis=2) / nz + 1j * sf.dct(h.imag, type=1, axis=2) / nz
            a[..., 0] *= 0.5
            a[..., nz] *= 0.5
            wz = np.full(nz + 1, 0.5)
            wz[0] = 1.0
            e += wz * np.abs(a) ** 2
        else:
            b = np.zeros(h.shape, dtype=complex)
            b[..., 1:nz] = (sf.dst(h.real[..., 1:nz], type=1, axis=2)
                            + 1j * sf.dst(h.imag[..., 1:nz], type=1, axis=2)) / nz
            e += 0.5 * np.abs(b) ** 2
    e *= 0.5
    shell = np.rint(K).astype(int)
    return np.bincount(shell.ravel(), weights=e.ravel())


def fitted_exponent(E, k0, kmin=2, kmax=12):
    k = np.arange(kmin, kmax + 1)
    y = np.log(E[k]) + 2.0 * (k / k0) ** 2
    return np.polyfit(np.log(k), y, 1)[0]


def test_shell_spectrum_sums_to_kinetic_energy():
    vel, _ = ics.random_phase_fields(32, 32, 32, k0=3.0, seed=4)
    wz = np.ones(33)
    wz[0] = wz[-1] = 0.5
    ke = 0.5 * ((vel ** 2).sum(0) * wz).sum() / (32 * 32 * 32)
    assert abs(shell_spectrum(vel).sum() - ke) < 1e-12 * ke


@pytest.mark.parametrize("k0,seed", [(4.0, 1), (3.0, 7), (5.0, 3)])
def test_random_phase_spectrum_k4(k0, seed):
    vel, vor = ics.random_phase_fields(64, 64, 64, k0=k0, seed=seed)
    E = shell_spectrum(vel)
    p = fitted_exponent(E, k0)
    assert abs(p - 4.0) < 0.4, p
    assert abs(int(np.argmax(E)) - k0) <= 1, (np.argmax(E), k0)
    # the vorticity returned is the curl of the velocity returned
    assert np.array_equal(vor, ics.random_phase_isotropic(64, 64, 64, k0=k0, seed=seed))


def test_random_phase_is_solenoidal_with_free_slip_lids():
    vel, vor = ics.random_phase_fields(32, 32, 32, k0=4.0, seed=2)
    assert np.abs(vel[2][..., 0]).max() < 1e-12 and np.abs(vel[2][..., -1]).max() < 1e-12
    k = np.fft.fftfreq(32, d=1.0 / 32)
    div = (np.real(np.fft.ifft(1j * k[:, None, None] * np.fft.fft(vel[0], axis=0), axis=0))
           + np.real(np.fft.ifft(1j * k[None, :, None] * np.fft.fft(vel[1], axis=1), axis=1)))
    # w is a sine series: its z-derivative by the same spectral route
    w = vel[2]
    b = sf.dst(w[..., 1:-1], type=1, axis=2) / 32
    m = np.arange(1, 32)
    wz = (b * m).dot(np.cos(np.outer(m, np.linspace(0, PI, 33))))
    assert np.abs(div + wz).max() < 1e-10 * np.abs(vel).max()
