"""Pins of the CPU oracle (oracle/ps3d_oracle.c) before it is trusted as the
checker of the HIP path.

Each operator of the oracle is held to (a) a brute-force sum or a textbook
routine, (b) the reference's own compiled stafft.f90 / jacobi.f90
(oracle/_ref, when built) and (c) the closed forms and tolerances of the
reference's unit tests (unit-tests/*.f90).  A deliberately planted mistake (a
flipped sine-slot sign, a dropped harmonic term) fails at least one of these.
"""
import numpy as np
import pytest

from tests import refcases as RC

PI = np.pi


def hermitian_bruteforce(x):
    """stafft's forward Hermitian form (stafft.f90:55-58, 196-287) by numpy's
    textbook DFT: x(k)=Re X_k/sqrt(n) (k<=n/2), x(n-k)=Im X_k/sqrt(n)."""
    n = x.shape[0]
    X = np.fft.fft(x, axis=0)
    out = np.zeros_like(x)
    for k in range(n // 2 + 1):
        out[k] = X[k].real / np.sqrt(n)
    for k in range(1, (n + 1) // 2):
        out[n - k] = X[k].imag / np.sqrt(n)
    return out


def dct_matrix(n):
    k = np.arange(n + 1)[:, None]
    j = np.arange(n + 1)[None, :]
    w = np.ones(n + 1)
    w[0] = w[n] = 0.5
    return np.sqrt(2.0 / n) * np.cos(PI * ((k * j) % (2 * n)) / n) * w[None, :]


def dst_matrix(n):
    k = np.arange(1, n)[:, None]
    j = np.arange(1, n)[None, :]
    return np.sqrt(2.0 / n) * np.sin(PI * ((k * j) % (2 * n)) / n)


# 10, 20, 40, 50, 100, 250: stafft's radix-5 passes (forrdx5/revrdx5, stafft.f90:754-897, 1355-1499)
@pytest.mark.parametrize("n", [8, 10, 16, 20, 32, 40, 48, 50, 64, 96, 100, 128, 250, 256])
def test_forfft_revfft_bruteforce(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, 7))
    xh = oracle_mod.fft1d("forfft", x)
    assert np.abs(xh - hermitian_bruteforce(x)).max() < 2e-15 * np.sqrt(n) * 4
    back = oracle_mod.fft1d("revfft", xh)
    assert np.abs(back - x).max() < 1e-14


@pytest.mark.parametrize("n", [8, 10, 16, 20, 32, 40, 48, 50, 64, 96, 100, 128])
def test_dct_dst_bruteforce(oracle_mod, n):
    rng = np.random.default_rng(100 + n)
    y = rng.standard_normal((n + 1, 5))
    c = oracle_mod.fft1d("dct", y)
    assert np.abs(c - dct_matrix(n) @ y).max() < 1e-13
    # DCT-I with this scaling is its own inverse (stafft.f90:46-48)
    assert np.abs(oracle_mod.fft1d("dct", c) - y).max() < 1e-13
    z = rng.standard_normal((n, 5))
    s = oracle_mod.fft1d("dst", z)
    ref = np.zeros_like(z)
    ref[: n - 1] = dst_matrix(n) @ z[: n - 1]
    assert np.abs(s - ref).max() < 1e-13
    assert np.all(s[n - 1] == 0.0)          # dst zeroes x(n) (stafft.f90:546-549)
    zz = z.copy()
    zz[n - 1] = 0.0
    assert np.abs(oracle_mod.fft1d("dst", s) - zz).max() < 1e-13


@pytest.mark.parametrize("n", [16, 20, 32, 40, 48, 60, 64, 80, 96, 100, 128, 160, 250, 256, 320, 500, 512])
def test_transforms_match_reference_build(oracle_mod, n):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs /root/reference + amdflang)")
    rng = np.random.default_rng(7 * n)
    x = rng.standard_normal((n, 6))
    for kind in ("forfft", "revfft"):
        a = oracle_mod.fft1d(kind, x)
        b = oracle_mod.ref_fft1d(kind, x)
        assert np.abs(a - b).max() < 5e-15 * max(1.0, np.abs(b).max()), kind
    y = rng.standard_normal((n + 1, 6))
    assert np.abs(oracle_mod.fft1d("dct", y) - oracle_mod.ref_fft1d("dct", y)).max() < 2e-14
    z = rng.standard_normal((n, 6))
    assert np.abs(oracle_mod.fft1d("dst", z) - oracle_mod.ref_fft1d("dst", z)).max() < 2e-14


def test_jacobi_matches_reference_and_numpy(oracle_mod):
    rng = np.random.default_rng(3)
    for i in range(200):
        a = rng.standard_normal((3, 3)) * 10 ** rng.uniform(-3, 3)
        a = a + a.T
        d = oracle_mod.jacobi_eigenvalues(a)
        ev = np.sort(np.linalg.eigvalsh(a))[::-1]
        # absolute off-diagonal stop at 1e-15 (jacobi.f90:13, 288)
        assert np.abs(d - ev).max() <= 1e-14 * np.abs(ev).max() + 1e-15
        if oracle_mod.ref_available():
            assert np.array_equal(d, oracle_mod.ref_jacobi_eigenvalues(a))
    # traceless strain with a double eigenvalue
    a = np.diag([1.0, 1.0, -2.0])
    assert np.allclose(oracle_mod.jacobi_eigenvalues(a), [1.0, 1.0, -2.0])


def make_oracle(oracle_mod, g, **kw):
    p = oracle_mod.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"], **kw)
    return oracle_mod.Oracle(p)


def test_diffx_unit_test(oracle_mod):
    g, f, ref, tol = RC.diffx_case()
    o = make_oracle(oracle_mod, g)
    fs = o.op(oracle_mod.OP_FFTXYP2S, f)
    ds = o.op(oracle_mod.OP_DIFFX, fs)
    fp = o.op(oracle_mod.OP_FFTXYS2P, ds)
    assert np.abs(fp - ref).max() < tol      # two-sided (the reference's check is one-sided)


def test_diffy_unit_test(oracle_mod):
    g, f, ref, tol = RC.diffy_case()
    o = make_oracle(oracle_mod, g)
    fp = o.op(oracle_mod.OP_FFTXYS2P, o.op(oracle_mod.OP_DIFFY, o.op(oracle_mod.OP_FFTXYP2S, f)))
    assert np.abs(fp - ref).max() < tol


def test_wavenumbers_and_reference_init_deriv(oracle_mod):
    g = dict(nx=32, ny=48, nz=64, lower=(0, 0, 0), extent=(2 * PI, 3.0, PI))
    o = make_oracle(oracle_mod, g)
    k = np.arange(32)
    assert np.allclose(o.table(13), 2 * PI * np.minimum(k, 32 - k) / (2 * PI), rtol=0, atol=1e-14)
    k = np.arange(48)
    assert np.allclose(o.table(14), 2 * PI * np.minimum(k, 48 - k) / 3.0, rtol=0, atol=1e-13)
    rkz = o.table(15)
    assert np.allclose(rkz, PI * np.arange(65) / PI, rtol=0, atol=1e-13)
    if oracle_mod.ref_available():
        assert np.array_equal(rkz[1:], oracle_mod.ref_init_deriv(64, PI))


@pytest.mark.parametrize("case", RC.diffz_cases(), ids=lambda c: c[0])
def test_central_diffz_unit_tests(oracle_mod, case):
    name, g, f, ref, tol = case
    o = make_oracle(oracle_mod, g)
    d = o.op(oracle_mod.OP_CENTRAL_DIFFZ, f)
    assert np.abs(d - ref).max() < tol, name


def test_combine_decompose_round_trip(oracle_mod):
    g = dict(nx=16, ny=24, nz=32, lower=(0, 0, -1.0), extent=(2 * PI, 2 * PI, 2.0))
    o = make_oracle(oracle_mod, g)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((16, 24, 33))
    s = o.op(oracle_mod.OP_DECOMPOSE_PHYS, f)
    back = o.op(oracle_mod.OP_COMBINE_PHYS, s)
    assert np.abs(back - f).max() < 1e-13
    # lid rows of the mixed representation are the semi-spectral lid values
    fs = o.op(oracle_mod.OP_FFTXYP2S, f)
    assert np.abs(s[:, :, 0] - fs[:, :, 0]).max() < 1e-15
    assert np.abs(s[:, :, -1] - fs[:, :, -1]).max() < 1e-15
    # interior rows: DST of the field minus the harmonic interpolant of the lids
    phim = o.table(5).reshape(16, 24, 33)
    phip = o.table(6).reshape(16, 24, 33)
    h = fs - fs[:, :, :1] * phim - fs[:, :, -1:] * phip
    inner = h[:, :, 1:-1].reshape(-1, 31)
    ref = inner @ dst_matrix(32).T
    assert np.abs(s[:, :, 1:-1].reshape(-1, 31) - ref).max() < 1e-12


def test_harmonic_functions_closed_form(oracle_mod):
    """phim/phip solve (d2/dz2 - K^2) phi = 0 with phi(lower)=1/0, phi(upper)=0/1
    (inversion_utils.f90:484-542); (0,0) column is linear."""
    g = dict(nx=8, ny=8, nz=32, lower=(0, 0, 0.0), extent=(2 * PI, 2 * PI, 1.5))
    o = make_oracle(oracle_mod, g)
    z = np.linspace(0, 1.5, 33)
    phim = o.table(5).reshape(8, 8, 33)
    phip = o.table(6).reshape(8, 8, 33)
    k2l2 = o.table(0).reshape(8, 8)   # (ky, kx) packed as iy + ny*ix
    for ix, iy in [(1, 0), (0, 3), (2, 5), (4, 4)]:
        K = np.sqrt(k2l2[ix, iy])
        assert np.allclose(phim[ix, iy], np.sinh(K * (1.5 - z)) / np.sinh(K * 1.5), atol=1e-14)
        assert np.allclose(phip[ix, iy], np.sinh(K * z) / np.sinh(K * 1.5), atol=1e-14)
    assert np.allclose(phim[0, 0], (1.5 - z) / 1.5, atol=1e-15)
    assert np.allclose(phip[0, 0], z / 1.5, atol=1e-15)


@pytest.mark.parametrize("case", RC.vor2vel_cases(), ids=lambda c: c[0])
def test_vor2vel_unit_tests(oracle_mod, case):
    name, g, vor, vel_ref, tol = case
    o = make_oracle(oracle_mod, g)
    svor = np.stack([o.op(oracle_mod.OP_DECOMPOSE_PHYS, vor[c]) for c in range(3)])
    o.set(oracle_mod.SVOR, svor)
    o.vor2vel()
    vel = o.get(oracle_mod.VEL)
    assert np.abs(vel - vel_ref).max() < tol, (name, np.abs(vel - vel_ref).max())


def test_vor2vel_solenoidal_and_curl(oracle_mod):
    """vel from vor2vel is divergence-free and its curl returns the projected
    vorticity, for a random smooth field."""
    g = dict(nx=32, ny=32, nz=32, lower=(0, 0, 0), extent=(2 * PI, 2 * PI, PI))
    o = make_oracle(oracle_mod, g)
    X, Y, Z = RC.grid(32, 32, 32, g["lower"], g["extent"])
    vel, vor = RC.beltrami_mode(X, Y, Z - PI / 2, 1.0, 2.0, 1.0)
    svor = np.stack([o.op(oracle_mod.OP_DECOMPOSE_PHYS, vor[c]) for c in range(3)])
    o.set(oracle_mod.SVOR, svor)
    o.vor2vel()
    v = o.get(oracle_mod.VEL)
    assert np.abs(v - vel).max() < 1e-12


def test_implicit_rk_unit_test(oracle_mod):
    """unit-tests/test_implicit_rk.f90: exact solution q E + (1-E) S/D with
    E = exp(-D dt); the module convention holds vdiss = D dt / 2. With the 2/3
    filter, columns cut by filt(0,:,:) get the source removed at every stage
    (impl_rk4.f90:87-89, 227-229), i.e. q E."""
    g, dts, tol = RC.implicit_rk_steps()
    o = make_oracle(oracle_mod, g, filtering="2/3-rule")
    shape = (3, 32, 32, 33)
    f0 = o.table(3).reshape(32, 32, 33)[:, :, 0]
    err = None
    for dt in dts:
        vdiss = np.full(32 * 32, 0.5 * dt)
        q = o.rk4_check(dt, vdiss, np.ones(shape), np.full(shape, 2.0))
        E = np.exp(-dt)
        ref = 2.0 * E + (1.0 - E) * f0[None, :, :, None]
        err = np.abs(q - ref).max()
    assert err < tol, err


def test_rk4_order_of_accuracy(oracle_mod):
    """The scheme's one-step error for constant source falls as dt^5."""
    g, dts, tol = RC.implicit_rk_steps()
    o = make_oracle(oracle_mod, dict(g, nx=8, ny=8, nz=8), filtering="2/3-rule")
    shape = (3, 8, 8, 9)
    errs = []
    for dt in [0.1, 0.05, 0.025]:
        q = o.rk4_check(dt, np.full(64, 0.5 * dt * 3.0), np.ones(shape), np.full(shape, 2.0))
        E = np.exp(-3.0 * dt)
        errs.append(np.abs(q[:, 0, 0, :] - (2.0 * E + (1 - E) / 3.0)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 4.7), rates


@pytest.mark.parametrize("n,mode", [(32, (1, 2, 2)), (48, (1, 1, 3))])
def test_pressure_and_divergence_beltrami_closed_form(oracle_mod, n, mode):
    """fields_derived.f90:67-182 on one Beltrami mode.  For u x omega = 0 the
    Euler pressure is p = -|u|^2/2 + C with dp/dz = 0 on the lids, and the
    (0,0,0) cosine coefficient is zeroed by green, so C is the trapezoid mean of
    |u|^2/2; the horizontal divergence is -dw/dz = -m cos(m z) cos(kx + ly)."""
    from ps3d_amd import ics
    k, l, m = mode
    p = oracle_mod.make_params(n, n, n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI),
                               filtering="2/3-rule", nnu=3, prediss=30.0)
    o = oracle_mod.Oracle(p)
    x, y, z = ics.grid(n, n, n, (0.0, 0.0, 0.0), (2 * PI, 2 * PI, PI))
    vel, vor = ics.beltrami_mode(x, y, z, k, l, m)
    o.set_vorticity(vor)
    o.vor2vel()
    o.derived()
    q = 0.5 * (vel ** 2).sum(0)
    wz = np.ones(n + 1)
    wz[0] = wz[-1] = 0.5
    c = (q * wz).sum() / (n * n * n)
    pres = o.get(oracle_mod.PRES)
    assert np.abs(pres + q - c).max() / np.abs(q).max() < 1e-13
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    ref = -m * np.cos(m * Z) * np.cos(k * X + l * Y)
    assert np.abs(o.get(oracle_mod.DELTA) - ref).max() / m < 1e-13
    o.close()
