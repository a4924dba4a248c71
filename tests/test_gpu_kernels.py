"""Kernel checks of the HIP path against the CPU oracle and the reference's own
unit tests (unit-tests/*.f90), through the C ABI.  Needs an MI355X."""
import numpy as np
import pytest

from tests import refcases as RC

pytestmark = pytest.mark.gpu

PI = np.pi


@pytest.fixture(scope="module")
def G():
    from ps3d_amd import gpu
    gpu.lib()
    return gpu


def rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


def pair(oracle_mod, G, g, **kw):
    po = oracle_mod.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"], **kw)
    pg = G.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"], **kw)
    return oracle_mod.Oracle(po), G.Ps3dGpu(pg)


@pytest.mark.parametrize("n", [10, 16, 20, 32, 40, 48, 50, 64, 80, 96, 100, 128, 160, 250, 256, 320, 500, 512])
@pytest.mark.parametrize("m", [1, 6, 17])
def test_fft1d_matches_oracle(oracle_mod, G, n, m):
    print("fft1d n=%d m=%d" % (n, m), flush=True)
    rng = np.random.default_rng(n * 100 + m)
    x = rng.standard_normal((n, m))
    for kind in ("forfft", "revfft"):
        assert rel(G.fft1d(kind, x), oracle_mod.fft1d(kind, x)) < 1e-14, kind
    y = rng.standard_normal((n + 1, m))
    assert rel(G.fft1d("dct", y), oracle_mod.fft1d("dct", y)) < 1e-14
    z = rng.standard_normal((n, m))
    assert rel(G.fft1d("dst", z), oracle_mod.fft1d("dst", z)) < 1e-14


GRIDS = [
    dict(nx=32, ny=32, nz=32, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI)),
    dict(nx=48, ny=16, nz=24, lower=(-1.0, 0.5, -0.3), extent=(3.0, 1.0, 1.7)),
    dict(nx=64, ny=96, nz=64, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 3 * PI, PI)),
    # long columns: the wave-per-column z kernels (nz = 128, 256, 512)
    dict(nx=16, ny=12, nz=128, lower=(0.0, 0.0, -0.5), extent=(2 * PI, 1.5, 2.0)),
    dict(nx=8, ny=16, nz=256, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI)),
    dict(nx=4, ny=6, nz=512, lower=(0.0, 0.0, 0.0), extent=(1.0, 2.0, 3.0)),
    # long columns on the quad-per-thread kernels (one transform per FFT call)
    dict(nx=4, ny=4, nz=768, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI)),
    dict(nx=4, ny=4, nz=1024, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI)),
]


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_single_operators_match_oracle(oracle_mod, G, gi):
    g = GRIDS[gi]
    o, d = pair(oracle_mod, G, g)
    rng = np.random.default_rng(gi)
    f = rng.standard_normal(d.shape)
    for op in range(11):
        print("op %d grid %d" % (op, gi), flush=True)
        a = d.op(op, f)
        b = o.op(op, f)
        assert rel(a, b) < 2e-14, (op, rel(a, b))
    d.close()


def test_diffx_diffy_unit_tests(oracle_mod, G):
    for case in (RC.diffx_case(), RC.diffy_case()):
        g, f, ref, tol = case
        print("diff case", g, flush=True)
        d = G.Ps3dGpu(G.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"]))
        fs = d.fftxyp2s(f)
        ds = d.diffx(fs) if g["nx"] == 64 else d.diffy(fs)
        fp = d.fftxys2p(ds)
        assert np.abs(fp - ref).max() < tol
        d.close()


@pytest.mark.parametrize("case", RC.diffz_cases(), ids=lambda c: c[0])
def test_central_diffz_unit_tests(G, case):
    name, g, f, ref, tol = case
    print(name, flush=True)
    d = G.Ps3dGpu(G.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"]))
    assert np.abs(d.central_diffz(f) - ref).max() < tol
    d.close()


@pytest.mark.parametrize("case", RC.vor2vel_cases(), ids=lambda c: c[0])
def test_vor2vel_unit_tests(oracle_mod, G, case):
    name, g, vor, vel_ref, tol = case
    print(name, flush=True)
    o, d = pair(oracle_mod, G, g)
    svor = np.stack([d.field_decompose_physical(vor[c]) for c in range(3)])
    d.set(G.SVOR, svor)
    d.vor2vel()
    vel = d.get(G.VEL)
    assert np.abs(vel - vel_ref).max() < tol, np.abs(vel - vel_ref).max()
    # and field for field against the oracle
    o.set(oracle_mod.SVOR, svor)
    o.vor2vel()
    for which in (G.VEL, G.VOR, G.SVEL, G.SVOR):
        a, b = d.get(which), o.get(which)
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max()), which
    d.close()


def test_implicit_rk_unit_test(oracle_mod, G):
    """unit-tests/test_implicit_rk.f90 as a kernel check of the fused RK stage
    update: source held at 1, svor = 2, dissipation rate 1, with the 2/3 filter."""
    g, dts, tol = RC.implicit_rk_steps()
    o, d = pair(oracle_mod, G, g, filtering="2/3-rule")
    shape = (3,) + d.shape
    f0 = o.table(3).reshape(d.shape)[:, :, 0]
    err = None
    for dt in dts:
        print("implicit_rk dt=%g" % dt, flush=True)
        vdiss = np.full(d.nx * d.ny, 0.5 * dt)
        d.set(G.SVOR, np.full(shape, 2.0))
        d.set(G.SVORTS, np.ones(shape))
        for stage in (1, 2, 3, 4):
            d.rk_substep(stage, dt, vdiss)
        q = d.get(G.SVOR)
        E = np.exp(-dt)
        ref = 2.0 * E + (1.0 - E) * f0[None, :, :, None]
        err = np.abs(q - ref).max()
        qo = o.rk4_check(dt, vdiss, np.ones(shape), np.full(shape, 2.0))
        assert np.abs(q - qo).max() < 1e-13
    assert err < tol, err
    d.close()


def test_resample_band_limited(G):
    """Zero padding 32^3 -> 64^3 keeps a band-limited field: its physical values
    on the fine grid at the coarse points equal the coarse ones; truncating back
    returns the original coefficients."""
    box = dict(lower=(0.0, 0.0, -0.5), extent=(2 * PI, 3.0, 1.7))
    pa = G.make_params(32, 32, 32, **box)
    pb = G.make_params(64, 64, 64, **box)
    a, b, c = G.Ps3dGpu(pa), G.Ps3dGpu(pb), G.Ps3dGpu(pa)
    rng = np.random.default_rng(7)
    f = rng.standard_normal((3, 32, 32, 33))
    f[:, 16, :, :] = 0.0      # Nyquist slots carry no resampled content
    f[:, :, 16, :] = 0.0
    a.set(G.SVORTS, f)
    G.resample(b, a, G.SVORTS)
    fb = b.get(G.SVORTS)
    for k in range(3):
        pa_ = a.op(G.OP_COMBINE_PHYS, f[k])
        pb_ = b.op(G.OP_COMBINE_PHYS, fb[k])
        assert rel(pb_[::2, ::2, ::2], pa_) < 1e-13, k
    G.resample(c, b, G.SVORTS)
    assert rel(c.get(G.SVORTS), f) < 1e-14
    for d in (a, b, c):
        d.close()
