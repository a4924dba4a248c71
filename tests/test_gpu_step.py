"""Whole-step parity of the HIP path with the CPU oracle: the reference's
advance (vor2vel, adapt, source, impl_rk4 / cn2 step) from the same seeded
initial vorticity, field for field and for the diagnostics series."""
import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = np.pi
# FP64 trajectory bar: relative max-norm over each field and each diagnostic
TOL = 1.0e-12


@pytest.fixture(scope="module")
def G():
    from ps3d_amd import gpu
    gpu.lib()
    return gpu


def box(n):
    return dict(nx=n, ny=n, nz=n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))


def config1(**kw):
    """The 32^3 Beltrami workload: impl-diff-rk4, 2/3-rule, nnu=3, prediss=30,
    pretype vorch, Kolmogorov, fixed dt = 0.01."""
    d = dict(filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0, pretype="vorch",
             length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.01, limit=1.0e30)
    d.update(kw)
    return d


def make_pair(oracle_mod, G, n, **kw):
    g = box(n)
    po = oracle_mod.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"], **kw)
    pg = G.make_params(g["nx"], g["ny"], g["nz"], lower=g["lower"], extent=g["extent"], **kw)
    return oracle_mod.Oracle(po), G.Ps3dGpu(pg)


def field_rel(a, b):
    return max(np.abs(a[c] - b[c]).max() / np.abs(b[c]).max() for c in range(a.shape[0]))


def diag_rel(da, db, cols):
    out = {}
    for j in cols:
        scale = max(np.abs(db[:, j]).max(), 1e-300)
        out[j] = np.abs(da[:, j] - db[:, j]).max() / scale
    return out


# diagnostics compared relatively: all but the vorticity means (O(1e-5) noise
# means whose relative error is dominated by roundoff of O(1) sums), which
# vormean_abs compares absolutely, on the scale of the rms vorticity
DIAG_COLS = [0, 1, 2, 3, 4, 5, 6, 7, 11, 12, 13, 14, 15, 16, 17, 18]


def vormean_abs(da, db):
    """Largest difference of the vorticity means (columns 8-10) over the series,
    in units of the largest rms vorticity (column 6)."""
    return np.abs(da[:, 8:11] - db[:, 8:11]).max() / np.abs(db[:, 6]).max()


def test_setup_matches_oracle(oracle_mod, G):
    o, d = make_pair(oracle_mod, G, 32, **config1())
    vor = ics.beltrami_multimode(32, 32, 32, seed=1)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    for which in (G.SVOR, G.VEL, G.VOR, G.SVEL):
        r = field_rel(d.get(which), o.get(which))
        print("setup field", which, r, flush=True)
        assert r < 1e-13, which
    d.close()


def test_api_path_three_steps(oracle_mod, G):
    """vor2vel / adapt / source / step through the C ABI, one call at a time,
    against the same calls on the oracle."""
    o, d = make_pair(oracle_mod, G, 32, **config1())
    vor = ics.beltrami_multimode(32, 32, 32, seed=1)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    t_o = t_d = 0.0
    for k in range(3):
        print("api step", k, flush=True)
        o.vor2vel()
        d.vor2vel()
        dto, rowo = o.adapt(t_o)
        dtd, rowd = d.adapt(t_d)
        r = diag_rel(rowd[None], rowo[None], DIAG_COLS)
        assert max(r.values()) < 1e-12, r
        o.source()
        d.source()
        assert field_rel(d.get(G.SVORTS), o.get(oracle_mod.SVORTS)) < 1e-12
        t_o = o.step(t_o, dto)
        t_d = d.step(t_d, dtd)
        assert abs(t_o - t_d) < 1e-15
        assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < 1e-12
    d.close()


@pytest.mark.parametrize("nsteps", [100])
def test_beltrami32_fixed_dt_trajectory(oracle_mod, G, nsteps):
    """32^3 Beltrami, fixed dt = 0.01, 100 steps: every field and the
    diagnostics series within 1e-12 relative max-norm of the oracle."""
    o, d = make_pair(oracle_mod, G, 32, **config1())
    vor = ics.beltrami_multimode(32, 32, 32, seed=1)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    print("oracle run", flush=True)
    do = o.advance_n(nsteps)
    print("gpu run", flush=True)
    dd = d.advance_n(nsteps)
    r = diag_rel(dd, do, DIAG_COLS)
    print("diag rel", r, "vormean", vormean_abs(dd, do), flush=True)
    assert max(r.values()) < TOL, r
    assert vormean_abs(dd, do) < TOL
    o.vor2vel()
    d.vor2vel()
    for which in (G.SVOR, G.VEL, G.VOR):
        fr = field_rel(d.get(which), o.get(which))
        print("field", which, fr, flush=True)
        assert fr < TOL, which
    d.close()


def test_advance_until_stops_like_the_time_loop(oracle_mod, G):
    """ps3d_gpu_advance_until runs ps3d's `do while (t < limit)` loop on the
    device: steps start only while t < min(t_stop, limit), the remaining queued
    steps do nothing, the last step is clipped to the limit; the state equals
    advance_n's bit for bit and the oracle's to the trajectory bar."""
    cfg = config1(limit=0.047)
    o, d = make_pair(oracle_mod, G, 32, **cfg)
    _, e = make_pair(oracle_mod, G, 32, **cfg)
    vor = ics.beltrami_multimode(32, 32, 32, seed=1)
    for x in (o, d, e):
        x.set_vorticity(vor)
    print("advance_until to 0.025", flush=True)
    r1 = d.advance_until(0.025, 8)          # steps at t = 0, 0.01, 0.02
    assert len(r1) == 3 and abs(d.t - 0.03) < 1e-15
    ref = e.advance_n(3)
    assert np.array_equal(r1, ref) and np.array_equal(d.get(G.SVOR), e.get(G.SVOR))
    print("advance_until to the limit", flush=True)
    r2 = d.advance_until(1.0e30, 8)         # t = 0.03, 0.04 (dt 0.007) -> 0.047
    assert len(r2) == 2 and abs(d.t - 0.047) < 1e-15, (len(r2), d.t)
    assert r2[-1, 1] == pytest.approx(0.007, abs=1e-15)
    assert len(d.advance_until(1.0e30, 4)) == 0 and d.t == pytest.approx(0.047, abs=1e-15)
    rows = np.concatenate([r1, r2])
    do = np.concatenate([o.advance_n(1) for _ in range(5)])
    assert max(diag_rel(rows, do, DIAG_COLS).values()) < TOL
    assert vormean_abs(rows, do) < TOL
    assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    for x in (o, d, e):
        x.close()


def test_closure_refuses_the_one_call_path(G):
    """With a closure attached the closure term exists only inside the fused
    step: source and step refuse (PS3D_E_UNSUPPORTED = 6), advance_n runs."""
    from ps3d_amd import closure
    p = G.make_params(16, 16, 16, **{k: v for k, v in box(16).items() if k in ("lower", "extent")}, **config1())
    d = G.Ps3dGpu(p)
    d.set_vorticity(ics.beltrami_multimode(16, 16, 16, seed=3))
    net = closure.ClosureNet(16, 16, 16)
    net.set_weights(closure.init_weights(1))
    net.attach(d)
    d.vor2vel()
    dt, _ = d.adapt(0.0)
    with pytest.raises(G.Ps3dError) as ei:
        d.source()
    assert ei.value.code == 6
    with pytest.raises(G.Ps3dError) as ei:
        d.advance_until(1.0, 2)
    assert ei.value.code == 6
    d.advance_n(2)
    closure.ClosureNet.detach(d)
    d.vor2vel()
    d.adapt(d.t)
    d.source()
    net.close()
    d.close()


def test_adaptive_hou_li_trajectory(oracle_mod, G):
    """ps3d's defaults: Hou & Li filter, adaptive dt, roll-mean-max-strain."""
    kw = dict(filtering="Hou & Li", pretype="roll-mean-max-strain", prediss=10.0, nnu=3, alpha=0.1,
              dt_fixed=0.0, limit=1.0e30)
    o, d = make_pair(oracle_mod, G, 32, **kw)
    vor = ics.beltrami_multimode(32, 32, 32, seed=2)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    do = o.advance_n(30)
    dd = d.advance_n(30)
    r = diag_rel(dd, do, DIAG_COLS)
    print("diag rel", r, flush=True)
    assert max(r.values()) < TOL, r
    assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    d.close()


def test_cn2_trajectory(oracle_mod, G):
    """The cn2 stepper of the shipped examples (examples/beltrami_32.config)."""
    kw = dict(filtering="Hou & Li", stepper="cn2", pretype="vorch", prediss=30.0, nnu=3, alpha=0.1,
              dt_fixed=0.0, limit=1.0e30)
    o, d = make_pair(oracle_mod, G, 32, **kw)
    vor = ics.beltrami_multimode(32, 32, 32, seed=3)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    do = o.advance_n(10)
    dd = d.advance_n(10)
    r = diag_rel(dd, do, DIAG_COLS)
    print("diag rel", r, flush=True)
    assert max(r.values()) < TOL, r
    assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    d.close()


def test_fused_advance_equals_api_path(G):
    """advance_n (fused, no host round trips) and the one-call-at-a-time path
    give the same state."""
    p = G.make_params(**box(32), **config1())
    a, b = G.Ps3dGpu(p), G.Ps3dGpu(p)
    vor = ics.beltrami_multimode(32, 32, 32, seed=4)
    a.set_vorticity(vor)
    b.set_vorticity(vor)
    a.advance_n(2)
    t = 0.0
    for k in range(2):
        b.vor2vel()
        dt, _ = b.adapt(t)
        b.source()
        t = b.step(t, dt)
    assert abs(a.t - t) < 1e-15
    assert field_rel(a.get(G.SVOR), b.get(G.SVOR)) < 1e-13
    a.close()
    b.close()


@pytest.mark.parametrize("n,nsteps", [(32, 0), (32, 3), (48, 2)])
def test_derived_pressure_divergence(oracle_mod, G, n, nsteps):
    """pressure + horizontal_divergence (fields_derived.f90:67-182) from the
    vor2vel state after nsteps steps, against the oracle."""
    o, d = make_pair(oracle_mod, G, n, **config1())
    vor = ics.beltrami_multimode(n, n, n, seed=2)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    if nsteps:
        o.advance_n(nsteps)
        d.advance_n(nsteps)
    print("derived", n, nsteps, flush=True)
    o.vor2vel()
    d.vor2vel()
    o.derived()
    d.derived()
    for which in (G.PRES, G.DELTA):
        a, b = d.get(which), o.get(which)
        r = np.abs(a - b).max() / np.abs(b).max()
        print("field", which, r, flush=True)
        assert r < TOL, which
    # source consumes the derivative fields: derived must refuse afterwards
    d.source()
    with pytest.raises(G.Ps3dError):
        d.derived()
    d.close()


@pytest.mark.parametrize("shape", [(24, 20, 64), (16, 12, 128), (8, 16, 256), (4, 6, 512), (4, 4, 768), (4, 4, 1024)])
@pytest.mark.parametrize("filtering", ["Hou & Li", "2/3-rule"])
def test_long_column_trajectory(oracle_mod, G, shape, filtering):
    """Adaptive steps on grids whose z columns take the wave-per-column
    kernels, against the oracle: fields and diagnostics within 1e-12."""
    nx, ny, nz = shape
    kw = dict(filtering=filtering, stepper="impl-diff-rk4", nnu=3, prediss=30.0, pretype="vorch",
              length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0, limit=1.0e30)
    g = dict(nx=nx, ny=ny, nz=nz, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
    o = oracle_mod.Oracle(oracle_mod.make_params(**g, **kw))
    d = G.Ps3dGpu(G.make_params(**g, **kw))
    vor = ics.beltrami_multimode(nx, ny, nz, seed=3)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    print("long column", shape, filtering, flush=True)
    do = o.advance_n(8)
    dd = d.advance_n(8)
    r = diag_rel(dd, do, DIAG_COLS)
    print("diag rel", max(r.values()), flush=True)
    assert max(r.values()) < TOL, r
    assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    d.close()
    o.close()


@pytest.mark.parametrize("shape", [(40, 30, 20), (50, 20, 40), (32, 80, 100), (160, 10, 50)])
def test_radix5_grid_trajectory(oracle_mod, G, shape):
    """Grids whose lengths carry factors of 5 (stafft's factorisen accepts
    2^a 3^b 5^c, stafft.f90:128-187; radix-5 passes forrdx5/revrdx5): adaptive
    steps against the oracle, fields and diagnostics within 1e-12."""
    nx, ny, nz = shape
    kw = dict(filtering="Hou & Li", stepper="impl-diff-rk4", nnu=3, prediss=30.0, pretype="vorch",
              length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0, limit=1.0e30)
    g = dict(nx=nx, ny=ny, nz=nz, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
    o = oracle_mod.Oracle(oracle_mod.make_params(**g, **kw))
    d = G.Ps3dGpu(G.make_params(**g, **kw))
    vor = ics.beltrami_multimode(nx, ny, nz, seed=4)
    o.set_vorticity(vor)
    d.set_vorticity(vor)
    print("radix 5", shape, flush=True)
    do = o.advance_n(5)
    dd = d.advance_n(5)
    r = diag_rel(dd, do, DIAG_COLS)
    assert max(r.values()) < TOL, r
    assert field_rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    d.close()
    o.close()


def test_grid_with_prime_factor_7_is_refused(G):
    with pytest.raises(G.Ps3dError, match="Factorisation"):
        G.Ps3dGpu(G.make_params(**dict(box(28), nz=16), **config1()))


@pytest.mark.parametrize("nz", [64, 128, 256])
@pytest.mark.parametrize("variant", ["1", "2", "3"])
def test_wave_and_quad_z_kernels_agree(G, monkeypatch, variant, nz):
    """The wave-per-column (both register budgets) and quad-per-thread z kernels
    give the same step."""
    p = G.make_params(**box(32), **config1())
    p.nz = nz
    vor = ics.beltrami_multimode(32, 32, nz, seed=5)
    monkeypatch.setenv("PS3D_ZWAVE", "0")
    a = G.Ps3dGpu(p)
    monkeypatch.setenv("PS3D_ZWAVE", variant)
    b = G.Ps3dGpu(p)
    a.set_vorticity(vor)
    b.set_vorticity(vor)
    a.advance_n(3)
    b.advance_n(3)
    assert field_rel(a.get(G.SVOR), b.get(G.SVOR)) < 1e-13
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 48, 32), (256, 96, 16), (512, 32, 8)])
@pytest.mark.parametrize("threads", ["512,1024,1024", "1024,512,1024", "1024,1024,512", "256,256,256"])
def test_register_line_block_sizes_agree(oracle_mod, G, monkeypatch, shape, threads):
    """Register line kernels at 512 and 256 threads per block on half- and
    quarter-depth tiles (PS3D_LINE_T = x products, adapt-stage y, x adapt pass)
    against the oracle and the 1024-thread build, with adaptive dt so the adapt
    reductions steer the run."""
    nx, ny, nz = shape
    kw = dict(nx=nx, ny=ny, nz=nz, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
    cfg = config1(dt_fixed=0.0, pretype="roll-mean-max-strain", filtering="Hou & Li")
    vor = ics.beltrami_multimode(nx, ny, nz, seed=6)
    monkeypatch.setenv("PS3D_LINE_T", "1024,1024,1024")
    a = G.Ps3dGpu(G.make_params(**kw, **cfg))
    monkeypatch.setenv("PS3D_LINE_T", threads)
    b = G.Ps3dGpu(G.make_params(**kw, **cfg))
    o = oracle_mod.Oracle(oracle_mod.make_params(**kw, **cfg))
    for m in (a, b, o):
        m.set_vorticity(vor)
    print("threads", threads, shape, flush=True)
    da, db, do = a.advance_n(3), b.advance_n(3), o.advance_n(3)
    assert field_rel(b.get(G.SVOR), a.get(G.SVOR)) < 1e-13
    assert field_rel(b.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    r = diag_rel(db, do, DIAG_COLS)
    assert max(r.values()) < TOL, r
    a.close()
    b.close()


def test_step_graph_replays_equal_eager_launches(G, monkeypatch):
    """advance_n replays one captured step; every replay, with eager work on the
    same stream between the calls, equals the kernel-by-kernel launches exactly."""
    p = G.make_params(**box(32), **config1(dt_fixed=0.0, filtering="Hou & Li",
                                          pretype="roll-mean-max-strain"))
    a, b = G.Ps3dGpu(p), G.Ps3dGpu(p)
    vor = ics.beltrami_multimode(32, 32, 32, seed=8)
    a.set_vorticity(vor)
    b.set_vorticity(vor)
    for rep in range(4):
        print("replay round", rep, flush=True)
        monkeypatch.setenv("PS3D_GRAPH", "0")
        da = a.advance_n(3)
        monkeypatch.delenv("PS3D_GRAPH")
        db = b.advance_n(3)
        assert np.array_equal(da, db)
        assert np.array_equal(a.get(G.SVOR), b.get(G.SVOR))
        for m in (a, b):      # eager work between the replays
            m.vor2vel()
            m.derived()
        assert np.array_equal(a.get(G.PRES), b.get(G.PRES))
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(32, 48, 16), (64, 96, 16), (128, 256, 8)])
@pytest.mark.parametrize("sched", ["0,0", "1,1", "2,2", "3,3", "4,5", "6,7"])
def test_line_fft_schedules_agree(oracle_mod, G, monkeypatch, shape, sched):
    """Every FFT schedule of the transform-only line passes (PS3D_LINE_SCHED = x,y:
    radix-8 or radix-16 first plan, Hermitian pair folded or in its own pass, 256 or
    512 threads per tile) gives the default schedule's step and the oracle's."""
    nx, ny, nz = shape
    kw = dict(nx=nx, ny=ny, nz=nz, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
    cfg = config1(dt_fixed=0.0)
    vor = ics.beltrami_multimode(nx, ny, nz, seed=9)
    a = G.Ps3dGpu(G.make_params(**kw, **cfg))
    monkeypatch.setenv("PS3D_LINE_SCHED", sched)
    b = G.Ps3dGpu(G.make_params(**kw, **cfg))
    o = oracle_mod.Oracle(oracle_mod.make_params(**kw, **cfg))
    for m in (a, b, o):
        m.set_vorticity(vor)
    print("sched", sched, shape, flush=True)
    da, db, do = a.advance_n(2), b.advance_n(2), o.advance_n(2)
    assert field_rel(b.get(G.SVOR), a.get(G.SVOR)) < 1e-13
    assert field_rel(b.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    r = diag_rel(db, do, DIAG_COLS)
    assert max(r.values()) < TOL, r
    a.close()
    b.close()
