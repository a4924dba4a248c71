"""Every workload configuration at its own size on the GPU.

* 128^3 Beltrami, adaptive dt: 20 steps against the oracle on the host cores
  (diagnostics series, vorticity means, final fields);
* register line kernels at the line lengths of the 512^3 and 1024^3 grids, on
  thin grids that the oracle steps in seconds (x and y lines of 512 and 1024,
  with adaptive dt so the adapt reductions steer the run);
* 512^3 and 1024^3: the whole grid in HBM, set up from the closed-form Beltrami
  vorticity of two modes sharing alpha and k^2 + l^2 (u x omega = 0): vor2vel
  against the closed-form velocity (the reference's test_vor2vel_1.f90 check),
  the kinetic energy against its closed form, and one step's decay factor
  against exp(-pref dt visc (k^2 + l^2)^nnu) (impl_rk4.f90:37-52);
* the 96-instance ensemble of 32^3, 48^3, 64^3 and 96^3 grids (38/29/19/10),
  every group's instances against single handles and one instance per size
  against the oracle;
* the closure net at 64^3 against its CPU restatement, attached to a 64^3 solver.

Host arrays at 1024^3 are 8.6 GB per volume: the closed forms are built as outer
products of 2-D (x, y) and 1-D (z) factors, one volume at a time.
"""
import math
import os

import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = math.pi
TOL = 1.0e-12
BOX = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))


@pytest.fixture(scope="module")
def G():
    from ps3d_amd import gpu
    gpu.lib()
    return gpu


def workload(**extra):
    """The plain workload's physics: 2/3-rule, impl-diff-rk4, nnu 3, prediss 30,
    vorch prefactor, Kolmogorov, adaptive dt with alpha 0.1."""
    d = dict(filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0, pretype="vorch",
             length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0, limit=1.0e30)
    d.update(extra)
    return d


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def diag_check(dd, do):
    cols = [j for j in range(19) if j not in (8, 9, 10)]
    for j in cols:
        scale = max(np.abs(do[:, j]).max(), 1e-300)
        assert np.abs(dd[:, j] - do[:, j]).max() <= TOL * scale, (j, dd[:, j], do[:, j])
    # vorticity means: absolute, on the scale of the rms vorticity
    assert np.abs(dd[:, 8:11] - do[:, 8:11]).max() <= TOL * np.abs(do[:, 6]).max()


def threads():
    return min(16, os.cpu_count() or 1)


def test_128_adaptive_trajectory(oracle_mod, G):
    n, nsteps = 128, 20
    g = dict(nx=n, ny=n, nz=n, **BOX)
    vor = ics.beltrami_multimode(n, n, n, seed=1)
    d = G.Ps3dGpu(G.make_params(**g, **workload()))
    d.set_vorticity(vor)
    print("gpu 128^3, %d steps" % nsteps, flush=True)
    dd = d.advance_n(nsteps)
    oracle_mod.set_threads(threads())
    o = oracle_mod.Oracle(oracle_mod.make_params(**g, **workload()))
    o.set_vorticity(vor)
    do = np.concatenate([o.advance_n(5) for _ in range(nsteps // 5)])
    print("oracle done, t = %.6f" % o.t, flush=True)
    diag_check(dd, do)
    o.vor2vel()
    d.vor2vel()
    for which in (G.SVOR, G.VEL, G.VOR):
        assert rel(d.get(which), o.get(which)) < TOL, which
    d.close()
    o.close()


@pytest.mark.parametrize("shape", [(512, 16, 8), (16, 512, 8), (1024, 8, 8), (8, 1024, 8), (128, 128, 8)])
def test_long_line_trajectories(oracle_mod, G, shape):
    """The register line passes (x products, adapt-stage y, x adapt) at the line
    lengths of the 128^3, 512^3 and 1024^3 grids, 6 adaptive steps."""
    nx, ny, nz = shape
    g = dict(nx=nx, ny=ny, nz=nz, **BOX)
    cfg = workload(pretype="roll-mean-max-strain", filtering="Hou & Li")
    vor = ics.beltrami_multimode(nx, ny, nz, seed=8)
    d = G.Ps3dGpu(G.make_params(**g, **cfg))
    o = oracle_mod.Oracle(oracle_mod.make_params(**g, **cfg))
    d.set_vorticity(vor)
    o.set_vorticity(vor)
    print("long lines", shape, flush=True)
    dd = d.advance_n(6)
    do = o.advance_n(6)
    diag_check(dd, do)
    assert rel(d.get(G.SVOR), o.get(oracle_mod.SVOR)) < TOL
    d.close()
    o.close()


def beltrami_factors(n, modes, phases):
    """Per mode: the 2-D factors sin(theta), cos(theta) on the (x, y) grid and the
    1-D z factors of u, v, w of ics.beltrami_mode, so that
    u = Au(z) sin(theta), v = Av(z) sin(theta), w = Aw(z) cos(theta)."""
    x, y, z = ics.grid(n, n, n, BOX["lower"], BOX["extent"])
    out = []
    for (k, l, m), ph in zip(modes, phases):
        a = math.sqrt(k * k + l * l + m * m)
        f = 1.0 / (k * k + l * l)
        th = k * x[:, None] + l * y[None, :] + ph
        sm, cm = np.sin(m * z), np.cos(m * z)
        out.append((np.sin(th), np.cos(th), f * (-k * m * cm - l * a * sm), f * (-l * m * cm + k * a * sm), sm, a))
    return out


def beltrami_component(fac, c, scale_alpha=False):
    """Component c of the summed velocity (vorticity with scale_alpha) as one
    (n, n, n+1) array."""
    out = None
    for s2, c2, au, av, aw, a in fac:
        two, one = (s2, au) if c == 0 else (s2, av) if c == 1 else (c2, aw)
        term = np.multiply.outer(two * (a if scale_alpha else 1.0), one)
        if out is None:
            out = term
        else:
            out += term
    return out


@pytest.mark.parametrize("n", [512, pytest.param(1024, marks=pytest.mark.timeout(900))])
def test_large_grid_closed_forms(G, n):
    """Two Beltrami modes with alpha = 9 and k^2 + l^2 = 80 (u x omega = 0) on the
    whole n^3 grid in HBM: vor2vel gives the closed-form velocity (<= 1e-12), the
    first record's kinetic energy is 2 * alpha^2 / (4 (k^2 + l^2)) per unit
    amplitude, and one step multiplies the state by exp(-pref dt visc (k^2+l^2)^3).
    That decay is 5e-8 at 512^3 and 6e-10 at 1024^3 with the workload's
    prediss; the projection's O(dz^2) leaves a nonlinear residual far below it."""
    modes = ((8, 4, 1), (4, 8, 1))
    fac = beltrami_factors(n, modes, (0.3, 1.7))
    p = G.make_params(n, n, n, **BOX, **workload())
    d = G.Ps3dGpu(p)
    print("set-up at %d^3" % n, flush=True)
    vor = np.empty((3, n, n, n + 1))
    for c in range(3):
        vor[c] = beltrami_component(fac, c, scale_alpha=True)
    d.set_vorticity(vor)
    del vor
    d.vor2vel()
    print("vor2vel at %d^3" % n, flush=True)
    vel = d.get(G.VEL)
    for c in range(3):
        vel[c] -= beltrami_component(fac, c)
    err = np.abs(vel).max()
    del vel
    print("velocity max error", err, flush=True)
    assert err < 1e-12
    s0 = d.get(G.SVOR)
    visc, _ = d.viscosity()
    print("one step at %d^3" % n, flush=True)
    row = d.advance_n(1)[0]
    ke_ref = 2 * 81.0 / (4 * 80.0)
    assert abs(row[2] - ke_ref) < 1e-12 * ke_ref, (row[2], ke_ref)
    s1 = d.get(G.SVOR)
    fac_ref = math.exp(-row[18] * row[1] * visc * 80.0 ** 3)
    num = sum(float((s0[c] * s1[c]).sum()) for c in range(3))
    den = sum(float((s0[c] * s0[c]).sum()) for c in range(3))
    ratio = num / den
    print("decay", 1.0 - ratio, "closed form", 1.0 - fac_ref, flush=True)
    assert 1.0 - fac_ref > 1e-10
    assert abs(ratio - fac_ref) < max(1e-11, 1e-2 * (1.0 - fac_ref))
    d.close()


ENS_SIZES = [32, 48, 64, 96]
ENS_COUNTS = [38, 29, 19, 10]


def test_ensemble_96_instances(oracle_mod, G):
    """The ensemble workload: per group, two instances against single handles
    (1e-14) and the first against the oracle (1e-12) after 3 steps; every other
    instance's state stays finite."""
    rng = np.random.default_rng(2024)
    k0s = rng.uniform(2.0, 6.0, size=sum(ENS_COUNTS))
    params = [G.make_params(n, n, n, **BOX, **workload()) for n in ENS_SIZES]
    ens = G.Ps3dEnsemble(params, ENS_COUNTS)
    ics_of = {}
    i = 0
    for gi, n in enumerate(ENS_SIZES):
        for b in range(ENS_COUNTS[gi]):
            v = ics.random_phase_isotropic(n, n, n, k0=float(k0s[i]), seed=1000 + i)
            ens.groups[gi].set_vorticity(b, v)
            if b in (0, ENS_COUNTS[gi] - 1):
                ics_of[(gi, b)] = v
            i += 1
    print("ensemble, 96 instances, 3 steps", flush=True)
    ens.advance(3, graph=True)
    oracle_mod.set_threads(threads())
    for gi, n in enumerate(ENS_SIZES):
        grp = ens.groups[gi]
        for b in range(ENS_COUNTS[gi]):
            assert np.isfinite(grp.get_svor(b)).all(), (n, b)
        for b in (0, ENS_COUNTS[gi] - 1):
            s = G.Ps3dGpu(params[gi])
            s.set_vorticity(ics_of[(gi, b)])
            s.advance_n(3)
            assert rel(grp.get_svor(b), s.get(G.SVOR)) < 1e-14, (n, b)
            s.close()
        o = oracle_mod.Oracle(oracle_mod.make_params(n, n, n, **BOX, **workload()))
        o.set_vorticity(ics_of[(gi, 0)])
        o.advance_n(3)
        err = rel(grp.get_svor(0), o.get(oracle_mod.SVOR))
        print("group", n, "vs oracle", err, flush=True)
        assert err < TOL
        o.close()
    ens.close()


def test_closure_at_64(G):
    """The closure net at the coarse grid's size: forward against the CPU
    restatement at bf16 storage precision, then attached to a 64^3 solver: the
    step with the net differs from the step without it, stays finite, and a
    zero net gives the plain step bit for bit."""
    from oracle import closure_oracle as N
    from ps3d_amd import closure
    closure.lib()
    n = 64
    net = closure.ClosureNet(n, n, n)
    rng = np.random.default_rng(3)
    params = N.init_params(3)
    params = [(W, rng.standard_normal(b.shape) * 0.1) for W, b in params]
    w = N.pack(params).astype(np.float32)
    net.set_weights(w)
    params = N.unpack(w.astype(np.float64))
    x = rng.standard_normal((n, n, n + 1, 6))
    print("closure forward at 64^3", flush=True)
    y = net.forward(x.reshape(-1, 6)).reshape(x.shape[:3] + (3,))
    ref = N.forward_bf16(params, x)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("forward rel err", err, flush=True)
    assert err < 2e-3
    p = G.make_params(n, n, n, **BOX, **workload())
    vor = ics.random_phase_isotropic(n, n, n, k0=4.0, seed=9)
    a, b, c = G.Ps3dGpu(p), G.Ps3dGpu(p), G.Ps3dGpu(p)
    for h in (a, b, c):
        h.set_vorticity(vor)
    net.in_scale[:] = 1.0
    net.out_scale[:] = 1e-3
    net.attach(a)
    zero = closure.ClosureNet(n, n, n)
    zero.set_weights(np.zeros_like(w))
    zero.attach(c)
    a.advance_n(3)
    b.advance_n(3)
    c.advance_n(3)
    sa, sb = a.get(G.SVOR), b.get(G.SVOR)
    assert np.isfinite(sa).all()
    assert rel(sa, sb) > 1e-8
    assert np.array_equal(c.get(G.SVOR), sb)
    for h in (a, b, c):
        h.close()
    net.close()
    zero.close()
