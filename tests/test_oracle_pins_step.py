"""Pins of the oracle's time-step functions (oracle/ps3d_oracle.c) that the
transform and unit-test pins of test_oracle_pins.py leave open.

Every check compares the oracle with something computed here independently:
numpy's FFT for horizontal derivatives, numpy's gradient for the central
difference in z, numpy.linalg.eigvalsh for the strain eigenvalues, closed-form
Beltrami fields and their energies, brute-force sums over the grid, and exact
solutions of the schemes for a field whose nonlinear term vanishes.

  vorticity_tendency  inversion.f90:298-371        check_tendency
  init_diffusion      inversion_utils.f90:157-184  check_viscosity, check_decay
  adapt (strain, CFL, vorch, maxima, means)
                      advance.f90:160-377          check_adapt
  rolling mean, get_diffusion_pre_factor
                      rolling_mean.f90:36-69, advance.f90:381-410
                                                   check_rolling_mean, check_prefactor
  helicity            plotting/nc_reader.py:94-101 check_adapt
  cn2                 cn2.f90:39-80, 92-181        check_decay (cn2), check_cn2_held_source
  impl_rk4 (nonlinear) impl_rk4.f90:76-364         check_time_order
  adjust_vorticity_mean
                      field_diagnostics.f90:584-619 check_mean_restore

The check_* functions raise AssertionError on a mismatch; tests/test_oracle_mutants.py
runs them against copies of the oracle with one planted mistake each and
requires every mistake to be caught.
"""
import numpy as np
import pytest

from ps3d_amd import ics

PI = np.pi
BOX = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))


# ---------------------------------------------------------------------------
# independent numerics
# ---------------------------------------------------------------------------
def dspec(f, axis, length):
    """Exact derivative of a band-limited periodic field along axis 0 (x) or 1
    (y) by numpy's FFT; the Nyquist mode is dropped, as ps3d's diffx does
    (sta3dfft.f90:321-335)."""
    n = f.shape[axis]
    k = 2 * PI * np.fft.fftfreq(n, d=1.0 / n) / length
    if n % 2 == 0:
        k[n // 2] = 0.0
    shape = [1] * f.ndim
    shape[axis] = n
    F = np.fft.fft(f, axis=axis) * (1j * k.reshape(shape))
    return np.real(np.fft.ifft(F, axis=axis))


def dz_central(f, dz):
    """Second-order central difference with first-order one-sided lids."""
    return np.gradient(f, dz, axis=-1, edge_order=1)


def trap_mean(f):
    """Volume mean with trapezoid weights in z (1/2 on the lids) over nx*ny*nz cells."""
    nz = f.shape[-1] - 1
    w = np.ones(nz + 1)
    w[0] = w[-1] = 0.5
    return float((f * w).sum() / (f.shape[0] * f.shape[1] * nz))


def beltrami_sum(n, modes, nzz=None):
    """Velocity and vorticity of a sum of Beltrami modes (k, l, m, phase) on the
    2 pi x 2 pi x pi box, each with w = 0 on both lids (ps3d_amd.ics)."""
    nz = n if nzz is None else nzz
    x, y, z = ics.grid(n, n, nz, BOX["lower"], BOX["extent"])
    vel = np.zeros((3, n, n, nz + 1))
    vor = np.zeros_like(vel)
    for k, l, m, ph in modes:
        v, w = ics.beltrami_mode(x, y, z, k, l, m, ph)
        vel += v
        vor += w
    return vel, vor


def mk(O, n, nz=None, **kw):
    p = O.make_params(n, n, n if nz is None else nz, lower=BOX["lower"], extent=BOX["extent"], **kw)
    return O.Oracle(p)


TWO_MODES = [(1, 2, 1, 0.4), (2, 1, 2, 1.3)]   # alpha = sqrt 6 and 3: u x omega != 0


# ---------------------------------------------------------------------------
# vorticity_tendency (inversion.f90:298-371)
# ---------------------------------------------------------------------------
def check_tendency(O, f_cor=(0.0, 0.0, 0.0), n=32):
    """svorts, combined back to physical space, against the flux form built
    here: r = u eta - v xi, q = w xi - u zeta, p = v zeta - w eta (with the
    planetary vorticity added), d/dx, d/dy by numpy FFT, d/dz by numpy's central
    difference; dxi/dt = r_y - q_z, deta/dt = p_z - r_x, dzeta/dt = q_x - p_y."""
    o = mk(O, n, filtering="2/3-rule", nnu=3, prediss=30.0, f_cor=f_cor)
    _, vor = beltrami_sum(n, TWO_MODES)
    o.set_vorticity(vor)
    u, v, w = o.get(O.VEL)
    xi, eta, zeta = o.get(O.VOR) + np.asarray(f_cor)[:, None, None, None]
    o.source()
    st = o.get(O.SVORTS)
    got = np.stack([o.op(O.OP_COMBINE_PHYS, st[c]) for c in range(3)])
    o.close()
    Lx, Ly, Lz = BOX["extent"]
    dz = Lz / n
    r = u * eta - v * xi
    q = w * xi - u * zeta
    p = v * zeta - w * eta
    ref = np.stack([dspec(r, 1, Ly) - dz_central(q, dz),
                    dz_central(p, dz) - dspec(r, 0, Lx),
                    dspec(q, 0, Lx) - dspec(p, 1, Ly)])
    scale = np.abs(ref).max()
    assert scale > 0.5, scale          # the nonlinear term is far from zero here
    err = np.abs(got - ref).max() / scale
    assert err < 1e-12, err
    return err


# ---------------------------------------------------------------------------
# init_diffusion / get_viscosity (inversion_utils.f90:157-184)
# ---------------------------------------------------------------------------
def shear_mode(n, k, l, phase=0.0):
    """A z-independent horizontal flow with stream function sin(kx + ly + phase):
    u = -l cos, v = k cos, w = 0, vorticity (0, 0, -K^2 sin), K^2 = k^2 + l^2.
    It is an exact steady solution of the discrete Euler equations here (its
    nonlinear term vanishes pointwise and the solenoidal projection leaves it
    unchanged), with KE = K^2 / 4 and enstrophy = K^4 / 4 in ps3d's trapezoid
    means (field_diagnostics.f90:85-206)."""
    x, y, z = ics.grid(n, n, n, BOX["lower"], BOX["extent"])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    th = k * X + l * Y + phase
    vel = np.stack([-l * np.cos(th), k * np.cos(th), 0 * th])
    vor = np.stack([0 * th, 0 * th, -(k * k + l * l) * np.sin(th)])
    return vel, vor


def check_viscosity(O, length_scale, nnu, prediss=30.0, n=32, mode=(1, 2)):
    """vvisc from the kinetic energy and enstrophy of a shear mode, known in
    closed form (te / en = 1 / K^2): Kolmogorov prediss (K2max te/en)^(1/3)
    K2max^-nnu, geophysical prediss K2max^-nnu, with K2max = (2 pi (n/2) / L)^2."""
    k, l = mode
    o = mk(O, n, filtering="2/3-rule", nnu=nnu, prediss=prediss, length_scale=length_scale)
    vel, vor = shear_mode(n, k, l, 0.3)
    o.set_vorticity(vor)
    assert np.abs(o.get(O.VEL) - vel).max() < 1e-13
    visc = o.viscosity()
    o.close()
    K2 = float(k * k + l * l)
    te, en = K2 / 4.0, K2 * K2 / 4.0
    assert abs(0.5 * trap_mean((vel ** 2).sum(0)) - te) < 1e-13
    assert abs(0.5 * trap_mean((vor ** 2).sum(0)) - en) < 1e-12
    k2max = (2 * PI * (n // 2) / BOX["extent"][0]) ** 2
    if length_scale == "Kolmogorov":
        ref = prediss * (k2max * te / en) ** (1.0 / 3.0) * k2max ** (-nnu)
    else:
        ref = prediss * k2max ** (-nnu)
    err = abs(visc - ref) / ref
    assert err < 1e-12, (visc, ref)
    return err


def amplitude_ratio(a, b):
    """Least-squares factor c with b = c a, and the residual of that fit."""
    c = float((a * b).sum() / (a * a).sum())
    return c, float(np.abs(b - c * a).max() / np.abs(b).max())


def filter_table(n, nz, filtering):
    """ps3d's dealiasing filter on the mixed representation (numpy shape
    (nx, ny, nz+1)), from its definition (inversion_utils.f90:376-455): per-axis
    factors of rkx = 2 pi min(k, n-k) / Lx and rkz = pi kz / Lz, 2/3 rule
    (1 where rk <= 2/3 max rk) or Hou & Li (exp(-36 (rk / max rk)^36)); lid rows
    take the horizontal product, interior rows also the kz factor; the (0, 0)
    column is 1 throughout (inversion_utils.f90:274-277)."""
    Lx, Ly, Lz = BOX["extent"]
    k = np.arange(n)
    rkx = 2 * PI * np.minimum(k, n - k) / Lx
    rky = 2 * PI * np.minimum(k, n - k) / Ly
    rkz = PI * np.arange(nz + 1) / Lz
    if filtering == "2/3-rule":
        f1 = [(r <= (2.0 / 3.0) * r.max()).astype(float) for r in (rkx, rky, rkz)]
    else:
        f1 = [np.exp(-36.0 * (r / r.max()) ** 36) for r in (rkx, rky, rkz)]
    F = np.empty((n, n, nz + 1))
    F[:] = (f1[0][:, None] * f1[1][None, :])[:, :, None]
    F[:, :, 1:nz] *= f1[2][None, None, 1:nz]
    F[0, 0, :] = 1.0
    return F


def check_decay(O, stepper, pretype, nnu=3, prediss=2.0e3, n=32, mode=(3, 4), dt=0.01, nsteps=3):
    """A shear mode has no nonlinear tendency, so an impl_rk4 step only applies
    the hyperdiffusion: column (k, l) decays by exp(-pref dt nu K^(2 nnu)) (the
    integrating factor, impl_rk4.f90:37-53, 87-89; its filter is the horizontal
    one, 1 here).  cn2 multiplies by the full 3-D filter and by
    1 / (1 + pref dt nu K^(2 nnu)) (cn2.f90:39-80, 92-181): the filter reshapes
    the mixed representation of the z-independent mode, so the filtered field has
    a small tendency (~1e-9 of the step's change with Hou & Li), and the check
    is to 1e-7 over the first step (later steps start from the filtered field,
    whose own tendency enters at O(1e-4)).  pref is get_diffusion_pre_factor's value for pretype
    (advance.f90:381-410), read from the diagnostics record; nu is pinned by
    check_viscosity."""
    k, l = mode
    filtering = "2/3-rule" if stepper == "impl-diff-rk4" else "Hou & Li"
    o = mk(O, n, filtering=filtering, nnu=nnu, prediss=prediss, stepper=stepper, pretype=pretype,
           dt_fixed=dt)
    _, vor = shear_mode(n, k, l, 0.7)
    o.set_vorticity(vor)
    nu = o.viscosity()
    K2 = float(k * k + l * l)
    F = filter_table(n, n, filtering)
    prev = o.get(O.SVOR)
    for s in range(nsteps if stepper == "impl-diff-rk4" else 1):
        d = o.advance_n(1)[0]
        cur = o.get(O.SVOR)
        pref = {"constant": 1.0, "max-strain": d[11], "vorch": d[7]}[pretype]
        rate = pref * d[1] * nu * K2 ** nnu
        assert rate > 1e-3, rate       # the decay is far above roundoff
        if stepper == "impl-diff-rk4":
            c, res = amplitude_ratio(prev, cur)
            assert res < 1e-12, res
            assert abs(c - np.exp(-rate)) < 1e-12, (s, c, np.exp(-rate))
        else:
            ref = F[None] * prev / (1.0 + rate)
            err = np.abs(cur - ref).max() / np.abs(cur).max()
            assert err < 1e-7, (s, err)
        prev = cur
    o.close()


# ---------------------------------------------------------------------------
# adapt (advance.f90:160-377): strain, CFL, maxima, vorticity statistics
# ---------------------------------------------------------------------------
def analytic_gradient(fun, X, Y, Z, h=1e-3):
    """d/dx_j of each component of fun(X, Y, Z) -> (3, 3, ...) by a 6th-order
    central difference of the analytic function (error ~1e-13 for these modes)."""
    c = [(1, 45.0 / 60), (2, -9.0 / 60), (3, 1.0 / 60)]
    out = []
    for d in range(3):
        acc = 0.0
        for s, wgt in c:
            e = [0.0, 0.0, 0.0]
            e[d] = s * h
            fp = fun(X + e[0], Y + e[1], Z + e[2])
            e[d] = -s * h
            fm = fun(X + e[0], Y + e[1], Z + e[2])
            acc = acc + wgt * (fp - fm)
        out.append(acc / h)
    return np.stack(out, axis=1)   # (component i, derivative j, ...)


def beltrami_ke(k, l, m):
    """Kinetic energy (ps3d's trapezoid mean of |u|^2/2) of one Beltrami mode of
    ps3d_amd.ics on the 2 pi x 2 pi x pi box: alpha^2 / (4 (k^2 + l^2))."""
    return (k * k + l * l + m * m) / (4.0 * (k * k + l * l))


def check_adapt(O, alpha, mode=(2, 1, 1), phase=0.9, n=24, nz=16, t=0.0, limit=1.0e30):
    """One Beltrami mode: every adapt output against numpy.  The velocity is the
    closed form (vor2vel reproduces it to roundoff); the vorticity is the state
    adapt reads, vor2vel's solenoidal projection of the closed form, which moves
    xi and eta by O(dz^2) (inversion.f90:35-76).
    * ggmax/usggmax/lsggmax = max |eigvalsh(S)|, S = (J + J^T)/2 with J the
      velocity gradient: horizontal derivatives of the closed form by a 6th-order
      difference, vertical ones from the vorticity, u_z = eta + w_x,
      v_z = w_y - xi, w_z = -(u_x + v_y) (advance.f90:233-257);
    * signed maxima of u, v, w; vortmax, vortrms, vorch from the z-midpoint
      averages (field_diagnostics.f90:501-545); vorticity means; KE, enstrophy,
      helicity = trapezoid means of |u|^2/2, |omega|^2/2, u.omega;
    * dt = min(alpha / ggmax, 0.8 min(dx_i / umax_i), limit - t) (advance.f90:325-333).
    Closed-form cross-checks: KE = alpha^2 / (4 (k^2 + l^2)), helicity and
    enstrophy within O(dz^2) of 2 alpha KE and alpha^2 KE."""
    k, l, m = mode
    a_b = np.sqrt(k * k + l * l + m * m)
    o = mk(O, n, nz=nz, filtering="2/3-rule", nnu=3, prediss=30.0, alpha=alpha, limit=limit,
           pretype="max-strain")
    x, y, z = ics.grid(n, n, nz, BOX["lower"], BOX["extent"])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")

    def velf(XX, YY, ZZ):
        # the mode of ics.beltrami_mode, evaluated pointwise
        f = 1.0 / (k * k + l * l)
        s, c = np.sin(m * ZZ), np.cos(m * ZZ)
        th = k * XX + l * YY + phase
        return np.stack([f * (-k * m * c - l * a_b * s) * np.sin(th),
                         f * (-l * m * c + k * a_b * s) * np.sin(th),
                         s * np.cos(th)])

    vel = velf(X, Y, Z)
    o.set_vorticity(a_b * vel)
    assert np.abs(o.get(O.VEL) - vel).max() < 1e-13
    vor = o.get(O.VOR)
    dt, d = o.adapt(t)
    o.close()
    J = analytic_gradient(velf, X, Y, Z)
    Jz = np.stack([vor[1] + J[2, 0], J[2, 1] - vor[0], -(J[0, 0] + J[1, 1])])
    assert np.abs(Jz - J[:, 2]).max() < 0.05 * np.abs(J).max()    # O(dz^2) from the projection
    J[:, 2] = Jz
    S = 0.5 * (J + J.transpose(1, 0, 2, 3, 4))
    ev = np.linalg.eigvalsh(np.moveaxis(S, (0, 1), (-2, -1)))
    lam = np.abs(ev).max(-1)
    gg, us, ls = lam.max(), lam[..., -1].max(), lam[..., 0].max()
    w2 = (vor ** 2).sum(0)
    vortrms = np.sqrt(trap_mean(w2))
    mid = 0.5 * np.abs(vor[..., 1:] + vor[..., :-1])
    tsum = mid.sum(0)
    sel = tsum > vortrms
    vorch = (mid ** 2).sum(0)[sel].sum() / (1e-12 + tsum[sel].sum())
    ke = 0.5 * trap_mean((vel ** 2).sum(0))
    en = 0.5 * trap_mean(w2)
    hel = trap_mean((vel * vor).sum(0))
    assert abs(ke - beltrami_ke(k, l, m)) < 1e-13
    assert abs(hel - 2 * a_b * ke) < 0.02 * hel and abs(en - a_b ** 2 * ke) < 0.02 * en
    dx = np.array(BOX["extent"]) / np.array([n, n, nz])
    umax = vel.reshape(3, -1).max(1)
    dtcfl = 0.8 * np.min(dx / (umax + 1e-12))
    dt_ref = min(alpha / (gg + 1e-12), alpha / 1e-12, dtcfl, limit - t)
    ref = {2: ke, 3: en, 4: hel, 5: np.sqrt(w2.max()), 6: vortrms, 7: vorch,
           8: trap_mean(vor[0]), 9: trap_mean(vor[1]), 10: trap_mean(vor[2]),
           11: gg, 12: umax[0], 13: umax[1], 14: umax[2], 15: us, 16: ls, 18: gg}
    for i, r in ref.items():
        tol = 1e-10 if i in (11, 15, 16, 18) else 1e-12
        assert abs(d[i] - r) <= tol * max(1.0, abs(r)), (O.DIAG_NAMES[i], d[i], r)
    assert abs(dt - dt_ref) <= 1e-10 * dt_ref, (dt, dt_ref)
    assert d[1] == dt and d[0] == t
    return dt, dtcfl, alpha / gg


# ---------------------------------------------------------------------------
# rolling mean (rolling_mean.f90:36-69) and prefactors (advance.f90:381-410)
# ---------------------------------------------------------------------------
def check_rolling_mean(O, win, nsteps=7):
    """rgmax after each step is the arithmetic mean of the last min(step, win)
    ggmax values of the same record series, and it is the prefactor of
    roll-mean-max-strain."""
    o = mk(O, 32, filtering="2/3-rule", nnu=3, prediss=30.0, pretype="roll-mean-max-strain",
           roll_mean_win_size=win, alpha=0.2)
    _, vor = beltrami_sum(32, TWO_MODES)
    o.set_vorticity(vor)
    d = o.advance_n(nsteps)
    o.close()
    gg = d[:, 11]
    assert np.ptp(gg) > 1e-3 * gg.max()       # the strain changes between steps
    for s in range(nsteps):
        ref = gg[max(0, s - win + 1): s + 1].mean()
        assert abs(d[s, 17] - ref) <= 1e-13 * ref, (s, d[s, 17], ref)
        assert d[s, 18] == d[s, 17]


def check_prefactor(O, pretype):
    o = mk(O, 32, filtering="2/3-rule", nnu=3, prediss=30.0, pretype=pretype, dt_fixed=0.01)
    _, vor = beltrami_sum(32, TWO_MODES)
    o.set_vorticity(vor)
    d = o.advance_n(2)
    o.close()
    col = {"constant": None, "vorch": 7, "bfmax": None, "max-strain": 11, "us-max-strain": 15,
           "roll-mean-max-strain": 17}[pretype]
    for s in range(2):
        ref = {"constant": 1.0, "bfmax": 0.0}.get(pretype, d[s, col] if col is not None else None)
        assert d[s, 18] == ref, (pretype, d[s, 18], ref)


# ---------------------------------------------------------------------------
# time order of the whole nonlinear step (impl_rk4, cn2)
# ---------------------------------------------------------------------------
def check_time_order(O, stepper, expect, T=0.16, dts=(0.04, 0.02, 0.01), n=16):
    """Inviscid (prediss = 0) two-mode flow with u x omega != 0 stepped to time T
    at fixed dt: the error against a run at dt/8 of the smallest dt falls as
    dt^expect (4 for the integrating-factor RK4).  A dropped or misplaced stage
    changes the order.  (ps3d's cn2 is not held to an order: it multiplies the
    state by the 3-D filter every step and vor2vel re-projects it, two
    operations that do not commute, so its runs at different dt converge to
    each other only at first order; check_cn2_held_source and check_decay pin
    its algebra instead.)"""
    def run(dt):
        o = mk(O, n, filtering="2/3-rule", nnu=3, prediss=0.0, stepper=stepper, pretype="constant",
               dt_fixed=dt)
        _, vor = beltrami_sum(n, TWO_MODES)
        o.set_vorticity(vor)
        o.advance_n(int(round(T / dt)))
        s = o.get(O.SVOR)
        o.close()
        return s
    fine = run(dts[-1] / 8)
    errs = np.array([np.abs(run(dt) - fine).max() for dt in dts])
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(np.abs(rates - expect) < 0.35), (stepper, errs, rates)
    return rates


def check_cn2_held_source(O, filtering, n=16, dt=0.05):
    """cn2 with the tendency held fixed (cn2.f90:92-181, as test_implicit_rk
    drives impl_rk4): every iteration then gives the same value, the closed
    form vdiss * F * (q + dt S), with F the 3-D filter of filter_table and
    vdiss(k, l) given (cn2_set_diffusion's 1 / (1 + dfac vhdis))."""
    o = mk(O, n, filtering=filtering, stepper="cn2")
    rng = np.random.default_rng(11)
    q = rng.standard_normal((3, n, n, n + 1))
    S = rng.standard_normal((3, n, n, n + 1))
    vd = 1.0 / (1.0 + rng.uniform(0.0, 3.0, size=n * n))
    got = o.cn2_check(dt, vd, S, q)
    o.close()
    ref = vd.reshape(n, n)[None, :, :, None] * filter_table(n, n, filtering)[None] * (q + dt * S)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


# ---------------------------------------------------------------------------
# adjust_vorticity_mean (field_diagnostics.f90:584-619)
# ---------------------------------------------------------------------------
def check_mean_restore(O, stepper, n=32, dt=0.02):
    """A Beltrami mode and a shear mode sharing (k, l) (their cross term
    -u zeta differs between the lids, so the volume mean of the xi and eta
    tendencies, (q(top) - q(bottom)) / dz with q = w xi - u zeta, is nonzero)
    plus a horizontally uniform xi(z), eta(z): after every step the trapezoid
    means of xi and eta equal the initial ones, while the tendency alone would
    move them by far more than the tolerance."""
    o = mk(O, n, filtering="2/3-rule", nnu=3, prediss=30.0, stepper=stepper, pretype="vorch",
           dt_fixed=dt)
    _, vor = beltrami_sum(n, [(1, 1, 1, 0.3)])
    vor += shear_mode(n, 1, 1, 0.0)[1] + shear_mode(n, 1, 2, 0.5)[1]
    z = ics.grid(n, n, n, BOX["lower"], BOX["extent"])[2]
    vor[0] += 0.3 + 0.2 * np.cos(z)
    vor[1] += -0.1 + 0.25 * np.cos(2 * z)
    o.set_vorticity(vor)
    v0 = o.get(O.VOR)
    m0 = [trap_mean(v0[c]) for c in range(2)]
    o.adapt(0.0)
    o.source()
    st = o.get(O.SVORTS)
    drift = [dt * trap_mean(o.op(O.OP_COMBINE_PHYS, st[c])) for c in range(2)]
    assert max(abs(x) for x in drift) > 1e-4, drift
    o.step(0.0, dt)
    for s in range(3):
        o.vor2vel()
        v = o.get(O.VOR)
        for c in range(2):
            assert abs(trap_mean(v[c]) - m0[c]) < 1e-13, (s, c, trap_mean(v[c]), m0[c])
        o.advance_n(1)
    o.close()


# ---------------------------------------------------------------------------
# pytest entry points
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("f_cor", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.5)])
def test_tendency_flux_form(oracle_mod, f_cor):
    check_tendency(oracle_mod, f_cor)


@pytest.mark.parametrize("filtering", ["2/3-rule", "Hou & Li"])
def test_filter_tables(oracle_mod, filtering):
    o = mk(oracle_mod, 24, nz=18, filtering=filtering)
    assert np.abs(o.table(3).reshape(24, 24, 19) - filter_table(24, 18, filtering)).max() < 1e-14
    o.close()


@pytest.mark.parametrize("ls,nnu", [("Kolmogorov", 3), ("Kolmogorov", 1), ("geophysical", 2)])
def test_viscosity_closed_form(oracle_mod, ls, nnu):
    check_viscosity(oracle_mod, ls, nnu)


@pytest.mark.parametrize("stepper,pretype,prediss", [("impl-diff-rk4", "constant", 2.0e3),
                                                     ("impl-diff-rk4", "max-strain", 100.0),
                                                     ("cn2", "constant", 2.0e3), ("cn2", "vorch", 100.0)])
def test_hyperdiffusion_decay(oracle_mod, stepper, pretype, prediss):
    # prediss keeps exp(vdiss) finite at the largest wavenumbers of the grid
    check_decay(oracle_mod, stepper, pretype, prediss=prediss)


def test_hyperdiffusion_decay_nnu1(oracle_mod):
    check_decay(oracle_mod, "impl-diff-rk4", "constant", nnu=1, prediss=200.0)


def test_adapt_strain_limited(oracle_mod):
    dt, dtcfl, dtgg = check_adapt(oracle_mod, alpha=0.05)
    assert dt == pytest.approx(dtgg) and dtgg < dtcfl


def test_adapt_cfl_limited(oracle_mod):
    dt, dtcfl, dtgg = check_adapt(oracle_mod, alpha=50.0, mode=(1, 2, 2), phase=0.2)
    assert dt == pytest.approx(dtcfl) and dtcfl < dtgg


def test_adapt_limit_clipped(oracle_mod):
    dt, dtcfl, dtgg = check_adapt(oracle_mod, alpha=0.5, t=0.999, limit=1.0)
    assert dt == pytest.approx(0.001, rel=1e-9)


@pytest.mark.parametrize("win", [1, 3, 10])
def test_rolling_mean_window(oracle_mod, win):
    check_rolling_mean(oracle_mod, win)


@pytest.mark.parametrize("pretype", ["constant", "vorch", "bfmax", "max-strain", "us-max-strain",
                                     "roll-mean-max-strain"])
def test_prefactor_by_pretype(oracle_mod, pretype):
    check_prefactor(oracle_mod, pretype)


def test_time_order_rk4(oracle_mod):
    check_time_order(oracle_mod, "impl-diff-rk4", 4)


@pytest.mark.parametrize("filtering", ["2/3-rule", "Hou & Li"])
def test_cn2_held_source(oracle_mod, filtering):
    check_cn2_held_source(oracle_mod, filtering)


@pytest.mark.parametrize("stepper", ["impl-diff-rk4", "cn2"])
def test_vorticity_mean_restore(oracle_mod, stepper):
    check_mean_restore(oracle_mod, stepper)
