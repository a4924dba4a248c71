"""Mutation tests of the oracle's pins.

Each case copies oracle/ps3d_oracle.c, plants one plausible mistake (a flipped
sign, a swapped derivative, a dropped factor, an off-by-one window), builds the
copy with the oracle's own compiler flags and runs the pin that should catch
it (tests/test_oracle_pins_step.py).  The pin must fail.  The mistakes are
planted in the oracle, on the CPU; the product library is not involved.

A mutation names the exact source line it changes; if the oracle is edited so
that the line no longer occurs exactly once, the case fails and must be
updated with it.
"""
import concurrent.futures as cf
import os
import subprocess

import numpy as np
import pytest

from tests import test_oracle_pins_step as P
from tests import test_oracle_stats as ST

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ps3d_oracle.c")
CFLAGS = ["-O2", "-march=x86-64-v2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=gnu11", "-w"]

# (name, line in the oracle, planted replacement, pin that must fail)
MUTANTS = [
    ("tendency r sign", "fp[i] = u[i] * eta[i] - v[i] * xi[i];", "fp[i] = v[i] * xi[i] - u[i] * eta[i];",
     lambda O: P.check_tendency(O)),
    ("tendency q sign", "fp[i] = w[i] * xi[i] - u[i] * zeta[i];", "fp[i] = u[i] * zeta[i] - w[i] * xi[i];",
     lambda O: P.check_tendency(O)),
    ("tendency p sign", "fp[i] = v[i] * zeta[i] - w[i] * eta[i];", "fp[i] = w[i] * eta[i] - v[i] * zeta[i];",
     lambda O: P.check_tendency(O)),
    ("tendency diffx->diffy", "diffx(o, q, t3);", "diffy(o, q, t3);", lambda O: P.check_tendency(O)),
    ("tendency eta order", "t2[i] = r[i] - t2[i];", "t2[i] = t2[i] - r[i];", lambda O: P.check_tendency(O)),
    ("tendency planetary vorticity dropped", "o->vor[nc * N + i] += o->p.f_cor[nc];",
     "o->vor[nc * N + i] += 0.0;", lambda O: P.check_tendency(O, (0.3, -0.2, 0.5))),
    ("viscosity te/en swapped", "pow(K2max * te / en, 1.0 / 3.0)", "pow(K2max * en / te, 1.0 / 3.0)",
     lambda O: P.check_viscosity(O, "Kolmogorov", 3)),
    ("viscosity K2max as |k|^2 max", "double K2max = pow(fmax(rkxmax, rkymax), 2);",
     "double K2max = rkxmax * rkxmax + rkymax * rkymax;", lambda O: P.check_viscosity(O, "geophysical", 2)),
    ("kinetic energy lid weights", "return (0.5 * s1 + 0.25 * s0 + 0.25 * sn) * o->ncelli;",
     "return (0.5 * s1 + 0.5 * s0 + 0.5 * sn) * o->ncelli;", lambda O: P.check_adapt(O, alpha=0.05)),
    ("mean lid weights", "sl += f[0] + f[nz];", "sl += 2.0 * (f[0] + f[nz]);",
     lambda O: P.check_adapt(O, alpha=0.05)),
    ("vorch threshold on squares", "if (t1 + t2 + t3 > vortrms)",
     "if (t1 * t1 + t2 * t2 + t3 * t3 > vortrms * vortrms)", lambda O: P.check_adapt(O, alpha=0.05)),
    ("vortmax without root", "double vortmax = sqrt(get_abs_max(o, xp));",
     "double vortmax = get_abs_max(o, xp);", lambda O: P.check_adapt(O, alpha=0.05)),
    ("strain (1,3) sign", "a[6] = dwdx[i] + 0.5 * o->vor[N + i];", "a[6] = dwdx[i] - 0.5 * o->vor[N + i];",
     lambda O: P.check_adapt(O, alpha=0.05)),
    ("strain (2,3) sign", "a[7] = dwdy[i] - 0.5 * o->vor[i];", "a[7] = dwdy[i] + 0.5 * o->vor[i];",
     lambda O: P.check_adapt(O, alpha=0.05)),
    ("upper-lid strain level", "if (iz == nz) us = fmax(us, lmax);", "if (iz == nz - 1) us = fmax(us, lmax);",
     lambda O: P.check_adapt(O, alpha=0.05)),
    ("helicity pairs", "h += o->vel[c * o->npts + i] * o->vor[c * o->npts + i];",
     "h += o->vel[c * o->npts + i] * o->vor[((c + 1) % 3) * o->npts + i];",
     lambda O: P.check_adapt(O, alpha=0.05)),
    ("cfl factor", "static const double CFLMAX = 0.8;", "static const double CFLMAX = 0.9;",
     lambda O: P.check_adapt(O, alpha=50.0, mode=(1, 2, 2), phase=0.2)),
    ("dt not clipped to the limit", "fmin(dtcfl, o->p.limit - t)", "dtcfl",
     lambda O: P.check_adapt(O, alpha=0.5, t=0.999, limit=1.0)),
    ("rolling mean divisor", "o->r_sma = o->r_sma + (vnew - vold) / (double)o->r_len;",
     "o->r_sma = o->r_sma + (vnew - vold) / (double)(o->r_len + 1);", lambda O: P.check_rolling_mean(O, 3)),
    ("rolling mean window fills late", "o->r_filled = (o->r_len == o->r_inew);",
     "o->r_filled = (o->r_len + 1 == o->r_inew);", lambda O: P.check_rolling_mean(O, 3)),
    ("prefactor us/ls swapped", "case PS3D_PRE_US_MAX_STRAIN: vval = o->usggmax; break;",
     "case PS3D_PRE_US_MAX_STRAIN: vval = o->lsggmax; break;", lambda O: P.check_prefactor(O, "us-max-strain")),
    ("rk4 dissipation factor", "double dfac = 0.5 * vval * dt;", "double dfac = vval * dt;",
     lambda O: P.check_decay(O, "impl-diff-rk4", "constant")),
    ("rk4 stage-3 factor not squared", "for (size_t c = 0; c < o->n2; ++c) o->emq[c] = o->emq[c] * o->emq[c];",
     "", lambda O: P.check_decay(O, "impl-diff-rk4", "constant")),
    ("rk4 final weight", "for (size_t i = 0; i < N; ++i) q[i] = qdf[i] + dt6 * sqs[i];",
     "for (size_t i = 0; i < N; ++i) q[i] = qdf[i] + dt3 * sqs[i];",
     lambda O: P.check_time_order(O, "impl-diff-rk4", 4)),
    ("cn2 prefactor ignored", "double dfac = (o->p.nnu == 1) ? dt : vval * dt;", "double dfac = dt;",
     lambda O: P.check_decay(O, "cn2", "vorch", prediss=100.0)),
    ("cn2 half step", "q[i] = o->filt[i] * (sm[i] + dt2 * ts[i]);", "q[i] = o->filt[i] * (sm[i] + dt * ts[i]);",
     lambda O: P.check_cn2_held_source(O, "Hou & Li")),
    ("cn2 predictor without source", "o->vortsm[i] = o->svor[i] + dt2 * o->svorts[i];",
     "o->vortsm[i] = o->svor[i];", lambda O: P.check_cn2_held_source(O, "2/3-rule")),
    ("cn2 filter dropped", "q[i] = o->filt[i] * (sm[i] + dt2 * ts[i]);", "q[i] = (sm[i] + dt2 * ts[i]);",
     lambda O: P.check_cn2_held_source(O, "Hou & Li")),
    ("mean restore bottom lid only", "s[o->nz] = s[o->nz] + o->ini_vor_mean[nc] - savg[nc];", "",
     lambda O: P.check_mean_restore(O, "impl-diff-rk4")),
    ("vorticity mean without 1/nz", "savg[nc] = 0.5 * (s[0] + s[nz]) + o->fnzi * sum;",
     "savg[nc] = 0.5 * (s[0] + s[nz]) + sum;", lambda O: P.check_mean_restore(O, "cn2")),
    ("stats lid weight", "a0 += f[0] * f[0];", "a0 += 2.0 * f[0] * f[0];",
     lambda O: ST.check_horizontal_vorticity_stats(O)),
    ("stats kez component", "st[9] = trapezoid_sq(o, o->vel, 2, 2);", "st[9] = trapezoid_sq(o, o->vel, 0, 0);",
     lambda O: ST.check_beltrami_stats(O)),
    ("stats enxy_max without root", "const double he = sqrt(xi[i] * xi[i] + eta[i] * eta[i]);",
     "const double he = xi[i] * xi[i] + eta[i] * eta[i];", lambda O: ST.check_horizontal_vorticity_stats(O)),
    ("stats lids swapped", "lid[2 * c] = fmax(lid[2 * c], w3[c][t]);", "lid[2 * c] = fmax(lid[2 * c], w3[c][b]);",
     lambda O: ST.check_horizontal_vorticity_stats(O)),
    ("stats usuhmax without root", "st[30] = sqrt(uh);", "st[30] = uh;", lambda O: ST.check_beltrami_stats(O)),
    ("stats uszrms over all cells", "st[33] = sqrt(sz / (double)(o->nx * o->ny));", "st[33] = sqrt(sz * o->ncelli);",
     lambda O: ST.check_vortex_stats(O)),
    ("stats rossby of the ratio", "st[38] = mn[2] / o->p.f_cor[2];", "st[38] = fmin(mn[2], mx[2]) / fabs(o->p.f_cor[2]);",
     lambda O: ST.check_vortex_stats(O)),
]


def _build(tmp, idx, old, new):
    src = open(SRC).read()
    n = src.count(old)
    if n != 1:
        return None, "mutated line occurs %d times in ps3d_oracle.c: %r" % (n, old)
    d = os.path.join(tmp, "m%02d" % idx)
    os.makedirs(d, exist_ok=True)
    c = os.path.join(d, "ps3d_oracle.c")
    with open(c, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(d, "libmutant.so")
    cmd = ["gcc"] + CFLAGS + ["-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", so, c, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stderr[-2000:]
    return so, None


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory, oracle_mod):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = [ex.submit(_build, tmp, i, m[1], m[2]) for i, m in enumerate(MUTANTS)]
        return [f.result() for f in futs]


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(oracle_mod, mutant_libs, idx):
    so, err = mutant_libs[idx]
    assert so is not None, err
    name, _, _, check = MUTANTS[idx]
    prev = oracle_mod.use_library(so)
    try:
        with np.errstate(all="ignore"):
            with pytest.raises((AssertionError, np.linalg.LinAlgError)):
                check(oracle_mod)
    finally:
        oracle_mod.use_library_object(prev)


def test_unmutated_copy_passes(oracle_mod, tmp_path):
    """The same build recipe without a mutation passes every pin used above."""
    so, err = _build(str(tmp_path), 99, "static const double CFLMAX = 0.8;", "static const double CFLMAX = 0.8;")
    assert so is not None, err
    prev = oracle_mod.use_library(so)
    try:
        P.check_tendency(oracle_mod, (0.3, -0.2, 0.5))
        P.check_viscosity(oracle_mod, "Kolmogorov", 3)
        P.check_adapt(oracle_mod, alpha=0.05)
        P.check_rolling_mean(oracle_mod, 3)
        P.check_decay(oracle_mod, "cn2", "vorch", prediss=100.0)
        P.check_cn2_held_source(oracle_mod, "Hou & Li")
        P.check_mean_restore(oracle_mod, "impl-diff-rk4")
        ST.check_beltrami_stats(oracle_mod)
        ST.check_horizontal_vorticity_stats(oracle_mod)
        ST.check_vortex_stats(oracle_mod)
    finally:
        oracle_mod.use_library_object(prev)
