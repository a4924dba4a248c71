"""ps3d's field_stats record on the GPU (ps3d_gpu_field_stats) against the oracle's
restatement of update_netcdf_field_diagnostics
(src/fields/field_diagnostics_netcdf.f90:257-439), after a few steps of the
plain workload, on a wave-kernel grid and a quad-kernel grid."""
import math

import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = math.pi
BOX = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
TOL = 1.0e-12


def params(mod, nx, ny, nz, fcor3):
    return mod.make_params(nx, ny, nz, **BOX, filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0,
                           pretype="vorch", length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.0, limit=1.0e30,
                           f_cor=(0.0, 0.0, fcor3))


@pytest.mark.parametrize("shape,fcor3", [((32, 32, 32), 0.7), ((48, 32, 128), 0.0), ((16, 24, 64), -1.3)])
def test_field_stats_match_the_oracle(oracle_mod, shape, fcor3):
    from ps3d_amd import gpu as G
    nx, ny, nz = shape
    vor = ics.beltrami_multimode(nx, ny, nz, seed=4)
    d = G.Ps3dGpu(params(G, nx, ny, nz, fcor3))
    o = oracle_mod.Oracle(params(oracle_mod, nx, ny, nz, fcor3))
    d.set_vorticity(vor)
    o.set_vorticity(vor)
    print("field stats", shape, flush=True)
    d.advance_n(3)
    o.advance_n(3)
    d.vor2vel()
    o.vor2vel()
    _, dg = d.adapt(d.t)
    _, do = o.adapt(o.t)
    o.derived()
    sg = d.field_stats(dg)          # computes the divergence itself
    with np.errstate(all="ignore"):
        so = o.field_stats(do)
    d.close()
    o.close()
    assert len(sg) == len(so) == G.NSTAT
    for k, name in enumerate(G.STAT_NAMES):
        a, b = sg[k], so[k]
        if not np.isfinite(b):
            assert a == b, (name, a, b)          # ro with f_cor(3) = 0
            continue
        scale = max(abs(b), 1e-300)
        if name in ("uozmax", "lozmax", "uszrms", "z_vormean", "x_vormean", "y_vormean"):
            scale = max(scale, abs(so[G.STAT_NAMES.index("vortrms")]))
        assert abs(a - b) <= TOL * scale, (name, a, b)


def test_field_stats_needs_vor2vel(oracle_mod):
    from ps3d_amd import gpu as G
    d = G.Ps3dGpu(params(G, 16, 16, 16, 0.0))
    d.set_vorticity(ics.beltrami_multimode(16, 16, 16, seed=1))
    d.advance_n(1)
    with pytest.raises(G.Ps3dError) as e:
        d.field_stats()
    assert e.value.code == 5
    d.close()
