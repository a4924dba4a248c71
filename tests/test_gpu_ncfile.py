"""Field and field_stats files from a run on the GPU: what the file holds equals
what the handle returns, and a restart from the file reproduces the state
(read_netcdf_fields + setup_fields, src/fields/field_netcdf.f90:276-433,
src/utils/utils.f90:136-184)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = np.pi


def params(G, n):
    return G.make_params(n, n, n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI),
                         filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0,
                         pretype="vorch", length_scale="Kolmogorov", alpha=0.1, dt_fixed=0.01,
                         limit=1.0e30)


def test_fields_and_stats_files_from_a_run(tmp_path):
    from ps3d_amd import gpu as G
    from ps3d_amd import ncio
    n = 32
    d = G.Ps3dGpu(params(G, n))
    d.set_vorticity(ics.beltrami_multimode(n, n, n, seed=3))
    base = str(tmp_path / "beltrami")
    ff = ncio.FieldFile.for_model(base, d, derived=True)
    sf = ncio.StatsFile.for_model(base, d)
    kept = []
    for chunk in range(2):
        print("chunk", chunk, flush=True)
        d.advance_n(3)
        d.vor2vel()
        _, drow = d.adapt(d.t)
        assert sf.write_model(d, drow)
        srow = ncio.stats_row(drow, d.field_stats(drow))
        assert ff.write_model(d)
        kept.append((d.t, d.get(G.VEL), d.get(G.VOR), d.get(G.PRES), d.get(G.DELTA), srow))
    with netcdf_file(base + "_fields.nc", "r", mmap=False) as f:
        np.testing.assert_array_equal(np.array(f.variables["t"].data), [k[0] for k in kept])
        for r, (t, vel, vor, pres, delta, _) in enumerate(kept):
            for c, x in enumerate("xyz"):
                np.testing.assert_array_equal(f.variables[x + "_velocity"][r], vel[c].T)
                np.testing.assert_array_equal(f.variables[x + "_vorticity"][r], vor[c].T)
            np.testing.assert_array_equal(f.variables["pressure"][r], pres.T)
            np.testing.assert_array_equal(f.variables["horizontal_divergence"][r], delta.T)
    with netcdf_file(base + "_field_stats.nc", "r", mmap=False) as f:
        allrows = np.array([k[5] for k in kept])
        for j, name in enumerate(ncio.STATS_ROW_NAMES):
            np.testing.assert_array_equal(np.array(f.variables[name].data), allrows[:, j])
        np.testing.assert_array_equal(np.array(f.variables["t"].data), [k[0] for k in kept])

    # restart from the last record: the vorticity read back sets up the same state
    t, vor = ncio.read_fields(base + "_fields.nc", n, n, n, step=-1)
    assert t == kept[-1][0]
    e = G.Ps3dGpu(params(G, n))
    e.set_vorticity(vor)
    e.vor2vel()
    for which in (G.VOR, G.VEL):
        a, b = e.get(which), d.get(which)
        rel = max(np.abs(a[c] - b[c]).max() / np.abs(b[c]).max() for c in range(3))
        print("restart", which, rel, flush=True)
        assert rel < 1.0e-12
    d.close()
    e.close()
