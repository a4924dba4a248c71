"""The Fortran host program (ps3d_amd/fortran/ps3d_driver.f90): ps3d's &PS3D
namelist config, field file input, time loop and NetCDF output on the library.

CPU tests run the failure paths that stop before any GPU call; the GPU test runs a
32^3 Beltrami config and compares every output record with ps3d's loop run on the
oracle."""
import os
import subprocess

import numpy as np
import pytest

from ps3d_amd import build, gpu

HAVE = os.path.exists(build.DRIVER) and os.path.exists(gpu.LIB_PATH)
PI = np.pi

CONFIG = """&PS3D
 field_file              = '{field}'
 vor_visc%nnu = 3
 vor_visc%prediss = 30.0
 vor_visc%pretype = 'vorch'
 vor_visc%length_scale = 'Kolmogorov'
 stepper = 'impl-diff-rk4'
 filtering = '2/3-rule'
 output%field_freq         = 0.02
 output%field_stats_freq   = 0.01
 output%write_fields       = .true.
 output%overwrite          = .true.
 output%basename           = '{base}'
 time%limit              = 0.05
 time%alpha              = 0.1
 time%precise_stop       = .false.
 time%dt_fixed           = 0.01
/
"""


def run_driver(cfg, timeout=120):
    return subprocess.run([build.DRIVER, "--config", cfg], capture_output=True, text=True, timeout=timeout)


@pytest.mark.skipif(not HAVE, reason="driver not built")
def test_driver_usage_and_missing_config(tmp_path):
    r = subprocess.run([build.DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Usage" in r.stdout
    r = run_driver(str(tmp_path / "none.config"))
    assert r.returncode != 0 and "does not exist" in r.stdout


@pytest.mark.skipif(not HAVE, reason="driver not built")
def test_driver_rejects_bad_namelist_and_missing_field_file(tmp_path):
    bad = tmp_path / "bad.config"
    bad.write_text("&PS3D\n no_such_entry = 1\n/\n")
    r = run_driver(str(bad))
    assert r.returncode != 0 and "invalid Namelist" in r.stdout
    cfg = tmp_path / "run.config"
    cfg.write_text(CONFIG.format(field=str(tmp_path / "missing.nc"), base=str(tmp_path / "out")))
    r = run_driver(str(cfg))
    assert r.returncode != 0 and "field file grid" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not HAVE, reason="driver not built")
@pytest.mark.parametrize("dt_fixed,limit", [(0.01, 0.047), (0.0, 0.1)])
def test_driver_run_matches_the_oracle(tmp_path, oracle_mod, dt_fixed, limit):
    """The driver's field and field_stats files against ps3d's loop run on the
    oracle (src/ps3d.f90:114-126: write_step after adapt, every step clipped to
    time%limit - t by adapt, advance.f90:325-333; write_last_step at the end),
    with a limit that is not a multiple of the step."""
    from scipy.io import netcdf_file

    from ps3d_amd import ics, ncio
    n = 32
    kw = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))
    vor0 = ics.beltrami_multimode(n, n, n, seed=1)
    # initial field file, as ps3d's generator writes it (vorticity; velocity unused)
    ic = ncio.FieldFile(str(tmp_path / "ic"), n, n, n, kw["lower"], kw["extent"], derived=False)
    ic.write(0.0, np.zeros_like(vor0), vor0)
    base = str(tmp_path / "beltrami_32")
    cfg = tmp_path / "beltrami_32.config"
    cfg.write_text(CONFIG.format(field=ic.path, base=base).replace(
        "time%limit              = 0.05", "time%limit              = " + repr(limit)).replace(
        "time%dt_fixed           = 0.01", "time%dt_fixed           = " + repr(dt_fixed)))
    print("driver run dt_fixed=%g limit=%g" % (dt_fixed, limit), flush=True)
    r = run_driver(str(cfg))
    assert r.returncode == 0, r.stdout + r.stderr

    p = oracle_mod.make_params(n, n, n, **kw, filtering="2/3-rule", stepper="impl-diff-rk4", nnu=3, prediss=30.0,
                               pretype="vorch", length_scale="Kolmogorov", alpha=0.1, dt_fixed=dt_fixed,
                               limit=limit)
    o = oracle_mod.Oracle(p)
    o.set_vorticity(vor0)
    eps = np.finfo(float).eps
    ref, recs = [], []
    t, nfw, nsfw = 0.0, 0, 0
    adapt_slots = {"vortmax": 5, "vortrms": 6, "vorch": 7, "x_vormean": 8, "y_vormean": 9, "z_vormean": 10,
                   "gmax": 11, "umax": 12, "vmax": 13, "wmax": 14, "usgmax": 15, "lsgmax": 16,
                   "rolling_mean_gmax": 17}

    def record(stats, row):
        # t, the 40 statistics with the adapt values of the step starting at t, then
        # helicity, dt, prefactor (the file's record layout)
        r = dict(zip(ncio.STATS_ROW_NAMES[1:41], stats))
        for name, j in adapt_slots.items():
            r[name] = row[j]
        r.update(t=row[0], helicity=row[4], dt=row[1], prefactor=row[18])
        return [r[n] for n in ncio.STATS_ROW_NAMES]

    while t < limit:
        wf = t + eps >= nfw * 0.02
        ws = t + eps >= nsfw * 0.01
        if wf or ws:
            o.vor2vel()
            o.derived()
        if wf:
            ref.append((t, o.get(oracle_mod.VOR), o.get(oracle_mod.VEL)))
            nfw += 1
        if ws:
            stats = o.field_stats()
            nsfw += 1
        row = o.advance_n(1)[0]
        if ws:
            recs.append(record(stats, row))
        t = o.t
    o.vor2vel()
    o.derived()
    ref.append((t, o.get(oracle_mod.VOR), o.get(oracle_mod.VEL)))
    stats = o.field_stats()
    recs.append(record(stats, o.adapt(t)[1]))
    o.close()
    assert abs(t - limit) < 1e-15
    with netcdf_file(base + "_fields.nc", "r", mmap=False) as f:
        times = np.array(f.variables["t"].data)
        np.testing.assert_allclose(times, [r[0] for r in ref], rtol=0, atol=1e-15)
        for k, (tk, vor, vel) in enumerate(ref):
            for c, x in enumerate("xyz"):
                np.testing.assert_allclose(f.variables[x + "_vorticity"][k], vor[c].T, rtol=0,
                                           atol=1e-12 * np.abs(vor[c]).max())
                np.testing.assert_allclose(f.variables[x + "_velocity"][k], vel[c].T, rtol=0,
                                           atol=1e-12 * np.abs(vel[c]).max())
    recs = np.array(recs)
    with netcdf_file(base + "_field_stats.nc", "r", mmap=False) as f:
        tt = np.array(f.variables["t"].data)
        assert len(tt) == len(recs)
        np.testing.assert_allclose(tt, recs[:, 0], rtol=0, atol=1e-15)
        assert list(f.variables) == ncio.STATS_ROW_NAMES
        rms = np.abs(recs[:, ncio.STATS_ROW_NAMES.index("vortrms")]).max()
        for j, name in enumerate(ncio.STATS_ROW_NAMES):
            got = np.array(f.variables[name].data)
            want = recs[:, j]
            if not np.all(np.isfinite(want)):
                np.testing.assert_array_equal(got, want, err_msg=name)     # ro with f_cor(3) = 0
                continue
            scale = max(np.abs(want).max(), rms if name in ("uozmax", "lozmax", "uszrms", "x_vormean",
                                                            "y_vormean", "z_vormean") else 0.0, 1e-300)
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale, err_msg=name)
