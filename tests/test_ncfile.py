"""Field and field_stats NetCDF files (include/ps3d_nc.h), host code only.

The files are read back with scipy.io.netcdf_file, an independent reader of the
classic formats, and files scipy writes are read with the library's reader: names,
attributes and the (t, z, y, x) order of src/fields/field_netcdf.f90:65-272 and
src/netcdf/netcdf_writer.f90:355-432, the resume rule of field_netcdf.f90:75-95
and the step rule of read_netcdf_fields (field_netcdf.f90:286-299)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from ps3d_amd import gpu

pytestmark = pytest.mark.skipif(not os.path.exists(gpu.LIB_PATH), reason="HIP library not built")

NX, NY, NZ = 6, 4, 5
LOWER = [-1.0, 0.5, -0.25]
EXTENT = [2.0, 3.0, 0.5]


def _fields(seed):
    rng = np.random.default_rng(seed)
    shp = (3, NX, NY, NZ + 1)
    return rng.standard_normal(shp), rng.standard_normal(shp), rng.standard_normal(shp[1:]), \
        rng.standard_normal(shp[1:])


def _read(path):
    with netcdf_file(path, "r", mmap=False) as f:
        out = {"dims": dict(f.dimensions), "attrs": dict(f._attributes), "vars": {}}
        for n, v in f.variables.items():
            out["vars"][n] = (np.array(v.data), v.dimensions, dict(v._attributes))
        return out


def test_field_file_layout_and_records(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=True)
    recs = [(0.0,) + _fields(1), (0.25,) + _fields(2)]
    for t, vel, vor, p, d in recs:
        assert ff.write(t, vel, vor, p, d)
    r = _read(base + "_fields.nc")
    assert r["dims"] == {"t": None, "z": NZ + 1, "y": NY, "x": NX}
    attrs = {k.decode() if isinstance(k, bytes) else k: v for k, v in r["attrs"].items()}
    assert attrs["file_type"] in (b"fields", "fields")
    assert list(attrs["ncells"]) == [NX, NY, NZ]
    np.testing.assert_array_equal(attrs["extent"], EXTENT)
    np.testing.assert_array_equal(attrs["origin"], LOWER)
    # coordinates: origin + i * extent / n, z including the upper boundary
    for a, (name, n) in enumerate((("x", NX), ("y", NY), ("z", NZ))):
        data, dims, at = r["vars"][name]
        m = n + 1 if name == "z" else n
        np.testing.assert_allclose(data, LOWER[a] + np.arange(m) * (EXTENT[a] / n), rtol=0, atol=1e-15)
        assert dims == (name,)
        assert at["axis"] in (b"XYZ"[a:a + 1], "XYZ"[a])
        assert at["units"] in (b"m", "m")
    t, dims, at = r["vars"]["t"]
    np.testing.assert_array_equal(t, [0.0, 0.25])
    assert at["units"] in (b"seconds since 1970-01-01", "seconds since 1970-01-01")
    names = ["x_velocity", "y_velocity", "z_velocity", "x_vorticity", "y_vorticity", "z_vorticity"]
    for k, (t0, vel, vor, p, d) in enumerate(recs):
        for c in range(3):
            for arr, nm in ((vel, names[c]), (vor, names[3 + c])):
                data, dims, at = r["vars"][nm]
                assert dims == ("t", "z", "y", "x")
                # file order (t, z, y, x) = transpose of the host (x, y, z) volume
                np.testing.assert_array_equal(data[k], arr[c].transpose(2, 1, 0))
        np.testing.assert_array_equal(r["vars"]["pressure"][0][k], p.transpose(2, 1, 0))
        np.testing.assert_array_equal(r["vars"]["horizontal_divergence"][0][k], d.transpose(2, 1, 0))
    assert r["vars"]["x_velocity"][2]["units"] in (b"m/s", "m/s")
    assert r["vars"]["z_vorticity"][2]["long_name"] in (b"z vorticity component", "z vorticity component")
    assert ncio.records(base + "_fields.nc") == (2, 0.25)


def test_field_file_read_back_steps(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=False)
    data = [(0.5 * k,) + _fields(10 + k)[:2] for k in range(3)]
    for t, vel, vor in data:
        ff.write(t, vel, vor)
    path = base + "_fields.nc"
    for step, k in ((1, 0), (2, 1), (3, 2), (-1, 2), (7, 2)):
        t, vor, vel = ncio.read_fields(path, NX, NY, NZ, step=step, velocity=True)
        assert t == data[k][0]
        np.testing.assert_array_equal(vel, data[k][1])
        np.testing.assert_array_equal(vor, data[k][2])
    with pytest.raises(ncio.NcError):
        ncio.read_fields(path, NX, NY, NZ, step=0)
    with pytest.raises(ncio.NcError):
        ncio.read_fields(path, NX + 2, NY, NZ)


def test_field_file_resume_skips_earlier_times(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=False)
    vel, vor, _, _ = _fields(3)
    ff.write(0.0, vel, vor)
    ff.write(1.0, vel, vor)
    # a restarted run reopens the file and replays from an earlier time
    ff2 = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=False)
    assert ff2.restart_time == 1.0
    assert not ff2.write(0.5, vel, vor)
    assert not ff2.write(1.0, vel, vor)
    assert ff2.write(1.5, 2 * vel, vor)
    assert ncio.records(base + "_fields.nc") == (3, 1.5)
    t, vor3, vel3 = ncio.read_fields(base + "_fields.nc", NX, NY, NZ, step=-1, velocity=True)
    np.testing.assert_array_equal(vel3, 2 * vel)


def test_file_with_no_records_is_recreated(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=True)
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=False)
    assert ff.restart_time == -1.0
    r = _read(base + "_fields.nc")
    assert "pressure" not in r["vars"]


def test_derived_file_requires_pressure(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=True)
    vel, vor, _, _ = _fields(4)
    with pytest.raises(ncio.NcError):
        ff.write(0.0, vel, vor)


def test_reader_reads_scipy_written_file(tmp_path):
    """A 64-bit offset file written by scipy (another writer) reads back through the
    library, including the reordering to the host layout."""
    from ps3d_amd import ncio
    path = str(tmp_path / "other_fields.nc")
    rng = np.random.default_rng(7)
    vols = {n: rng.standard_normal((2, NZ + 1, NY, NX)) for n in
            ("x_vorticity", "y_vorticity", "z_vorticity", "x_velocity", "y_velocity", "z_velocity")}
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("t", None)
        f.createDimension("z", NZ + 1)
        f.createDimension("y", NY)
        f.createDimension("x", NX)
        tv = f.createVariable("t", "d", ("t",))
        tv[:] = [0.1, 0.2]
        for n, v in vols.items():
            var = f.createVariable(n, "d", ("t", "z", "y", "x"))
            var[:] = v
    assert ncio.records(path) == (2, 0.2)
    t, vor, vel = ncio.read_fields(path, NX, NY, NZ, step=1, velocity=True)
    assert t == 0.1
    for c, x in enumerate("xyz"):
        np.testing.assert_array_equal(vor[c], vols[x + "_vorticity"][0].transpose(2, 1, 0))
        np.testing.assert_array_equal(vel[c], vols[x + "_velocity"][0].transpose(2, 1, 0))


def test_field_record_over_4gib_is_rejected(tmp_path):
    from ps3d_amd import ncio
    with pytest.raises(ncio.NcError, match="4 GiB"):
        ncio.FieldFile(str(tmp_path / "big"), 1024, 1024, 1024, LOWER, EXTENT)
    assert not os.path.exists(str(tmp_path / "big_fields.nc"))


def test_stats_file(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    sf = ncio.StatsFile(base)
    rows = np.random.default_rng(5).standard_normal((4, gpu.NDIAG))
    rows[:, 0] = [0.0, 0.1, 0.2, 0.3]
    assert sf.write(rows[:3]) == 3
    # resume: rows at or before the last time are skipped
    sf2 = ncio.StatsFile(base)
    assert sf2.restart_time == 0.2
    assert sf2.write(rows) == 1
    r = _read(base + "_field_stats.nc")
    assert r["dims"] == {"t": None}
    names = ["t", "dt", "ke", "en", "helicity", "vortmax", "vortrms", "vorch", "x_vormean",
             "y_vormean", "z_vormean", "gmax", "umax", "vmax", "wmax", "usgmax", "lsgmax",
             "rolling_mean_gmax", "prefactor"]
    assert sorted(r["vars"]) == sorted(names)
    for k, n in enumerate(names):
        np.testing.assert_array_equal(r["vars"][n][0], rows[:, k])
    assert r["vars"]["ke"][2]["long_name"] in (b"domain-averaged kinetic energy", "domain-averaged kinetic energy")
    assert r["vars"]["en"][2]["units"] in (b"1/s^2", "1/s^2")


def test_grid_attributes_read_back(tmp_path):
    from ps3d_amd import ncio
    ncio.FieldFile(str(tmp_path / "g"), NX, NY, NZ, LOWER, EXTENT, derived=False)
    n, e, o = ncio.grid(str(tmp_path / "g_fields.nc"))
    assert n == [NX, NY, NZ]
    assert e == EXTENT and o == LOWER


def test_kinetic_energy_weights_the_lids_by_half():
    """get_kinetic_energy (field_diagnostics.f90:85-122): a uniform |u| = 1 gives
    1/2; energy only on the lids counts a quarter per lid level."""
    from ps3d_amd.genspec import kinetic_energy
    vel = np.zeros((3, 4, 6, 9))
    vel[0] = 1.0
    assert abs(kinetic_energy(vel) - 0.5) < 1e-15
    vel[:] = 0.0
    vel[2, :, :, 0] = 2.0
    assert abs(kinetic_energy(vel) - 0.25 * 4.0 / 8) < 1e-15
