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


class Cdf:
    """A reader of the NetCDF classic formats written from the published format
    specification (CDF-2: 4-byte NON_NEG counts and sizes; CDF-5: 8-byte), for the
    64-bit data files scipy cannot open: header fields and record data."""

    def __init__(self, path):
        self.b = open(path, "rb").read()
        assert self.b[:3] == b"CDF"
        self.ver = self.b[3]
        self.o = 4
        self.numrecs = self.nn()
        self.dims = self._list(0x0A, lambda: (self.name(), self.nn()))
        self.gatts = dict(self._list(0x0C, self.att))
        self.vars = {}
        for _ in range(self._tag(0x0B)):
            name = self.name()
            dimids = [self.nn() for _ in range(self.nn())]
            atts = dict(self._list(0x0C, self.att))
            typ, vsize, begin = self.u32(), self.nn(), self.u64()
            self.vars[name] = dict(dimids=dimids, atts=atts, type=typ, vsize=vsize, begin=begin)
        rec = [v for v in self.vars.values() if v["dimids"] and self.dims[v["dimids"][0]][1] == 0]
        self.recsize = sum(v["vsize"] for v in rec)

    def _int(self, n):
        v = int.from_bytes(self.b[self.o:self.o + n], "big")
        self.o += n
        return v

    def u32(self):
        return self._int(4)

    def u64(self):
        return self._int(8)

    def nn(self):
        return self.u64() if self.ver == 5 else self.u32()

    def name(self):
        n = self.nn()
        s = self.b[self.o:self.o + n].decode()
        self.o += (n + 3) // 4 * 4
        return s

    def att(self):
        name, typ, n = self.name(), self.u32(), None
        n = self.nn()
        if typ == 2:
            v = self.b[self.o:self.o + n].decode()
            self.o += (n + 3) // 4 * 4
        elif typ == 4:
            v = [int.from_bytes(self.b[self.o + 4 * i:self.o + 4 * i + 4], "big", signed=True) for i in range(n)]
            self.o += 4 * n
        else:
            v = list(np.frombuffer(self.b[self.o:self.o + 8 * n], dtype=">f8"))
            self.o += 8 * n
        return name, v

    def _tag(self, tag):
        t = self.u32()
        n = self.nn()
        assert t in (0, tag)
        return n if t == tag else 0

    def _list(self, tag, item):
        return [item() for _ in range(self._tag(tag))]

    def record(self, name, rec, shape):
        v = self.vars[name]
        off = v["begin"] + rec * self.recsize
        return np.frombuffer(self.b[off:off + v["vsize"]], dtype=">f8").reshape(shape)


def test_cdf2_reader_agrees_with_scipy(tmp_path):
    """The spec reader above, on a CDF-2 file, reads what scipy reads."""
    from ps3d_amd import ncio
    base = str(tmp_path / "two")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=False)
    t, vel, vor, _, _ = (0.5,) + _fields(7)
    ff.write(t, vel, vor)
    c = Cdf(base + "_fields.nc")
    assert c.ver == 2 and c.numrecs == 1
    r = _read(base + "_fields.nc")
    for name in ("x_velocity", "z_vorticity"):
        np.testing.assert_array_equal(c.record(name, 0, (NZ + 1, NY, NX)), r["vars"][name][0][0])


def test_cdf5_field_file(tmp_path):
    """A field file in the 64-bit data format: header fields as the specification
    lays them out, records in (t, z, y, x) order, and the library reads it back."""
    from ps3d_amd import ncio
    base = str(tmp_path / "five")
    ff = ncio.FieldFile(base, NX, NY, NZ, LOWER, EXTENT, derived=True, cdf5=True)
    recs = [(0.0,) + _fields(11), (0.5,) + _fields(12)]
    for t, vel, vor, p, d in recs:
        assert ff.write(t, vel, vor, p, d)
    c = Cdf(base + "_fields.nc")
    assert c.ver == 5 and c.numrecs == 2
    assert c.dims == [("t", 0), ("z", NZ + 1), ("y", NY), ("x", NX)]
    assert c.gatts["file_type"] == "fields" and c.gatts["ncells"] == [NX, NY, NZ]
    np.testing.assert_array_equal(c.gatts["extent"], EXTENT)
    assert c.vars["x_velocity"]["vsize"] == 8 * NX * NY * (NZ + 1)
    assert c.vars["x_velocity"]["atts"]["long_name"] == "x velocity component"
    for k, (t, vel, vor, p, d) in enumerate(recs):
        assert c.record("t", k, ())[()] == t
        for comp, x in enumerate("xyz"):
            np.testing.assert_array_equal(c.record(x + "_velocity", k, (NZ + 1, NY, NX)), vel[comp].T)
            np.testing.assert_array_equal(c.record(x + "_vorticity", k, (NZ + 1, NY, NX)), vor[comp].T)
        np.testing.assert_array_equal(c.record("pressure", k, (NZ + 1, NY, NX)), p.T)
        tt, vor_b, vel_b = ncio.read_fields(base + "_fields.nc", NX, NY, NZ, step=k + 1, velocity=True)
        assert tt == t
        np.testing.assert_array_equal(vor_b, vor)
        np.testing.assert_array_equal(vel_b, vel)
    assert ncio.records(base + "_fields.nc") == (2, 0.5)


def test_1024_cubed_field_file_is_cdf5(tmp_path):
    """ps3d's 1024^3 grid: each field record is 8.6 GB, beyond CDF-2's 32-bit size
    field, so the file is CDF-5 with the full 64-bit vsize; the header alone is
    written (no record)."""
    from ps3d_amd import ncio
    base = str(tmp_path / "big")
    ncio.FieldFile(base, 1024, 1024, 1024, LOWER, EXTENT)
    c = Cdf(base + "_fields.nc")
    assert c.ver == 5 and c.numrecs == 0
    assert c.vars["z_vorticity"]["vsize"] == 8 * 1024 * 1024 * 1025
    assert c.vars["x"]["vsize"] == 8 * 1024 and c.vars["z"]["vsize"] == 8 * 1025
    np.testing.assert_allclose(np.frombuffer(c.b[c.vars["z"]["begin"]:c.vars["z"]["begin"] + 8 * 1025], ">f8"),
                               LOWER[2] + np.arange(1025) * (EXTENT[2] / 1024), rtol=0, atol=1e-15)
    # 512^3 records fit CDF-2
    ncio.FieldFile(str(tmp_path / "mid"), 512, 512, 512, LOWER, EXTENT)
    assert Cdf(str(tmp_path / "mid") + "_fields.nc").ver == 2


def test_stats_file(tmp_path):
    from ps3d_amd import ncio
    base = str(tmp_path / "run")
    sf = ncio.StatsFile(base, NX, NY, NZ, LOWER, EXTENT)
    rows = np.random.default_rng(5).standard_normal((4, ncio.STATS_ROW))
    rows[:, 0] = [0.0, 0.1, 0.2, 0.3]
    assert sf.write(rows[:3]) == 3
    # resume: rows at or before the last time are skipped
    sf2 = ncio.StatsFile(base, NX, NY, NZ, LOWER, EXTENT)
    assert sf2.restart_time == 0.2
    assert sf2.write(rows) == 1
    r = _read(base + "_field_stats.nc")
    assert r["dims"] == {"t": None}
    # ps3d's 40 statistics in its order (field_diagnostics_netcdf.f90:36-75), then the extensions
    names = ["t", "ke", "en", "vortmax", "vortrms", "vorch", "x_vormean", "y_vormean", "z_vormean",
             "kexy", "kez", "enxy", "enz", "x_vormin", "y_vormin", "z_vormin", "x_vormax", "y_vormax",
             "z_vormax", "enxy_max", "gmax", "bfmax", "umax", "vmax", "wmax", "uoxmax", "loxmax", "uoymax",
             "loymax", "uozmax", "lozmax", "usuhmax", "usgmax", "lsgmax", "uszrms", "usdeltarms",
             "rolling_mean_gmax", "rolling_mean_bfmax", "ri_min", "ro_min", "ro_max",
             "helicity", "dt", "prefactor"]
    assert sorted(r["vars"]) == sorted(names)
    with netcdf_file(base + "_field_stats.nc", "r", mmap=False) as f:
        assert list(f.variables) == names
    for k, n in enumerate(names):
        np.testing.assert_array_equal(r["vars"][n][0], rows[:, k])
    assert r["vars"]["ke"][2]["long_name"] in (b"domain-averaged kinetic energy", "domain-averaged kinetic energy")
    assert r["vars"]["en"][2]["units"] in (b"1/s^2", "1/s^2")
    assert r["vars"]["usdeltarms"][2]["long_name"] in (b"rms of upper surface horizontal divergence",
                                                       "rms of upper surface horizontal divergence")
    attrs = {k.decode() if isinstance(k, bytes) else k: v for k, v in r["attrs"].items()}
    assert attrs["file_type"] in (b"field_stats", "field_stats")
    assert list(attrs["ncells"]) == [NX, NY, NZ]


def test_stats_row_takes_adapt_values_from_the_adapt_record():
    """ps3d_nc_stats_row: t and the entries adapt sets (advance.f90:188-193,
    315-321, 366) come from the adapt record, the rest from the field statistics."""
    from ps3d_amd import ncio
    diag = 1000.0 + np.arange(gpu.NDIAG)
    stats = -1.0 - np.arange(gpu.NSTAT)
    row = ncio.stats_row(diag, stats)
    d = dict(zip(ncio.STATS_ROW_NAMES, row))
    adapt = {"vortmax": "vortmax", "vortrms": "vortrms", "vorch": "vorch", "x_vormean": "vormean_x",
             "y_vormean": "vormean_y", "z_vormean": "vormean_z", "gmax": "ggmax", "umax": "umax", "vmax": "vmax",
             "wmax": "wmax", "usgmax": "usggmax", "lsgmax": "lsggmax", "rolling_mean_gmax": "rgmax",
             "t": "t", "helicity": "hel", "dt": "dt", "prefactor": "prefactor"}
    for name in ncio.STATS_ROW_NAMES:
        if name in adapt:
            assert d[name] == diag[gpu.DIAG_NAMES.index(adapt[name])], name
        else:
            assert d[name] == stats[gpu.STAT_NAMES.index(name)], name


def test_grid_attributes_read_back(tmp_path):
    from ps3d_amd import ncio
    ncio.FieldFile(str(tmp_path / "g"), NX, NY, NZ, LOWER, EXTENT, derived=False)
    n, e, o = ncio.grid(str(tmp_path / "g_fields.nc"))
    assert n == [NX, NY, NZ]
    assert e == EXTENT and o == LOWER
