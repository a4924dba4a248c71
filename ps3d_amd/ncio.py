"""ps3d's field and field_stats NetCDF files (classic data model: the 64-bit
offset format, or the 64-bit data format CDF-5 where a field record exceeds
4 GiB), written by the library's host code (include/ps3d_nc.h).

FieldFile mirrors src/fields/field_netcdf.f90 (create_netcdf_field_file,
write_netcdf_fields, read_netcdf_fields) and StatsFile mirrors
src/fields/field_diagnostics_netcdf.f90 (create_netcdf_field_stats_file,
write_netcdf_field_stats): file names basename + '_fields.nc' /
'_field_stats.nc'; an existing file with records is resumed, and a write at a
time not later than its last record is skipped, as ps3d does after a restart.
"""
import ctypes as C
import os

import numpy as np

from .gpu import DELTA, PRES, STAT_NAMES, VEL, VOR, lib

# one field_stats record: t, ps3d's 40 statistics, then helicity, dt, prefactor
STATS_ROW_NAMES = ["t"] + STAT_NAMES + ["helicity", "dt", "prefactor"]
STATS_ROW = len(STATS_ROW_NAMES)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ready = False


class NcError(RuntimeError):
    pass


def _L():
    global _ready
    L = lib()
    if not _ready:
        L.ps3d_nc_last_error.restype = C.c_char_p
        L.ps3d_nc_create_fields.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.ps3d_nc_append_fields.argtypes = [C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                            _dp, _dp, C.c_void_p, C.c_void_p]
        L.ps3d_nc_read_fields.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int,
                                          C.c_int, C.c_void_p, _dp]
        L.ps3d_nc_records.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ps3d_nc_grid.argtypes = [C.c_char_p, _ip, _dp, _dp]
        L.ps3d_nc_create_stats.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.ps3d_nc_stats_row.argtypes = [_dp, _dp, _dp]
        L.ps3d_nc_append_stats.argtypes = [C.c_char_p, _dp, C.c_int]
        _ready = True
    return L


def _check(st):
    if st != 0:
        raise NcError("ps3d_nc error %d: %s" % (st, _L().ps3d_nc_last_error().decode()))


def _b(path):
    return os.fsencode(path)


def records(path):
    """(number of records, time of the last record)."""
    n = C.c_int()
    t = C.c_double()
    _check(_L().ps3d_nc_records(_b(path), C.byref(n), C.byref(t)))
    return n.value, t.value


def grid(path):
    """(ncells, extent, origin) of a field file (get_netcdf_box)."""
    n = np.zeros(3, dtype=np.int32)
    e = np.zeros(3)
    o = np.zeros(3)
    _check(_L().ps3d_nc_grid(_b(path), n, e, o))
    return [int(v) for v in n], list(e), list(o)


def _resume(path):
    """restart_time of an existing file with records, else None (field_netcdf.f90:75-95)."""
    if os.path.exists(path):
        n, t = records(path)
        if n > 0:
            return t
        os.remove(path)
    return None


class FieldFile:
    """basename_fields.nc: velocity, vorticity and, with derived=True, pressure and
    horizontal divergence, one record per write."""

    def __init__(self, basename, nx, ny, nz, lower, extent, derived=True, cdf5=False):
        """cdf5=True writes the 64-bit data format even where every record fits
        the 64-bit offset one (which grids of 4 GiB records always use)."""
        self.path = basename + "_fields.nc"
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.derived = bool(derived)
        t = _resume(self.path)
        self.restart_time = -1.0 if t is None else t
        if t is None:
            _check(_L().ps3d_nc_create_fields(
                _b(self.path), self.nx, self.ny, self.nz,
                np.ascontiguousarray(lower, dtype=np.float64),
                np.ascontiguousarray(extent, dtype=np.float64), int(self.derived) | (2 if cdf5 else 0)))

    @classmethod
    def for_model(cls, basename, model, derived=True):
        p = model.p
        return cls(basename, p.nx, p.ny, p.nz, list(p.lower), list(p.extent), derived)

    def write(self, t, vel, vor, pres=None, delta=None):
        """write_netcdf_fields (field_netcdf.f90:177-254); returns False when skipped."""
        if t <= self.restart_time:
            return False
        shp = (3, self.nx, self.ny, self.nz + 1)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(shp)
        vor = np.ascontiguousarray(vor, dtype=np.float64).reshape(shp)
        pp = dd = None
        if pres is not None:
            pres = np.ascontiguousarray(pres, dtype=np.float64).reshape(shp[1:])
            pp = pres.ctypes.data
        if delta is not None:
            delta = np.ascontiguousarray(delta, dtype=np.float64).reshape(shp[1:])
            dd = delta.ctypes.data
        _check(_L().ps3d_nc_append_fields(_b(self.path), float(t), self.nx, self.ny, self.nz,
                                          vel, vor, pp, dd))
        return True

    def write_model(self, model, t=None):
        """Write the fields of a Ps3dGpu after vor2vel (and derived() when the file
        holds pressure and divergence)."""
        t = model.t if t is None else t
        if t <= self.restart_time:
            return False
        vel, vor = model.get(VEL), model.get(VOR)
        if self.derived:
            model.derived()
            return self.write(t, vel, vor, model.get(PRES), model.get(DELTA))
        return self.write(t, vel, vor)


def read_fields(path, nx, ny, nz, step=-1, velocity=False):
    """read_netcdf_fields (field_netcdf.f90:276-433): (t, vor) or (t, vor, vel) of
    record `step` (from 1; -1 or past the end: the last)."""
    vor = np.zeros((3, nx, ny, nz + 1))
    vel = np.zeros((3, nx, ny, nz + 1)) if velocity else None
    t = C.c_double()
    _check(_L().ps3d_nc_read_fields(_b(path), int(step), C.byref(t), int(nx), int(ny), int(nz),
                                    vel.ctypes.data if velocity else None, vor))
    return (t.value, vor, vel) if velocity else (t.value, vor)


def stats_row(diag, stats):
    """A field_stats record from an adapt record (NDIAG) and the field statistics
    of the same state (NSTAT, Ps3dGpu.field_stats): ps3d_nc_stats_row."""
    row = np.zeros(STATS_ROW)
    _check(_L().ps3d_nc_stats_row(np.ascontiguousarray(diag, dtype=np.float64),
                                  np.ascontiguousarray(stats, dtype=np.float64), row))
    return row


class StatsFile:
    """basename_field_stats.nc: one record (STATS_ROW doubles, t first) per write."""

    def __init__(self, basename, nx, ny, nz, lower, extent):
        self.path = basename + "_field_stats.nc"
        t = _resume(self.path)
        self.restart_time = -1.0 if t is None else t
        if t is None:
            _check(_L().ps3d_nc_create_stats(_b(self.path), int(nx), int(ny), int(nz),
                                             np.ascontiguousarray(lower, dtype=np.float64),
                                             np.ascontiguousarray(extent, dtype=np.float64)))

    @classmethod
    def for_model(cls, basename, model):
        p = model.p
        return cls(basename, p.nx, p.ny, p.nz, list(p.lower), list(p.extent))

    def write(self, rows):
        """write_netcdf_field_stats (field_diagnostics_netcdf.f90:211-252) for each
        record whose time is later than the file's last one; returns records written."""
        rows = np.atleast_2d(np.asarray(rows, dtype=np.float64))
        if rows.shape[1] != STATS_ROW:
            raise ValueError("field_stats records must have %d entries" % STATS_ROW)
        rows = np.ascontiguousarray(rows[rows[:, 0] > self.restart_time])
        if len(rows):
            _check(_L().ps3d_nc_append_stats(_b(self.path), rows, len(rows)))
        return len(rows)

    def write_model(self, model, diag):
        """The record of a Ps3dGpu's state after vor2vel, with diag the adapt record
        of the same time; returns False when skipped."""
        if diag[0] <= self.restart_time:
            return False
        return self.write(stats_row(diag, model.field_stats(diag))) == 1
