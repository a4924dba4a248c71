"""Run parameters: the ps3d_params struct of include/ps3d_gpu.h as a ctypes
Structure, built from the &PS3D namelist vocabulary of src/utils/options.f90."""
import ctypes as C
import math

FILTERING = {"Hou & Li": 0, "2/3-rule": 1}
STEPPER = {"impl-diff-rk4": 0, "cn2": 1}
PRETYPE = {"constant": 0, "vorch": 1, "bfmax": 2, "roll-mean-max-strain": 3,
           "max-strain": 4, "us-max-strain": 5}
LENGTH_SCALE = {"Kolmogorov": 0, "geophysical": 1}


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lower", C.c_double * 3), ("extent", C.c_double * 3),
                ("filtering", C.c_int), ("stepper", C.c_int), ("nnu", C.c_int),
                ("prediss", C.c_double), ("pretype", C.c_int),
                ("roll_mean_win_size", C.c_int), ("length_scale", C.c_int),
                ("alpha", C.c_double), ("limit", C.c_double),
                ("dt_fixed", C.c_double), ("f_cor", C.c_double * 3)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k in ("lower", "extent", "f_cor") else getattr(self, k))
                for k, _ in self._fields_}


def _lookup(table, v, what):
    if isinstance(v, str):
        if v not in table:
            raise ValueError("unknown %s %r (expected one of %s)" % (what, v, sorted(table)))
        return table[v]
    return int(v)


def make_params(nx, ny, nz, lower=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0),
                filtering="Hou & Li", stepper="impl-diff-rk4", nnu=3, prediss=10.0,
                pretype="roll-mean-max-strain", roll_mean_win_size=1000,
                length_scale="Kolmogorov", alpha=0.1, limit=1.0e30, dt_fixed=0.0,
                f_cor=(0.0, 0.0, 0.0)):
    """Defaults follow src/utils/options.f90:28-91 (nnu=3 and prediss=10 being the
    values it recommends at :52-53)."""
    p = Params()
    p.nx, p.ny, p.nz = int(nx), int(ny), int(nz)
    p.lower[:] = [float(v) for v in lower]
    p.extent[:] = [float(v) for v in extent]
    p.filtering = _lookup(FILTERING, filtering, "filtering")
    p.stepper = _lookup(STEPPER, stepper, "stepper")
    p.nnu = int(nnu)
    p.prediss = float(prediss)
    p.pretype = _lookup(PRETYPE, pretype, "pretype")
    p.roll_mean_win_size = int(roll_mean_win_size)
    p.length_scale = _lookup(LENGTH_SCALE, length_scale, "length_scale")
    p.alpha = float(alpha)
    p.limit = float(limit)
    p.dt_fixed = float(dt_fixed)
    p.f_cor[:] = [float(v) for v in f_cor]
    return p


def beltrami_box(n):
    """The 2 pi x 2 pi x pi box of the Beltrami workloads, lower corner at 0."""
    return dict(nx=n, ny=n, nz=n, lower=(0.0, 0.0, 0.0), extent=(2 * math.pi, 2 * math.pi, math.pi))
