"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference's own
Fortran transform core compiled into oracle/_ref (libps3d_ref_core.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker; the product package never
imports it.

Arrays use ps3d's field layout (src/fields/fields.f90:59-65): numpy shape
(nx, ny, nz+1), C order, i.e. Fortran (0:nz, 0:ny-1, 0:nx-1); vector fields
(3, nx, ny, nz+1).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libps3d_ref_core.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# field / op / diag enums of include/ps3d_gpu.h
SVOR, VOR, VEL, SVEL, SVORTS, PRES, DELTA = range(7)
(OP_FFTXYP2S, OP_FFTXYS2P, OP_DIFFX, OP_DIFFY, OP_FFTSINE, OP_FFTCOSINE,
 OP_CENTRAL_DIFFZ, OP_COMBINE_SEMI, OP_DECOMPOSE_SEMI, OP_COMBINE_PHYS,
 OP_DECOMPOSE_PHYS) = range(11)
NDIAG = 19
DIAG_NAMES = ["t", "dt", "ke", "en", "hel", "vortmax", "vortrms", "vorch",
              "vormean_x", "vormean_y", "vormean_z", "ggmax", "umax", "vmax",
              "wmax", "usggmax", "lsggmax", "rgmax", "prefactor"]


class Params(C.Structure):
    """Mirror of ps3d_params (include/ps3d_gpu.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lower", C.c_double * 3), ("extent", C.c_double * 3),
                ("filtering", C.c_int), ("stepper", C.c_int), ("nnu", C.c_int),
                ("prediss", C.c_double), ("pretype", C.c_int),
                ("roll_mean_win_size", C.c_int), ("length_scale", C.c_int),
                ("alpha", C.c_double), ("limit", C.c_double),
                ("dt_fixed", C.c_double), ("f_cor", C.c_double * 3)]


FILTERING = {"Hou & Li": 0, "2/3-rule": 1}
STEPPER = {"impl-diff-rk4": 0, "cn2": 1}
PRETYPE = {"constant": 0, "vorch": 1, "bfmax": 2, "roll-mean-max-strain": 3,
           "max-strain": 4, "us-max-strain": 5}
LENGTH_SCALE = {"Kolmogorov": 0, "geophysical": 1}


def make_params(nx, ny, nz, lower=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0),
                filtering="Hou & Li", stepper="impl-diff-rk4", nnu=3,
                prediss=10.0, pretype="roll-mean-max-strain",
                roll_mean_win_size=1000, length_scale="Kolmogorov", alpha=0.1,
                limit=1.0e30, dt_fixed=0.0, f_cor=(0.0, 0.0, 0.0)):
    p = Params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.lower[:] = list(lower)
    p.extent[:] = list(extent)
    p.filtering = FILTERING.get(filtering, filtering)
    p.stepper = STEPPER.get(stepper, stepper)
    p.nnu = nnu
    p.prediss = prediss
    p.pretype = PRETYPE.get(pretype, pretype)
    p.roll_mean_win_size = roll_mean_win_size
    p.length_scale = LENGTH_SCALE.get(length_scale, length_scale)
    p.alpha = alpha
    p.limit = limit
    p.dt_fixed = dt_fixed
    p.f_cor[:] = list(f_cor)
    return p


_lib = None


def use_library(path):
    """Load the oracle from another build of ps3d_oracle.c (the mutation tests
    build copies with one planted mistake each) and return the library that was
    in use, so that the caller can restore it with use_library_object()."""
    global _lib
    prev = _lib
    _lib = _bind(C.CDLL(path))
    if _threads is not None:
        _lib.orc_set_threads(_threads)
    return prev


def use_library_object(L):
    global _lib
    _lib = L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = _bind(C.CDLL(_LIB))
    return _lib


def _bind(L):
    for name in ("orc_forfft", "orc_revfft", "orc_dct", "orc_dst"):
        getattr(L, name).argtypes = [C.c_int, C.c_int, _dp]
        getattr(L, name).restype = C.c_int
    L.orc_jacobi_eigenvalues.argtypes = [_dp, _dp]
    L.orc_create.argtypes = [C.POINTER(Params)]
    L.orc_create.restype = C.c_void_p
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_last_error.restype = C.c_char_p
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_op.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.orc_set_vorticity.argtypes = [C.c_void_p, _dp]
    L.orc_vor2vel.argtypes = [C.c_void_p]
    L.orc_source.argtypes = [C.c_void_p]
    L.orc_adapt.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_void_p]
    L.orc_step.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double]
    L.orc_advance_n.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_void_p]
    L.orc_derived.argtypes = [C.c_void_p]
    L.orc_field_stats.argtypes = [C.c_void_p, C.c_void_p, _dp]
    L.orc_rk4_check.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
    L.orc_cn2_check.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
    L.orc_get.argtypes = [C.c_void_p, C.c_int, _dp]
    L.orc_set.argtypes = [C.c_void_p, C.c_int, _dp]
    L.orc_table.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.orc_table.restype = C.c_long
    L.orc_viscosity.argtypes = [C.c_void_p]
    L.orc_viscosity.restype = C.c_double
    return L


def _check(st):
    if st != 0:
        raise RuntimeError("oracle error %d: %s" % (st, lib().orc_last_error().decode()))


_threads = None   # the thread count last set, applied to libraries loaded later


def set_threads(n):
    global _threads
    _threads = int(n)
    lib().orc_set_threads(_threads)


def fft1d(kind, x):
    """stafft transform of x with shape (n_or_n+1, m) viewed as Fortran x(m, n).

    x is given as a numpy array of shape (len, m) in C order, which is the memory
    of Fortran x(m, len). kind: 'forfft', 'revfft', 'dct', 'dst'."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    ln, m = x.shape
    n = ln - 1 if kind == "dct" else ln
    f = {"forfft": lib().orc_forfft, "revfft": lib().orc_revfft,
         "dct": lib().orc_dct, "dst": lib().orc_dst}[kind]
    _check(f(m, n, x))
    return x


def jacobi_eigenvalues(a):
    a = np.asfortranarray(np.array(a, dtype=np.float64)).ravel(order="F").copy()
    d = np.zeros(3)
    lib().orc_jacobi_eigenvalues(a, d)
    return d


class Oracle:
    """CPU restatement of ps3d's solver state and operators."""

    def __init__(self, params):
        self.p = params
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.shape = (self.nx, self.ny, self.nz + 1)
        self.h = lib().orc_create(C.byref(params))
        if not self.h:
            raise RuntimeError("oracle create failed: " + lib().orc_last_error().decode())
        self.t = 0.0

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def op(self, op, fin):
        fin = np.ascontiguousarray(fin, dtype=np.float64)
        out = np.zeros(self.shape)
        _check(lib().orc_op(self.h, op, fin.reshape(self.shape).copy(), out))
        return out

    def set_vorticity(self, vor):
        _check(lib().orc_set_vorticity(self.h, np.ascontiguousarray(vor, dtype=np.float64)))

    def vor2vel(self):
        _check(lib().orc_vor2vel(self.h))

    def source(self):
        _check(lib().orc_source(self.h))

    def adapt(self, t):
        dt = C.c_double()
        diag = np.zeros(NDIAG)
        _check(lib().orc_adapt(self.h, t, C.byref(dt), diag.ctypes.data))
        return dt.value, diag

    def step(self, t, dt):
        tt = C.c_double(t)
        _check(lib().orc_step(self.h, C.byref(tt), dt))
        return tt.value

    def advance_n(self, nsteps):
        tt = C.c_double(self.t)
        diag = np.zeros((nsteps, NDIAG))
        _check(lib().orc_advance_n(self.h, C.byref(tt), nsteps, diag.ctypes.data))
        self.t = tt.value
        return diag

    def derived(self):
        _check(lib().orc_derived(self.h))

    def field_stats(self, diag=None):
        """The 40 field_stats values (NC_KE .. NC_ROMAX) after vor2vel and derived."""
        out = np.zeros(40)
        d = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        _check(lib().orc_field_stats(self.h, None if d is None else d.ctypes.data, out))
        return out

    def rk4_check(self, dt, vdiss, src, svor):
        svor = np.ascontiguousarray(svor, dtype=np.float64).copy()
        _check(lib().orc_rk4_check(self.h, dt, np.ascontiguousarray(vdiss, dtype=np.float64),
                                   np.ascontiguousarray(src, dtype=np.float64), svor))
        return svor

    def cn2_check(self, dt, vdiss, src, svor):
        svor = np.ascontiguousarray(svor, dtype=np.float64).copy()
        _check(lib().orc_cn2_check(self.h, dt, np.ascontiguousarray(vdiss, dtype=np.float64),
                                   np.ascontiguousarray(src, dtype=np.float64), svor))
        return svor

    def get(self, which):
        nv = 1 if which in (PRES, DELTA) else 3
        out = np.zeros((nv,) + self.shape) if nv == 3 else np.zeros(self.shape)
        _check(lib().orc_get(self.h, which, out))
        return out

    def set(self, which, arr):
        _check(lib().orc_set(self.h, which, np.ascontiguousarray(arr, dtype=np.float64)))

    def table(self, which):
        n = lib().orc_table(self.h, which, None)
        out = np.zeros(n)
        lib().orc_table(self.h, which, out.ctypes.data)
        return out

    def viscosity(self):
        return lib().orc_viscosity(self.h)


# --------------------------------------------------------------------------
# genspec's spectrum and nc_reader's derived means (numpy)
# --------------------------------------------------------------------------
def kinetic_energy(vel):
    """get_kinetic_energy (src/fields/field_diagnostics.f90:85-122): half the sum
    of |u|^2 with the lid levels weighted 1/2, over nx ny nz cells."""
    e = np.sum(vel ** 2, axis=0)                     # (nx, ny, nz+1)
    nx, ny, nz1 = e.shape
    tot = 0.5 * e[:, :, 1:-1].sum() + 0.25 * (e[:, :, 0].sum() + e[:, :, -1].sum())
    return tot / (nx * ny * (nz1 - 1))


def moments(vel, vor):
    """Kinetic energy, and the means of the helicity u.w and of the cross-helicity
    magnitude |u x w| (plotting/nc_reader.py:94-120) with the lids weighted 1/2,
    over nx ny nz cells; and the largest |u x w|."""
    nx, ny, nz1 = vel.shape[1:]
    w = np.ones(nz1)
    w[0] = w[-1] = 0.5
    hel = np.sum(vel * vor, axis=0)
    cr = np.linalg.norm(np.cross(vel, vor, axis=0), axis=0)
    n = nx * ny * (nz1 - 1)
    return {"ke": kinetic_energy(vel), "helicity": float((hel * w).sum() / n),
            "cross_helicity": float((cr * w).sum() / n), "cross_helicity_max": float(cr.max())}


def _nint(x):
    """Fortran nint for x >= 0: halves round away from zero (numpy's rint rounds
    them to even)."""
    f = np.floor(x)
    return np.where(x - f >= 0.5, f + 1.0, f)


def spectrum(o, vel):
    """genspec (src/genspec.f90:44-124) with the oracle's transforms, on the grid of
    Oracle o: (shell wavenumbers m dk, spectrum, point counts), m = 0..kmax."""
    nx, ny, nz = o.nx, o.ny, o.nz
    ke = kinetic_energy(vel)                                          # :46
    s = [o.op(OP_FFTXYP2S, vel[c]) for c in range(3)]                 # :49-52
    s[0] = o.op(OP_FFTCOSINE, s[0])                                   # :55 dct(u)
    s[1] = o.op(OP_FFTCOSINE, s[1])                                   # :56 dct(v)
    s[2] = o.op(OP_FFTSINE, s[2])                                     # :57 dst(w(1:nz))
    rkx, rky, rkz = o.table(13), o.table(14), o.table(15)
    k2 = (rkx[:, None, None] ** 2 + rky[None, :, None] ** 2) + rkz[None, None, :] ** 2
    kmag = _nint(np.sqrt(k2)).astype(np.int64)                        # :62-68
    kmax = int(kmag.max())                                            # :70
    dk = kmax / np.sqrt((0.5 * nx) ** 2 + (0.5 * ny) ** 2 + float(nz) ** 2)   # :83
    dki = 1.0 / dk
    m = (kmag * dki).astype(np.int64)                                 # :93
    e = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    keep = m <= kmax       # the reference's spec(0:kmax) has no room for m > kmax
    spec = np.bincount(m[keep], weights=e[keep], minlength=kmax + 1)[:kmax + 1]
    num = np.bincount(m[keep], minlength=kmax + 1)[:kmax + 1]
    pref = 4.0 / 3.0 * np.pi * (dk * dk * dk)                         # :104
    mm = np.arange(kmax + 1, dtype=np.int64)
    ok = num > 0
    spec = np.where(ok, spec * pref * ((mm + 1) ** 3 - mm ** 3).astype(np.float64) / np.maximum(num, 1), 0.0)
    spec = spec * (ke / np.sum(spec * dk))                            # :112-117
    return mm * dk, spec, num


# --------------------------------------------------------------------------
# The reference's own compiled transform core (oracle/_ref)
# --------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(_REF)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(_REF)
    return _ref


def _ip(v):
    return C.byref(C.c_int(v))


def ref_fft1d(kind, x):
    """Same contract as fft1d() but through the reference's stafft.f90."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    ln, m = x.shape
    n = ln - 1 if kind == "dct" else ln
    trig = np.zeros(2 * n)
    fac = np.zeros(5, dtype=np.int32)
    R = ref()
    R._QMstafftPinitfft(_ip(n), fac.ctypes.data_as(C.c_void_p), trig.ctypes.data_as(C.c_void_p))
    f = {"forfft": R._QMstafftPforfft, "revfft": R._QMstafftPrevfft,
         "dct": R._QMstafftPdct, "dst": R._QMstafftPdst}[kind]
    f(_ip(m), _ip(n), x.ctypes.data_as(C.c_void_p), trig.ctypes.data_as(C.c_void_p),
      fac.ctypes.data_as(C.c_void_p))
    return x


def ref_jacobi_eigenvalues(a):
    a = np.asfortranarray(np.array(a, dtype=np.float64)).ravel(order="F").copy()
    d = np.zeros(3)
    ref()._QMjacobiPjacobi_eigenvalues(a.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p))
    return d


def ref_init_deriv(n, length):
    k = np.zeros(n)
    ref()._QMderiv1dPinit_deriv(_ip(n), C.byref(C.c_double(length)), k.ctypes.data_as(C.c_void_p))
    return k
