"""Host side of the MI355X pseudo-spectral step: ctypes binding of the C ABI
(include/ps3d_gpu.h) with the reference's operator names.

The entry points mirror ps3d's module procedures (src/inversion/inversion.f90,
src/stepper/advance.f90, src/stepper/impl_rk4.f90, src/fft/sta3dfft.f90):
vor2vel, source, adapt, step, advance, fftxyp2s, diffx, ...  Arrays use ps3d's
field layout: numpy (nx, ny, nz+1) C order = Fortran (0:nz, 0:ny-1, 0:nx-1),
vector fields (3, nx, ny, nz+1).  Errors raise Ps3dError carrying the C status
and message, in place of the reference's mpi_stop.
"""
import ctypes as C
import os

import numpy as np

from .params import Params, make_params  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libps3d_gpu.so")

SVOR, VOR, VEL, SVEL, SVORTS, PRES, DELTA = range(7)
(OP_FFTXYP2S, OP_FFTXYS2P, OP_DIFFX, OP_DIFFY, OP_FFTSINE, OP_FFTCOSINE,
 OP_CENTRAL_DIFFZ, OP_COMBINE_SEMI, OP_DECOMPOSE_SEMI, OP_COMBINE_PHYS,
 OP_DECOMPOSE_PHYS) = range(11)
NDIAG = 19
DIAG_NAMES = ["t", "dt", "ke", "en", "hel", "vortmax", "vortrms", "vorch",
              "vormean_x", "vormean_y", "vormean_z", "ggmax", "umax", "vmax",
              "wmax", "usggmax", "lsggmax", "rgmax", "prefactor"]

# ps3d's field_stats variables in the order of NC_KE .. NC_ROMAX
# (src/fields/field_diagnostics_netcdf.f90:36-75), enum ps3d_stat
STAT_NAMES = ["ke", "en", "vortmax", "vortrms", "vorch", "x_vormean", "y_vormean", "z_vormean",
              "kexy", "kez", "enxy", "enz", "x_vormin", "y_vormin", "z_vormin",
              "x_vormax", "y_vormax", "z_vormax", "enxy_max", "gmax", "bfmax", "umax", "vmax", "wmax",
              "uoxmax", "loxmax", "uoymax", "loymax", "uozmax", "lozmax", "usuhmax", "usgmax", "lsgmax",
              "uszrms", "usdeltarms", "rolling_mean_gmax", "rolling_mean_bfmax", "ri_min", "ro_min", "ro_max"]
NSTAT = len(STAT_NAMES)
# enum ps3d_moment
MOMENT_NAMES = ["ke", "helicity", "cross_helicity", "cross_helicity_max"]

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Ps3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ps3d_gpu error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load libps3d_gpu.so; fails loudly when the HIP library is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library %s is not built: run `python -m ps3d_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.ps3d_gpu_last_error.restype = C.c_char_p
        L.ps3d_gpu_version.restype = C.c_char_p
        L.ps3d_gpu_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.ps3d_gpu_destroy.argtypes = [C.c_void_p]
        L.ps3d_gpu_set_vorticity.argtypes = [C.c_void_p, _dp]
        for n in ("ps3d_gpu_vor2vel", "ps3d_gpu_source", "ps3d_gpu_sync", "ps3d_gpu_derived"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.ps3d_gpu_adapt.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_void_p]
        L.ps3d_gpu_step.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double]
        L.ps3d_gpu_advance_n.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_void_p]
        L.ps3d_gpu_advance_until.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_int, C.c_void_p,
                                             C.POINTER(C.c_int)]
        L.ps3d_gpu_get.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ps3d_gpu_set.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ps3d_gpu_op.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.ps3d_gpu_fft1d.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.ps3d_gpu_rk_substep.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
        L.ps3d_gpu_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.ps3d_gpu_profile_kernels.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ps3d_gpu_resample.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.ps3d_gpu_viscosity.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ps3d_gpu_field_stats.argtypes = [C.c_void_p, C.c_void_p, _dp]
        L.ps3d_gpu_moments.argtypes = [C.c_void_p, _dp]
        L.ps3d_gpu_spectrum.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        L.ps3d_gpu_create_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.ps3d_gpu_batch_size.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.ps3d_gpu_set_vorticity_batch.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ps3d_gpu_get_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.ps3d_gpu_advance_n_batch.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p]
        L.ps3d_gpu_ensemble_create.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.ps3d_gpu_ensemble_group.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.ps3d_gpu_ensemble_advance.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.ps3d_gpu_ensemble_tune.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ps3d_gpu_ensemble_set_choice.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ps3d_gpu_ensemble_line_choice.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ps3d_gpu_ensemble_fft_schedule.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ps3d_gpu_ensemble_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise Ps3dError(st, lib().ps3d_gpu_last_error().decode())


FFT1D_KIND = {"forfft": 0, "revfft": 1, "dct": 2, "dst": 3}


def resample(dst, src, which=SVOR):
    """Spectral truncation / zero padding of src's svor (or svorts) onto dst's grid
    (same box); with SVOR, dst is then set up as by set_vorticity."""
    _check(lib().ps3d_gpu_resample(dst.h, src.h, int(which)))
    if which == SVOR:
        dst.t = 0.0


def fft1d(kind, x):
    """stafft transform (src/fft/stafft.f90) of Fortran x(m, n) given as a numpy
    array of shape (n, m) (dct: (n+1, m)) in C order."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    ln, m = x.shape
    n = ln - 1 if kind == "dct" else ln
    _check(lib().ps3d_gpu_fft1d(FFT1D_KIND[kind], m, n, x))
    return x


class Ps3dGpu:
    """One grid resident on the current HIP device (a ps3d_handle)."""

    def __init__(self, params):
        if not isinstance(params, Params):
            params = _convert(params)
        self.p = params
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.shape = (self.nx, self.ny, self.nz + 1)
        self.h = C.c_void_p()
        _check(lib().ps3d_gpu_create(C.byref(params), C.byref(self.h)))
        self.t = 0.0

    def close(self):
        if self.h:
            lib().ps3d_gpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- setup (src/utils/utils.f90:136-184) ---
    def set_vorticity(self, vor):
        vor = np.ascontiguousarray(vor, dtype=np.float64)
        if vor.shape != (3,) + self.shape:
            raise ValueError("vorticity must have shape %s" % ((3,) + self.shape,))
        _check(lib().ps3d_gpu_set_vorticity(self.h, vor))
        self.t = 0.0

    # --- operators ---
    def vor2vel(self):
        _check(lib().ps3d_gpu_vor2vel(self.h))

    def derived(self):
        """pressure + horizontal_divergence (fields_derived.f90:67-182) from the last
        vor2vel; read the results with get(PRES) / get(DELTA)."""
        _check(lib().ps3d_gpu_derived(self.h))

    def source(self):
        _check(lib().ps3d_gpu_source(self.h))

    def field_stats(self, diag=None):
        """ps3d's field_stats record (STAT_NAMES order) of the state of the last
        vor2vel; diag: the adapt record of the same time (adapt() or the step's
        row of advance_n), whose entries ps3d's adapt hands to the stats."""
        out = np.zeros(NSTAT)
        d = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        _check(lib().ps3d_gpu_field_stats(self.h, None if d is None else d.ctypes.data, out))
        return out

    def moments(self):
        """{ke, helicity, cross_helicity, cross_helicity_max} of the state of the
        last vor2vel: domain means of |u|^2/2, u.w and |u x w|, and the largest |u x w|."""
        out = np.zeros(len(MOMENT_NAMES))
        _check(lib().ps3d_gpu_moments(self.h, out))
        return dict(zip(MOMENT_NAMES, out.tolist()))

    def spectrum(self, vel=None):
        """genspec's spectrum: (shell wavenumbers m dk, spectrum, point counts), of the
        velocity of the last vor2vel or of vel (3 volumes), which then replaces the
        handle's physical velocity."""
        ns, dk = C.c_int(), C.c_double()
        _check(lib().ps3d_gpu_spectrum(self.h, None, 0, C.byref(ns), C.byref(dk), None, None))
        spec = np.zeros(ns.value)
        num = np.zeros(ns.value, dtype=np.int64)
        v = None
        if vel is not None:
            v = np.ascontiguousarray(vel, dtype=np.float64)
            if v.shape != (3, self.nx, self.ny, self.nz + 1):
                raise ValueError("vel must have shape (3, nx, ny, nz + 1)")
        _check(lib().ps3d_gpu_spectrum(self.h, None if v is None else v.ctypes.data, ns.value, C.byref(ns),
                                       C.byref(dk), spec.ctypes.data, num.ctypes.data))
        return np.arange(ns.value) * dk.value, spec, num

    def viscosity(self):
        """(visc, prefactor of the last adapt): column (k, l) decays at the rate
        pref * visc * (k^2 + l^2)^nnu."""
        v, p = C.c_double(), C.c_double()
        _check(lib().ps3d_gpu_viscosity(self.h, C.byref(v), C.byref(p)))
        return v.value, p.value

    def adapt(self, t):
        dt = C.c_double()
        diag = np.zeros(NDIAG)
        _check(lib().ps3d_gpu_adapt(self.h, t, C.byref(dt), diag.ctypes.data))
        return dt.value, diag

    def step(self, t, dt):
        tt = C.c_double(t)
        _check(lib().ps3d_gpu_step(self.h, C.byref(tt), dt))
        return tt.value

    def advance(self):
        """One ps3d advance() (src/stepper/advance.f90:77-104); returns its diag row."""
        return self.advance_n(1)[0]

    def advance_n(self, nsteps):
        tt = C.c_double(self.t)
        diag = np.zeros((max(1, nsteps), NDIAG))
        _check(lib().ps3d_gpu_advance_n(self.h, C.byref(tt), int(nsteps), diag.ctypes.data))
        self.t = tt.value
        return diag[:nsteps]

    def advance_until(self, t_stop, nmax):
        """ps3d's `do while (t < limit)` loop (src/ps3d.f90:114-116) run on the
        device up to nmax steps, stopping before the first step that would start
        at t >= min(t_stop, limit); returns the records of the steps taken."""
        tt = C.c_double(self.t)
        n = C.c_int()
        diag = np.zeros((max(1, nmax), NDIAG))
        _check(lib().ps3d_gpu_advance_until(self.h, C.byref(tt), float(t_stop), int(nmax), diag.ctypes.data,
                                            C.byref(n)))
        self.t = tt.value
        return diag[:n.value]

    def sync(self):
        _check(lib().ps3d_gpu_sync(self.h))

    def last_timing(self):
        ms = C.c_double()
        n = C.c_int()
        _check(lib().ps3d_gpu_last_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def get(self, which):
        nv = 1 if which in (PRES, DELTA) else 3
        out = np.zeros((3,) + self.shape) if nv == 3 else np.zeros(self.shape)
        _check(lib().ps3d_gpu_get(self.h, which, out))
        return out

    def set(self, which, arr):
        _check(lib().ps3d_gpu_set(self.h, which, np.ascontiguousarray(arr, dtype=np.float64)))

    def op(self, op, fin):
        fin = np.ascontiguousarray(fin, dtype=np.float64).reshape(self.shape)
        out = np.zeros(self.shape)
        _check(lib().ps3d_gpu_op(self.h, op, fin, out))
        return out

    def rk_substep(self, stage, dt, vdiss):
        _check(lib().ps3d_gpu_rk_substep(self.h, int(stage), float(dt),
                                         np.ascontiguousarray(vdiss, dtype=np.float64)))

    # reference operator names (src/fft/sta3dfft.f90, src/inversion/inversion_utils.f90)
    def fftxyp2s(self, fp):
        return self.op(OP_FFTXYP2S, fp)

    def fftxys2p(self, fs):
        return self.op(OP_FFTXYS2P, fs)

    def diffx(self, fs):
        return self.op(OP_DIFFX, fs)

    def diffy(self, fs):
        return self.op(OP_DIFFY, fs)

    def fftsine(self, fs):
        return self.op(OP_FFTSINE, fs)

    def fftcosine(self, fs):
        return self.op(OP_FFTCOSINE, fs)

    def central_diffz(self, fs):
        return self.op(OP_CENTRAL_DIFFZ, fs)

    def field_combine_semi_spectral(self, f):
        return self.op(OP_COMBINE_SEMI, f)

    def field_decompose_semi_spectral(self, f):
        return self.op(OP_DECOMPOSE_SEMI, f)

    def field_combine_physical(self, f):
        return self.op(OP_COMBINE_PHYS, f)

    def field_decompose_physical(self, f):
        return self.op(OP_DECOMPOSE_PHYS, f)


class Ps3dBatch:
    """nbatch instances of one grid stepped by the same kernel launches (a batched
    ps3d_handle).  Owned (closed) unless it belongs to an ensemble."""

    def __init__(self, params=None, nbatch=1, _handle=None):
        if _handle is not None:
            self.h, self.owned = _handle, False
            nb = C.c_int()
            _check(lib().ps3d_gpu_batch_size(self.h, C.byref(nb)))
            self.nbatch = nb.value
        else:
            if not isinstance(params, Params):
                params = _convert(params)
            self.h, self.owned = C.c_void_p(), True
            _check(lib().ps3d_gpu_create_batch(C.byref(params), int(nbatch), C.byref(self.h)))
            self.nbatch = int(nbatch)
            self.set_shape(params.nx, params.ny, params.nz)
        self.p = params
        self.t = np.zeros(self.nbatch)

    def set_shape(self, nx, ny, nz):
        self.shape = (nx, ny, nz + 1)

    def set_vorticity(self, b, vor):
        vor = np.ascontiguousarray(vor, dtype=np.float64)
        _check(lib().ps3d_gpu_set_vorticity_batch(self.h, int(b), vor))
        self.t[b] = 0.0

    def get_svor(self, b):
        out = np.zeros((3,) + self.shape)
        _check(lib().ps3d_gpu_get_batch(self.h, int(b), SVOR, out))
        return out

    def advance_n(self, nsteps, series=False):
        diag = np.zeros((self.nbatch, max(1, nsteps), NDIAG)) if series else None
        tt = np.ascontiguousarray(self.t, dtype=np.float64)
        _check(lib().ps3d_gpu_advance_n_batch(self.h, tt, int(nsteps),
                                              diag.ctypes.data if series else None))
        self.t = tt
        return diag[:, :nsteps] if series else None

    def close(self):
        if self.owned and self.h:
            lib().ps3d_gpu_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ps3dEnsemble:
    """Several grid sizes, each a batched group, stepped concurrently; one step of
    every instance can replay as one captured HIP graph."""

    def __init__(self, params_list, counts):
        n = len(params_list)
        arr = (Params * n)()
        for i, p in enumerate(params_list):
            arr[i] = p if isinstance(p, Params) else _convert(p)
        cnt = (C.c_int * n)(*[int(c) for c in counts])
        self.e = C.c_void_p()
        _check(lib().ps3d_gpu_ensemble_create(n, arr, cnt, C.byref(self.e)))
        self.groups = []
        for g in range(n):
            h = C.c_void_p()
            _check(lib().ps3d_gpu_ensemble_group(self.e, g, C.byref(h)))
            b = Ps3dBatch(_handle=h)
            b.p = arr[g]
            b.set_shape(arr[g].nx, arr[g].ny, arr[g].nz)
            self.groups.append(b)

    def advance(self, nsteps, graph=True):
        ms = C.c_double()
        _check(lib().ps3d_gpu_ensemble_advance(self.e, int(nsteps), 1 if graph else 0, C.byref(ms)))
        return ms.value

    def tune(self, steps=2):
        """z launch shape, then line-pass block sizes, per group; returns the z shape
        codes (line_choices() gives the block sizes)."""
        ch = (C.c_int * len(self.groups))()
        _check(lib().ps3d_gpu_ensemble_tune(self.e, int(steps), ch))
        return list(ch)

    def line_choices(self):
        """(x products, adapt-stage y, x adapt) threads per block of each group."""
        out = []
        for g in range(len(self.groups)):
            t = (C.c_int * 3)()
            _check(lib().ps3d_gpu_ensemble_line_choice(self.e, g, t))
            out.append(tuple(t))
        return out

    def fft_schedules(self):
        """(x lines, y lines) FFT schedule of each group's transform-only passes as
        names: radix-16 or radix-8 first plan, Hermitian pair folded into the first
        pass or formed in a pass of its own, +t512 for 512 threads per tile."""
        out = []
        for g in range(len(self.groups)):
            t = (C.c_int * 2)()
            _check(lib().ps3d_gpu_ensemble_fft_schedule(self.e, g, t))
            out.append(tuple(("r8" if v & 1 else "r16") + ("+fold" if v & 2 else "+pairpass") + ("+t512" if v & 4 else "")
                             for v in t))
        return out

    def set_choice(self, g, code):
        _check(lib().ps3d_gpu_ensemble_set_choice(self.e, int(g), int(code)))

    def close(self):
        if self.e:
            for b in self.groups:
                b.h = C.c_void_p()
            lib().ps3d_gpu_ensemble_destroy(self.e)
            self.e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _convert(p):
    """Copy a ps3d_params-like ctypes structure (same fields) into Params."""
    q = Params()
    for name, _ in Params._fields_:
        v = getattr(p, name)
        if name in ("lower", "extent", "f_cor"):
            getattr(q, name)[:] = list(v)
        else:
            setattr(q, name, v)
    return q
