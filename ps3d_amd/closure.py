"""Learned subgrid closure for a coarse grid: the 3x3x3 convolution net of
ps3d_closure.h (bf16 matrix cores), its training data from filtered fine-grid
runs, training, and the genspec-style energy spectrum used to evaluate it.

Training target: tau = R(S_f) - S_c(R(omega_f)), where S is ps3d's vorticity
tendency (src/inversion/inversion.f90:298-388), R the sharp spectral truncation
of a fine grid onto the coarse one (ps3d_gpu_resample) and omega_f a fine-grid
state, optionally restricted to the modes the coarse step's 2/3-rule mask passes
(the experiment driver does so); input: the coarse u, v, w, xi, eta, zeta.  Both
are physical fields on the coarse points, scaled to unit rms per group (velocity,
vorticity, tendency).
"""
import ctypes as C
import math

import numpy as np

from . import gpu as G
from .gpu import _check as _gcheck
from .gpu import lib as _glib

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _glib()
        L.ps3d_cnn_last_error.restype = C.c_char_p
        L.ps3d_cnn_nparams.argtypes = [C.POINTER(C.c_int)]
        L.ps3d_cnn_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.ps3d_cnn_destroy.argtypes = [C.c_void_p]
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.ps3d_cnn_set_weights.argtypes = [C.c_void_p, fp]
        L.ps3d_cnn_get_weights.argtypes = [C.c_void_p, fp]
        L.ps3d_cnn_forward.argtypes = [C.c_void_p, fp, fp]
        L.ps3d_cnn_grad.argtypes = [C.c_void_p, fp, fp, C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.ps3d_cnn_train_step.argtypes = [C.c_void_p, fp, fp, C.c_float, C.POINTER(C.c_float)]
        L.ps3d_cnn_dataset_reserve.argtypes = [C.c_void_p, C.c_int]
        L.ps3d_cnn_dataset_set.argtypes = [C.c_void_p, C.c_int, fp, fp]
        L.ps3d_cnn_train_epoch.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                           C.c_int, C.c_float, C.POINTER(C.c_float)]
        L.ps3d_cnn_timing.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.ps3d_gpu_set_closure.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise G.Ps3dError(st, lib().ps3d_cnn_last_error().decode())


LAYERS = [(6, 32), (32, 32), (32, 32), (32, 3)]
GFLOP_FWD_PER_POINT = 2.0 * 27 * sum(ci * co for ci, co in LAYERS) / 1e9   # algorithmic (unpadded)


def nparams():
    n = C.c_int()
    _check(lib().ps3d_cnn_nparams(C.byref(n)))
    return n.value


def init_weights(seed=0):
    """Same initialisation as the oracle: W ~ U(-s, s), s = sqrt(6/(27 cin)), b = 0
    (the last layer is scaled by 0.1 so the untrained closure is small)."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (cin, cout) in enumerate(LAYERS):
        s = math.sqrt(6.0 / (27 * cin)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        out.append(rng.uniform(-s, s, size=27 * cout * cin))
        out.append(np.zeros(cout))
    return np.concatenate(out).astype(np.float32)


class ClosureNet:
    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.npts = nx * ny * (nz + 1)
        self.h = C.c_void_p()
        _check(lib().ps3d_cnn_create(nx, ny, nz, C.byref(self.h)))
        self.in_scale = np.ones(6, dtype=np.float32)
        self.out_scale = np.ones(3, dtype=np.float32)

    def close(self):
        if self.h:
            lib().ps3d_cnn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w):
        _check(lib().ps3d_cnn_set_weights(self.h, np.ascontiguousarray(w, dtype=np.float32)))

    def get_weights(self):
        w = np.zeros(nparams(), dtype=np.float32)
        _check(lib().ps3d_cnn_get_weights(self.h, w))
        return w

    def forward(self, x):
        y = np.zeros((self.npts, 3), dtype=np.float32)
        _check(lib().ps3d_cnn_forward(self.h, np.ascontiguousarray(x, dtype=np.float32).reshape(self.npts, 6), y))
        return y

    def grad(self, x, t, want_gx=False):
        loss = C.c_float()
        g = np.zeros(nparams(), dtype=np.float32)
        gx = np.zeros((self.npts, 6), dtype=np.float32) if want_gx else None
        _check(lib().ps3d_cnn_grad(self.h, np.ascontiguousarray(x, dtype=np.float32).reshape(self.npts, 6),
                                   np.ascontiguousarray(t, dtype=np.float32).reshape(self.npts, 3), C.byref(loss),
                                   g.ctypes.data, gx.ctypes.data if want_gx else None))
        return loss.value, g, gx

    def train_step(self, x, t, lr):
        loss = C.c_float()
        _check(lib().ps3d_cnn_train_step(self.h, np.ascontiguousarray(x, dtype=np.float32).reshape(self.npts, 6),
                                         np.ascontiguousarray(t, dtype=np.float32).reshape(self.npts, 3),
                                         float(lr), C.byref(loss)))
        return loss.value

    def reserve(self, n):
        _check(lib().ps3d_cnn_dataset_reserve(self.h, int(n)))

    def set_sample(self, i, x, t):
        _check(lib().ps3d_cnn_dataset_set(self.h, int(i), np.ascontiguousarray(x, dtype=np.float32).reshape(self.npts, 6),
                                          np.ascontiguousarray(t, dtype=np.float32).reshape(self.npts, 3)))

    def train_epoch(self, order, lr):
        loss = C.c_float()
        order = np.ascontiguousarray(order, dtype=np.int32)
        _check(lib().ps3d_cnn_train_epoch(self.h, order, len(order), float(lr), C.byref(loss)))
        return loss.value

    def timing(self, reps=20):
        f, t = C.c_float(), C.c_float()
        _check(lib().ps3d_cnn_timing(self.h, int(reps), C.byref(f), C.byref(t)))
        return f.value, t.value

    def attach(self, solver):
        """Add this net's tendency correction to every RK stage of solver.advance_n."""
        ins = np.ascontiguousarray(self.in_scale, dtype=np.float32)
        outs = np.ascontiguousarray(self.out_scale, dtype=np.float32)
        self._keep = (ins, outs)
        _gcheck(lib().ps3d_gpu_set_closure(solver.h, self.h, ins.ctypes.data, outs.ctypes.data))

    @staticmethod
    def detach(solver):
        _gcheck(lib().ps3d_gpu_set_closure(solver.h, None, None, None))


# --------------------------------------------------------------------------
# data: channels-last point arrays from solver fields
# --------------------------------------------------------------------------
def to_points(f):
    """(nc, nx, ny, nz+1) solver-layout components -> (npts, nc), p = (ix*ny+iy)*(nz+1)+iz."""
    return np.ascontiguousarray(np.moveaxis(f, 0, -1).reshape(-1, f.shape[0]))


def hyperdiffusion_rate(solver):
    """Decay rate pref * visc * (k^2 + l^2)^nnu of every column (nx, ny) of solver
    (impl_rk4.f90:37-53 integrating factors; visc and pref of its last set-up and
    adapt)."""
    visc, pref = solver.viscosity()
    nx, ny = solver.nx, solver.ny
    k = 2 * np.pi / solver.p.extent[0] * np.minimum(np.arange(nx), nx - np.arange(nx))
    l = 2 * np.pi / solver.p.extent[1] * np.minimum(np.arange(ny), ny - np.arange(ny))
    return pref * visc * (k[:, None] ** 2 + l[None, :] ** 2) ** solver.p.nnu


def dealias_mask(solver):
    """The 2/3-rule mask the coarse step multiplies every tendency by (filx fily
    filz, the lids taking filx fily), on the host layout (nx, ny, nz+1)."""
    nx, ny, nz = solver.nx, solver.ny, solver.nz
    kx = np.minimum(np.arange(nx), nx - np.arange(nx)) * (2 * np.pi / solver.p.extent[0])
    ky = np.minimum(np.arange(ny), ny - np.arange(ny)) * (2 * np.pi / solver.p.extent[1])
    kz = np.arange(nz + 1) * (np.pi / solver.p.extent[2])
    fx = (kx <= 2.0 / 3.0 * kx.max()).astype(float)
    fy = (ky <= 2.0 / 3.0 * ky.max()).astype(float)
    fz = (kz <= 2.0 / 3.0 * kz.max()).astype(float)
    fz[0] = fz[-1] = 1.0
    return fx[:, None, None] * fy[None, :, None] * fz[None, None, :]


def sample_from_fine(fine, coarse, tcoarse, diffusion=False, dealias=False):
    """One training pair at the current state of `fine` (its vor2vel and source
    run here): input = coarse (u, v, w, omega) of R(omega_f); target = R(S_f) -
    S_c(R(omega_f)) in physical space on the coarse points.  diffusion=True adds
    the difference of the two grids' hyperdiffusion at the coarse wavenumbers,
    (rate_c - rate_f) R(omega_f), so the target is the whole missing tendency.
    dealias=True keeps only the modes the coarse step passes (its 2/3-rule mask)."""
    fine.vor2vel()
    if diffusion:
        fine.adapt(fine.t)
    fine.source()
    G.resample(coarse, fine)                 # coarse state = truncated vorticity (set up, vor2vel)
    vel = coarse.get(G.VEL)
    vor = coarse.get(G.VOR)
    if diffusion:
        coarse.adapt(fine.t)
    coarse.source()
    sc = coarse.get(G.SVORTS)
    G.resample(tcoarse, fine, G.SVORTS)      # R(S_f)
    sf = tcoarse.get(G.SVORTS)
    d = sf - sc
    if diffusion:
        rate_f = hyperdiffusion_rate(fine)[np.ix_(_fold(fine.nx, coarse.nx), _fold(fine.ny, coarse.ny))]
        rate = hyperdiffusion_rate(coarse) - rate_f
        d = d + rate[None, :, :, None] * coarse.get(G.SVOR)
    if dealias:
        d = d * dealias_mask(coarse)[None]
    tau = np.stack([coarse.op(G.OP_COMBINE_PHYS, d[c]) for c in range(3)])
    x = to_points(np.concatenate([vel, vor]))
    return x.astype(np.float32), to_points(tau).astype(np.float32)


def _fold(nf, nc):
    """Fine-grid slot of each coarse slot with the same wavenumber magnitude."""
    i = np.arange(nc)
    return np.where(i <= nc // 2, i, nf - (nc - i))


# --------------------------------------------------------------------------
# genspec energy spectrum (src/genspec.f90:50-124)
# --------------------------------------------------------------------------
def energy_spectrum(solver):
    """genspec's spectrum (src/genspec.f90:44-124) of the solver's current physical
    velocity (after vor2vel), binned on the GPU: (shell wavenumbers m dk, spectrum),
    normalised so that sum(spec) dk is the kinetic energy."""
    k, spec, _ = solver.spectrum()
    return k, spec


def spectrum_error(spec, ref, kcut=None):
    """Relative L2 difference of two spectra over the shells below kcut."""
    n = min(len(spec), len(ref)) if kcut is None else kcut
    return float(np.linalg.norm(spec[:n] - ref[:n]) / np.linalg.norm(ref[:n]))
