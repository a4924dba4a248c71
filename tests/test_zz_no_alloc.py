"""ps3d_gpu_field_stats_noalloc against ps3d_gpu_field_stats through the C ABI.

The allocation-free entry point writes the field_stats record into device memory
the caller owns (allocated here through the HIP runtime the library itself uses)
and keeps its per-block partials in the library's own per-GPU array.  Two calls in a row, the second on a smaller grid,
must each give exactly the record of the allocating entry point (same kernels,
same partition of the columns over the blocks)."""
import ctypes as C
import math

import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = math.pi


def _hip():
    """The HIP runtime the library is linked against (already loaded with it)."""
    from ps3d_amd import gpu as G
    G.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DeviceBuffer:
    def __init__(self, n):
        self.hip = _hip()
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def fill(self, value):
        h = np.full(self.n, value)
        assert self.hip.hipMemcpy(self.p, h.ctypes.data, 8 * self.n, 1) == 0     # host to device

    def get(self):
        h = np.empty(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data, self.p, 8 * self.n, 2) == 0     # device to host
        return h

    def free(self):
        self.hip.hipFree(self.p)


def _state(G, n, seed):
    d = G.Ps3dGpu(G.make_params(n, n, n, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI),
                                filtering="2/3-rule", nnu=3, prediss=30.0, pretype="vorch",
                                f_cor=(0.0, 0.0, 0.8)))
    d.set_vorticity(ics.beltrami_multimode(n, n, n, seed=seed))
    d.advance_n(2)
    d.vor2vel()
    _, diag = d.adapt(d.t)
    d.derived()
    return d, diag


def test_noalloc_field_stats_matches_the_allocating_entry_point():
    from ps3d_amd import gpu as G
    L = G.lib()
    L.ps3d_gpu_field_stats_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    out = DeviceBuffer(G.NSTAT)
    out.fill(np.nan)
    for n, seed in ((48, 2), (24, 5)):          # the second call on a smaller grid
        print("noalloc field stats", n, flush=True)
        d, diag = _state(G, n, seed)
        dg = np.ascontiguousarray(diag)
        st = L.ps3d_gpu_field_stats_noalloc(d.h, dg.ctypes.data, out.p)
        assert st == 0, L.ps3d_gpu_last_error().decode()
        d.sync()
        got = out.get()
        ref = d.field_stats(diag)
        assert np.array_equal(got, ref), (n, got, ref)
        d.close()
    out.free()


def test_noalloc_field_stats_refuses_without_derived():
    from ps3d_amd import gpu as G
    L = G.lib()
    L.ps3d_gpu_field_stats_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    out = DeviceBuffer(G.NSTAT)
    d, _ = _state(G, 16, 1)
    d.vor2vel()                                  # a new state: its divergence is not formed yet
    st = L.ps3d_gpu_field_stats_noalloc(d.h, None, out.p)
    assert st == 5
    d.close()
    out.free()
