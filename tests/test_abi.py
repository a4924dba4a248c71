"""The C-ABI library builds for gfx950, loads without a GPU and exports every
entry point the headers under include/ declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(ROOT, "include")
    syms = set()
    for name in sorted(os.listdir(inc)):
        if name.endswith(".h"):
            src = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, name)).read(), flags=re.S)
            syms.update(re.findall(r"\b(ps3d_\w+)\s*\(", src))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ps3d_gpu_create", "ps3d_gpu_vor2vel", "ps3d_gpu_source", "ps3d_gpu_adapt",
              "ps3d_gpu_step", "ps3d_gpu_advance_n", "ps3d_gpu_op", "ps3d_gpu_fft1d",
              "ps3d_cnn_forward", "ps3d_nc_create_fields", "ps3d_nc_append_stats"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from ps3d_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_params_struct_matches_header_defaults():
    from ps3d_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = gpu.lib()
    p = gpu.Params()
    lib.ps3d_params_default(ctypes.byref(p))
    # options.f90 defaults: Hou & Li, impl-diff-rk4, roll-mean-max-strain, 1000, Kolmogorov, alpha 0.1
    assert (p.filtering, p.stepper, p.pretype, p.roll_mean_win_size, p.length_scale) == (0, 0, 3, 1000, 0)
    assert p.alpha == 0.1 and p.nnu == 3
    # natural C alignment of ps3d_params on x86-64
    assert ctypes.sizeof(gpu.Params) == 152
    assert gpu.Params.lower.offset == 16 and gpu.Params.prediss.offset == 80 and gpu.Params.f_cor.offset == 128


def test_make_params_rejects_unknown_names():
    from ps3d_amd import make_params
    with pytest.raises(ValueError):
        make_params(8, 8, 8, filtering="spectral")
    with pytest.raises(ValueError):
        make_params(8, 8, 8, pretype="unknown")
