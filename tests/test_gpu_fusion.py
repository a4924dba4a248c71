"""The RK4 tendency pass of stages 1-3 fused with the next stage's vor2vel z pass
(k_ztend_zinv_w: the stage result never leaves the registers) against the two
passes launched apart (PS3D_ZFUSE=0, read when a handle is created), on every
wave-kernel depth."""
import math
import os

import numpy as np
import pytest

from ps3d_amd import ics

pytestmark = pytest.mark.gpu

PI = math.pi
BOX = dict(lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI))


def run(G, shape, env, nsteps=5, seed=6):
    nx, ny, nz = shape
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        d = G.Ps3dGpu(G.make_params(nx, ny, nz, **BOX, filtering="2/3-rule", nnu=3, prediss=30.0,
                                    pretype="vorch", alpha=0.1))
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    d.set_vorticity(ics.beltrami_multimode(nx, ny, nz, seed=seed))
    rows = d.advance_n(nsteps)
    s = d.get(G.SVOR)
    d.close()
    return rows, s


@pytest.mark.parametrize("shape", [(32, 32, 64), (32, 32, 128), (24, 16, 256), (16, 16, 512)])
def test_fused_tendency_and_inversion_match_separate_passes(shape):
    from ps3d_amd import gpu as G
    print("fusion", shape, flush=True)
    ra, sa = run(G, shape, {"PS3D_ZFUSE": "1", "PS3D_ZWAVE": "2"})
    rb, sb = run(G, shape, {"PS3D_ZFUSE": "0", "PS3D_ZWAVE": "2"})
    for c in range(3):
        assert np.abs(sa[c] - sb[c]).max() <= 1e-13 * np.abs(sb[c]).max()
    np.testing.assert_allclose(ra, rb, rtol=1e-13, atol=1e-13 * np.abs(rb).max())
