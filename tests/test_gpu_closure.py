"""The closure net on the matrix cores against its CPU restatement
(oracle/closure_oracle.py): forward at bf16 storage precision, gradients,
Adam, and the hook into the solver step."""
import numpy as np
import pytest

from oracle import closure_oracle as N
from ps3d_amd import ics

pytestmark = pytest.mark.gpu
PI = np.pi


@pytest.fixture(scope="module")
def CL():
    from ps3d_amd import closure
    closure.lib()
    return closure


def setup_net(CL, shape, seed):
    nx, ny, nz = shape
    net = CL.ClosureNet(nx, ny, nz)
    rng = np.random.default_rng(seed)
    params = N.init_params(seed)
    params = [(W, rng.standard_normal(b.shape) * 0.1) for W, b in params]
    w = N.pack(params).astype(np.float32)
    net.set_weights(w)
    params = N.unpack(w.astype(np.float64))
    x = rng.standard_normal((nx, ny, nz + 1, 6))
    t = rng.standard_normal((nx, ny, nz + 1, 3))
    return net, params, x, t


@pytest.mark.parametrize("shape", [(8, 6, 5), (4, 4, 16), (16, 16, 7)])
def test_forward_matches_oracle(CL, shape):
    net, params, x, _ = setup_net(CL, shape, 1)
    print("closure forward", shape, flush=True)
    y = net.forward(x.reshape(-1, 6)).reshape(x.shape[:3] + (3,))
    ref = N.forward_bf16(params, x)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("rel err", err, flush=True)
    assert err < 2e-3
    net.close()


def test_gradients_match_oracle(CL):
    net, params, x, t = setup_net(CL, (8, 6, 5), 2)
    loss, g, gx = net.grad(x.reshape(-1, 6), t.reshape(-1, 3), want_gx=True)
    # oracle at the GPU's input / weight precision; hidden activations fp64
    bparams = [(N.bf16_round(W), b) for W, b in params]
    lref, gref, gxref = N.loss_and_grad(bparams, N.bf16_round(x), t)
    assert abs(loss - lref) < 1e-3 * lref
    gvec = N.pack(gref)
    assert np.abs(g - gvec).max() / np.abs(gvec).max() < 3e-2
    # per-layer direction agreement
    o = 0
    for cin, cout in N.LAYERS:
        n = 27 * cout * cin + cout
        a, b = g[o:o + n], gvec[o:o + n]
        assert np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b) > 0.999
        o += n
    gxr = gxref.reshape(-1, 6)
    assert np.abs(gx - gxr).max() / np.abs(gxr).max() < 3e-2
    net.close()


def test_weight_gradient_on_matrix_cores_matches_fp32_path(CL, monkeypatch):
    """The bf16 matrix-core weight gradient against the fp32 reduction
    (PS3D_CNN_WGRAD=fp32, read when a net is created) on a grid whose point count
    leaves a partial last chunk: per-layer direction and magnitude."""
    shape = (12, 10, 20)
    monkeypatch.setenv("PS3D_CNN_WGRAD", "fp32")
    ref_net, params, x, t = setup_net(CL, shape, 5)
    monkeypatch.delenv("PS3D_CNN_WGRAD")
    net, _, _, _ = setup_net(CL, shape, 5)
    _, g_ref, _ = ref_net.grad(x.reshape(-1, 6), t.reshape(-1, 3))
    _, g, _ = net.grad(x.reshape(-1, 6), t.reshape(-1, 3))
    o = 0
    for cin, cout in N.LAYERS:
        nw = 27 * cout * cin
        a, b = g[o:o + nw], g_ref[o:o + nw]
        cos = np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)
        print("layer", cin, cout, "cos", cos, "norm ratio", np.linalg.norm(a) / np.linalg.norm(b), flush=True)
        assert cos > 0.9999
        assert abs(np.linalg.norm(a) / np.linalg.norm(b) - 1.0) < 1e-2
        ab, bb = g[o + nw:o + nw + cout], g_ref[o + nw:o + nw + cout]
        assert np.abs(ab - bb).max() <= 1e-4 * np.abs(bb).max()
        o += nw + cout
    net.close()
    ref_net.close()


def test_adam_step_matches_oracle(CL):
    net, params, x, t = setup_net(CL, (8, 6, 5), 3)
    w0 = net.get_weights().astype(np.float64)
    _, g, _ = net.grad(x.reshape(-1, 6), t.reshape(-1, 3))
    net.set_weights(w0.astype(np.float32))      # resets the Adam moments
    net.train_step(x.reshape(-1, 6), t.reshape(-1, 3), 1e-3)
    w1 = net.get_weights().astype(np.float64)
    ref, _, _ = N.adam_step(w0, g.astype(np.float64), np.zeros_like(w0), np.zeros_like(w0), 1, 1e-3)
    assert np.abs(w1 - ref).max() < 1e-6
    net.close()


def test_training_reduces_loss(CL):
    net, params, x, t = setup_net(CL, (8, 8, 7), 4)
    # a learnable target: the oracle net's own output for other weights
    tgt = N.forward(N.init_params(9), x).reshape(-1, 3).astype(np.float32)
    l0 = net.grad(x.reshape(-1, 6), tgt)[0]
    for _ in range(60):
        net.train_step(x.reshape(-1, 6), tgt, 3e-3)
    l1 = net.grad(x.reshape(-1, 6), tgt)[0]
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.5 * l0
    net.close()


def test_zero_closure_leaves_the_step_unchanged(CL):
    from ps3d_amd import gpu as G
    p = G.make_params(32, 32, 32, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI), filtering="2/3-rule",
                      nnu=3, prediss=30.0, pretype="vorch", alpha=0.1)
    vor = ics.random_phase_isotropic(32, 32, 32, k0=3.0, seed=5)
    a, b = G.Ps3dGpu(p), G.Ps3dGpu(p)
    a.set_vorticity(vor)
    b.set_vorticity(vor)
    net = CL.ClosureNet(32, 32, 32)          # all weights zero: output exactly 0
    net.attach(b)
    a.advance_n(3)
    b.advance_n(3)
    assert np.array_equal(a.get(G.SVOR), b.get(G.SVOR))
    CL.ClosureNet.detach(b)
    net.close()
    a.close()
    b.close()


def test_closure_changes_the_step_by_its_tendency(CL):
    """A closure whose output is a constant c_z on the vorticity-z channel adds,
    per step, the tendency of that constant: the (0,0) mean of zeta grows by
    c_z * dt per step (the mean mode is undamped and unfiltered)."""
    from ps3d_amd import gpu as G
    p = G.make_params(16, 16, 16, lower=(0.0, 0.0, 0.0), extent=(2 * PI, 2 * PI, PI), filtering="2/3-rule",
                      nnu=3, prediss=30.0, pretype="vorch", alpha=0.1, dt_fixed=0.01)
    vor = ics.random_phase_isotropic(16, 16, 16, k0=3.0, seed=6)
    a, b = G.Ps3dGpu(p), G.Ps3dGpu(p)
    a.set_vorticity(vor)
    b.set_vorticity(vor)
    net = CL.ClosureNet(16, 16, 16)
    w = np.zeros(CL.nparams(), dtype=np.float32)
    w[-1] = 1.0                               # bias of output channel 2 (zeta)
    net.set_weights(w)
    net.out_scale = np.array([1.0, 1.0, 0.5], dtype=np.float32)
    net.attach(b)
    da = a.advance_n(2)
    db = b.advance_n(2)
    za, zb = a.get(G.SVOR)[2], b.get(G.SVOR)[2]
    # (0,0) column: lid values are horizontal means times sqrt(nx ny) (stafft scaling)
    d = (zb[0, 0, 0] - za[0, 0, 0]) / 16.0
    assert abs(d - 0.5 * 2 * 0.01) < 1e-9, d
    assert np.array_equal(da[:, 1], db[:, 1])
    net.close()
    a.close()
    b.close()
