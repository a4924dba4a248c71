"""Pins of the closure network's CPU restatement (oracle/closure_oracle.py):
the vectorised convolution against a brute-force loop, and every gradient of
the loss against central finite differences."""
import numpy as np
import pytest

from oracle import closure_oracle as N


def small_params(seed=0):
    return N.init_params(seed)


def test_conv3d_matches_bruteforce():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 3, 5, 6))
    W = rng.standard_normal((27, 5, 6))
    b = rng.standard_normal(5)
    assert np.abs(N.conv3d(x, W, b) - N.conv3d_bruteforce(x, W, b)).max() < 1e-12


def test_shift_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 6, 3))
    g = rng.standard_normal((4, 5, 6, 3))
    for dx, dy, dz in N.tap_offsets():
        lhs = np.sum(g * N.shift(x, dx, dy, dz))
        rhs = np.sum(N.shift_adjoint(g, dx, dy, dz) * x)
        assert abs(lhs - rhs) < 1e-12


def test_gelu_grad_finite_difference():
    x = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (N.gelu(x + h) - N.gelu(x - h)) / (2 * h)
    assert np.abs(fd - N.gelu_grad(x)).max() < 1e-8


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_finite_difference(seed):
    rng = np.random.default_rng(seed)
    params = small_params(seed)
    x = rng.standard_normal((4, 4, 5, 6))
    t = rng.standard_normal((4, 4, 5, 3))
    loss, grads, gx = N.loss_and_grad(params, x, t)
    vec = N.pack(params)
    gvec = N.pack(grads)
    h = 1e-6
    for idx in rng.choice(vec.size, size=12, replace=False):
        vp, vm = vec.copy(), vec.copy()
        vp[idx] += h
        vm[idx] -= h
        fd = (N.loss_and_grad(N.unpack(vp), x, t)[0] - N.loss_and_grad(N.unpack(vm), x, t)[0]) / (2 * h)
        assert abs(fd - gvec[idx]) < 1e-7 * max(1.0, abs(fd)), idx
    for _ in range(6):
        i = tuple(rng.integers(0, s) for s in x.shape)
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (N.loss_and_grad(params, xp, t)[0] - N.loss_and_grad(params, xm, t)[0]) / (2 * h)
        assert abs(fd - gx[i]) < 1e-7 * max(1.0, abs(fd)), i


def test_adam_first_step_is_signed_lr():
    """With bias correction the first Adam step moves every weight by lr * sign(g)
    (up to eps)."""
    rng = np.random.default_rng(3)
    w = rng.standard_normal(50)
    g = rng.standard_normal(50)
    w1, m, v = N.adam_step(w, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # exact step lr g / (|g| + eps)
    assert np.all(np.abs((w - w1) - 1e-3 * np.sign(g)) <= 1e-3 * 1e-8 / np.abs(g) + 1e-15)


def test_bf16_round():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0])
    r = N.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] - (-3.140625)) < 1e-12
