"""CPU restatement of the learned closure network (test infrastructure only).

The closure has no counterpart in ps3d; this module is the checker for the GPU
network in ps3d_amd/csrc/closure.hip: a 4-layer 3x3x3 convolution net
6 -> 32 -> 32 -> 32 -> 3 channels with GELU between layers, on a ps3d grid of
nx x ny x (nz+1) points (x, y periodic; z zero-padded beyond the lids), its
mean-square loss, the exact backward pass and one Adam update, in float64.

Conventions shared with the GPU code:
* tensors are (nx, ny, nz+1, C), channels last;
* a layer's weights are W[t, cout, cin] with tap t = (dx+1)*9 + (dy+1)*3 + (dz+1),
  dx, dy, dz in {-1, 0, 1}; out[p] = b + sum_t W[t] @ in[p + (dx, dy, dz)];
* GELU(x) = x/2 (1 + erf(x / sqrt 2));
* loss = mean over points and the 3 output channels of (y - target)^2;
* Adam (Kingma & Ba 2015) with bias correction: m = b1 m + (1-b1) g,
  v = b2 v + (1-b2) g^2, w -= lr * (m / (1-b1^k)) / (sqrt(v / (1-b2^k)) + eps).

Parity pins (tests/test_closure_oracle.py): the vectorised convolution against a
brute-force loop; every gradient against central finite differences of the loss.
"""
import math

import numpy as np
from scipy.special import erf

LAYERS = [(6, 32), (32, 32), (32, 32), (32, 3)]


def tap_offsets():
    return [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def shift(x, dx, dy, dz):
    """xs[p] = x[p + (dx, dy, dz)]: periodic in x, y; zero beyond the lids in z."""
    xs = np.roll(np.roll(x, -dx, axis=0), -dy, axis=1)
    out = np.zeros_like(xs)
    nz1 = x.shape[2]
    if dz == 0:
        out[:] = xs
    elif dz > 0:
        out[:, :, :nz1 - dz] = xs[:, :, dz:]
    else:
        out[:, :, -dz:] = xs[:, :, :nz1 + dz]
    return out


def shift_adjoint(g, dx, dy, dz):
    """Adjoint of shift: sum_p g[p] . xs[p] = sum_q shift_adjoint(g)[q] . x[q]."""
    gs = np.zeros_like(g)
    nz1 = g.shape[2]
    if dz == 0:
        gs[:] = g
    elif dz > 0:
        gs[:, :, dz:] = g[:, :, :nz1 - dz]
    else:
        gs[:, :, :nz1 + dz] = g[:, :, -dz:]
    return np.roll(np.roll(gs, dx, axis=0), dy, axis=1)


def conv3d(x, W, b):
    out = np.zeros(x.shape[:3] + (W.shape[1],))
    for t, (dx, dy, dz) in enumerate(tap_offsets()):
        out += shift(x, dx, dy, dz) @ W[t].T
    return out + b


def conv3d_bruteforce(x, W, b):
    nx, ny, nz1, _ = x.shape
    out = np.zeros((nx, ny, nz1, W.shape[1]))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz1):
                acc = b.astype(np.float64).copy()
                for t, (dx, dy, dz) in enumerate(tap_offsets()):
                    kk = k + dz
                    if 0 <= kk < nz1:
                        acc += W[t] @ x[(i + dx) % nx, (j + dy) % ny, kk]
                out[i, j, k] = acc
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def init_params(seed=0):
    """He-style uniform initialisation: W ~ U(-s, s), s = sqrt(6 / (27 cin)), b = 0."""
    rng = np.random.default_rng(seed)
    params = []
    for cin, cout in LAYERS:
        s = math.sqrt(6.0 / (27 * cin))
        params.append((rng.uniform(-s, s, size=(27, cout, cin)), np.zeros(cout)))
    return params


def pack(params):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params])


def unpack(vec):
    params, o = [], 0
    for cin, cout in LAYERS:
        nw = 27 * cout * cin
        W = vec[o:o + nw].reshape(27, cout, cin)
        o += nw
        b = vec[o:o + cout]
        o += cout
        params.append((W.copy(), b.copy()))
    return params


def forward(params, x, keep=False):
    acts, pres = [x], []
    h = x
    for li, (W, b) in enumerate(params):
        z = conv3d(h, W, b)
        pres.append(z)
        h = gelu(z) if li < len(params) - 1 else z
        acts.append(h)
    return (h, acts, pres) if keep else h


def loss_and_grad(params, x, target):
    """Loss, parameter gradients (same structure as params) and the input gradient."""
    y, acts, pres = forward(params, x, keep=True)
    n = y.size
    loss = float(np.sum((y - target) ** 2) / n)
    g = 2.0 * (y - target) / n
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        W, b = params[li]
        if li < len(params) - 1:
            g = g * gelu_grad(pres[li])
        h = acts[li]
        gW = np.zeros_like(W)
        gin = np.zeros_like(h)
        g2 = g.reshape(-1, g.shape[-1])
        for t, (dx, dy, dz) in enumerate(tap_offsets()):
            gW[t] = g2.T @ shift(h, dx, dy, dz).reshape(-1, h.shape[-1])
            gin += shift_adjoint(g @ W[t], dx, dy, dz)
        grads[li] = (gW, g2.sum(axis=0))
        g = gin
    return loss, grads, g


def adam_step(w, g, m, v, k, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the packed parameter vector (step count k >= 1)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** k)
    vh = v / (1.0 - b2 ** k)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(params, x):
    """forward() with the GPU's storage precisions: input, weights and hidden
    activations rounded to bfloat16 (biases and accumulation exact)."""
    h = bf16_round(x)
    for li, (W, b) in enumerate(params):
        z = conv3d(h, bf16_round(W), b)
        h = bf16_round(gelu(z)) if li < len(params) - 1 else z
    return h
