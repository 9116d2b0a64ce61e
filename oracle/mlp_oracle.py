"""ORACLE -- test infrastructure only.

Float64 restatement of the plan-cost MLP (16 -> 64 -> 64 -> 2, ReLU, mean squared error)
used to check the HIP forward, backward and Adam kernels of
tensororder_amd/csrc/costmodel.hip.  Parameter layout: W1[64][16], b1[64], W2[64][64],
b2[64], W3[2][64], b3[2].  There is no counterpart in the reference (whose time model is
performance_factor * FLOPs, planning.py:168-172); the checks that pin this oracle are the
finite-difference gradient test and the closed-form Adam step in tests/test_costmodel.py.
"""

import numpy as np

SIZES = [(64, 16), (64,), (64, 64), (64,), (2, 64), (2,)]


def unpack(params):
    out, off = [], 0
    for shp in SIZES:
        n = int(np.prod(shp))
        out.append(np.asarray(params[off:off + n], dtype=np.float64).reshape(shp))
        off += n
    return out


def forward(params, x):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    y = h2 @ W3.T + b3
    return h1, h2, y


def loss_and_grad(params, x, target):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(target, dtype=np.float64)
    h1, h2, y = forward(params, x)
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / y.size
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad


def adam(params, grad, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
