"""CPU oracle of the learned control proposal's network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and the training check of
bench.py), never by the product.

The network has no counterpart in the reference (kranzowe/Kino-PAX-Fork draws
every control uniformly, statePropagatorSpatialHash.cu:236-241); this file
restates the build's own definition in numpy float64, rounding to bf16 at the
points where the GPU feeds its matrix cores:

    h1 = bf(X) bf(W1) + b1,  a1 = bf(gelu(h1))
    h2 = a1 bf(W2) + b2,     a2 = bf(gelu(h2))
    out = a2 bf(W3) + b3     (4 means, then 4 log sigmas)

loss = mean over rows of sum_j 0.5 ((y_j - mu_j) / sigma_j)^2 + log sigma_j,
log sigma floored at -4 (zero gradient below the floor).  Backward:
g = dloss/dout (fp32), dA2 = bf(g) bf(W3)^T, dH2 = dA2 * gelu'(h2),
dA1 = bf(dH2) bf(W2)^T, dH1 = dA1 * gelu'(h1); weight gradients from the
bf16 operands (bf(X)^T bf(dH1), a1^T bf(dH2), a2^T bf(g)), bias gradients
from the unrounded dH1, dH2 and g.
"""
import numpy as np

IN, H, OUT = 80, 128, 8
W1, B1 = 0, IN * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * OUT
NPARAMS = B3 + OUT


def bf16(x):
    """Round float values to bf16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(IN, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, OUT),
            p[B3:B3 + OUT])


K = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K * (1.0 + 3 * 0.044715 * x * x)


def _ident(x):
    return np.asarray(x, np.float64)


def forward(params, X, exact=False):
    """exact=True skips the bf16 roundings (the plain float64 network)."""
    bf16 = _ident if exact else globals()["bf16"]
    w1, b1, w2, b2, w3, b3 = unpack(params)
    xb = bf16(X)
    h1 = xb @ bf16(w1) + b1
    a1 = bf16(gelu(h1))
    h2 = a1 @ bf16(w2) + b2
    a2 = bf16(gelu(h2))
    out = a2 @ bf16(w3) + b3
    return out, (xb, h1, a1, h2, a2)


def loss_and_grads(params, X, Y, exact=False):
    bf16 = _ident if exact else globals()["bf16"]
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    n = X.shape[0]
    w1, b1, w2, b2, w3, b3 = unpack(params)
    out, (xb, h1, a1, h2, a2) = forward(params, X, exact)
    mu, s = out[:, :4], out[:, 4:]
    sc = np.maximum(s, -4.0)
    iv = np.exp(-2.0 * sc)
    d = Y - mu
    loss = float(np.sum(0.5 * d * d * iv + sc) / n)
    g = np.concatenate([-d * iv / n, np.where(s < -4.0, 0.0, (1.0 - d * d * iv) / n)], 1)
    gb = bf16(g)
    dh2 = (gb @ bf16(w3).T) * gelu_grad(h2)
    dh2b = bf16(dh2)
    dh1 = (dh2b @ bf16(w2).T) * gelu_grad(h1)
    dh1b = bf16(dh1)
    grads = np.zeros(NPARAMS)
    grads[W1:B1] = (xb.T @ dh1b).ravel()
    grads[B1:W2] = dh1.sum(0)
    grads[W2:B2] = (a1.T @ dh2b).ravel()
    grads[B2:W3] = dh2.sum(0)
    grads[W3:B3] = (a2.T @ gb).ravel()
    grads[B3:] = g.sum(0)
    return loss, grads
