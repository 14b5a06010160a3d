"""CPU restatement of the warm-start MLP (test infrastructure only).

THIS MODULE IS THE CHECKER, NOT THE PRODUCT. Only ``tests/`` may import it.

Per-face MLP 12 -> 128 -> 128 -> 128 -> 4 with GELU (erf form) on the hidden layers, mean
squared loss, written with plain numpy matrix products. ``bf16=True`` rounds the operands of
every product to BF16 (round to nearest even) at the same points as the device kernels, so the
device results can be compared at FP32-accumulation level; ``bf16=False`` is exact float64 and
is pinned by central finite differences in tests/test_mlp.py.
"""
import numpy as np
from scipy.special import erf

DIMS = (12, 128, 128, 128, 4)


def round_bf16(x):
    """float -> nearest-even BF16 value (as float64)."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(params):
    """Flat parameter vector -> [(W1, b1), ..., (W4, b4)], W row-major (fan_in, fan_out)."""
    out, o = [], 0
    for a, b in zip(DIMS[:-1], DIMS[1:]):
        W = params[o:o + a * b].reshape(a, b)
        o += a * b
        bb = params[o:o + b]
        o += b
        out.append((W, bb))
    return out


def _mm(a, b, bf16):
    return round_bf16(a) @ round_bf16(b) if bf16 else a @ b


def forward(params, X, bf16=False):
    """Returns Y and the cache (inputs, pre-activations) of every layer."""
    layers = unpack(np.asarray(params, np.float64))
    h = np.asarray(X, np.float64)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = _mm(h, W, bf16) + b
        cache.append((h, z))
        h = gelu(z) if li < len(layers) - 1 else z
    return h, cache


def loss_and_grad(params, X, T, bf16=False):
    """Mean squared loss over all outputs and its gradient (flat, same layout as params)."""
    Y, cache = forward(params, X, bf16)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    d = 2.0 * (Y - T) / n
    layers = unpack(np.asarray(params, np.float64))
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        h, z = cache[li]
        W, _ = layers[li]
        dz = d if li == len(layers) - 1 else d * gelu_grad(z)
        grads.append((_mm(h.T, dz, bf16), dz.sum(axis=0)))
        d = _mm(dz, W.T, bf16)
    g = []
    for dW, db in reversed(grads):
        g.append(dW.ravel())
        g.append(db)
    return loss, np.concatenate(g)


def finite_difference(params, X, T, idx, h=1e-6):
    """Central differences of the float64 loss at parameter indices idx."""
    out = []
    for i in idx:
        p = np.array(params, np.float64)
        p[i] += h
        lp = loss_and_grad(p, X, T)[0]
        p[i] -= 2 * h
        lm = loss_and_grad(p, X, T)[0]
        out.append((lp - lm) / (2 * h))
    return np.array(out)
