"""Pins of the warm-start MLP oracle (oracle/mlp_oracle.py): central finite differences of the
float64 loss, BF16 rounding rules. CPU only."""
import numpy as np

from oracle import mlp_oracle as MO


def _params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in zip(MO.DIMS[:-1], MO.DIMS[1:]):
        parts.append(rng.standard_normal(a * b) * np.sqrt(2.0 / a))
        parts.append(0.1 * rng.standard_normal(b))
    return np.concatenate(parts)


def test_parameter_count():
    assert _params().size == 35204


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    X = rng.standard_normal((17, 12))
    T = rng.standard_normal((17, 4))
    loss, g = MO.loss_and_grad(p, X, T)
    # indices in every weight matrix and bias vector
    offs, o = [], 0
    for a, b in zip(MO.DIMS[:-1], MO.DIMS[1:]):
        offs += [o, o + a * b // 2, o + a * b - 1, o + a * b, o + a * b + b - 1]
        o += a * b + b
    fd = MO.finite_difference(p, X, T, offs)
    assert np.max(np.abs(fd - g[offs]) / (np.abs(g[offs]) + 1e-3 * np.abs(g).max())) < 1e-6


def test_gelu_gradient():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    assert np.max(np.abs((MO.gelu(z + h) - MO.gelu(z - h)) / (2 * h) - MO.gelu_grad(z))) < 1e-8


def test_bf16_rounding():
    # representable values stay, ties go to even, the step at 1 is 2^-7
    assert MO.round_bf16(np.array([1.0, -2.5, 0.0]))[0] == 1.0
    assert MO.round_bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0            # tie -> even (1.0)
    assert MO.round_bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6   # tie -> even (upper)
    assert MO.round_bf16(np.array([1.0 + 2.0 ** -7]))[0] == 1.0 + 2.0 ** -7
    x = np.random.default_rng(2).standard_normal(1000)
    assert np.max(np.abs(MO.round_bf16(x) - x) / np.abs(x)) <= 2.0 ** -8
