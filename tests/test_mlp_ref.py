"""Pins of the surrogate's CPU reference (oracle/mlp_ref.py): the gradient
against central finite differences in float64, GELU against closed-form
values, bf16 rounding against exact cases."""
import numpy as np

from oracle import mlp_ref as M


def test_gradient_matches_finite_differences():
    dims = [5, 7, 6, 3]
    rng = np.random.default_rng(1)
    p = M.init_params(dims, seed=2).astype(np.float64)
    p += rng.normal(scale=0.1, size=p.size)        # non-zero biases too
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 3))
    _, g = M.loss_grad(p, dims, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        lp = M.loss_grad(p + e, dims, x, t)[0]
        lm = M.loss_grad(p - e, dims, x, t)[0]
        fd[k] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_gelu_values_and_derivative():
    # tanh-form GELU: gelu(0) = 0, gelu(x) - gelu(-x) = x, gelu(1) = 0.8411920
    x = np.linspace(-4, 4, 81)
    assert np.allclose(M.gelu(x) - M.gelu(-x), x, atol=1e-12)
    assert abs(M.gelu(1.0) - 0.8411919906082768) < 1e-12
    h = 1e-6
    assert np.allclose((M.gelu(x + h) - M.gelu(x - h)) / (2 * h), M.gelu_grad(x), atol=1e-8)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38, 1e-40])
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie to even
    assert r[2] == 1.0 + 2.0 ** -7          # above the tie
    assert r[3] == -2.5
    rel = np.abs(M.bf16_round(np.linspace(0.1, 10, 1000)) / np.linspace(0.1, 10, 1000) - 1)
    assert rel.max() <= 2.0 ** -8           # 8 significant bits


def test_pack_roundtrip():
    dims = [4, 3, 2]
    p = M.init_params(dims, seed=5)
    assert np.array_equal(M.pack(M.unpack(p, dims)).astype(np.float32), p)
    assert p.size == 4 * 3 + 3 + 3 * 2 + 2
