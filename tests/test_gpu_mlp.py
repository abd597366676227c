"""The surrogate's BF16 MFMA kernels (axonexa_amd/csrc/mlp.hip) against the
float64 CPU reference (oracle/mlp_ref.py), whose own gradient is pinned by
finite differences (tests/test_mlp_ref.py).

Tolerances: the reference rounds what the GPU keeps in bf16; what remains
is fp32 accumulation order (linear layers: 1e-4 relative) and, for
gradients, the bf16 storage of the backward activations (3 % relative per
layer in the 2-norm; cosine > 0.999).
"""
import numpy as np
import pytest

from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu


def _engine():
    from axonexa_amd import engine
    return engine


@pytest.mark.parametrize("dims,n", [([384, 256], 256), ([268, 30], 300), ([1000, 130], 77)])
def test_linear_layer_is_a_bf16_gemm(dims, n):
    eng = _engine()
    rng = np.random.default_rng(dims[0])
    p = R.init_params(dims, seed=1)
    p[-dims[1]:] = rng.normal(size=dims[1])                 # biases
    x = rng.normal(size=(n, dims[0])).astype(np.float32)
    with eng.MLP(dims, max_batch=512) as m:
        m.set_params(p)
        y = m.forward(x)
        assert np.array_equal(m.get_params(), p)
    ref, _ = R.forward(p, dims, x, bf16=True)
    err = np.abs(y - ref) / (np.abs(ref) + 1e-3 * np.abs(ref).max())
    print(f"[mlp] linear {dims} n={n}: max rel err {err.max():.2e}", flush=True)
    assert err.max() < 1e-4


def test_forward_matches_reference():
    eng = _engine()
    dims = [268, 1024, 1024, 1024, 30]
    rng = np.random.default_rng(3)
    p = R.init_params(dims, seed=3)
    x = rng.uniform(-1, 1, size=(300, 268)).astype(np.float32)
    with eng.MLP(dims, max_batch=1024) as m:
        m.set_params(p)
        y = m.forward(x)
    ref, _ = R.forward(p, dims, x, bf16=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"[mlp] forward 4 layers: max err / max |y| = {err:.2e}", flush=True)
    assert err < 2e-2


def test_gradients_match_reference():
    eng = _engine()
    dims = [268, 512, 512, 30]
    rng = np.random.default_rng(4)
    p = R.init_params(dims, seed=4)
    p += rng.normal(scale=0.02, size=p.size).astype(np.float32)
    x = rng.uniform(-1, 1, size=(700, 268)).astype(np.float32)
    t = rng.uniform(0, 1, size=(700, 30)).astype(np.float32)
    with eng.MLP(dims, max_batch=1024) as m:
        m.set_params(p)
        loss, g = m.loss_grad(x, t)
    rloss, rg = R.loss_grad(p, dims, x, t, bf16=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    o = 0
    for a, b in zip(dims[:-1], dims[1:]):
        for part, size in (("W", a * b), ("b", b)):
            gg, rr = g[o:o + size].astype(np.float64), rg[o:o + size]
            o += size
            rel = np.linalg.norm(gg - rr) / np.linalg.norm(rr)
            cos = gg @ rr / (np.linalg.norm(gg) * np.linalg.norm(rr))
            print(f"[mlp] grad {part} {a}->{b}: rel {rel:.2e} cos {cos:.6f}", flush=True)
            assert rel < 3e-2 and cos > 0.999


def test_training_reduces_loss_and_is_deterministic():
    eng = _engine()
    dims = [32, 256, 256, 8]
    rng = np.random.default_rng(5)
    A = rng.normal(size=(32, 8))
    x = rng.uniform(-1, 1, size=(4096, 32)).astype(np.float32)
    t = np.tanh(x @ A / 3.0).astype(np.float32)     # smooth target, O(1) arguments
    p0 = R.init_params(dims, seed=6)
    runs = []
    for _ in range(2):
        with eng.MLP(dims, max_batch=1024) as m:
            m.set_params(p0)
            m.set_data(x, t)
            trace = m.train(300, batch=512, lr=2e-3, seed=7)
            runs.append((trace, m.get_params()))
    trace = runs[0][0]
    print(f"[mlp] training loss {trace[0]:.4f} -> {trace[-10:].mean():.4f}", flush=True)
    assert trace[-10:].mean() < 0.1 * trace[:10].mean()
    assert np.array_equal(runs[0][1], runs[1][1])
