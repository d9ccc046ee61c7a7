"""Parity of the surrogate network's BF16 MFMA kernels (wt_mlp_* through the C ABI) against the
FP64 oracle (oracle/mlp_oracle.py) with the GPU's BF16 roundings emulated.

Bars: a linear layer (no rounding between products) within 1e-5 relative L2 for outputs and
1e-4 for gradients; the full 48-256-256-256-24 network within 5e-4 (outputs) and 5e-3
(gradients): FP32 accumulation moves a few BF16 activations across a rounding boundary
(each such flip is 2^-8 relative). One Adam step within 1e-6 of the oracle's step from the same
gradient; training is bitwise reproducible and lowers the loss."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu

SURR = [48, 256, 256, 256, 24]


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _data(sizes, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, sizes[0])).astype(np.float32)
    T = rng.standard_normal((n, sizes[-1])).astype(np.float32)
    M = (rng.random((n, sizes[-1])) > 0.25).astype(np.float32)
    return X, T, M


@pytest.mark.parametrize("sizes,n", [([48, 24], 70), ([40, 96], 128), ([64, 64], 64)])
def test_linear_layer_forward_and_gradient(sizes, n):
    from fastweavturb_amd.mlp import MLP
    flat = MO.init_params(sizes, 11).astype(np.float32)
    X, T, M = _data(sizes, n, 12)
    m = MLP(sizes)
    m.set_params(flat)
    np.testing.assert_array_equal(m.get_params(), flat)
    y = m.forward(X)
    yo, _ = MO.forward(sizes, flat, X)
    assert rel(y, yo) <= 1e-5
    loss, g = m.loss_grad(X, T, M)
    lo, go = MO.loss_grad(sizes, flat, X, T, M)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert rel(g, go) <= 1e-4


@pytest.mark.parametrize("n", [37, 256])
def test_surrogate_network_forward_and_gradient(n):
    from fastweavturb_amd.mlp import MLP
    flat = MO.init_params(SURR, 21).astype(np.float32)
    X, T, M = _data(SURR, n, 22)
    m = MLP(SURR)
    m.set_params(flat)
    y = m.forward(X)
    yo, _ = MO.forward(SURR, flat, X)
    assert rel(y, yo) <= 5e-4
    loss, g = m.loss_grad(X, T, M)
    lo, go = MO.loss_grad(SURR, flat, X, T, M)
    assert loss == pytest.approx(lo, rel=1e-3)
    assert rel(g, go) <= 5e-3
    # per layer, so that no layer hides behind a larger one
    Wg, bg = MO.unflatten(SURR, g)
    Wo, bo = MO.unflatten(SURR, go)
    for l in range(len(Wg)):
        assert rel(Wg[l], Wo[l]) <= 1e-2 and rel(bg[l], bo[l]) <= 1e-2


def test_adam_step_matches_oracle():
    from fastweavturb_amd.mlp import MLP
    flat = MO.init_params(SURR, 31).astype(np.float32)
    X, T, M = _data(SURR, 128, 32)
    m = MLP(SURR)
    m.set_params(flat)
    _, g = m.loss_grad(X, T, M)
    m.train_step(X, T, M, lr=1e-3)
    w1 = m.get_params()
    ref = MO.Adam(flat.size).step(flat.astype(np.float64), g.astype(np.float64), 1e-3)
    np.testing.assert_allclose(w1, ref, rtol=1e-6, atol=1e-9)


def test_training_lowers_loss_and_is_reproducible():
    from fastweavturb_amd.mlp import MLP
    rng = np.random.default_rng(41)
    X = rng.standard_normal((3000, 48)).astype(np.float32)
    A = rng.standard_normal((48, 24)) / np.sqrt(48)
    T = np.tanh(X @ A).astype(np.float32)
    runs = []
    for _ in range(2):
        m = MLP(SURR, seed=5)
        m.set_data(X, T)
        hist = m.train(30, batch=256, lr=2e-3, lr_final=1e-4, seed=9)
        runs.append((hist, m.get_params()))
    hist, w = runs[0]
    np.testing.assert_array_equal(w, runs[1][1])
    base = float(np.mean((T - T.mean(axis=0)) ** 2))
    assert hist[-1] < 0.15 * base
    assert hist[-1] < 0.2 * hist[0]
