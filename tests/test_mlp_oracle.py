"""Pins of the surrogate-network oracle (oracle/mlp_oracle.py) on the CPU: gradients against
torch autograd (float64) and central finite differences, BF16 rounding against torch's
conversion, GELU against erf closed forms, Adam against torch.optim.Adam (eps = 0)."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as MO

SIZES = [5, 7, 6, 3]


def _batch(seed=0, n=9, sizes=SIZES):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, sizes[0]))
    T = rng.standard_normal((n, sizes[-1]))
    M = (rng.random((n, sizes[-1])) > 0.3).astype(np.float64)
    return X, T, M


def test_gradients_match_finite_differences():
    flat = MO.init_params(SIZES, 1)
    X, T, M = _batch(1)
    loss, g = MO.loss_grad(SIZES, flat, X, T, M, emulate=False)
    h = 1e-6
    num = np.zeros_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp, _ = MO.loss_grad(SIZES, flat + e, X, T, M, emulate=False)
        lm, _ = MO.loss_grad(SIZES, flat - e, X, T, M, emulate=False)
        num[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_gradients_match_torch_autograd():
    torch = pytest.importorskip("torch")
    flat = MO.init_params(SIZES, 2)
    X, T, M = _batch(2)
    loss, g = MO.loss_grad(SIZES, flat, X, T, M, emulate=False)
    Ws, bs = MO.unflatten(SIZES, flat)
    tw = [torch.tensor(W, dtype=torch.float64, requires_grad=True) for W in Ws]
    tb = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for b in bs]
    a = torch.tensor(X, dtype=torch.float64)
    for l in range(len(tw)):
        a = a @ tw[l].T + tb[l]
        if l < len(tw) - 1:
            a = torch.nn.functional.gelu(a)           # exact (erf) GELU
    tm = torch.tensor(M)
    tl = (tm * (a - torch.tensor(T)) ** 2).sum() / tm.sum()
    tl.backward()
    ref = np.concatenate([np.concatenate([w.grad.numpy().ravel(), b.grad.numpy()]) for w, b in zip(tw, tb)])
    assert loss == pytest.approx(float(tl.detach()), rel=1e-14)
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-14)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(5000) * 10 ** rng.uniform(-8, 8, 5000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 65504.0]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(MO.bf16(x), ref)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert MO.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 6.0, -6.0])
    g = MO.gelu(z)
    assert g[0] == 0.0
    assert g[1] == pytest.approx(0.5 * (1 + math.erf(1 / math.sqrt(2))), rel=1e-15)
    assert g[1] - g[2] == pytest.approx(1.0, rel=1e-15)        # gelu(z) - gelu(-z) = z
    assert g[3] == pytest.approx(6.0, rel=1e-8) and abs(g[4]) < 1e-7
    assert MO.gelu_grad(np.array([0.0]))[0] == 0.5
    zz = np.linspace(-4, 4, 33)
    h = 1e-6
    np.testing.assert_allclose(MO.gelu_grad(zz), (MO.gelu(zz + h) - MO.gelu(zz - h)) / (2 * h), atol=1e-8)


def test_adam_matches_torch_without_eps():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    w0 = rng.standard_normal(17)
    grads = [rng.standard_normal(17) for _ in range(5)]
    opt = MO.Adam(17, eps=0.0)
    w = w0.copy()
    tw = torch.tensor(w0, requires_grad=True)
    topt = torch.optim.Adam([tw], lr=0.01, betas=(0.9, 0.999), eps=0.0)
    for g in grads:
        w = opt.step(w, g, 0.01)
        tw.grad = torch.tensor(g)
        topt.step()
    np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-12)


def test_emulated_mode_rounds_products_only():
    # with BF16-representable inputs and weights, emulation changes only the hidden activations
    # and dZ; a one-layer (linear) network is then identical in both modes
    sizes = [4, 3]
    flat = MO.bf16(MO.init_params(sizes, 5))
    X, T, M = _batch(5, 6, sizes)
    X = MO.bf16(X)
    y1, _ = MO.forward(sizes, flat, X, emulate=True)
    y0, _ = MO.forward(sizes, flat, X, emulate=False)
    np.testing.assert_array_equal(y1, y0)
