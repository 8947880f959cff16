"""GPU tests of the degree-predicting MLP: the HIP forward pass and gradients against the
NumPy oracle (oracle/mlp.py), and training on the GPU."""
import sys

import numpy as np
import pytest

from oracle import mlp as O

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_oracle():
    from hfmm0_amd.learned import DegreeMLP, init_params
    print("case: mlp forward/gradient vs oracle, 300 samples", flush=True)
    rng = np.random.default_rng(4)
    params = init_params(7)
    X = rng.normal(size=(300, 6)).astype(np.float32)
    y = rng.normal(size=300).astype(np.float32)
    net = DegreeMLP(params)
    got = net.forward(X)
    ref, _ = O.forward(params, X, np.float64)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
    loss, g = net.gradient(X, y)
    rloss, rg = O.loss_and_grad(params, X, y, np.float64)
    assert abs(loss - rloss) <= 1e-4 * rloss
    np.testing.assert_allclose(g, rg, rtol=2e-3, atol=2e-5 * np.abs(rg).max())


def test_training_matches_oracle_steps_and_learns():
    from hfmm0_amd.learned import DegreeMLP, init_params
    print("case: mlp training on a smooth target, 512 samples", flush=True)
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(512, 6)).astype(np.float32)
    y = (0.5 * np.tanh(2 * X[:, 0]) + 0.3 * X[:, 1] * X[:, 2]).astype(np.float32)
    params = init_params(8)
    # one full-batch Adam step on the GPU equals the oracle's step
    net = DegreeMLP(params)
    net.train(X, y, epochs=1, batch=512, lr=1e-3, seed=0)
    _, g = O.loss_and_grad(params, X, y, np.float64)
    want = O.adam_step(params.copy(), np.zeros_like(params), np.zeros_like(params), g.astype(np.float32), 1, 1e-3)
    np.testing.assert_allclose(net.params, want, rtol=0, atol=2e-6)
    # and training drives the loss down
    net2 = DegreeMLP(params)
    hist = net2.train(X, y, epochs=150, batch=64, lr=3e-3, seed=1)
    print(f"  loss {hist[0]:.4f} -> {hist[-1]:.5f}", flush=True)
    assert hist[-1] < 0.05 * hist[0]
