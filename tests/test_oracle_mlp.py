"""Pins for the MLP oracle: gradients against finite differences, GELU derivative."""
import numpy as np

from oracle import mlp


def test_gelu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mlp.gelu(z + h, np.float64) - mlp.gelu(z - h, np.float64)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_grad(z, np.float64), fd, atol=1e-8)
    # GELU(z) -> z for large z, -> 0 for very negative z
    assert abs(mlp.gelu(np.array([8.0]), np.float64)[0] - 8.0) < 1e-12
    assert abs(mlp.gelu(np.array([-8.0]), np.float64)[0]) < 1e-12


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 0.0], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                  # tie 1 + 2^-8 rounds to even (1.0)
    assert r[2] == np.float32(1.0078125)                # above the tie rounds up
    assert r[3] == np.float32(-3.140625)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    flat = mlp.bf16(mlp.init_params(1)).astype(np.float64)   # already on the BF16 grid
    X = rng.normal(size=(17, mlp.D_IN))
    y = rng.normal(size=17)
    loss, g = mlp.loss_and_grad(flat, X, y, np.float64)
    idx = rng.choice(mlp.N_PARAMS, 40, replace=False)
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = 1e-6
        # evaluate the network on exactly these weights (no re-rounding of the nudge)
        lp = _loss_no_round(flat + e, X, y)
        lm = _loss_no_round(flat - e, X, y)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def _loss_no_round(flat, X, y):
    P = mlp.unpack(flat)
    h = X
    for k in (1, 2, 3):
        h = mlp.gelu(h @ P[f"W{k}"].T + P[f"b{k}"], np.float64)
    pred = h @ P["W4"] + P["b4"][0]
    return float(np.mean((pred - y) ** 2))


def test_adam_reduces_loss():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(64, mlp.D_IN))
    y = np.tanh(X[:, 0]) + 0.5 * X[:, 1]
    flat = mlp.init_params(3)
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    l0, _ = mlp.loss_and_grad(flat, X, y)
    for t in range(1, 201):
        loss, g = mlp.loss_and_grad(flat, X, y)
        flat = mlp.adam_step(flat, m, v, g, t, 3e-3)
    assert loss < 0.2 * l0


def _gelu_tanh_mp(z):
    """tanh-form GELU of Hendrycks & Gimpel (2016), 0.5 z (1 + tanh(sqrt(2/pi)(z + 0.044715 z^3))),
    at 50 digits."""
    import mpmath as mp
    mp.mp.dps = 50
    z = mp.mpf(float(z))
    return float(mp.mpf("0.5") * z * (1 + mp.tanh(mp.sqrt(2 / mp.pi) * (z + mp.mpf("0.044715") * z ** 3))))


def test_gelu_values_pinned_by_mpmath():
    """GELU at fixed points against a 50-digit evaluation of the tanh form, in FP64 and in the
    FP32 the kernels use; the exact (erf) GELU differs by up to ~2e-4 near |z| = 2, so a
    wrong constant or the erf form fails the FP64 bound."""
    import mpmath as mp
    z = np.array([-6.0, -3.0, -2.0, -1.5, -0.7, -0.1, 0.0, 0.05, 0.4, 1.0, 1.7, 2.0, 3.5, 6.0])
    ref = np.array([_gelu_tanh_mp(v) for v in z])
    got64 = mlp.gelu(z, np.float64)
    np.testing.assert_allclose(got64, ref, rtol=1e-14, atol=1e-16)
    got32 = mlp.gelu(z.astype(np.float32), np.float32).astype(np.float64)
    np.testing.assert_allclose(got32, ref, rtol=3e-6, atol=1e-7)
    erf_form = np.array([float(0.5 * v * (1 + mp.erf(v / mp.sqrt(2)))) for v in z])
    assert np.max(np.abs(erf_form - ref)) > 1e-4


def test_adam_first_steps_closed_form():
    """With a constant gradient every bias-corrected Adam step equals lr g / (|g| + eps)
    (m_hat = g and v_hat = g^2 at every t), so after t steps the weights moved by
    t lr g / (|g| + eps); a missing bias correction or a swapped beta moves them by a
    different amount already at t = 1."""
    rng = np.random.default_rng(5)
    flat0 = rng.normal(size=257).astype(np.float32)
    g = rng.normal(size=257).astype(np.float32)
    g[:3] = [1e-12, -2.5e-9, 0.0]
    lr, eps = np.float32(1e-3), 1e-8
    m = np.zeros(257)
    v = np.zeros(257)
    flat = flat0.copy()
    step = lr * g.astype(np.float64) / (np.abs(g.astype(np.float64)) + eps)
    for t in range(1, 6):
        flat = mlp.adam_step(flat, m, v, g.astype(np.float64), t, lr)
        np.testing.assert_allclose(flat.astype(np.float64), flat0 - t * step, rtol=0, atol=t * 2e-7)
    assert flat[2] == flat0[2]      # zero gradient: no move
