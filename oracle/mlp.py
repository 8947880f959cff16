"""Oracle (test infrastructure only): NumPy restatement of the degree-predicting MLP.

The network is a project component with no counterpart in the reference (the role it
takes is that of ``jaxfmm/tune.py``'s certify-and-ladder order choice, ``tune.py:375-516``).
This module is the CPU checker of the HIP kernels in ``hfmm0_amd/csrc/mlp.hip``: same
parameter layout, same BF16 rounding of the weights (round to nearest even), FP32
accumulation, the tanh form of GELU, mean-squared-error loss and Adam.

Parameter layout (float32, 8833 values): W1[64][6], b1[64], W2[64][64], b2[64],
W3[64][64], b3[64], W4[64], b4.

Pinned by ``tests/test_oracle_mlp.py``: gradients against central finite differences
of the loss (FP64 evaluation of the same BF16-rounded network), the closed-form GELU
derivative against finite differences, GELU values against a 50-digit mpmath evaluation of
the tanh form (Hendrycks & Gimpel 2016), BF16 rounding cases, and the Adam step against its
closed form for a constant gradient (every bias-corrected step is lr g / (|g| + eps)).
"""
import numpy as np

D_IN, H = 6, 64
SIZES = [("W1", (H, D_IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (H,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
C_GELU = np.float32(np.sqrt(2.0 / np.pi))


def unpack(flat):
    out, pos = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = flat[pos:pos + n].reshape(shape)
        pos += n
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and back to float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(z, dtype=np.float32):
    z = z.astype(dtype)
    c = dtype(np.sqrt(2.0 / np.pi))
    return dtype(0.5) * z * (dtype(1) + np.tanh(c * (z + dtype(0.044715) * z * z * z)))


def gelu_grad(z, dtype=np.float32):
    z = z.astype(dtype)
    c = dtype(np.sqrt(2.0 / np.pi))
    u = c * (z + dtype(0.044715) * z * z * z)
    t = np.tanh(u)
    du = c * (dtype(1) + dtype(3 * 0.044715) * z * z)
    return dtype(0.5) * (dtype(1) + t) + dtype(0.5) * z * (dtype(1) - t * t) * du


def forward(flat, X, dtype=np.float32):
    """Returns (prediction (n,), cache).  Weights are BF16-rounded, sums in dtype."""
    P = unpack(bf16(flat).astype(dtype))
    X = X.astype(dtype)
    z1 = X @ P["W1"].T + P["b1"]
    h1 = gelu(z1, dtype)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = gelu(z2, dtype)
    z3 = h2 @ P["W3"].T + P["b3"]
    h3 = gelu(z3, dtype)
    y = h3 @ P["W4"] + P["b4"][0]
    return y, dict(P=P, X=X, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3)


def loss_and_grad(flat, X, y_true, dtype=np.float32):
    """MSE loss mean((y - t)^2) and its gradient with respect to the (rounded) weights."""
    y, c = forward(flat, X, dtype)
    n = len(X)
    r = y - y_true.astype(dtype)
    loss = float(np.mean(r.astype(np.float64) ** 2))
    dy = (dtype(2) / dtype(n)) * r
    P = c["P"]
    g = {}
    g["W4"] = c["h3"].T @ dy
    g["b4"] = np.array([dy.sum()], dtype=dtype)
    dz3 = np.outer(dy, P["W4"]) * gelu_grad(c["z3"], dtype)
    g["W3"] = dz3.T @ c["h2"]
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ P["W3"]) * gelu_grad(c["z2"], dtype)
    g["W2"] = dz2.T @ c["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * gelu_grad(c["z1"], dtype)
    g["W1"] = dz1.T @ c["X"]
    g["b1"] = dz1.sum(0)
    flat_g = np.concatenate([np.asarray(g[name], dtype=dtype).reshape(-1) for name, _ in SIZES])
    return loss, flat_g


def adam_step(flat, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the FP32 master weights (t counts from 1)."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    flat -= (lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return flat


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("b"):
            parts.append(np.zeros(int(np.prod(shape)), np.float32))
        else:
            fan_in = shape[-1] if len(shape) == 2 else H
            a = 1.0 / np.sqrt(fan_in)
            parts.append(rng.uniform(-a, a, int(np.prod(shape))).astype(np.float32))
    return np.concatenate(parts)
