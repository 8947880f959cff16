"""Synthetic inputs of the benchmark workloads (seeded, as in the reference harness:
``default_rng(seed + N)``, positions drawn first, then charges; test.py:40-44)."""
import numpy as np


def uniform_cube(n, complex_charges=True, seed=42):
    """N points uniform in [0,1]^3; charges N(0,1) (+ i N(0,1))."""
    rng = np.random.default_rng(seed + n)
    x = rng.uniform(0.0, 1.0, (n, 3))
    q = rng.normal(size=n)
    if complex_charges:
        q = q + 1j * rng.normal(size=n)
    return x, q


def sample_targets(n, n_samp):
    """Sampled-target set for error figures: default_rng(0).choice(N, n, replace=False)."""
    return np.sort(np.random.default_rng(0).choice(n, min(n_samp, n), replace=False))


def f_m2l(p):
    """Algorithmic flops of one rotate / coaxial / rotate-back M2L at degree p:
    8 S_rot + 8 S_z + 12 (p+1)^2 with S_rot = sum_n (2n+1)^2 (two real rotations of a
    complex vector) and S_z = sum_m (p+1-|m|)^2 (complex coaxial translation)."""
    s_rot = (p + 1) * (2 * p + 1) * (2 * p + 3) // 3
    s_z = sum((p + 1 - abs(m)) ** 2 for m in range(-p, p + 1))
    return 8 * s_rot + 8 * s_z + 12 * (p + 1) ** 2


P2P_FLOPS_PER_PAIR = 23      # 3 sub, 5 r^2, rsqrt, r, kr, sin, cos, 2 scale, 8 complex MAC


def mixed_problem(n, k_range=(1.0, 60.0), seed=0):
    """One seeded problem of the learned-degree training distribution: N points that are
    uniform in the unit cube, on the sphere |x - c| = 0.45 (c = (0.5, 0.5, 0.5)), in
    Gaussian clusters (sigma 0.02, centres uniform in [0.1, 0.9]^3), or a mix of the
    three; complex N(0,1) charges; k uniform in k_range."""
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(0, 4))
    k = float(rng.uniform(*k_range))

    def sphere(m):
        v = rng.normal(size=(m, 3))
        return 0.5 + 0.45 * v / np.linalg.norm(v, axis=1, keepdims=True)

    def clusters(m):
        c = rng.uniform(0.1, 0.9, (8, 3))
        return c[rng.integers(0, 8, m)] + rng.normal(scale=0.02, size=(m, 3))

    if kind == 0:
        x = rng.uniform(0.0, 1.0, (n, 3))
    elif kind == 1:
        x = sphere(n)
    elif kind == 2:
        x = clusters(n)
    else:
        a, b = n // 3, n // 3
        x = np.vstack([rng.uniform(0.0, 1.0, (a, 3)), sphere(b), clusters(n - a - b)])
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    return x, q, k, ("uniform", "sphere", "clusters", "mixed")[kind]


def ellipsoid_problem(n, k_range=(1.0, 50.0), seed=0):
    """One problem of the batched configuration: N points uniform in the solid
    axis-aligned ellipsoid centred at (0.5, 0.5, 0.5) with semi-axes uniform in
    [0.2, 0.5]; k uniform in k_range; complex N(0,1) charges rounded to complex64.
    default_rng(seed) draws the semi-axes, k, the points, then the charges."""
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.2, 0.5, 3)
    k = float(rng.uniform(*k_range))
    v = rng.normal(size=(n, 3))
    v *= (rng.uniform(size=n) ** (1.0 / 3.0) / np.linalg.norm(v, axis=1))[:, None]
    x = 0.5 + v * axes
    q = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64).astype(np.complex128)
    return x, q, k, axes


def batch_problems(n_problems=256, n_range=(2_000, 200_000), k_range=(1.0, 50.0), seed=4000):
    """The batched configuration: problem i has N log-uniform in n_range (default_rng(seed)
    draws all N first) and the ellipsoid, k and charges of ellipsoid_problem(N, k_range,
    seed + 1 + i)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_problems)).astype(int)
    return [ellipsoid_problem(int(n), k_range, seed + 1 + i) for i, n in enumerate(ns)]


def sphere_clusters(n_sphere, n_clusters, seed):
    """Points on the sphere |x - c| = 0.45, c = (0.5, 0.5, 0.5), then points in 8 Gaussian
    clusters (sigma 0.02, centres uniform in [0.1, 0.9]^3), from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n_sphere, 3))
    s = 0.5 + 0.45 * v / np.linalg.norm(v, axis=1, keepdims=True)
    c = rng.uniform(0.1, 0.9, (8, 3))
    g = c[rng.integers(0, 8, n_clusters)] + rng.normal(scale=0.02, size=(n_clusters, 3))
    return rng, np.vstack([s, g])


def c3_problem(n=2_000_000):
    """Adaptive configuration: N/2 points on the sphere and N/2 in 8 Gaussian clusters,
    k = 100, complex N(0,1) charges; default_rng(42 + N) draws the positions, then the
    charges."""
    rng, x = sphere_clusters(n // 2, n - n // 2, 42 + n)
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    return x, q, 100.0


def c5_problem(n=4_000_000):
    """Mixed configuration: a third uniform in the unit cube, a third on the sphere, the
    rest in 8 Gaussian clusters, k = 60, complex N(0,1) charges; default_rng(42 + N)."""
    rng = np.random.default_rng(42 + n)
    a = n // 3
    u = rng.uniform(0.0, 1.0, (a, 3))
    _, sg = sphere_clusters(a, n - 2 * a, int(rng.integers(0, 2**31)))
    x = np.vstack([u, sg])
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    return x, q, 60.0
