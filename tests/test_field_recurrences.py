"""The gradient recurrences used by the field path, checked by the mathematics alone:
for the scaled regular basis B_nm = jhat_n(kr) (r/R)^n Y_n^m (Y with the Condon-Shortley
phase, as scipy's), central differences of B_nm equal the stated two-term combinations
of B_{n-1,.} and B_{n+1,.} (d/dz and d/dx +- i d/dy), at k > 0 and k = 0."""
import numpy as np
import pytest
from scipy.special import spherical_jn

try:
    from scipy.special import sph_harm_y

    def Y(n, m, az, pol):
        return sph_harm_y(n, m, pol, az)
except ImportError:   # older scipy
    from scipy.special import sph_harm

    def Y(n, m, az, pol):
        return sph_harm(m, n, az, pol)


def dfact(n):
    return float(np.prod(np.arange(1, 2 * n + 2, 2)))


def basis(n, m, x, k, R):
    if n < 0 or abs(m) > n:
        return np.zeros(len(x), complex)
    r = np.linalg.norm(x, axis=1)
    pol, az = np.arccos(x[:, 2] / r), np.arctan2(x[:, 1], x[:, 0])
    jh = np.ones_like(r) if k == 0 else spherical_jn(n, k * r) * dfact(n) / (k * r) ** n
    return jh * (r / R) ** n * Y(n, m, az, pol)


def lo(which, n, m, R):
    num = [(n + m) * (n - m), (n - m - 1) * (n - m), (n + m - 1) * (n + m)][which]
    v = (2 * n + 1) / R * np.sqrt(num / ((2 * n - 1) * (2 * n + 1))) if n >= 1 else 0.0
    return -v if which == 2 else v


def hi(which, n, m, k, R):
    num = [(n + 1 + m) * (n + 1 - m), (n + m + 1) * (n + m + 2), (n - m + 1) * (n - m + 2)][which]
    v = k * k * R / (2 * n + 3) * np.sqrt(num / ((2 * n + 1) * (2 * n + 3)))
    return v if which == 1 else -v


@pytest.mark.parametrize("k", [0.0, 3.0])
def test_gradient_recurrences(k):
    R, h = 0.7, 1e-5
    x = np.random.default_rng(0).normal(size=(30, 3)) * 0.3
    for n in range(0, 7):
        for m in range(-n, n + 1):
            g = []
            for a in range(3):
                e = np.zeros(3)
                e[a] = h
                g.append((basis(n, m, x + e, k, R) - basis(n, m, x - e, k, R)) / (2 * h))
            scale = max(np.abs(basis(n + 1, m, x, k, R)).max(), np.abs(g[2]).max(), 1.0)
            for which, (dm, fd) in enumerate([(0, g[2]), (1, g[0] + 1j * g[1]), (-1, g[0] - 1j * g[1])]):
                pred = lo(which, n, m, R) * basis(n - 1, m + dm, x, k, R) + hi(which, n, m, k, R) * basis(n + 1, m + dm, x, k, R)
                assert np.abs(pred - fd).max() <= 1e-6 * scale, (n, m, which)
