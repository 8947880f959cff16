"""Host FP64 operator tables of libhfmm, pinned without a GPU.

* jhat/hhat against scipy's spherical Bessel functions;
* rotation blocks against scipy's spherical harmonics: Y(R_y(theta) u) = E Y(u), and E
  orthogonal;
* rotate -> coaxial -> rotate-back chains built from the library's own tables reproduce
  direct sums: M2L (multipole of a box evaluated in a well-separated box), M2M
  (shifted multipole equals the multipole formed directly about the parent centre) and
  L2L, at k = 0 (the 1/r kernel), k = 3 and k = 20.  Expansions on the NumPy side use
  scipy.special.sph_harm_y, independent of the library's Legendre code.
"""
import ctypes

import numpy as np
import pytest
import scipy.special as sp

from hfmm0_amd import _lib as L


def dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def jhat(x, p):
    out = np.empty(p + 1)
    L.check(L.lib().hfmm_host_jhat(float(x), p, L.ptr(out)))
    return out


def hhat(x, p):
    re, im = np.empty(p + 1), np.empty(p + 1)
    L.check(L.lib().hfmm_host_hhat(float(x), p, L.ptr(re), L.ptr(im)))
    return re + 1j * im


def rot_offset(n):
    return n * (2 * n - 1) * (2 * n + 1) // 3


def rotation(theta, p):
    out = np.empty(rot_offset(p + 1))
    L.check(L.lib().hfmm_host_rotation(float(theta), p, L.ptr(out)))
    return [out[rot_offset(n):rot_offset(n + 1)].reshape(2 * n + 1, 2 * n + 1) for n in range(p + 1)]


def coax(kind, k, t, rs, rt, ps, pt):
    size = sum((pt + 1 - m) * (ps + 1 - m) for m in range(min(ps, pt) + 1))
    out = np.empty(2 * size)
    L.check(L.lib().hfmm_host_coax(kind, float(k), float(t), float(rs), float(rt), ps, pt, L.ptr(out)))
    c = out[0::2] + 1j * out[1::2]
    blocks, pos = [], 0
    for m in range(min(ps, pt) + 1):
        sz = (pt + 1 - m) * (ps + 1 - m)
        blocks.append(c[pos:pos + sz].reshape(pt + 1 - m, ps + 1 - m))
        pos += sz
    return blocks


def idx(n, m):
    return n * n + n + m


def ylm_all(v, p):
    r = np.linalg.norm(v)
    th = np.arccos(np.clip(v[2] / r, -1, 1)) if r > 0 else 0.0
    ph = np.arctan2(v[1], v[0])
    return np.array([sp.sph_harm_y(n, m, th, ph) for n in range(p + 1) for m in range(-n, n + 1)])


def radial_in(r, k, R, p):
    return jhat(k * r, p) * (r / R) ** np.arange(p + 1)


def radial_out(r, k, R, p):
    n = np.arange(p + 1)
    return 4j * np.pi * hhat(k * r, p) * (R / r) ** n / ((2 * n + 1) * r)


def per_coef(v, p):
    return np.concatenate([np.full(2 * n + 1, v[n]) for n in range(p + 1)])


def p2m(pts, q, c, k, R, p):
    return sum(qq * per_coef(radial_in(np.linalg.norm(y - c), k, R, p), p) * np.conj(ylm_all(y - c, p))
               for y, qq in zip(pts, q))


def p2l(pts, q, c, k, R, p):
    return sum(qq * per_coef(radial_out(np.linalg.norm(y - c), k, R, p), p) * np.conj(ylm_all(y - c, p))
               for y, qq in zip(pts, q))


def l2p(Lc, x, c, k, R, p):
    return np.sum(Lc * per_coef(radial_in(np.linalg.norm(x - c), k, R, p), p) * ylm_all(x - c, p))


def direct(pts, q, x, k):
    r = np.linalg.norm(pts - x, axis=1)
    return np.sum(q * np.exp(1j * k * r) / r)


def translate(vec, t, kind, k, rs, rt, ps, pt):
    """rotate (t -> +z), coaxial, rotate back, as the GPU kernels do."""
    r = np.linalg.norm(t)
    th = np.arccos(np.clip(t[2] / r, -1, 1))
    ph = np.arctan2(t[1], t[0])
    E = rotation(th, max(ps, pt))
    T = coax(kind, k, r, rs, rt, ps, pt)
    a = np.zeros((ps + 1) ** 2, complex)
    for n in range(ps + 1):
        ms = np.arange(-n, n + 1)
        a[n * n:(n + 1) ** 2] = E[n].T @ (vec[n * n:(n + 1) ** 2] * np.exp(1j * ms * ph))
    b = np.zeros((pt + 1) ** 2, complex)
    for m in range(-min(ps, pt), min(ps, pt) + 1):
        am = abs(m)
        src = np.array([a[idx(n, m)] for n in range(am, ps + 1)])
        res = T[am] @ src
        for i, n in enumerate(range(am, pt + 1)):
            b[idx(n, m)] = res[i]
    out = np.zeros_like(b)
    for n in range(pt + 1):
        ms = np.arange(-n, n + 1)
        out[n * n:(n + 1) ** 2] = (E[n] @ b[n * n:(n + 1) ** 2]) * np.exp(-1j * ms * ph)
    return out


@pytest.mark.parametrize("x", [1e-6, 1e-3, 0.3, 1.0, 5.0, 20.0, 43.0, 120.0])
def test_bessel(x):
    p = 36
    n = np.arange(p + 1)
    jn, yn = sp.spherical_jn(n, x), sp.spherical_yn(n, x)
    df1 = np.array([dfact(2 * j + 1) for j in n])
    dfm = np.array([dfact(2 * j - 1) for j in n])
    ok = np.abs(jn) > 1e-280
    ref_j = jn * df1 / x ** n
    np.testing.assert_allclose(jhat(x, p)[ok], ref_j[ok], rtol=2e-12, atol=0)
    ref_h = (jn + 1j * yn) * x ** (n + 1) / dfm
    fin = np.isfinite(ref_h)
    np.testing.assert_allclose(hhat(x, p)[fin], ref_h[fin], rtol=2e-12)
    # k -> 0 limits
    np.testing.assert_array_equal(jhat(0.0, 5), np.ones(6))
    np.testing.assert_allclose(hhat(0.0, 5), -1j * np.ones(6), rtol=0, atol=0)


@pytest.mark.parametrize("theta", [0.0, 0.3, np.arccos(1 / np.sqrt(3)), 1.9, np.pi])
def test_rotation_against_scipy(theta):
    p = 30
    E = rotation(theta, p)
    rng = np.random.default_rng(3)
    Ry = np.array([[np.cos(theta), 0, np.sin(theta)], [0, 1, 0], [-np.sin(theta), 0, np.cos(theta)]])
    for _ in range(3):
        u = rng.normal(size=3)
        Yr, Yu = ylm_all(Ry @ u, p), ylm_all(u, p)
        for n in range(p + 1):
            np.testing.assert_allclose(Yr[n * n:(n + 1) ** 2], E[n] @ Yu[n * n:(n + 1) ** 2], atol=2e-12)
    for n in range(p + 1):
        np.testing.assert_allclose(E[n] @ E[n].T, np.eye(2 * n + 1), atol=1e-12)


@pytest.mark.parametrize("k", [0.0, 3.0, 20.0])
def test_translation_chains(k):
    rng = np.random.default_rng(5)
    p = 14
    h = 0.05
    R = np.sqrt(3) * h
    cs = np.array([0.3, 0.3, 0.3])
    pts = cs + rng.uniform(-h, h, (12, 3))
    q = rng.normal(size=12) + 1j * rng.normal(size=12)
    tol = {0.0: 1e-10, 3.0: 1e-9, 20.0: 1e-6}[k]
    # M2L over displacement (2,-3,1) boxes
    ct = cs + np.array([2, -3, 1]) * 2 * h
    M = p2m(pts, q, cs, k, R, p)
    Lt = translate(M, ct - cs, 0, k, R, R, p, p)
    for x in ct + rng.uniform(-h, h, (3, 3)):
        d = direct(pts, q, x, k)
        assert abs(l2p(Lt, x, ct, k, R, p) - d) / abs(d) < tol
    # M2M child -> parent, different degrees at the two levels
    hc = h / 2
    Rc = np.sqrt(3) * hc
    cc = cs + np.array([1, -1, 1]) * hc
    ptc = cc + rng.uniform(-hc, hc, (8, 3))
    qc = rng.normal(size=8)
    Mc = p2m(ptc, qc, cc, k, Rc, p + 2)
    Mp = translate(Mc, cs - cc, 1, k, Rc, R, p + 2, p)
    ref = p2m(ptc, qc, cs, k, R, p)
    assert np.max(np.abs(Mp - ref)) / np.max(np.abs(ref)) < 50 * tol
    # L2L parent -> child
    far = ct + rng.uniform(-h, h, (10, 3))
    qf = rng.normal(size=10)
    Lp = p2l(far, qf, cs, k, R, p)
    Lc = translate(Lp, cc - cs, 2, k, R, Rc, p, p + 2)
    for x in cc + rng.uniform(-hc, hc, (3, 3)):
        d = direct(far, qf, x, k)
        assert abs(l2p(Lc, x, cc, k, Rc, p + 2) - d) / abs(d) < tol
