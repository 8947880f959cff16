"""Pins for the direct-sum oracle: closed forms, identities, and C/NumPy agreement."""
import numpy as np
import pytest

from oracle import direct
from oracle import direct as D


def test_two_points_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.4, 0.0]])          # r = 0.5
    q = np.array([2.0, -1.0 + 0.5j])
    k = 3.0
    g = np.exp(1j * k * 0.5) / 0.5
    want = np.array([g * q[1], g * q[0]])
    for f in (direct.direct_numpy, direct.direct_c):
        got = f(x, q, k)
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)


def test_coincident_points_excluded():
    x = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [1.1, 0.2, 0.3]])
    q = np.array([1.0, 5.0, 1.0])
    for f in (direct.direct_numpy, direct.direct_c):
        got = f(x, q, 2.0)
        g = np.exp(2j) / 1.0
        np.testing.assert_allclose(got, [g, g, 6 * g], rtol=1e-15)


def test_near_coincident_points_below_the_floor_excluded():
    # a pair 1e-15 apart: test.py's rule keeps its 1e15 term, the evaluate's near-field
    # floor (r <= floor) drops it; pairs above the floor are unchanged
    x = np.array([[0.1, 0.2, 0.3], [0.1 + 1e-15, 0.2, 0.3], [1.1, 0.2, 0.3]])
    q = np.array([1.0, 5.0, 1.0])
    k = 2.0
    g = np.exp(2j) / 1.0
    got = direct.direct_numpy(x, q, k, floor=1e-12)
    np.testing.assert_allclose(got, [g, g, 6 * g], rtol=1e-14)
    raw = direct.direct_numpy(x, q, k)
    r01 = np.linalg.norm(x[0] - x[1])
    np.testing.assert_allclose(raw[0] - got[0], 5.0 * np.exp(1j * k * r01) / r01, rtol=1e-12)


def test_polygon_coulomb_k0():
    # n unit charges on a unit circle: every point sees sum_{j=1}^{n-1} 1/(2 sin(pi j/n))
    n = 12
    t = 2 * np.pi * np.arange(n) / n
    x = np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1)
    want = sum(1.0 / (2 * np.sin(np.pi * j / n)) for j in range(1, n))
    for f in (direct.direct_numpy, direct.direct_c):
        got = f(x, np.ones(n), 0.0)
        np.testing.assert_allclose(got.real, want, rtol=1e-14)
        assert np.all(got.imag == 0)


def test_polygon_helmholtz():
    n = 9
    k = 4.5
    t = 2 * np.pi * np.arange(n) / n
    x = np.stack([np.cos(t), np.sin(t), np.full(n, 0.7)], 1)
    want = sum(np.exp(1j * k * 2 * np.sin(np.pi * j / n)) / (2 * np.sin(np.pi * j / n)) for j in range(1, n))
    got = direct.direct_c(x, np.ones(n), k)
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_linearity_and_reciprocity():
    rng = np.random.default_rng(7)
    x = rng.uniform(size=(300, 3))
    q1 = rng.normal(size=300) + 1j * rng.normal(size=300)
    q2 = rng.normal(size=300)
    k = 7.0
    a = direct.direct_c(x, q1, k)
    b = direct.direct_c(x, q2, k)
    c = direct.direct_c(x, 2 * q1 - 3 * q2, k)
    np.testing.assert_allclose(c, 2 * a - 3 * b, rtol=1e-12, atol=1e-12)
    # reciprocity: sum_t q2_t phi_1(t) == sum_s q1_s phi_2(s) (symmetric kernel)
    np.testing.assert_allclose(np.dot(q2, a), np.dot(q1, b), rtol=1e-12)


def test_c_matches_numpy_and_samples():
    x, q = direct.uniform_problem(2000)
    idx = direct.sample_targets(2000, 100)
    a = direct.direct_numpy(x, q, 10.0, idx)
    b = direct.direct_c(x, q, 10.0, idx)
    assert direct.relative_error(b, a) < 1e-14
    assert len(np.unique(idx)) == 100


def test_negative_k_is_conjugate():
    x, q = direct.uniform_problem(400)
    a = direct.direct_c(x, q, 5.0)
    b = direct.direct_c(x, np.conj(q), -5.0)
    np.testing.assert_allclose(b, np.conj(a), rtol=1e-13, atol=1e-13)


def test_field_matches_central_differences_of_phi():
    rng = np.random.default_rng(11)
    x = rng.uniform(size=(60, 3))
    q = rng.normal(size=60) + 1j * rng.normal(size=60)
    k, h = 7.0, 1e-6
    E = D.direct_field_numpy(x, q, k, np.arange(5))
    for t in range(5):
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            tp, tm = x[t] + e, x[t] - e

            def phi_at(p):
                r = np.linalg.norm(p[None, :] - np.delete(x, t, 0), axis=1)
                return np.sum(np.delete(q, t) * np.exp(1j * k * r) / r)
            fd = -(phi_at(tp) - phi_at(tm)) / (2 * h)
            assert abs(E[t, a] - fd) <= 1e-6 * abs(fd) + 1e-8


def test_field_coulomb_closed_form_and_newton_third_law():
    x = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    q = np.array([2.0, -0.5])
    E = D.direct_field_numpy(x, q, 0.0)
    d = x[0] - x[1]
    r = np.linalg.norm(d)
    np.testing.assert_allclose(E[0], q[1] * d / r ** 3, rtol=1e-14)
    # forces q_t E_t on the two points are equal and opposite (real kernel at k = 0)
    np.testing.assert_allclose(q[0] * E[0], -q[1] * E[1], rtol=1e-14)
    rng = np.random.default_rng(12)
    y = rng.uniform(size=(300, 3))
    qq = rng.normal(size=300) + 1j * rng.normal(size=300)
    np.testing.assert_allclose(D.direct_field_c(y, qq, 5.0, np.arange(40)),
                               D.direct_field_numpy(y, qq, 5.0, np.arange(40)), rtol=1e-12, atol=1e-12)
