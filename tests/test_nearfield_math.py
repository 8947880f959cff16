"""Error bounds of the near-field weight arithmetic (hfmm0_amd/csrc/p2p_math.h), built for
the host from the same header the GPU kernels include, against NumPy FP64 on dense grids.

The bounds: the Remez polynomials' approximation error (4.9e-8 sin, 1.8e-8 cos, printed by
hfmm0_amd/csrc/sincos_remez.py) plus the binary32 roundings of the Horner steps (sin(pi y)
reaches 1.7e-7 at |y| = 1/2, where the polynomial's value is near pi / 2 and its last
rounding is half an ulp of 1.6); a weight's relative error
adds the rounding of its phase argument in FP32 (about u 2^-24 for u = k r / pi, or k r
2^-24 in radians), which the FP64-phase forms remove.  A wrong coefficient, sign or
reduction step fails these by orders of magnitude.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "nearfield_host.cpp")


@pytest.fixture(scope="module")
def nf():
    d = tempfile.mkdtemp(prefix="nf_")
    so = os.path.join(d, "libnf.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", so],
                   check=True)
    return ctypes.CDLL(so)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _call(fn, *arrays, scalars=()):
    fn(*[_p(a) if isinstance(a, np.ndarray) else a for a in arrays])


def test_sincospi_half(nf):
    y = np.linspace(-0.5, 0.5, 200001).astype(np.float32)
    s = np.empty_like(y)
    c = np.empty_like(y)
    nf.nf_sincospi_half(_p(y), ctypes.c_int(len(y)), _p(s), _p(c))
    yd = y.astype(np.float64)
    es = np.max(np.abs(s - np.sin(np.pi * yd)))
    ec = np.max(np.abs(c - np.cos(np.pi * yd)))
    assert es < 2e-7 and ec < 1.2e-7, (es, ec)
    # odd / even symmetry and the exact values at 0
    assert np.array_equal(s[::-1], -s) and np.array_equal(c[::-1], c)
    i0 = len(y) // 2
    assert s[i0] == 0.0 and c[i0] == 1.0


def test_halfturn_weight(nf):
    # u = k r / pi over the range of the FP32-phase near field (k r <= 8 -> u <= 2.6) and beyond
    u = np.concatenate([np.geomspace(1e-4, 1.0, 50000), np.linspace(1.0, 30.0, 200000)])
    u2 = (u * u).astype(np.float32)
    wr = np.empty_like(u2)
    wi = np.empty_like(u2)
    nf.nf_halfturn(_p(u2), ctypes.c_int(len(u2)), _p(wr), _p(wi))
    ue = np.sqrt(u2.astype(np.float64))       # the u the function sees
    w = (wr + 1j * wi) * ue
    ref = np.exp(1j * np.pi * ue)
    err = np.abs(w - ref)
    bound = 3e-7 + 1.5 * np.pi * ue * 2.0 ** -23
    assert np.all(err <= bound), float(np.max(err / bound))
    assert np.max(err[ue <= 8 / np.pi]) < 9e-7


def test_halfturn_weight_phase64(nf):
    u = np.concatenate([np.geomspace(1e-3, 1.0, 20000), np.linspace(1.0, 5000.0, 400000)])
    u2 = u * u
    wr = np.empty(len(u), np.float32)
    wi = np.empty(len(u), np.float32)
    nf.nf_halfturn_phase64(_p(u2), ctypes.c_int(len(u)), _p(wr), _p(wi))
    w = (wr.astype(np.float64) + 1j * wi) * u
    err = np.abs(w - np.exp(1j * np.pi * u))
    # the phase is exact to FP64 rounding: the error does not grow with u
    assert np.max(err) < 3e-7, float(np.max(err))


@pytest.mark.parametrize("ph", [0, 1])
def test_pair_weight_floor(nf, ph):
    # at or below the floor (r = 0 included) the weight is exactly zero, never NaN; above
    # it the weight is the half-turn form's
    fl2 = 1e-20
    r2 = np.array([0.0, 1e-300, 1e-40, 1e-21, fl2, 1.000001e-20, 1e-10, 0.25, 2.0, 50.0])
    wr = np.empty(len(r2), np.float32)
    wi = np.empty(len(r2), np.float32)
    nf.nf_pair_weight(_p(r2), ctypes.c_int(len(r2)), ctypes.c_double(fl2), ctypes.c_int(ph), _p(wr), _p(wi))
    assert np.all(np.isfinite(wr)) and np.all(np.isfinite(wi))
    below = r2 <= fl2
    assert np.all(wr[below] == 0) and np.all(wi[below] == 0)
    u = np.sqrt(r2[~below])
    w = (wr[~below] + 1j * wi[~below].astype(np.float64)) * u
    assert np.max(np.abs(w - np.exp(1j * np.pi * u))) < 3e-6


def test_sincos_halfpi(nf):
    y = np.linspace(-np.pi / 2, np.pi / 2, 200001).astype(np.float32)
    s = np.empty_like(y)
    c = np.empty_like(y)
    nf.nf_sincos_halfpi(_p(y), ctypes.c_int(len(y)), _p(s), _p(c))
    yd = y.astype(np.float64)
    es = np.max(np.abs(s - np.sin(yd)))
    ec = np.max(np.abs(c - np.cos(yd)))
    assert es < 1.2e-7 and ec < 1.6e-7, (es, ec)


@pytest.mark.parametrize("k", [1.0, 40.0, 100.0])
def test_radian_weight(nf, k):
    kr = np.linspace(1e-3, 60.0, 300000)
    r2 = ((kr / k) ** 2).astype(np.float32)
    wr = np.empty_like(r2)
    wi = np.empty_like(r2)
    nf.nf_radian(_p(r2), ctypes.c_int(len(r2)), ctypes.c_float(k), _p(wr), _p(wi))
    r = np.sqrt(r2.astype(np.float64))
    w = (wr + 1j * wi) * r
    err = np.abs(w - np.exp(1j * np.float32(k) * r))
    bound = 3e-7 + 4 * (np.float32(k) * r) * 2.0 ** -24
    assert np.all(err <= bound), float(np.max(err / bound))


def test_radian_weight_phase64(nf):
    k = 100.0
    kr = np.linspace(1e-3, 2000.0, 400000)
    r2 = (kr / k) ** 2
    n = len(r2)
    wr, wi, rinv, krf = (np.empty(n, np.float32) for _ in range(4))
    nf.nf_radian_phase64(_p(r2), ctypes.c_int(n), ctypes.c_double(k), _p(wr), _p(wi), _p(rinv), _p(krf))
    r = np.sqrt(r2)
    err = np.abs((wr + 1j * wi.astype(np.float64)) * r - np.exp(1j * k * r))
    assert np.max(err) < 3e-7, float(np.max(err))
    assert np.max(np.abs(rinv * r - 1)) < 2.5e-7
    assert np.max(np.abs(krf / (k * r) - 1)) < 1.2e-7
