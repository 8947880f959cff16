"""Oracle (test infrastructure only): FP64 direct Helmholtz sums and the error metric.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker or the timed CPU baseline.

* :func:`direct_numpy` restates ``test.py:11-18`` (256-row chunks, complex128,
  ``exp(ikr)/r`` with the r = 0 term excluded, no 1/(4 pi)); with ``floor`` > 0 it
  excludes every pair with r <= floor instead, the rule of the reference evaluate's near
  field for near-coincident points (``jaxfmm/near.py:252-261``).
* :func:`direct_c` is the same sum in C with OpenMP (``oracle/direct.c``), used for
  larger sample sets.
* :func:`direct_field_numpy` / :func:`direct_field_c`: the field E = -grad phi of the
  same sum, with the reference kernel bundle's field weight (``HFMM.py:29-37``:
  (cos kr + kr sin kr)/r^3 and (sin kr - kr cos kr)/r^3 times (x_t - x_s)).
* :func:`relative_error` restates ``test.py:20-22``.
* :func:`sample_targets` is the sampled-target set: ``default_rng(0).choice(N, n,
  replace=False)`` (seed as ``jaxfmm/tune.py:90``).

Pinned by ``tests/test_oracle_direct.py``: closed-form two- and three-point values,
the k=0 Coulomb sum of a regular polygon, symmetry and linearity identities, and
agreement between the NumPy and C implementations; the field by central differences
of the potential, the k=0 Coulomb field in closed form and Newton's third law.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def direct_numpy(positions, charges, k, targets=None, floor=0.0):
    """phi at ``positions[targets]`` (all points if None) from all sources; pairs with
    r <= floor contribute nothing (floor 0: only r = 0 is excluded, as ``test.py:16``)."""
    positions = np.asarray(positions, dtype=np.float64)
    charges = np.asarray(charges).astype(np.complex128)
    tgt = positions if targets is None else positions[np.asarray(targets)]
    result = np.empty(len(tgt), dtype=np.complex128)
    for start in range(0, len(tgt), 256):
        x = tgt[start:start + 256]
        r = np.linalg.norm(x[:, None, :] - positions[None, :, :], axis=-1)
        ok = r > floor
        w = np.where(ok, np.exp(1j * k * r) / np.where(ok, r, 1.0), 0.0)
        result[start:start + len(x)] = w @ charges
    return result


def direct_field_numpy(positions, charges, k, targets=None):
    """E = -grad phi at ``positions[targets]`` from all sources, (n, 3) complex128."""
    positions = np.asarray(positions, dtype=np.float64)
    charges = np.asarray(charges).astype(np.complex128)
    tgt = positions if targets is None else positions[np.asarray(targets)]
    result = np.empty((len(tgt), 3), dtype=np.complex128)
    for start in range(0, len(tgt), 256):
        x = tgt[start:start + 256]
        d = x[:, None, :] - positions[None, :, :]
        r = np.linalg.norm(d, axis=-1)
        ok = r > 0
        rs = np.where(ok, r, 1.0)
        w = np.where(ok, (1 - 1j * k * rs) * np.exp(1j * k * rs) / rs ** 3, 0.0)
        result[start:start + len(x)] = np.einsum("tsa,ts,s->ta", d, w, charges)
    return result


def build(force=False):
    """Compile ``liboracle_direct.so`` with the committed Makefile."""
    so = os.path.join(_HERE, "liboracle_direct.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return so


def _lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle_direct.so")
        if not os.path.exists(so):
            build()
        lib = ctypes.CDLL(so)
        lib.oracle_direct.restype = None
        lib.oracle_direct.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        lib.oracle_direct_field.restype = None
        lib.oracle_direct_field.argtypes = lib.oracle_direct.argtypes
        _LIB = lib
    return _LIB


def direct_c(positions, charges, k, targets=None):
    """Same sum as :func:`direct_numpy`, in C/OpenMP."""
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    q = np.ascontiguousarray(np.asarray(charges).astype(np.complex128))
    tgt = positions if targets is None else np.ascontiguousarray(positions[np.asarray(targets)])
    out = np.empty(len(tgt), dtype=np.complex128)
    _lib().oracle_direct(positions.ctypes.data, len(positions), q.ctypes.data, float(k),
                         tgt.ctypes.data, len(tgt), out.ctypes.data)
    return out


def direct_field_c(positions, charges, k, targets=None):
    """Same field as :func:`direct_field_numpy`, in C/OpenMP."""
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    q = np.ascontiguousarray(np.asarray(charges).astype(np.complex128))
    tgt = positions if targets is None else np.ascontiguousarray(positions[np.asarray(targets)])
    out = np.empty((len(tgt), 3), dtype=np.complex128)
    _lib().oracle_direct_field(positions.ctypes.data, len(positions), q.ctypes.data, float(k),
                               tgt.ctypes.data, len(tgt), out.ctypes.data)
    return out


def relative_error(actual, reference):
    """``test.py:20-22``: ||a - r|| / ||r|| (1 if ||r|| == 0)."""
    norm = np.linalg.norm(reference)
    return float(np.linalg.norm(np.asarray(actual) - reference) / (norm if norm else 1.0))


def sample_targets(n, n_samp):
    """Sampled-target index set used for every error figure."""
    return np.sort(np.random.default_rng(0).choice(n, min(n_samp, n), replace=False))


def uniform_problem(n, complex_charges=True):
    """Synthetic input of the uniform-cube workloads: seed 42+N, positions uniform in
    [0,1]^3 drawn first, then charges N(0,1) (+ i N(0,1))."""
    rng = np.random.default_rng(42 + n)
    x = rng.uniform(0.0, 1.0, (n, 3))
    q = rng.normal(size=n)
    if complex_charges:
        q = q + 1j * rng.normal(size=n)
    return x, q
