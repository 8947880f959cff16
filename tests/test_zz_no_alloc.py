"""hfmm_direct_noalloc (device buffers, no device allocation, partial sums in a fixed
per-GPU scratch array) against hfmm_direct (host buffers, allocating), through the C ABI,
on two calls in a row, the second with fewer points and targets (the scratch array then
holds the first call's sums, which the second call must clear).  Both are FP64 direct sums
combined with FP64 atomics, so they agree to rounding (relative 1e-13)."""
import ctypes

import numpy as np
import pytest

from oracle import direct as D

pytestmark = pytest.mark.gpu


def _noalloc(x, q, k, idx):
    import torch
    from hfmm0_amd import _lib as L
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    qd = torch.from_numpy(np.ascontiguousarray(q.astype(np.complex128))).cuda()
    td = torch.from_numpy(np.ascontiguousarray(idx.astype(np.int64))).cuda()
    od = torch.full((len(idx),), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    st = L.lib().hfmm_direct_noalloc(ctypes.c_void_p(xd.data_ptr()), len(x), ctypes.c_void_p(qd.data_ptr()),
                                     float(k), ctypes.c_void_p(td.data_ptr()), len(idx),
                                     ctypes.c_void_p(od.data_ptr()), None)
    L.check(st)
    torch.cuda.synchronize()
    return od.cpu().numpy()


def test_noalloc_direct_matches_allocating_direct_twice():
    import hfmm0_amd as hf
    from hfmm0_amd import _lib as L
    for n, nt, k in ((200_000, 4096, 25.0), (30_000, 700, 3.0)):
        print(f"case: direct no-alloc n={n} targets={nt} k={k}", flush=True)
        x, q = D.uniform_problem(n)
        idx = D.sample_targets(n, nt)
        ref = hf.direct(x, q, k, idx)
        got = _noalloc(x, q, k, idx)
        e = D.relative_error(got, ref)
        print(f"  rel diff vs hfmm_direct {e:.2e}", flush=True)
        assert np.isfinite(got).all() and e < 1e-13
    # the limit is enforced before any device work
    assert L.lib().hfmm_direct_noalloc(ctypes.c_void_p(8), 10, ctypes.c_void_p(8), 1.0, ctypes.c_void_p(8),
                                       65537, ctypes.c_void_p(8), None) == L.HFMM_ERR_ARG
    assert b"HFMM_DIRECT_NOALLOC_MAX_TARGETS" in L.lib().hfmm_last_error()
