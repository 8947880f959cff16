"""Batches of independent problems evaluated together (libhfmm ``hfmm_batch_*``).

One plan per problem.  An evaluate is one captured hipGraph: the upward and downward
passes of the problems forked over a few streams, and the matrix-core M2L and the near
field of all problems as persistent launches whose resident workgroups take (problem,
item) work from device queues, so that many small problems fill the GPU together.  The reference has no batch API:
its tuner driver (``jaxfmm/tune.py:32-172``) loops over problems.
"""
import ctypes

import numpy as np

from . import _lib as L
from .api import HFMMState, _check_problem, device_charge_type, make_opts


class Batch:
    """``Batch(positions_list, k_list, p=..., N_max=..., precision=...)``; every keyword of
    :func:`hfmm0_amd.setup` is accepted and applies to all problems (``p_level`` may be
    a list of per-problem lists)."""

    def __init__(self, positions_list, k_list, p=4, N_max=16, max_depth=None, *, precision="fp64",
                 p_level=None, m2l_kernel="auto", p2p_tile=0):
        if len(positions_list) != len(k_list) or not positions_list:
            raise ValueError("need one k per problem and at least one problem")
        xs, ks, opts = [], [], (L.Opts * len(positions_list))()
        for i, (x, k) in enumerate(zip(positions_list, k_list)):
            x, k = _check_problem(x, k)
            pl = p_level[i] if (p_level is not None and np.ndim(p_level[0]) > 0) else p_level
            opts[i] = make_opts(p, N_max, max_depth, precision=precision, p_level=pl, m2l_kernel=m2l_kernel,
                                p2p_tile=p2p_tile, _k=k)
            xs.append(x)
            ks.append(float(k))
        self._xs = xs                      # keep the position arrays alive during create
        ptrs = (ctypes.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        ns = np.array([len(x) for x in xs], dtype=np.int64)
        karr = np.array(ks, dtype=np.float64)
        h = ctypes.c_void_p()
        L.check(L.lib().hfmm_batch_create(len(xs), ptrs, L.ptr(ns), L.ptr(karr), opts, ctypes.byref(h)))
        self._h = h
        self.n = [int(v) for v in ns]
        self.k = ks

    def __len__(self):
        return len(self.n)

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("the batch has been released")
        return self._h

    def plan(self, i):
        """State view of problem i (owned by the batch; valid while the batch lives)."""
        h = ctypes.c_void_p()
        L.check(L.lib().hfmm_batch_plan(self.handle, int(i), ctypes.byref(h)))
        st = HFMMState(h, self.n[i], self.k[i], None)
        st.close = lambda: None            # the batch owns it
        return st

    def evaluate(self, charges_list):
        """phi of every problem (complex128, input order), host arrays in and out."""
        if len(charges_list) != len(self):
            raise ValueError("need one charge vector per problem")
        qs, complex_ = [], any(np.iscomplexobj(q) for q in charges_list)
        for q, n in zip(charges_list, self.n):
            q = np.asarray(q)
            if q.shape != (n,):
                raise ValueError("charges must be a real array of shape (N,)")
            qs.append(np.ascontiguousarray(q, dtype=np.complex128 if complex_ else np.float64))
        outs = [np.empty(n, dtype=np.complex128) for n in self.n]
        qp = (ctypes.c_void_p * len(qs))(*[q.ctypes.data for q in qs])
        op = (ctypes.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        L.check(L.lib().hfmm_batch_evaluate(self.handle, qp, L.C128 if complex_ else L.F64, op, L.HOST, None))
        return outs

    def _dev_ptrs(self, q_devs, out_devs):
        if len(q_devs) != len(self) or len(out_devs) != len(self):
            raise ValueError("need one charge vector and one output per problem")
        types = {device_charge_type(q, o, n) for q, o, n in zip(q_devs, out_devs, self.n)}
        if len(types) != 1:
            raise ValueError("all charge vectors of a batch must have the same dtype")
        qp = (ctypes.c_void_p * len(self))(*[q.data_ptr() for q in q_devs])
        op = (ctypes.c_void_p * len(self))(*[o.data_ptr() for o in out_devs])
        return qp, op, types.pop()

    def evaluate_device(self, q_devs, out_devs):
        """All problems with device-resident torch tensors (one graph replay)."""
        import torch
        qp, op, qt = self._dev_ptrs(q_devs, out_devs)
        stream = torch.cuda.current_stream().cuda_stream
        L.check(L.lib().hfmm_batch_evaluate(self.handle, qp, qt, op, L.DEVICE, ctypes.c_void_p(stream)))
        return out_devs

    def benchmark(self, q_devs, out_devs, warmup=2, reps=10):
        """Median milliseconds of one batch evaluate (hipEvents around graph replays)."""
        qp, op, qt = self._dev_ptrs(q_devs, out_devs)
        med = ctypes.c_double()
        L.check(L.lib().hfmm_batch_benchmark(self.handle, qp, qt, op, int(warmup), int(reps), ctypes.byref(med)))
        return med.value

    def benchmark_phases(self, q_devs, out_devs, reps=5):
        """Mean milliseconds of the parts of one batch evaluate run eagerly: upward passes,
        persistent M2L, downward passes, persistent near field."""
        qp, op, qt = self._dev_ptrs(q_devs, out_devs)
        ph = (ctypes.c_double * 4)()
        L.check(L.lib().hfmm_batch_benchmark_phases(self.handle, qp, qt, op, int(reps), ph))
        return dict(zip(("up", "m2l", "down", "p2p"), list(ph)))

    def certify(self, charges_list, tol=1e-6, n_samp=256, seed=0, max_steps=4):
        """Evaluate and certify every problem on n_samp sampled targets against the GPU FP64
        direct sum; a problem that misses tol gets one more degree at every level (within
        its table degrees, through ``set_degrees``: no new setup) and the batch is evaluated
        again, at most max_steps times -- the batch counterpart of the planned evaluator's
        climb (the role of the reference's ``jaxfmm/tune.py:375-516`` ladder).  Returns
        (phi list, per-problem degrees of levels 1..depth, per-problem errors)."""
        from .learned import sampled_error
        plans = [self.plan(i) for i in range(len(self))]
        phis = self.evaluate(charges_list)
        errs = [sampled_error(x, q, k, phi, n_samp, seed)
                for x, q, k, phi in zip(self._xs, charges_list, self.k, phis)]
        for _ in range(max_steps):
            bad = [i for i, e in enumerate(errs) if e > tol]
            raised = []
            for i in bad:
                cur, tab = plans[i].p_level, plans[i].p_table
                new = [min(c + 1, t) for c, t in zip(cur, tab)]
                if new != cur:
                    plans[i].set_degrees(new)
                    raised.append(i)
            if not raised:
                break
            phis = self.evaluate(charges_list)
            for i in raised:
                errs[i] = sampled_error(self._xs[i], charges_list[i], self.k[i], phis[i], n_samp, seed)
        return phis, [pl.p_level for pl in plans], errs

    def close(self):
        if self._h is not None:
            L.lib().hfmm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
