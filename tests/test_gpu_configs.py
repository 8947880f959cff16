"""Full-size GPU tests of the benchmark configurations, each against the FP64 direct sum on
4096 sampled targets (default_rng(0).choice(N, 4096)):

* uniform cube N=1M, k=40, leaf 128: the GPU FP64 direct sum is first pinned to the C
  oracle on 256 of the targets (< 1e-13); the plain FP64 FMM at p=16 and p=17 (p=17 is
  the smallest fixed degree reaching 1e-6), and the mixed form at fixed per-level degrees;
* sphere + clusters N=2M, k=100, depth <= 10 (adaptive tree with W and X lists): the
  learned degrees with the high-frequency floor, mixed and FP64, both within 1e-6;
* uniform + sphere + clusters N=4M, k=60: the learned degrees, mixed and FP64, within 1e-6;
* 256 problems in ellipsoids (N log-uniform in [2k, 200k], k in [1, 50]) in one batch of
  persistent launches, each problem on 512 sampled targets.

Each case prints its name and sizes before touching the GPU.
"""
import sys

import numpy as np
import pytest

from oracle import direct as D

pytestmark = pytest.mark.gpu
TOL = 1e-6
SAMPLES = 4096


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev_eval(hf, st, q_dev, o_dev, idx_dev):
    import torch
    hf.evaluate_device(st, q_dev, o_dev)
    torch.cuda.synchronize()
    return o_dev[idx_dev].cpu().numpy()


def test_uniform_1m_k40():
    import torch
    import hfmm0_amd as hf
    from hfmm0_amd.workloads import sample_targets, uniform_cube
    n, k, leaf = 1_000_000, 40.0, 128
    say(f"case: uniform cube N={n} k={k} leaf={leaf}")
    x, q = uniform_cube(n)
    q = q.astype(np.complex64).astype(np.complex128)   # complex64 charges, as the workload states
    idx = sample_targets(n, SAMPLES)
    ref = hf.direct(x, q, k, idx)
    pin = D.direct_c(x, q, k, idx[:256])
    e_pin = D.relative_error(ref[:256], pin)
    say(f"  GPU FP64 direct vs C oracle on 256 targets: {e_pin:.2e}")
    assert e_pin < 1e-13
    q_dev = torch.from_numpy(q.astype(np.complex64)).cuda()
    o_dev = torch.empty(n, dtype=torch.complex128, device="cuda")
    idx_dev = torch.from_numpy(idx).cuda()
    st = hf.setup(x, k, p=17, N_max=leaf)
    assert st.depth == 5 and st.info.n_leaves == 32768
    e17 = D.relative_error(_dev_eval(hf, st, q_dev, o_dev, idx_dev), ref)
    st.set_degrees([16] * st.depth)
    e16 = D.relative_error(_dev_eval(hf, st, q_dev, o_dev, idx_dev), ref)
    st.close()
    say(f"  FP64 p=17 err {e17:.3e}, p=16 err {e16:.3e}")
    assert e17 <= TOL
    assert e17 < e16 < 20 * TOL
    # mixed precision at fixed per-level degrees (levels 1..5)
    for pl in ([18, 18, 14, 11, 11], [17] * 5):
        st = hf.setup(x, k, p=max(pl), N_max=leaf, p_level=pl, precision="mixed")
        e = D.relative_error(_dev_eval(hf, st, q_dev, o_dev, idx_dev), ref)
        say(f"  mixed {pl} err {e:.3e} (FP32 M2L levels {st.m2l_fp32_levels})")
        assert st.m2l_fp32_levels == [2, 3, 4, 5]
        assert e <= TOL
        st.close()


def _learned_config(make, leaf, max_depth, name):
    import torch
    import hfmm0_amd as hf
    from hfmm0_amd.learned import DegreeMLP, predict_p_level
    from hfmm0_amd.workloads import sample_targets
    x, q, k = make()
    n = len(x)
    say(f"case: {name} N={n} k={k} leaf={leaf} max_depth={max_depth}")
    idx = sample_targets(n, SAMPLES)
    ref = hf.direct(x, q, k, idx)
    net = DegreeMLP.load()
    pl = predict_p_level(net, x, q, k, n_max=leaf, margin=0.0, max_depth=max_depth)
    net.close()
    say(f"  learned degrees {list(pl)}")
    q_dev = torch.from_numpy(q).cuda()
    o_dev = torch.empty(n, dtype=torch.complex128, device="cuda")
    idx_dev = torch.from_numpy(idx).cuda()
    res = {}
    for prec in ("mixed", "fp64"):
        st = hf.setup(x, k, p=max(pl), N_max=leaf, p_level=list(pl), precision=prec, max_depth=max_depth,
                      tree_device=True)
        e = D.relative_error(_dev_eval(hf, st, q_dev, o_dev, idx_dev), ref)
        info = st.info
        say(f"  {prec}: depth {st.depth} err {e:.3e} W {info.n_w} X {info.n_x}")
        res[prec] = (e, st.depth, int(info.n_w), int(info.n_x))
        st.close()
    return pl, res


def test_sphere_clusters_2m_k100():
    from hfmm0_amd.workloads import c3_problem
    pl, res = _learned_config(c3_problem, 128, 10, "sphere + clusters")
    assert 8 <= min(pl[1:]) and max(pl) <= 40
    for prec, (e, depth, nw, nx) in res.items():
        assert depth <= 10 and nw > 0 and nx > 0
        assert e <= TOL, (prec, e)


def test_mixed_distribution_4m_k60():
    from hfmm0_amd.workloads import c5_problem
    pl, res = _learned_config(c5_problem, 128, None, "uniform + sphere + clusters")
    for prec, (e, depth, nw, nx) in res.items():
        assert e <= TOL, (prec, e)


def test_batched_256_ellipsoids():
    """256 problems (uniform in ellipsoids with semi-axes in [0.2, 0.5], N log-uniform in
    [2k, 200k], k in [1, 50], complex64 charges) in one batch, leaf 128, mixed precision:
    degree 12 raised to the high-frequency floor, tables built 4 degrees higher, and
    Batch.certify raising the degrees of every problem that misses 1e-6 on its 512
    sampled targets against the FP64 direct sum.  All 256 must certify; the batch of
    persistent launches equals its own plans evaluated one by one."""
    import torch
    import hfmm0_amd as hf
    from hfmm0_amd.autotune import batch_degrees
    from hfmm0_amd.batch import Batch
    from hfmm0_amd.workloads import batch_problems
    probs = batch_problems(256)
    xs, qs, ks = [x for x, _, _, _ in probs], [q for _, q, _, _ in probs], [k for _, _, k, _ in probs]
    say(f"case: batch of {len(probs)} ellipsoid problems, {sum(len(x) for x in xs)} points, leaf 128, mixed")
    pls = [batch_degrees(x, k) for x, k in zip(xs, ks)]
    tabs = [[min(40, v + 4) for v in pl] for pl in pls]
    B = Batch(xs, ks, p=max(max(t) for t in tabs), N_max=128, p_level=tabs, precision="mixed")
    for i, pl in enumerate(pls):
        st = B.plan(i)
        st.set_degrees(pl[:st.depth])
    phis, cert, errs = B.certify(qs, tol=TOL, n_samp=512, max_steps=6)
    raised = sum(list(c) != list(p)[:len(c)] for c, p in zip(cert, pls))
    worst = int(np.argmax(errs))
    say(f"  certified rel err max {max(errs):.2e} (problem {worst}: n={len(xs[worst])} k={ks[worst]:.1f}) "
        f"median {np.median(errs):.2e}; {raised} problems raised; {sum(e <= TOL for e in errs)}/256 within {TOL:g}")
    qd = [torch.from_numpy(q).cuda() for q in qs]
    od = [torch.empty(len(q), dtype=torch.complex128, device="cuda") for q in qs]
    ms = B.benchmark(qd, od, warmup=1, reps=5)
    say(f"  batch evaluate {ms:.1f} ms: {sum(len(x) for x in xs) / ms * 1e-3:.3e} evals/s (tables 4 degrees up)")
    torch.cuda.synchronize()
    for i in (0, worst, 255):
        np.testing.assert_allclose(od[i].cpu().numpy(), phis[i], rtol=0, atol=1e-12 * np.abs(phis[i]).max())
        st = B.plan(i)
        one = torch.empty_like(od[i])
        hf.evaluate_device(st, qd[i], one)
        torch.cuda.synchronize()
        np.testing.assert_allclose(od[i].cpu().numpy(), one.cpu().numpy(), rtol=0, atol=1e-12 * np.abs(phis[i]).max())
    assert max(errs) <= TOL
    B.close()
