"""Full-size GPU tests of the benchmark configurations, each against the FP64 direct sum on
4096 sampled targets (default_rng(0).choice(N, 4096)):

* uniform cube N=1M, k=40, leaf 128: the GPU FP64 direct sum is first pinned to the C
  oracle on 256 of the targets (< 1e-13); the plain FP64 FMM at p=16 and p=17 (p=17 is
  the smallest fixed degree reaching 1e-6), and the mixed form at fixed per-level degrees;
* sphere + clusters N=2M, k=100, depth <= 10 (adaptive tree with W and X lists): the
  learned degrees with the high-frequency floor, mixed and FP64, both within 1e-6;
* uniform + sphere + clusters N=4M, k=60: the learned degrees, mixed and FP64, within 1e-6.

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
