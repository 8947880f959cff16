"""GPU tests of the batch API: every problem of a batch gives what its own plan gives,
through host and device pointers and across graph replays, and meets the error of a
single-problem evaluate against the FP64 direct sum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(m, seed0):
    from hfmm0_amd.workloads import mixed_problem
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(m):
        n = int(rng.integers(2_000, 30_000))
        x, q, k, _ = mixed_problem(n, (1.0, 30.0), seed=seed0 + i)
        out.append((x, q, k))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_single_plans(precision):
    import torch
    import hfmm0_amd as hf
    from hfmm0_amd.batch import Batch
    from hfmm0_amd.learned import sampled_error
    probs = _problems(7, 700)
    print(f"case: batch of {len(probs)} problems, {precision}", flush=True)
    B = Batch([x for x, _, _ in probs], [k for _, _, k in probs], p=12, N_max=48, precision=precision)
    outs = B.evaluate([q for _, q, _ in probs])
    for i, (x, q, k) in enumerate(probs):
        st = hf.setup(x, k, p=12, N_max=48, precision=precision)
        ref = hf.evaluate(st, q)
        np.testing.assert_allclose(outs[i], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        assert B.plan(i).p_level == st.p_level
        e_b, e_s = sampled_error(x, q, k, outs[i], n_samp=200), sampled_error(x, q, k, ref, n_samp=200)
        print(f"  problem {i}: n={len(x)} k={k:.1f} err {e_b:.2e}", flush=True)
        assert e_b <= 1.001 * e_s + 1e-12 and e_b < 1e-4
        st.close()
    # device pointers, three replays of the captured graph, eager work in between
    qd = [torch.from_numpy(q).cuda() for _, q, _ in probs]
    od = [torch.empty(len(q), dtype=torch.complex128, device="cuda") for _, q, _ in probs]
    for rep in range(3):
        B.evaluate_device(qd, od)
        torch.cuda.synchronize()
        for i in range(len(probs)):
            got = od[i].cpu().numpy()
            np.testing.assert_allclose(got, outs[i], rtol=0, atol=1e-12 * np.abs(outs[i]).max())
        _ = (qd[0] * 2).sum().item()
    ms = B.benchmark(qd, od, warmup=1, reps=3)
    print(f"  batch evaluate {ms:.2f} ms for {sum(B.n)} points", flush=True)
    assert ms > 0
    B.close()


def test_batch_argument_errors():
    from hfmm0_amd.batch import Batch
    x = np.random.default_rng(0).uniform(size=(100, 3))
    with pytest.raises(ValueError):
        Batch([x], [1.0, 2.0])
    with pytest.raises(ValueError):
        Batch([x[:, :2]], [1.0])
    B = Batch([x, x], [1.0, 2.0], p=4)
    with pytest.raises(ValueError):
        B.evaluate([np.ones(100)])
    with pytest.raises(ValueError):
        B.evaluate([np.ones(100), np.ones(99)])
    B.close()


def test_batch_recaptures_after_set_degrees():
    """set_degrees on a member plan makes the batch capture its graph again: the next
    batch evaluate equals a batch built at the new degrees, and the other problem is
    unchanged (device pointers, so that the stale graph would otherwise be replayed)."""
    import torch
    from hfmm0_amd.batch import Batch
    probs = _problems(2, 900)
    xs, qs, ks = [x for x, _, _ in probs], [q for _, q, _ in probs], [k for _, _, k in probs]
    print(f"case: batch recapture, n={[len(x) for x in xs]}", flush=True)
    B = Batch(xs, ks, p=12, N_max=48)
    qd = [torch.from_numpy(q).cuda() for q in qs]
    od = [torch.empty(len(q), dtype=torch.complex128, device="cuda") for q in qs]
    B.evaluate_device(qd, od)
    torch.cuda.synchronize()
    before = [o.cpu().numpy() for o in od]
    st = B.plan(1)
    low = [8] * st.depth
    st.set_degrees(low)
    B.evaluate_device(qd, od)
    torch.cuda.synchronize()
    after = [o.cpu().numpy() for o in od]
    R = Batch(xs, ks, p=12, N_max=48, p_level=[[12] * B.plan(0).depth, low])
    ref = R.evaluate(qs)
    np.testing.assert_allclose(after[0], before[0], rtol=0, atol=1e-12 * np.abs(before[0]).max())
    np.testing.assert_allclose(after[1], ref[1], rtol=0, atol=1e-12 * np.abs(ref[1]).max())
    assert np.abs(after[1] - before[1]).max() > 1e-9 * np.abs(before[1]).max()
    R.close()
    B.close()


def test_batch_certify_raises_the_missing_problem():
    """Batch.certify raises the degrees of the problem that misses the tolerance (within
    its table degrees) until it meets it, and leaves the others."""
    from hfmm0_amd.batch import Batch
    probs = _problems(3, 950)
    xs, qs, ks = [x for x, _, _ in probs], [q for _, q, _ in probs], [k for _, _, k in probs]
    print(f"case: batch certify, n={[len(x) for x in xs]} k={[round(k, 1) for k in ks]}", flush=True)
    B = Batch(xs, ks, p=14, N_max=48)
    st = B.plan(2)
    st.set_degrees([5] * st.depth)
    _, pls, errs = B.certify(qs, tol=1e-5, n_samp=200)
    print(f"  degrees {pls} errors {[f'{e:.1e}' for e in errs]}", flush=True)
    assert all(e <= 1e-5 for e in errs)
    assert pls[0] == [14] * B.plan(0).depth and pls[1] == [14] * B.plan(1).depth
    assert 5 < pls[2][0] <= 14
    B.close()
