"""GPU test of the autotuner: the planned evaluator certifies the tolerance on sampled
targets, records its configuration in the wisdom file, and a second plan of the same
problem class reuses it without searching."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_plan_tunes_certifies_and_reuses_wisdom(tmp_path, monkeypatch):
    import hfmm0_amd.autotune as A
    from hfmm0_amd.learned import sampled_error
    from hfmm0_amd.workloads import mixed_problem
    monkeypatch.setattr(A, "wisdom_path", lambda: str(tmp_path / "wisdom.json"))
    print("case: autotune plan n=25000 k=15 tol=1e-6", flush=True)
    x, q, _, _ = mixed_problem(25_000, (15.0, 15.0), seed=800_001)
    fmm = A.plan(x, 15.0, tol=1e-6, verbose=True)
    phi = fmm(q)
    assert fmm.tuned
    err = sampled_error(x, q, 15.0, phi, n_samp=1024, seed=3)
    print(f"  config {fmm.config} err {err:.2e}", flush=True)
    assert err <= 2e-6
    w = A.load_wisdom()
    assert fmm.key in w and w[fmm.key]["N_max"] in (32, 64, 128)
    fmm.close()
    x2, q2, _, _ = mixed_problem(25_000, (15.0, 15.0), seed=800_002)
    fmm2 = A.plan(x2, 15.0, tol=1e-6)
    phi2 = fmm2(q2)
    assert not fmm2.tuned and fmm2.key == fmm.key or not fmm2.tuned
    assert fmm2.report[-1][1] <= 1e-6
    fmm2.close()


def test_refine_degrees_certifies_and_is_not_slower():
    """From a generous seed (the network's degrees plus one), the local search returns
    degrees whose error on targets the search never saw meets the tolerance, that were
    measured no slower than the seed, and whose history starts with the seed."""
    import hfmm0_amd.autotune as A
    from hfmm0_amd.learned import DegreeMLP, predict_p_level, sampled_error
    import hfmm0_amd as hf
    from hfmm0_amd.workloads import uniform_cube
    print("case: refine_degrees n=200000 k=20 leaf=64", flush=True)
    x, q = uniform_cube(200_000)
    net = DegreeMLP.load()
    seed = [p + 1 for p in predict_p_level(net, x, q, 20.0, n_max=64, margin=0.0)]
    net.close()
    pl, ms, err, hist = A.refine_degrees(x, q, 20.0, seed, tol=1e-6, leaf=64, reps=5, verbose=True)
    assert hist[0][0] == seed
    assert ms <= hist[0][1]
    assert err <= 0.8e-6
    st = hf.setup(x, 20.0, p=max(pl), N_max=64, p_level=pl, precision="mixed")
    held_out = sampled_error(x, q, 20.0, hf.evaluate(st, q), n_samp=2048, seed=11)
    st.close()
    print(f"  seed {seed} -> {pl}: {hist[0][1]:.3f} -> {ms:.3f} ms, held-out err {held_out:.2e}", flush=True)
    assert held_out <= 1e-6


def test_tune_batch_picks_leaf_and_tile():
    from hfmm0_amd.autotune import tune_batch
    from hfmm0_amd.workloads import batch_problems
    probs = batch_problems(10, n_range=(2_000, 20_000), seed=4100)
    print(f"case: batch autotune, {len(probs)} problems", flush=True)
    leaf, tile, t = tune_batch([p[0] for p in probs], [p[2] for p in probs], [p[1] for p in probs],
                               leaves=(32, 64), tiles=(2, 4), subsample=6, reps=2)
    print(f"  leaf {leaf} tile {tile} timings {t}", flush=True)
    assert leaf in (32, 64) and tile in (0, 2, 4)
    assert t[f"leaf{leaf}_tile{tile}"] == min(v for kk, v in t.items() if kk.startswith(f"leaf{leaf}_"))
