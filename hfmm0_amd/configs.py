"""Runs the adaptive parity configurations at full size on one GPU and prints one JSON line:
the tuned form (mixed precision, learned degrees, optionally the certified local search)
certified against the FP64 direct sum on 4096 sampled targets, beside the FP64 form at the
same degrees.

    python -m hfmm0_amd.configs C3      # 2M sphere + clusters, k = 100
    python -m hfmm0_amd.configs C5      # 4M uniform + sphere + clusters, k = 60
    python -m hfmm0_amd.configs C4      # 256 problems, N in [2k, 200k], k in [1, 50]: one batch
"""
import argparse
import json
import time

import numpy as np

from . import api
from .learned import DegreeMLP, predict_p_level
from .workloads import c3_problem, c5_problem, mixed_problem, sample_targets

CONFIGS = {"C3": (c3_problem, 128, 10), "C5": (c5_problem, 128, None)}   # (problem, leaf, max_depth)


def run(name, n=None, refine=False, tol=1e-6, samples=4096, reps=10):
    import torch
    make, leaf, max_depth = CONFIGS[name]
    x, q, k = make(n) if n else make()
    n = len(x)
    idx = sample_targets(n, samples)
    t0 = time.perf_counter()
    ref = api.direct(x, q, k, idx)
    direct_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    net = DegreeMLP.load()
    pl = predict_p_level(net, x, q, k, n_max=leaf, margin=0.0, max_depth=max_depth)
    net.close()
    predict_s = time.perf_counter() - t0
    out = {"config": name, "n": n, "k": k, "leaf": leaf, "learned_p_level": [int(v) for v in pl],
           "predict_s": predict_s, "direct_s": direct_s}
    if refine:
        from .autotune import refine_degrees
        t0 = time.perf_counter()
        pl, _, _, hist = refine_degrees(x, q, k, pl, tol=tol, leaf=leaf)
        out["refine"] = {"p_level": [int(v) for v in pl], "configurations": len(hist),
                         "seconds": time.perf_counter() - t0}
    q_dev = torch.from_numpy(q).cuda()
    o_dev = torch.empty(n, dtype=torch.complex128, device="cuda")
    for prec in ("mixed", "fp64"):
        t0 = time.perf_counter()
        st = api.setup(x, k, p=max(pl), N_max=leaf, p_level=list(pl), precision=prec, max_depth=max_depth,
                       tree_device=True)
        setup_s = time.perf_counter() - t0
        api.evaluate_device(st, q_dev, o_dev)
        torch.cuda.synchronize()
        phi = o_dev[torch.from_numpy(idx).cuda()].cpu().numpy()
        err = float(np.linalg.norm(phi - ref) / np.linalg.norm(ref))
        ms = api.benchmark(st, q_dev, o_dev, warmup=2, reps=reps)
        out[prec] = {"ms": ms, "evals_per_s": n / (ms * 1e-3), "rel_err": err, "setup_s": setup_s,
                     "depth": st.depth, "p_level": list(st.p_level), "certified": err <= tol}
        st.close()
    return out


def c4_problems(n_problems=256, n_range=(2_000, 200_000), k_range=(1.0, 50.0), seed=4000):
    """The batched configuration: problem i has N log-uniform in n_range (default_rng(seed)
    draws all N first) and the geometry, k and charges of mixed_problem(N, k_range,
    seed + 1 + i)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_problems)).astype(int)
    return [mixed_problem(int(n), k_range, seed + 1 + i) for i, n in enumerate(ns)]


def run_c4(n_problems=256, leaf=64, tol=1e-6, samples=256, reps=10, headroom=2):
    """All problems in one batch (mixed precision, learned degrees raised to the
    high-frequency floor).  The tables are built `headroom` degrees above the prediction;
    Batch.certify evaluates at the predicted degrees, checks every problem on `samples`
    targets against the FP64 direct sum and raises the degrees of those that miss `tol`.
    The batch is then built again at the certified degrees and its evaluate timed against
    the sum of the problems evaluated one by one."""
    import torch
    from .batch import Batch
    t0 = time.perf_counter()
    probs = c4_problems(n_problems)
    gen_s = time.perf_counter() - t0
    xs, qs, ks = [x for x, _, _, _ in probs], [q for _, q, _, _ in probs], [k for _, _, k, _ in probs]
    t0 = time.perf_counter()
    net = DegreeMLP.load()
    pls = [predict_p_level(net, x, q, k, n_max=leaf, margin=0.0) for x, q, k in zip(xs, qs, ks)]
    net.close()
    predict_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tabs = [[min(40, v + headroom) for v in pl] for pl in pls]
    batch = Batch(xs, ks, p=max(max(t) for t in tabs), N_max=leaf, p_level=tabs, precision="mixed")
    for i, pl in enumerate(pls):
        st = batch.plan(i)
        st.set_degrees(list(pl)[:st.depth])
    _, certified_pl, errs = batch.certify(qs, tol=tol, n_samp=samples)
    certify_s = time.perf_counter() - t0
    raised = sum(list(c) != list(p)[:len(c)] for c, p in zip(certified_pl, pls))
    batch.close()
    t0 = time.perf_counter()
    batch = Batch(xs, ks, p=max(max(pl) for pl in certified_pl), N_max=leaf, p_level=certified_pl,
                  precision="mixed")
    setup_s = time.perf_counter() - t0
    q_dev = [torch.from_numpy(q).cuda() for q in qs]
    o_dev = [torch.empty(len(q), dtype=torch.complex128, device="cuda") for q in qs]
    ms = batch.benchmark(q_dev, o_dev, warmup=2, reps=reps)
    torch.cuda.synchronize()
    phi = [o.cpu().numpy() for o in o_dev]
    from .learned import sampled_error
    errs2 = [sampled_error(x, q, k, f, samples, 0) for x, q, k, f in zip(xs, qs, ks, phi)]
    seq = sum(api.benchmark(batch.plan(i), q_dev[i], o_dev[i], warmup=1, reps=3) for i in range(len(probs)))
    n_total = sum(len(q) for q in qs)
    kinds = [kind for _, _, _, kind in probs]
    out = {"config": "C4", "problems": len(probs), "points": n_total, "leaf": leaf,
           "n_min": min(len(q) for q in qs), "n_max": max(len(q) for q in qs),
           "k_min": min(ks), "k_max": max(ks), "kinds": {kd: kinds.count(kd) for kd in sorted(set(kinds))},
           "batch_ms": ms, "evals_per_s": n_total / (ms * 1e-3), "sequential_ms": seq, "batch_speedup": seq / ms,
           "raised": raised, "max_rel_err": max(errs2), "median_rel_err": float(np.median(errs2)),
           "certified": sum(e <= tol for e in errs2), "certify_max_err": max(errs), "gen_s": gen_s,
           "predict_s": predict_s, "certify_s": certify_s, "setup_s": setup_s}
    batch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=sorted(CONFIGS) + ["C4"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--refine", action="store_true")
    ap.add_argument("--problems", type=int, default=256)
    args = ap.parse_args()
    if args.name == "C4":
        print(json.dumps(run_c4(n_problems=args.problems)), flush=True)
    else:
        print(json.dumps(run(args.name, n=args.n, refine=args.refine)), flush=True)


if __name__ == "__main__":
    main()
