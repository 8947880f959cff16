"""Runs the adaptive parity configurations at full size on one GPU and prints one JSON line:
the tuned form (mixed precision, learned degrees, optionally the certified local search)
certified against the FP64 direct sum on 4096 sampled targets, beside the FP64 form at the
same degrees.

    python -m hfmm0_amd.configs C3      # 2M sphere + clusters, k = 100
    python -m hfmm0_amd.configs C5      # 4M uniform + sphere + clusters, k = 60
    python -m hfmm0_amd.configs C4      # 256 ellipsoids, N in [2k, 200k], k in [1, 50]: one batch
"""
import argparse
import json
import time

import numpy as np

from . import api
from .learned import DegreeMLP, predict_p_level
from .workloads import c3_problem, c5_problem, sample_targets

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


def run_c4(n_problems=256, tol=1e-6, samples=512, reps=10, tune=True, headroom=4):
    """All problems in one batch (workloads.batch_problems: ellipsoids, N log-uniform in
    [2k, 200k], k in [1, 50]; complex64 charges; mixed precision).  The leaf size and the
    near-field tile come from the batch autotuner (autotune.tune_batch, on 24 of the
    problems).  Degrees: 12, raised to the high-frequency floor where k D needs more
    (autotune.batch_degrees); tables `headroom` degrees higher, and Batch.certify raises
    the degrees of every problem that misses `tol` on `samples` targets against the FP64
    direct sum.  The batch is then built again at the certified degrees and its evaluate
    timed against the problems evaluated one by one."""
    import torch
    from .autotune import batch_degrees, tune_batch
    from .batch import Batch
    from .learned import sampled_error
    from .workloads import batch_problems
    t0 = time.perf_counter()
    probs = batch_problems(n_problems)
    gen_s = time.perf_counter() - t0
    xs, qs, ks = [x for x, _, _, _ in probs], [q for _, q, _, _ in probs], [k for _, _, k, _ in probs]
    t0 = time.perf_counter()
    leaf, tile, tune_ms = tune_batch(xs, ks, qs) if tune else (128, 0, {})
    tune_s = time.perf_counter() - t0
    pls = [batch_degrees(x, k) for x, k in zip(xs, ks)]
    t0 = time.perf_counter()
    tabs = [[min(40, v + headroom) for v in pl] for pl in pls]
    batch = Batch(xs, ks, p=max(max(t) for t in tabs), N_max=leaf, p_level=tabs, precision="mixed", p2p_tile=tile)
    for i, pl in enumerate(pls):
        st = batch.plan(i)
        st.set_degrees(pl[:st.depth])
    _, cert, errs0 = batch.certify(qs, tol=tol, n_samp=samples, max_steps=6)
    certify_s = time.perf_counter() - t0
    raised = sum(list(c) != list(p)[:len(c)] for c, p in zip(cert, pls))
    batch.close()
    t0 = time.perf_counter()
    batch = Batch(xs, ks, p=max(max(c) for c in cert), N_max=leaf, p_level=cert, precision="mixed", p2p_tile=tile)
    setup_s = time.perf_counter() - t0
    q_dev = [torch.from_numpy(q).cuda() for q in qs]
    o_dev = [torch.empty(len(q), dtype=torch.complex128, device="cuda") for q in qs]
    ms = batch.benchmark(q_dev, o_dev, warmup=2, reps=reps)
    phases = batch.benchmark_phases(q_dev, o_dev, reps=3)
    torch.cuda.synchronize()
    phi = [o.cpu().numpy() for o in o_dev]
    errs = [sampled_error(x, q, k, f, samples, 0) for x, q, k, f in zip(xs, qs, ks, phi)]
    seq = sum(api.benchmark(batch.plan(i), q_dev[i], o_dev[i], warmup=1, reps=3) for i in range(len(probs)))
    n_total = sum(len(q) for q in qs)
    out = {"config": "C4", "problems": len(probs), "points": n_total, "leaf": leaf, "p2p_tile": tile,
           "n_min": min(len(q) for q in qs), "n_max": max(len(q) for q in qs),
           "k_min": min(ks), "k_max": max(ks), "degree_max": max(max(c) for c in cert), "raised": raised,
           "batch_ms": ms, "evals_per_s": n_total / (ms * 1e-3), "phases_ms": phases, "sequential_ms": seq,
           "batch_speedup": seq / ms, "max_rel_err": max(errs), "median_rel_err": float(np.median(errs)),
           "within_tol": sum(e <= tol for e in errs), "tol": tol, "certify_max_err": max(errs0), "gen_s": gen_s,
           "tune_s": tune_s, "tune_ms": tune_ms, "certify_s": certify_s, "setup_s": setup_s}
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
