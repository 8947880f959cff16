"""Runs the adaptive parity configurations at full size on one GPU and prints one JSON line:
the tuned form (mixed precision, learned degrees, optionally the certified local search)
certified against the FP64 direct sum on 4096 sampled targets, beside the FP64 form at the
same degrees.

    python -m hfmm0_amd.configs C3      # 2M sphere + clusters, k = 100
    python -m hfmm0_amd.configs C5      # 4M uniform + sphere + clusters, k = 60
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=sorted(CONFIGS))
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--refine", action="store_true")
    args = ap.parse_args()
    print(json.dumps(run(args.name, n=args.n, refine=args.refine)), flush=True)


if __name__ == "__main__":
    main()
