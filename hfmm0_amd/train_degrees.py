"""Generates the learned-degree training data with the FMM itself and trains the MLP on
the GPU.  Run on a GPU:  python -m hfmm0_amd.train_degrees --problems 200 --out DIR
(--seed0 S labels the problems S, S+1, ...; --label-only stops after writing the data;
--data A.npz,B.npz trains on those labelled sets, in order, after the ones generated).

For every seeded problem (hfmm0_amd.workloads.mixed_problem, N log-uniform in [10k,
200k], k uniform in [1, 80], leaf size drawn from 32/64/128) it builds one plan at the
table degree, labels every level by the cost-optimal degree vector for the total budget
(hfmm0_amd.learned.label_levels_joint: greedy descent from the table degrees, the level
saving the most measured M2L time per unit of added error first; --calibrate measures the
per-pair time of every degree first) and records the level features.  The network is then
trained by the HIP kernels (hfmm_mlp_train) on 80% of the rows; the rest is held out.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import api
from .learned import P_LO, P_SPAN, DegreeMLP, label_levels_joint, level_features
from .tree import build_tree
from .workloads import mixed_problem


def calibrate(path, n=1_000_000, k=40.0, leaf=128, degrees=range(4, 24), reps=5):
    """Measured M2L time per V pair against the degree: the uniform cube of n points,
    mixed precision, every level at degree p, the finest level's matrix-core launch timed
    with hipEvents (hfmm_benchmark_m2l_levels); nanoseconds per pair -> `path`."""
    import torch
    from .workloads import uniform_cube
    x, q = uniform_cube(n)
    q_dev = torch.from_numpy(q).cuda()
    o_dev = torch.empty(n, dtype=torch.complex128, device="cuda")
    table = {}
    for p in degrees:
        st = api.setup(x, k, p=p, N_max=leaf, precision="mixed")
        t = st.tree()
        lv = t["level"][t["V"][:, 0]]
        pairs = int((lv == st.depth).sum())
        ms = api.benchmark_m2l_levels(st, q_dev, o_dev, reps=reps)
        table[p] = ms[st.depth] * 1e6 / pairs
        print(f"calibrate: p={p} level {st.depth} {pairs} pairs {ms[st.depth]:.3f} ms -> {table[p]:.3f} ns/pair",
              flush=True)
        st.close()
    with open(path, "w") as f:
        json.dump({"workload": f"uniform cube N={n}, k={k:g}, leaf {leaf}, mixed precision, finest level",
                   "ns_per_pair": table}, f, indent=1)
    return table


def generate(n_problems, seed0=1000, n_range=(10_000, 200_000), k_range=(1.0, 80.0), leaves=(32, 64, 128),
             p_tab=23, budget=6e-7, log=None):
    feats, labels, meta = [], [], []
    for i in range(n_problems):
        rng = np.random.default_rng(seed0 + i)
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        leaf = int(leaves[int(rng.integers(0, len(leaves)))])
        x, q, k, kind = mixed_problem(n, k_range, seed=seed0 + i)
        t0 = time.time()
        st = api.setup(x, k, p=p_tab, N_max=leaf)
        tree = st.tree()
        lab = label_levels_joint(st, q, budget=budget)
        f = level_features(tree, q, k)
        st.close()
        feats.append(f)
        labels.append(lab)
        meta += [(i, n, k, kind, leaf, lv) for lv in range(2, len(lab) + 2)]
        if log:
            print(f"problem {i}: n={n} k={k:.1f} {kind} leaf={leaf} depth={len(lab) + 1} labels={lab.tolist()} "
                  f"({time.time() - t0:.2f}s)", file=log, flush=True)
    return np.vstack(feats), np.concatenate(labels), meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=200)
    ap.add_argument("--out", default="degree_training")
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget", type=float, default=6e-7, help="total error budget of the labels")
    ap.add_argument("--calibrate", action="store_true", help="measure the per-pair M2L cost table first")
    ap.add_argument("--seed0", type=int, default=1000, help="seed of the first generated problem")
    ap.add_argument("--label-only", action="store_true", help="write the labelled data, do not train")
    ap.add_argument("--data", default="", help="comma-separated degree_data.npz files to train on as well")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    if args.calibrate:
        calibrate(os.path.join(args.out, "m2l_cost.json"))
        import shutil
        from .learned import DEFAULT_COSTS
        shutil.copy(os.path.join(args.out, "m2l_cost.json"), DEFAULT_COSTS)
    Xs, ps = [], []
    n_problems = 0
    if args.problems > 0:
        X, p, meta = generate(args.problems, seed0=args.seed0, budget=args.budget, log=sys.stdout)
        np.savez_compressed(os.path.join(args.out, "degree_data.npz"), X=X, p=p,
                            meta=np.array([json.dumps(m) for m in meta]))
        Xs.append(X)
        ps.append(p)
        n_problems += args.problems
    if args.label_only:
        return
    for path in [d for d in args.data.split(",") if d]:
        with np.load(path, allow_pickle=False) as z:
            Xs.append(z["X"])
            ps.append(z["p"])
            n_problems += len({json.loads(m)[0] for m in z["meta"]})
    X, p = np.vstack(Xs), np.concatenate(ps)
    y = ((p - P_LO) / P_SPAN).astype(np.float32)
    n_train = int(0.8 * len(X))
    net = DegreeMLP(seed=args.seed)
    hist = net.train(X[:n_train], y[:n_train], epochs=args.epochs, batch=64, lr=2e-3, seed=args.seed)
    pred = net.predict_degrees(X[n_train:], margin=0.0, p_min=2, p_max=40)
    resid = pred - p[n_train:]
    print(f"train loss {hist[0]:.4f} -> {hist[-1]:.5f}; held-out degree error: mean {resid.mean():+.2f}, "
          f"under-predicted {np.mean(resid < 0):.1%}, |err|<=1 {np.mean(np.abs(resid) <= 1):.1%}", flush=True)
    net.save(os.path.join(args.out, "degree_mlp.npz"))
    with open(os.path.join(args.out, "degree_mlp.json"), "w") as f:
        json.dump({"problems": n_problems, "rows": int(len(X)), "train_rows": n_train, "epochs": args.epochs,
                   "budget_total": args.budget, "k_range": [1.0, 80.0],
                   "labels": "cost-optimal joint degrees (greedy), priced by measured M2L time per pair", "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                   "heldout_mean_resid": float(resid.mean()), "heldout_under": float(np.mean(resid < 0)),
                   "heldout_within1": float(np.mean(np.abs(resid) <= 1))}, f, indent=1)


if __name__ == "__main__":
    main()
