"""Writes the golden fixtures of the reference evaluate (restated by oracle/kifmm.py).

Each fixture holds the inputs (positions, charges, k, reference order p_ref, leaf size),
the restated reference result phi_ref and the FP64 direct sum phi_direct on all points.
Run from the repository root: python tests/golden/make_kifmm_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import direct as D  # noqa: E402
from oracle import kifmm  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def case(name, x, q, k, p, leaf):
    ref = kifmm.evaluate(x, q, k, p=p, n_max=leaf)
    dire = D.direct_c(x, q, k)
    np.savez_compressed(os.path.join(HERE, f"kifmm_{name}.npz"), positions=x, charges=q, k=k, p_ref=p,
                        leaf=leaf, phi_ref=ref, phi_direct=dire)
    print(name, "ref err vs direct", D.relative_error(ref, dire))


def main():
    for n, k, p in ((640, 1.0, 4), (640, 3.0, 6), (640, 0.0, 5)):
        rng = np.random.default_rng(42 + n)                 # test.py:40-44 draws
        x = rng.uniform(0.0, 1.0, (n, 3))
        q = rng.normal(size=n)
        case(f"n{n}_k{k:g}_p{p}", x, q, k, p, 16)
    rng = np.random.default_rng(42 + 1500)                  # test.py --clustered layout
    x = rng.uniform(0.0, 1.0, (1500, 3))
    x[:750] = rng.uniform(0.1, 0.25, (750, 3))
    q = rng.normal(size=1500)
    case("clustered_n1500_k2_p5", x, q, 2.0, 5, 16)


if __name__ == "__main__":
    main()
