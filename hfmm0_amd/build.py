"""Compile libhfmm.so (hipcc, gfx950) in-tree with the committed Makefile."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhfmm.so")


def build(jobs=4, verbose=False):
    """Always recompiles stale objects; returns the path of libhfmm.so."""
    cmd = ["make", f"-j{jobs}", "-C", os.path.join(HERE, "csrc")]
    res = subprocess.run(cmd, capture_output=not verbose, text=True)
    if res.returncode != 0:
        raise RuntimeError("building libhfmm.so failed:\n" + (res.stdout or "") + (res.stderr or ""))
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
