"""The C ABI: libhfmm.so loads without a GPU and exports every function that
include/hfmm.h declares; the ctypes binding declares exactly that set; error paths
that need no device report through status codes and hfmm_last_error."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "hfmm.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfmm_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from hfmm0_amd import _lib as L
    names = header_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(L.EXPORTED) == names


def test_argument_errors_without_device():
    from hfmm0_amd import _lib as L
    lib = L.lib()
    h = ctypes.c_void_p()
    o = L.default_opts()
    x = np.zeros((4, 3))
    assert lib.hfmm_plan_create(L.ptr(x), 4, 1.0, None, ctypes.byref(h)) == L.HFMM_ERR_ARG
    assert b"null" in lib.hfmm_last_error()
    o.p = 1
    assert lib.hfmm_plan_create(L.ptr(x), 4, 1.0, ctypes.byref(o), ctypes.byref(h)) == L.HFMM_ERR_ARG
    assert b"p must be an integer >=2" in lib.hfmm_last_error()
    o = L.default_opts()
    o.m2l_kernel = 7
    assert lib.hfmm_plan_create(L.ptr(x), 4, 1.0, ctypes.byref(o), ctypes.byref(h)) == L.HFMM_ERR_ARG
    assert b"m2l_kernel" in lib.hfmm_last_error()
    assert lib.hfmm_batch_create(0, None, None, None, None, ctypes.byref(h)) == L.HFMM_ERR_ARG
    assert lib.hfmm_evaluate(None, None, 0, None, 0, None) == L.HFMM_ERR_ARG


def test_noalloc_direct_rejects_sizes_beyond_its_scratch():
    # the limit is checked before any device call (no GPU needed)
    from hfmm0_amd import _lib as L
    lib = L.lib()
    dummy = ctypes.c_void_p(8)   # never dereferenced: the size check fails first
    assert lib.hfmm_direct_noalloc(dummy, 10, dummy, 1.0, dummy, 65537, dummy, None) == L.HFMM_ERR_ARG
    assert b"HFMM_DIRECT_NOALLOC_MAX_TARGETS" in lib.hfmm_last_error()
    assert lib.hfmm_direct_noalloc(None, 10, dummy, 1.0, dummy, 5, dummy, None) == L.HFMM_ERR_ARG


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of hfmm_opts and hfmm_plan_info have the header's size and field
    offsets (a shorter mirror would let the library write past it)."""
    import shutil
    import subprocess
    from hfmm0_amd import _lib as L
    if shutil.which("gcc") is None:
        import pytest
        pytest.skip("gcc not available")
    structs = {"hfmm_opts": L.Opts, "hfmm_plan_info": L.PlanInfo}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "hfmm.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
