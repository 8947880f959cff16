"""GPU parity tests: the HIP FMM (through the C ABI) against the CPU oracle.

Every case prints its name, sizes and seed before touching the GPU.  Bounds:
* GPU FP64 direct sum vs the C oracle: relative error < 1e-13 (same FP64 sum);
* FMM vs the FP64 direct oracle on sampled targets: the truncation tolerance stated
  per case (degree-p solid-harmonic expansions);
* invariants of the reference harness (test.py:63-72): complex128 and finite output,
  linearity evaluate(-0.5 q) == -0.5 evaluate(q) to 1e-10, k = 0 gives an exactly real
  result, errors fall as p grows.
"""
import sys

import numpy as np
import pytest

from oracle import direct as D
from oracle import tree as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hf():
    import hfmm0_amd
    return hfmm0_amd


def test_gpu_direct_matches_oracle(hf):
    say("case: gpu direct N=3000 k=7")
    x, q = D.uniform_problem(3000)
    idx = D.sample_targets(3000, 300)
    got = hf.direct(x, q, 7.0, idx)
    ref = D.direct_c(x, q, 7.0, idx)
    assert D.relative_error(got, ref) < 1e-13


def _err(hf, x, q, k, p, nmax, **kw):
    st = hf.setup(x, k, p=p, N_max=nmax, **kw)
    out = hf.evaluate(st, q)
    idx = D.sample_targets(len(x), 400)
    ref = D.direct_c(x, q, k, idx)
    return D.relative_error(out[idx], ref), out, st


@pytest.mark.parametrize("k,p,tol", [(0.0, 8, 1e-4), (1.0, 8, 1e-4), (10.0, 10, 1e-4), (10.0, 16, 2e-7)])
def test_fmm_uniform(hf, k, p, tol):
    say(f"case: fmm uniform N=6000 k={k} p={p} N_max=32")
    x, q = D.uniform_problem(6000)
    e, out, st = _err(hf, x, q, k, p, 32)
    say(f"  rel err {e:.3e} depth {st.depth}")
    assert out.dtype == np.complex128 and np.isfinite(out).all()
    assert e < tol


def test_fmm_adaptive_clustered(hf):
    say("case: fmm clustered N=4000 k=3 p=12 N_max=16")
    rng = np.random.default_rng(42 + 4000)
    x = rng.uniform(size=(4000, 3))
    x[:2000] = rng.uniform(0.1, 0.25, (2000, 3))
    q = rng.normal(size=4000)
    e, out, st = _err(hf, x, q, 3.0, 12, 16)
    info = st.info
    say(f"  rel err {e:.3e} depth {st.depth} W {info.n_w} X {info.n_x}")
    assert info.n_w > 0 and info.n_x > 0
    assert e < 1e-5


def test_tree_of_plan_matches_oracle(hf):
    say("case: plan tree vs oracle N=5000 N_max=16 clustered")
    rng = np.random.default_rng(9)
    x = rng.uniform(size=(5000, 3))
    x[:2500] = rng.uniform(0.3, 0.4, (2500, 3))
    st = hf.setup(x, 2.0, p=4, N_max=16)
    got = st.tree()
    ref = T.build_tree(x, 16)
    for key in ("level", "code", "count", "pstart", "parent", "is_leaf", "perm"):
        np.testing.assert_array_equal(np.asarray(got[key]).astype(np.int64), np.asarray(ref[key]).astype(np.int64))


def test_reference_harness_invariants(hf):
    """test.py defaults: n in {256, 640}, k in {0, 1, 3}, leaf 16, real charges."""
    for n in (256, 640):
        rng = np.random.default_rng(42 + n)
        x = rng.uniform(0.0, 1.0, (n, 3))
        q = rng.normal(size=n)
        for k in (0.0, 1.0, 3.0):
            ref = D.direct_numpy(x, q, k)
            base = None
            for p in (3, 5, 8):
                say(f"case: harness n={n} k={k} p={p}")
                st = hf.setup(x, k, p=p, N_max=16)
                out = hf.evaluate(st, q)
                assert out.dtype == np.complex128 and np.isfinite(out).all()
                errs = np.array([D.relative_error(a, b) for a, b in
                                 ((out, ref), (out.real, ref.real), (out.imag, ref.imag))])
                if k == 0:
                    errs = errs[:2]
                    assert np.count_nonzero(out.imag) == 0
                assert max(errs) < (5e-2 if p <= 3 else 5e-3), errs
                np.testing.assert_allclose(hf.evaluate(st, -0.5 * q), -0.5 * out, rtol=1e-10, atol=1e-10)
                if base is not None:
                    assert np.all(errs < np.maximum(0.8 * base, 2e-6)), (errs, base)
                base = errs if base is None else base


def test_complex_charges_are_linear_combination(hf):
    say("case: complex charges N=3000 k=5 p=10")
    x, q = D.uniform_problem(3000)
    st = hf.setup(x, 5.0, p=10, N_max=32)
    a = hf.evaluate(st, q)
    b = hf.evaluate(st, q.real) + 1j * hf.evaluate(st, q.imag)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_negative_k_is_conjugate(hf):
    say("case: negative k N=2000 k=-4")
    x, q = D.uniform_problem(2000)
    a = hf.evaluate(hf.setup(x, 4.0, p=10, N_max=32), q)
    b = hf.evaluate(hf.setup(x, -4.0, p=10, N_max=32), np.conj(q))
    np.testing.assert_allclose(b, np.conj(a), rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_mixed_precision_near_field(hf):
    say("case: mixed precision N=8000 k=20 p=14")
    x, q = D.uniform_problem(8000)
    e64, out64, _ = _err(hf, x, q, 20.0, 14, 64)
    e32, out32, _ = _err(hf, x, q, 20.0, 14, 64, precision="mixed")
    say(f"  fp64 {e64:.3e} mixed {e32:.3e}")
    assert e32 < 2 * e64 + 1e-6
    assert D.relative_error(out32, out64) < 1e-6


def test_graph_replay_matches_stream_run(hf):
    say("case: graph vs eager N=4000 k=6 p=9")
    x, q = D.uniform_problem(4000)
    a = hf.setup(x, 6.0, p=9, N_max=32, use_graph=True)
    b = hf.setup(x, 6.0, p=9, N_max=32, use_graph=False)
    ra = [hf.evaluate(a, q) for _ in range(3)]
    rb = hf.evaluate(b, q)
    # coarse-level M2L work items add into the local expansions atomically, so the
    # summation order (and the last bits) may differ between runs
    for r in ra:
        np.testing.assert_allclose(r, rb, rtol=0, atol=1e-13 * np.abs(rb).max())


def test_device_pointer_path(hf):
    import torch
    say("case: device tensors N=5000 k=8 p=10")
    x, q = D.uniform_problem(5000)
    st = hf.setup(x, 8.0, p=10, N_max=32)
    host = hf.evaluate(st, q)
    qd = torch.from_numpy(q).to("cuda")
    od = torch.empty(5000, dtype=torch.complex128, device="cuda")
    hf.evaluate_device(st, qd, od)
    torch.cuda.synchronize()
    np.testing.assert_allclose(od.cpu().numpy(), host, rtol=0, atol=1e-13 * np.abs(host).max())


def test_per_level_degrees(hf):
    say("case: per-level p N=20000 k=10 p_level=[10,12,10,9]")
    x, q = D.uniform_problem(20000)
    e, _, st = _err(hf, x, q, 10.0, 10, 64, p_level=[10, 12, 10, 9])
    say(f"  rel err {e:.3e} levels {st.p_level}")
    assert st.p_level == [10, 12, 10]
    assert e < 1e-4


@pytest.mark.parametrize("k,p", [(0.0, 8), (10.0, 12), (25.0, 18)])
def test_matrix_core_m2l_matches_generic_fp64(hf, k, p):
    say(f"case: mfma m2l fp64 vs generic N=20000 k={k} p={p}")
    x, q = D.uniform_problem(20000)
    a = hf.evaluate(hf.setup(x, k, p=p, N_max=32), q)
    b = hf.evaluate(hf.setup(x, k, p=p, N_max=32, m2l_kernel="generic"), q)
    d = D.relative_error(a, b)
    say(f"  rel diff {d:.3e}")
    assert d < 1e-13


@pytest.mark.parametrize("p_level", [[12, 14, 10, 13], [15, 11, 13, 9]])
def test_block_translations_match_generic_per_level(hf, p_level):
    """Degrees that rise and fall between levels give non-square M2M and L2L blocks (child
    degree != parent degree); in FP64 the matrix-core translations equal the generic
    thread-per-coefficient kernels to rounding."""
    say(f"case: block vs generic fp64, per-level p={p_level} N=30000 k=15")
    x, q = D.uniform_problem(30000)
    a = hf.evaluate(hf.setup(x, 15.0, p=max(p_level), N_max=32, p_level=p_level), q)
    b = hf.evaluate(hf.setup(x, 15.0, p=max(p_level), N_max=32, p_level=p_level, m2l_kernel="generic"), q)
    d = D.relative_error(a, b)
    say(f"  rel diff {d:.3e}")
    assert d < 1e-13


def test_matrix_core_m2l_fp32(hf):
    say("case: mfma m2l fp32 N=30000 k=20 p=14")
    x, q = D.uniform_problem(30000)
    idx = D.sample_targets(30000, 500)
    ref = D.direct_c(x, q, 20.0, idx)
    st = hf.setup(x, 20.0, p=14, N_max=32, precision="mixed")
    # k D <= 32 at every level: M2L, and with it L2L and the M2M of levels with at least
    # 512 blocks of 8 parents (none at this size; those with fewer split and stay FP64), in FP32
    assert st.m2l_fp32_levels == list(range(2, st.depth + 1))
    assert st.l2l_fp32_levels == list(range(2, st.depth)) and st.m2m_fp32_levels == []
    a = hf.evaluate(st, q)
    b = hf.evaluate(hf.setup(x, 20.0, p=14, N_max=32, m2l_kernel="generic"), q)
    ea, eb = D.relative_error(a[idx], ref), D.relative_error(b[idx], ref)
    say(f"  mixed {ea:.3e} fp64 {eb:.3e} diff {D.relative_error(a, b):.3e}")
    assert D.relative_error(a, b) < 3e-6
    assert ea < eb + 3e-6


def test_mixed_precision_per_level(hf):
    """M2L precision per level under precision="mixed": the automatic rule keeps levels with
    k D_l > MIXED_FP32_MAX_KD (D_l the box diagonal) in FP64; listing every level in
    m2l_fp64_levels leaves only the FP32 near field, so the result then sits within the
    FP32 P2P rounding of the FP64 plan, and a single FP32 level sits between the two."""
    from hfmm0_amd import api
    say("case: mixed per-level precision N=30000 k=80 p=14 leaf=32")
    x, q = D.uniform_problem(30000)
    k = 80.0
    base = dict(p=14, N_max=32)
    auto = hf.setup(x, k, precision="mixed", **base)
    ext = auto.info.extent
    levels = [lv for lv in range(2, auto.depth + 1)]
    want = [lv for lv in levels if k * 2 * np.sqrt(3) * ext * 0.5 ** (lv + 1) <= api.MIXED_FP32_MAX_KD]
    say(f"  depth {auto.depth} fp32 levels {auto.m2l_fp32_levels} want {want}")
    assert auto.m2l_fp32_levels == want and 2 not in want
    # M2M and L2L take FP32 only where every M2L level from 2 on is FP32: none here
    assert auto.l2l_fp32_levels == [] and auto.m2m_fp32_levels == []
    ref = hf.evaluate(hf.setup(x, k, precision="fp64", **base), q)
    all64 = hf.setup(x, k, precision="mixed", m2l_fp64_levels=levels, **base)
    assert all64.m2l_fp32_levels == []
    # neighbour leaves of width 1/16 reach k r = 80 * 2 sqrt(3) / 16 = 17 > 8: every leaf
    # takes the FP64 phase
    assert all64.info.n_leaves_phase64 == all64.info.n_leaves > 0
    d_all64 = D.relative_error(hf.evaluate(all64, q), ref)
    one = hf.setup(x, k, precision="mixed", m2l_fp64_levels=[lv for lv in levels if lv != 3], **base)
    assert one.m2l_fp32_levels == [3] and one.l2l_fp32_levels == []
    d_one = D.relative_error(hf.evaluate(one, q), ref)
    none64 = hf.setup(x, k, precision="mixed", m2l_fp64_levels=[], **base)
    assert none64.m2l_fp32_levels == levels and none64.l2l_fp32_levels == levels[:-1]
    d_none = D.relative_error(hf.evaluate(none64, q), ref)
    say(f"  vs fp64 plan: all fp64 {d_all64:.2e} level 3 fp32 {d_one:.2e} all fp32 {d_none:.2e}")
    assert d_all64 < 3e-7
    assert d_all64 < d_one < d_none < 3e-6
    with pytest.raises(ValueError):
        hf.setup(x, k, precision="mixed", m2l_fp64_levels=[0], **base)


def test_fp32_phase_near_field(hf):
    """The FP32 near field with an FP32 phase (neighbour leaves of width 1/16 at k = 20: k r
    at most 20 * 2 sqrt(3) / 16 = 4.3, below the FP64-phase bound of 8): with every M2L level
    kept in FP64 the result sits within FP32 weight rounding of the FP64 plan."""
    say("case: fp32-phase near field N=30000 k=20 p=12 leaf=32")
    x, q = D.uniform_problem(30000)
    ref = hf.setup(x, 20.0, p=12, N_max=32)
    st = hf.setup(x, 20.0, p=12, N_max=32, precision="mixed", m2l_fp64_levels=list(range(2, ref.depth + 1)))
    assert st.info.n_leaves_phase64 == 0 and st.m2l_fp32_levels == []
    d = D.relative_error(hf.evaluate(st, q), hf.evaluate(ref, q))
    say(f"  vs fp64 plan {d:.2e}")
    assert d < 2e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_set_degrees_equals_dedicated_plan(hf, precision):
    say(f"case: set_degrees {precision} N=20000 k=12 tables p=14 -> [12, 11, 9]")
    x, q = D.uniform_problem(20000)
    st = hf.setup(x, 12.0, p=14, N_max=64, precision=precision)
    full = hf.evaluate(st, q)
    st.set_degrees([12, 12, 11, 9])
    assert st.p_level == [12, 12, 11] and st.p_table == [14, 14, 14]
    a = hf.evaluate(st, q)
    b = hf.evaluate(hf.setup(x, 12.0, p=12, N_max=64, p_level=[12, 12, 11, 9], precision=precision), q)
    d = D.relative_error(a, b)
    say(f"  truncated vs dedicated rel diff {d:.3e} (vs full {D.relative_error(a, full):.3e})")
    assert d < (1e-13 if precision == "fp64" else 2e-6)
    st.set_degrees([0, 0, 0])
    np.testing.assert_allclose(hf.evaluate(st, q), full, rtol=0, atol=1e-12 * np.abs(full).max())


def test_set_degrees_generic_kernel_levels(hf):
    say("case: set_degrees above the block-kernel degree N=8000 k=30 tables p=26 -> 20")
    x, q = D.uniform_problem(8000)
    st = hf.setup(x, 30.0, p=26, N_max=32)
    st.set_degrees([20] * st.depth)
    a = hf.evaluate(st, q)
    b = hf.evaluate(hf.setup(x, 30.0, p=20, N_max=32), q)
    assert D.relative_error(a, b) < 1e-13
    with pytest.raises(ValueError):
        st.set_degrees([30] * st.depth)


@pytest.mark.parametrize("precision,tb", [("fp64", 8), ("fp64", 16), ("fp64", 32), ("mixed", 8), ("mixed", 16),
                                         ("mixed", 32), ("mixed", 64)])
def test_matrix_core_block_sizes(hf, precision, tb):
    say(f"case: mfma m2l {tb} pairs per operator {precision} N=30000 k=15 p=13 vs generic")
    x, q = D.uniform_problem(30000)
    a = hf.evaluate(hf.setup(x, 15.0, p=13, N_max=32, precision=precision, m2l_kernel=tb), q)
    b = hf.evaluate(hf.setup(x, 15.0, p=13, N_max=32, precision=precision, m2l_kernel="generic"), q)
    d = D.relative_error(a, b)
    say(f"  rel diff {d:.3e}")
    assert d < (1e-13 if precision == "fp64" else 3e-6)


def test_charge_and_output_types(hf):
    say("case: float32/complex64 charges and complex64 output N=12000 k=8 p=10")
    x, q = D.uniform_problem(12000)
    st = hf.setup(x, 8.0, p=10, N_max=32)
    q32 = q.astype(np.complex64)
    ref = hf.evaluate(st, q32.astype(np.complex128))
    scale = np.abs(ref).max()
    np.testing.assert_allclose(hf.evaluate(st, q32), ref, rtol=0, atol=1e-13 * scale)
    r32 = q32.real.copy()
    np.testing.assert_allclose(hf.evaluate(st, r32), hf.evaluate(st, r32.astype(np.float64)), rtol=0,
                               atol=1e-13 * scale)
    o64 = hf.evaluate(st, q32, out_dtype=np.complex64)
    assert o64.dtype == np.complex64
    np.testing.assert_allclose(o64, ref.astype(np.complex64), rtol=0, atol=2e-7 * scale)
    st0 = hf.setup(x, 0.0, p=10, N_max=32)
    assert np.all(hf.evaluate(st0, r32).imag == 0)
    assert np.all(hf.evaluate(st0, r32, out_dtype=np.complex64).imag == 0)


def test_device_benchmark_narrow_charge_types(hf):
    """benchmark / benchmark_m2l_levels / evaluate_device read float32 and complex64 device
    tensors as their own width (they were once sent as float64 / complex128)."""
    import torch
    say("case: device benchmark with float32 / complex64 charges N=12000 k=8 p=10")
    x, q = D.uniform_problem(12000)
    st = hf.setup(x, 8.0, p=10, N_max=32)
    od = torch.empty(12000, dtype=torch.complex128, device="cuda")
    for qn in (q.astype(np.complex64), q.real.astype(np.float32)):
        ref = hf.evaluate(st, qn.astype(np.complex128 if np.iscomplexobj(qn) else np.float64))
        qd = torch.from_numpy(qn).cuda()
        ms = hf.benchmark(st, qd, od, warmup=1, reps=3)
        lv = hf.benchmark_m2l_levels(st, qd, od, reps=2)
        hf.evaluate_device(st, qd, od)
        torch.cuda.synchronize()
        assert ms > 0 and lv
        np.testing.assert_allclose(od.cpu().numpy(), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_near_coincident_points_follow_the_reference_floor(hf, precision):
    # pairs 1e-15 apart lie below the reference's coincidence floor (16 eps (max|x_t| +
    # max|x_s|), jaxfmm/near.py:252-261) and contribute nothing; without the floor their
    # 1e15 terms would swamp the result
    say(f"case: near-coincident pairs N=4000 k=5 p=14 N_max=32 {precision}")
    x, q = D.uniform_problem(4000)
    x = x.copy()
    x[1::40] = x[0::40][: len(x[1::40])] + np.array([1e-15, 0.0, 0.0])
    st = hf.setup(x, 5.0, p=14, N_max=32, precision=precision)
    out = hf.evaluate(st, q)
    idx = np.unique(np.concatenate([np.arange(0, 400, 1), D.sample_targets(len(x), 200)]))
    ref = D.direct_numpy(x, q, 5.0, idx, floor=1e-12)
    e = D.relative_error(out[idx], ref)
    say(f"  rel err {e:.3e} against the direct sum with the floor")
    assert np.isfinite(out).all()
    assert e < (2e-7 if precision == "fp64" else 2e-6)
    st.close()


@pytest.mark.parametrize("k", [0.0, 1e-7, 30.0])
def test_mixed_near_field_tiles_agree(hf, k):
    # targets per lane 1, 2 and 4 evaluate the same sums in another order; k = 0 and
    # k = 1e-7 take the low-frequency weights 1/r + ik
    say(f"case: mixed near-field tiles N=20000 k={k} p=12 N_max=64")
    x, q = D.uniform_problem(20000)
    idx = D.sample_targets(len(x), 300)
    ref = D.direct_c(x, q, k, idx)
    outs = []
    for tile in (1, 2, 4):
        st = hf.setup(x, k, p=14, N_max=64, precision="mixed", p2p_tile=tile)
        outs.append(hf.evaluate(st, q))
        st.close()
        e = D.relative_error(outs[-1][idx], ref)
        say(f"  tile {tile}: rel err {e:.3e}")
        assert e < 2e-6
    for o in outs[1:]:
        assert D.relative_error(o, outs[0]) < 1e-6
    if k == 0.0:
        st = hf.setup(x, k, p=14, N_max=64, precision="mixed")
        assert np.count_nonzero(hf.evaluate(st, q.real).imag) == 0
        st.close()


def test_fp32_coordinate_leaves(hf):
    """Mixed near field: FP32-phase leaves whose estimated relative error from FP32
    coordinate differences is within p2p_coord32_tol take them.  On a uniform problem most
    leaves do, and the result stays within that error of the FP64-coordinate one; points
    placed 1e-9 from others keep their leaves in FP64."""
    say("case: fp32 coordinate leaves N=200000 k=20 p=12 leaf 64")
    x, q = D.uniform_problem(200000)
    idx = D.sample_targets(200000, 400)
    ref = D.direct_c(x, q, 20.0, idx)
    st = hf.setup(x, 20.0, p=12, N_max=64, precision="mixed")
    st64 = hf.setup(x, 20.0, p=12, N_max=64, precision="mixed", p2p_coord32_tol=-1)
    n_leaves = st.info.n_leaves
    say(f"  leaves {n_leaves}: FP32 coordinates at {st.n_leaves_coord32}, with tol < 0 at {st64.n_leaves_coord32}")
    assert st64.n_leaves_coord32 == 0
    assert st.n_leaves_coord32 > 0.5 * n_leaves
    a, b = hf.evaluate(st, q), hf.evaluate(st64, q)
    d = D.relative_error(a, b)
    ea, eb = D.relative_error(a[idx], ref), D.relative_error(b[idx], ref)
    say(f"  FP32 vs FP64 coordinates {d:.2e}; vs direct {ea:.3e} / {eb:.3e}")
    assert d < 2e-7
    assert ea < eb + 1e-7
    x2 = x.copy()
    x2[:2000] = x[2000:4000] + 1e-9   # near duplicates, far above the coincidence floor
    st2 = hf.setup(x2, 20.0, p=12, N_max=64, precision="mixed")
    say(f"  with 2000 near duplicates: FP32 coordinates at {st2.n_leaves_coord32} of {st2.info.n_leaves}")
    assert st2.n_leaves_coord32 < st.n_leaves_coord32 - 500
