"""CPU tests of the C ABI: the library loads and exports every symbol the
headers declare; handles, tensor infos, executors and products behave as the
reference's (crates/tprims-capi/tests/tapp_rejections.rs, tapp_conformance.rs,
c/abi_test.c); every validation error fires before any GPU work; and the
planner's index analysis matches the reference's analysis goldens
(crates/tprims-contract/src/plan/analysis.rs:323-388)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import tprims_amd as tp
from tprims_amd import _capi
from tprims_amd._capi import TprimsError, i64_array, intptr, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for root, _, files in os.walk(os.path.join(ROOT, "include")):
        for f in files:
            if f.endswith(".h"):
                text = open(os.path.join(root, f)).read()
                names.update(re.findall(r"TAPP_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    out = subprocess.check_output(["nm", "-D", "--defined-only", _capi.LIB_PATH]).decode()
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = header_symbols()
    assert len(declared) >= 45
    missing = declared - exported
    assert not missing, missing
    assert set(_capi.EXPORTED) <= exported


def test_identity_and_parts():
    l = lib()
    assert l.tprims_abi_version() == 100
    assert l.tprims_has_part(b"tapp") and l.tprims_has_part(b"core") and l.tprims_has_part(b"hip")
    assert not l.tprims_has_part(b"blas") and not l.tprims_has_part(None)
    assert l.TAPP_implementation_name() and l.TAPP_implementation_version() == b"0.1.0"
    buf = ctypes.create_string_buffer(64)
    n = l.TAPP_explain_error(_capi.TPRIMS_BUSY, 64, buf)
    assert n == len(buf.value) > 0
    assert l.TAPP_explain_error(_capi.TPRIMS_BUSY, 0, None) == n
    small = ctypes.create_string_buffer(4)
    assert l.TAPP_explain_error(_capi.TPRIMS_ERR_ALIASED, 4, small) > 3 and len(small.value) == 3
    assert l.TAPP_check_success(0) and not l.TAPP_check_success(1)


def test_handles_attrs_status():
    l = lib()
    h = intptr(0)
    assert l.TAPP_create_handle(ctypes.byref(h)) == 0 and h.value != 0
    h2 = intptr(0)
    assert l.TAPP_create_handle(ctypes.byref(h2)) == 0 and h2.value != h.value
    assert l.TAPP_create_handle(None) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    v = ctypes.c_void_p(123)
    assert l.TAPP_attr_get(h.value, 0, ctypes.byref(v)) == _capi.TPRIMS_ERR_UNSUPPORTED and v.value is None
    assert l.TAPP_attr_set(h.value, 0, None) == _capi.TPRIMS_ERR_UNSUPPORTED
    assert l.TAPP_attr_clear(h.value, 0) == _capi.TPRIMS_ERR_UNSUPPORTED
    assert l.tprims_last_error() == b"no attribute keys are defined"
    assert l.TAPP_destroy_status(12345) == 0 and l.tprims_last_error() == b""
    assert l.TAPP_destroy_handle(h.value) == 0 and l.TAPP_destroy_handle(h2.value) == 0
    assert l.TAPP_destroy_handle(0) == _capi.TPRIMS_ERR_INVALID_ARGUMENT


def test_tensor_info_lifecycle():
    l = lib()
    info = intptr(0)
    assert l.TAPP_create_tensor_info(ctypes.byref(info), _capi.TAPP_C64, 3, i64_array([2, 3, 4]),
                                     i64_array([1, 2, 6])) == 0
    assert l.TAPP_get_nmodes(info.value) == 3
    e = (ctypes.c_int64 * 3)()
    l.TAPP_get_extents(info.value, e)
    assert list(e) == [2, 3, 4]
    assert l.TAPP_set_strides(info.value, i64_array([3, 1, 9])) == 0
    s = (ctypes.c_int64 * 3)()
    l.TAPP_get_strides(info.value, s)
    assert list(s) == [3, 1, 9]
    assert l.TAPP_set_nmodes(info.value, 4) == 0
    e4, s4 = (ctypes.c_int64 * 4)(), (ctypes.c_int64 * 4)()
    l.TAPP_get_extents(info.value, e4)
    l.TAPP_get_strides(info.value, s4)
    assert list(e4) == [2, 3, 4, 1] and list(s4) == [3, 1, 9, 0]
    assert l.TAPP_set_nmodes(info.value, -1) == _capi.TPRIMS_ERR_SHAPE
    assert l.TAPP_set_extents(info.value, None) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert l.TAPP_get_nmodes(0) == -1
    assert l.TAPP_destroy_tensor_info(info.value) == 0
    bad = intptr(0)
    assert l.TAPP_create_tensor_info(ctypes.byref(bad), _capi.TAPP_F16, 1, i64_array([2]),
                                     i64_array([1])) == _capi.TPRIMS_ERR_DTYPE
    assert l.TAPP_create_tensor_info(ctypes.byref(bad), _capi.TAPP_BF16, 0, None, None) == _capi.TPRIMS_ERR_DTYPE
    assert l.TAPP_create_tensor_info(ctypes.byref(bad), _capi.TAPP_F32, -1, None, None) == _capi.TPRIMS_ERR_SHAPE
    assert l.TAPP_create_tensor_info(ctypes.byref(bad), _capi.TAPP_F32, 2, None, None) == \
        _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert l.TAPP_create_tensor_info(ctypes.byref(bad), _capi.TAPP_F32, 0, None, None) == 0


def test_executor_thread_settings():
    # crates/tprims-capi/tests/c/abi_test.c executor section
    l = lib()
    pool, serial = intptr(0), intptr(0)
    assert l.tprims_tapp_executor_create_rayon(ctypes.byref(pool), 4, None) == 0 and pool.value
    ps, bd = ctypes.c_size_t(99), ctypes.c_size_t(99)
    assert l.tprims_tapp_executor_get_threads(pool.value, ctypes.byref(ps), ctypes.byref(bd)) == 0
    assert (ps.value, bd.value) == (4, 4)
    assert l.tprims_tapp_executor_set_budget(pool.value, 2) == 0
    l.tprims_tapp_executor_get_threads(pool.value, ctypes.byref(ps), ctypes.byref(bd))
    assert (ps.value, bd.value) == (4, 2)
    assert l.tprims_tapp_executor_set_budget(pool.value, 0) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert l.tprims_tapp_executor_set_budget(pool.value, 99) == 0
    l.tprims_tapp_executor_get_threads(pool.value, ctypes.byref(ps), ctypes.byref(bd))
    assert (ps.value, bd.value) == (4, 4)
    assert l.TAPP_create_executor(ctypes.byref(serial)) == 0
    l.tprims_tapp_executor_get_threads(serial.value, ctypes.byref(ps), ctypes.byref(bd))
    assert (ps.value, bd.value) == (0, 1)
    l.tprims_tapp_executor_get_threads(0, ctypes.byref(ps), ctypes.byref(bd))
    assert (ps.value, bd.value) == (0, 1)
    assert l.tprims_tapp_executor_create_rayon(ctypes.byref(pool), 0, None) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert l.TAPP_destroy_executor(pool.value) == 0
    assert l.TAPP_destroy_executor(serial.value) == 0
    assert l.TAPP_destroy_executor(0) == 0


def _plan(dtype, a, la, b, lb, d, ld, c=None, lc=None, conj=(0, 0, 0, 0), prec=_capi.TAPP_DEFAULT_PREC,
          dtypes=None):
    l = lib()
    h = intptr(0)
    l.TAPP_create_handle(ctypes.byref(h))
    infos = []
    dts = dtypes or [dtype] * 4
    for (ext, st), dt in zip([a, b, c or d, d], dts):
        i = intptr(0)
        assert l.TAPP_create_tensor_info(ctypes.byref(i), dt, len(ext), i64_array(ext), i64_array(st)) == 0
        infos.append(i.value)
    p = intptr(0)
    status = l.TAPP_create_tensor_product(ctypes.byref(p), h.value, conj[0], infos[0], i64_array(la), conj[1],
                                          infos[1], i64_array(lb), conj[2], infos[2], i64_array(lc or ld), conj[3],
                                          infos[3], i64_array(ld), prec)
    for i in infos:
        l.TAPP_destroy_tensor_info(i)
    return status, p.value


M44 = ([4, 4], [1, 4])


def test_product_rejections():
    F64, F32 = _capi.TAPP_F64, _capi.TAPP_F32
    g = ([2, 3], [1, 2]), [0, 2], ([3, 2], [1, 3]), [2, 1], ([2, 2], [1, 2]), [0, 1]
    st, _ = _plan(F64, *g, dtypes=[F64, F32, F64, F64])
    assert st == _capi.TPRIMS_ERR_DTYPE
    st, _ = _plan(F64, *g, prec=_capi.TAPP_F32F32_ACCUM_F32)
    assert st == _capi.TPRIMS_ERR_UNSUPPORTED
    st, _ = _plan(F64, *g, prec=_capi.TAPP_BF16BF16_ACCUM_F32)
    assert st == _capi.TPRIMS_ERR_UNSUPPORTED
    st, _ = _plan(F64, *g, conj=(2, 0, 0, 0))
    assert st == _capi.TPRIMS_ERR_UNSUPPORTED
    # an output-only label (TAPP case 5) is declined
    st, _ = _plan(F64, ([2], [1]), [0], ([2], [1]), [0], ([2, 3], [1, 2]), [0, 9])
    assert st == _capi.TPRIMS_ERR_UNSUPPORTED
    # a label with two extents
    st, _ = _plan(F64, ([2, 3], [1, 2]), [0, 2], ([4, 2], [1, 4]), [2, 1], ([2, 2], [1, 2]), [0, 1])
    assert st == _capi.TPRIMS_ERR_SHAPE
    # C and D labelled differently
    st, _ = _plan(F64, *g, c=([2, 2], [1, 2]), lc=[0, 5])
    assert st == _capi.TPRIMS_ERR_LABELS
    # an output that overlaps itself
    st, _ = _plan(F64, ([2, 3], [1, 2]), [0, 2], ([3, 2], [1, 3]), [2, 1], ([2, 2], [1, 1]), [0, 1])
    assert st == _capi.TPRIMS_ERR_ALIASED
    # negative extent
    st, _ = _plan(F64, ([2, -3], [1, 2]), [0, 2], ([3, 2], [1, 3]), [2, 1], ([2, 2], [1, 2]), [0, 1])
    assert st == _capi.TPRIMS_ERR_SHAPE
    st, p = _plan(F64, *g)
    assert st == 0
    assert lib().TAPP_destroy_tensor_product(p) == 0
    assert lib().TAPP_destroy_tensor_product(0) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    # zero handle
    l = lib()
    i = intptr(0)
    l.TAPP_create_tensor_info(ctypes.byref(i), F64, 2, i64_array([2, 2]), i64_array([1, 2]))
    p = intptr(0)
    assert l.TAPP_create_tensor_product(ctypes.byref(p), 0, 0, i.value, i64_array([0, 2]), 0, i.value,
                                        i64_array([2, 1]), 0, i.value, i64_array([0, 1]), 0, i.value,
                                        i64_array([0, 1]), -1) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert l.TAPP_create_tensor_product(ctypes.byref(p), 1, 0, i.value, None, 0, i.value, i64_array([2, 1]), 0,
                                        i.value, i64_array([0, 1]), 0, i.value, i64_array([0, 1]),
                                        -1) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    l.TAPP_destroy_tensor_info(i.value)


def test_execute_validation_fires_before_any_device_work():
    """Null pointers, a null C with beta != 0, and output aliasing are refused
    from the pointers alone (execute.rs:30-47, plan/mod.rs:384-428)."""
    F64 = _capi.TAPP_F64
    l = lib()
    st, p = _plan(F64, ([2, 3], [3, 1]), [0, 2], ([3, 2], [1, 3]), [2, 1], ([2, 2], [1, 2]), [0, 1])
    assert st == 0
    one, zero = ctypes.c_double(1.0), ctypes.c_double(0.0)
    a = (ctypes.c_double * 6)(1, 3, 5, 2, 4, 6)
    b = (ctypes.c_double * 6)(1, 3, 5, 2, 4, 6)
    d = (ctypes.c_double * 4)()
    status = intptr(-1)
    r = l.TAPP_execute_product(p, 0, ctypes.byref(status), ctypes.byref(one), a, b, ctypes.byref(zero), None, a)
    assert r == _capi.TPRIMS_ERR_ALIASED and status.value == 0
    assert l.tprims_last_error()
    r = l.TAPP_execute_product(p, 0, None, ctypes.byref(one), None, b, ctypes.byref(zero), None, d)
    assert r == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    r = l.TAPP_execute_product(p, 0, None, ctypes.byref(one), a, b, ctypes.byref(one), None, d)
    assert r == _capi.TPRIMS_ERR_UNSUPPORTED
    r = l.TAPP_execute_product(p, 0, None, None, a, b, ctypes.byref(zero), None, d)
    assert r == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    # a separate C that partially overlaps D
    big = (ctypes.c_double * 8)()
    c_ptr = ctypes.addressof(big) + 8
    r = l.TAPP_execute_product(p, 0, None, ctypes.byref(one), a, b, ctypes.byref(one), c_ptr, big)
    assert r == _capi.TPRIMS_ERR_ALIASED
    # batched: a negative count, null arrays
    assert l.TAPP_execute_batched_product(p, 0, None, -1, ctypes.byref(one), None, None, ctypes.byref(zero), None,
                                          None) == _capi.TPRIMS_ERR_SHAPE
    assert l.TAPP_execute_batched_product(p, 0, None, 1, ctypes.byref(one), None, None, ctypes.byref(zero), None,
                                          None) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    arr = (ctypes.c_void_p * 1)
    assert l.TAPP_execute_batched_product(p, 0, None, 0, ctypes.byref(one), arr(), arr(), ctypes.byref(zero), None,
                                          arr()) == 0
    l.TAPP_destroy_tensor_product(p)


def _report(dims_a, la, dims_b, lb, dims_d, ld, strides=None):
    s = strides or (tp.col_major(dims_a), tp.col_major(dims_b), tp.col_major(dims_d))
    A = tp.OperandSpec(tp.LayoutSpec(dims_a, s[0]))
    B = tp.OperandSpec(tp.LayoutSpec(dims_b, s[1]))
    D = tp.OperandSpec(tp.LayoutSpec(dims_d, s[2]))
    prob = tp.Problem.from_labels(tp.DType.F64, A, B, tp.CSpec.separate(D), D, tp.Labels(la, lb, ld, ld))
    return tp.Plan(prob).report()


def test_analysis_goldens():
    # analysis.rs:327-338 plain matmul folds to a pure GEMM (4, 6, 5, 1)
    r = _report([4, 5], [0, 2], [5, 6], [2, 1], [4, 6], [0, 1])
    assert (r["m"], r["n"], r["k"], r["batch"]) == (4, 6, 5, 1)
    assert (r["m_axes"], r["n_axes"], r["k_axes"], r["h_axes"]) == (1, 1, 1, 0)
    # analysis.rs:341-352 adjacent free indices fold
    r = _report([4, 3, 5], [0, 1, 3], [5, 6], [3, 2], [4, 3, 6], [0, 1, 2])
    assert r["m_axes"] == 1 and r["m"] == 12
    # analysis.rs:355-364 Hadamard index recognised
    r = _report([2, 4, 5], [9, 0, 3], [2, 5, 6], [9, 3, 1], [2, 4, 6], [9, 0, 1])
    assert (r["m"], r["n"], r["k"], r["batch"]) == (4, 6, 5, 2)
    # analysis.rs:367-377 isolated index becomes a zero-stride contraction: k = 15
    r = _report([4, 5, 3], [0, 2, 7], [5, 6], [2, 1], [4, 6], [0, 1])
    assert r["k"] == 15
    # analysis.rs:380-387 repeated label takes the diagonal: k = 4
    r = _report([4, 4], [0, 0], [4, 6], [0, 1], [6], [1])
    assert (r["k"], r["m"], r["n"]) == (4, 1, 6)


def test_benchmark_shapes_reduce_as_the_reference():
    """The reduced roles of the benchmark contractions (folding included)."""
    e = 24
    r = _report([e] * 4, [0, 4, 1, 5], [e] * 4, [5, 2, 4, 3], [e] * 4, [0, 1, 2, 3])
    assert (r["m"], r["n"], r["k"], r["batch"]) == (576, 576, 576, 1)
    assert r["m_axes"] == 2 and r["n_axes"] == 2 and r["k_axes"] == 2
    # complex benchmark: A (j,i,l,n) as every 2nd element of a 2x-padded buffer
    j, i, l, n, k = 64, 512, 64, 16, 512
    sa = [2, 2 * j, 2 * j * i, 2 * j * i * l]
    A = tp.OperandSpec(tp.LayoutSpec([j, i, l, n], sa), tp.Op.CONJUGATE)
    B = tp.OperandSpec(tp.LayoutSpec([l, k, j, n], tp.col_major([l, k, j, n])))
    D = tp.OperandSpec(tp.LayoutSpec([i, k, n], tp.col_major([i, k, n])))
    J, I, L, N, KK = 1, 2, 3, 4, 5
    prob = tp.Problem.from_labels(tp.DType.C64, A, B, tp.CSpec.separate(D), D,
                                  tp.Labels([J, I, L, N], [L, KK, J, N], [I, KK, N], [I, KK, N]))
    r = tp.Plan(prob).report()
    assert (r["m"], r["n"], r["k"], r["batch"]) == (512, 512, 4096, 16)
    assert r["a_k_major"] == 1 and r["b_k_major"] == 1 and r["d_m_major"] == 1
    assert r["flops"] == 8 * 512 * 512 * 4096 * 16


def test_dot_general_mirror():
    # dot_general.rs:14-18: bmm out dims [5, 2, 3]
    dg = tp.DotGeneral([2], [1], [0], [0])
    assert dg.validate([3, 5, 4], [3, 4, 2])[2] == [5, 2, 3]
    with pytest.raises(TprimsError):
        tp.DotGeneral([2], [1], [0], [0]).validate([3, 5, 4], [3, 5, 2])
    A = tp.OperandSpec(tp.LayoutSpec([3, 5, 4], tp.col_major([3, 5, 4])))
    B = tp.OperandSpec(tp.LayoutSpec([3, 4, 2], tp.col_major([3, 4, 2])))
    D = tp.OperandSpec(tp.LayoutSpec([5, 2, 3], tp.col_major([5, 2, 3])))
    p = tp.Plan(tp.Problem.from_dot_general(tp.DType.F64, A, B, D, dg))
    r = p.report()
    assert (r["m"], r["n"], r["k"], r["batch"]) == (5, 2, 4, 3)


def test_planner_picks_elementwise_for_all_batch_products():
    """Hadamard-type products (tiny M x N and K, many batch entries) take the
    elementwise strategy (crates/tprims-contract/src/strategy/elementwise.rs:194);
    GEMM-shaped ones the MFMA tile kernels."""
    import tprims_amd as tp
    from harness import T, make_plan
    n = 1 << 16
    had = make_plan(tp.DType.F64, T([n], [1], [0]), T([n], [1], [0]), T([n], [1], [0])).report()
    assert had["strategy"] == 3 and had["kernel"] == "elementwise.f64" and had["flops"] == 2 * n
    # a batched 2x2 product with K = 3 is still elementwise; a 64x64 GEMM is not
    bat = make_plan(tp.DType.C32, T([2, 3, n], [1, 2, 6], [0, 2, 3]), T([3, 2, n], [1, 3, 6], [2, 1, 3]),
                    T([2, 2, n], [1, 2, 4], [0, 1, 3])).report()
    assert bat["strategy"] == 3
    gemm = make_plan(tp.DType.F32, T([64, 64], [1, 64], [0, 2]), T([64, 64], [1, 64], [2, 1]),
                     T([64, 64], [1, 64], [0, 1])).report()
    assert gemm["strategy"] == 0


def test_register_direct_tile_choice():
    """f64 in 16-byte runs with K <= 1024 where split K would run: 32x32 tiles
    on eight K groups at any tile count, 32x48 when more than 256 tiles of
    32x32 become at most 256 of 32x48 (B gathered along K); split K beyond
    K = 1024 and for f32; an explicit tile or split wins."""
    import tprims_amd as tp
    from harness import T, make_plan

    def plan(M, K, N, dt=tp.DType.F64, **cfg):
        a, b, d = T([M, K], [1, M], [0, 2]), T([K, N], [1, K], [2, 1]), T([M, N], [1, M], [0, 1])
        return make_plan(dt, a, b, d, config=tp.PlanConfig(**cfg) if cfg else None).report()

    assert plan(202, 300, 130)["tile"] == 6  # 35 tiles
    assert plan(512, 512, 512)["tile"] == 6  # 256 tiles: one per CU already
    assert plan(576, 576, 576)["tile"] == 7  # 324 -> 216
    assert plan(640, 1024, 640)["tile"] == 6  # 400 -> 280: 32x48 would need a second round
    assert ".splitk" in plan(384, 2048, 384)["kernel"]
    assert ".splitk" in plan(202, 300, 130, tp.DType.F32)["kernel"]
    assert plan(576, 576, 576, tile=6)["kernel"] == "f64.mfma_f64.32x32x16.regs16.kgroup8"
    assert plan(576, 576, 576, split_k=3)["kernel"].endswith(".splitk3")
    # 32x48 with B gathered along N runs (and reports) the 32x32 form
    a, b, d = T([576, 576], [1, 576], [0, 2]), T([576, 576], [576, 1], [2, 1]), T([576, 576], [1, 576], [0, 1])
    r = make_plan(tp.DType.F64, a, b, d, config=tp.PlanConfig(tile=7)).report()
    assert r["b_k_major"] == 0 and r["kernel"] == "f64.mfma_f64.32x32x16.regs16.kgroup8", r


def test_split_k_choice_and_override():
    """Few 64x64 output tiles with a long K: the planner takes the
    register-direct K-group tile for f64 in 16-byte runs (on 32x48 tiles for
    the f64 24^6 contraction, whose 324 tiles of 32x32 would put two on some
    CUs), else splits K (the f32 form of the same contraction, whose A
    gathers along M); split_k = 1 keeps one K range, n > 1 asks for n;
    out-of-range values are refused."""
    import tprims_amd as tp
    from harness import T, make_plan
    e = 24
    sp = [1, e, e * e, e ** 3]
    args = (T([e] * 4, sp, [0, 4, 1, 5]), T([e] * 4, sp, [5, 2, 4, 3]), T([e] * 4, sp, [0, 1, 2, 3]))
    assert make_plan(tp.DType.F64, *args).report()["kernel"] == "f64.mfma_f64.32x48x16.regs16.kgroup8"
    assert make_plan(tp.DType.F64, *args).report()["tile"] == 7
    assert make_plan(tp.DType.F32, *args).report()["kernel"] == "f32.mfma_f32.64x64x16.dma16.splitk3"
    assert ".splitk" not in make_plan(tp.DType.F64, *args, config=tp.PlanConfig(split_k=1)).report()["kernel"]
    assert make_plan(tp.DType.F64, *args, config=tp.PlanConfig(split_k=5)).report()["kernel"].endswith(".splitk5")
    with pytest.raises(tp.TprimsError):
        make_plan(tp.DType.F64, *args, config=tp.PlanConfig(split_k=-1))


def test_tile_choice_follows_the_busiest_cu_estimate():
    """Large products take the 128x128 DMA tile; small ones a tile that does
    not leave one CU computing an oversized tile (estimated MFMA time of the
    busiest CU: tiles per CU x tile area over the SIMDs their waves occupy)."""
    import tprims_amd as tp
    from harness import T, make_plan

    def gemm(dt, m, n, k, h=1, a_k=False):
        a = T([m, k, h], [k, 1, m * k] if a_k else [1, m, m * k], [0, 2, 3])
        return make_plan(dt, a, T([k, n, h], [1, k, k * n], [2, 1, 3]),
                         T([m, n, h], [1, m, m * n], [0, 1, 3]), config=tp.PlanConfig(split_k=1)).report()

    for dt in (tp.DType.F32, tp.DType.F64, tp.DType.C32, tp.DType.C64):
        # A contiguous along K for c64: with an M-major A it takes 64x64 (the
        # complex rule below); c32 in 16-byte runs takes 64x64 either way
        big = gemm(dt, 4096, 4096, 256, a_k=dt.is_complex)
        if dt == tp.DType.C32:
            assert big["tile"] == 1 and big["kernel"].endswith("64x64x16.dma16"), big["kernel"]
            continue
        # f32 and f64 gather 16-byte runs (column-major operands qualify)
        suffix = "128x128x16.dma" if dt == tp.DType.C64 else "128x128x16.dma16"
        assert big["tile"] == 4 and big["kernel"].endswith(suffix), big["kernel"]
        if dt.is_complex:
            # A gathered along M: the 64x64 tile (TCCG ij-ik-kj c32 2904^3:
            # 1.67 ms on 64x64 against 2.19 on 128x128)
            assert gemm(dt, 4096, 4096, 256)["tile"] == 1
    # 512 x 512 x 16 batches over K = 4096 with A and B contiguous along K
    # fill the GPU with 128x128 tiles (the headline shape); with K = 64 the
    # short-K rule takes 64x64
    assert gemm(tp.DType.C64, 512, 512, 4096, 16, a_k=True)["tile"] == 4
    assert gemm(tp.DType.C64, 512, 512, 64, 16, a_k=True)["tile"] == 1
    # A a stride-2 view along K still runs along K (the headline's A)
    for dt in (tp.DType.C32, tp.DType.C64):
        m, k, n, h = 512, 4096, 512, 16
        a2 = T([m, k, h], [2 * k, 2, 2 * m * k], [0, 2, 3])
        rep = make_plan(dt, a2, T([k, n, h], [1, k, k * n], [2, 1, 3]), T([m, n, h], [1, m, m * n], [0, 1, 3]),
                        config=tp.PlanConfig(split_k=1)).report()
        assert rep["tile"] == 4, rep
    # one 64x64 output: a 64x64 tile on 4 SIMDs, not a 128x128 one
    small = gemm(tp.DType.C64, 64, 64, 64)
    assert small["tile"] == 1 and small["kernel"].endswith("64x64x16.dma"), small["kernel"]
    # 128 tiles of 128x128 would leave half the CUs idle: 64x64 spreads over all
    assert gemm(tp.DType.C64, 256, 256, 1024, 32)["tile"] == 1
    # an explicit tile is kept
    p = make_plan(tp.DType.F64, T([64, 64], [1, 64], [0, 2]), T([64, 64], [1, 64], [2, 1]),
                  T([64, 64], [1, 64], [0, 1]), config=tp.PlanConfig(tile=4)).report()
    assert p["tile"] == 4


def test_scatter_vectors_match_the_reference_examples():
    """build_scatter / build_block_scatter against the reference's own
    examples (crates/tprims-kernel/src/pack/scatter.rs:62-68, 119-128,
    138-146, 285-331, 360-366)."""
    import tprims_amd as tp
    assert tp.build_scatter([2, 3], [1, 10]) == [0, 1, 10, 11, 20, 21]
    assert tp.build_scatter([], []) == [0]
    assert tp.build_scatter([2, 3], [1, 2]) == [0, 1, 2, 3, 4, 5]
    assert tp.build_scatter([3, 2], [1, 3]) == [0, 1, 2, 3, 4, 5]
    assert tp.build_scatter([2, 3], [3, 1]) == [0, 3, 1, 4, 2, 5]
    assert tp.build_scatter([4], [0]) == [0, 0, 0, 0]
    assert tp.build_scatter([2, 0, 3], [1, 2, 2]) == []
    s = tp.build_scatter([2, 3], [1, 10])
    assert tp.build_block_scatter(s, 2) == [1, 1, 1]
    assert tp.build_block_scatter(s, 3) == [tp.IRREGULAR, tp.IRREGULAR]
    assert tp.build_block_scatter([0, 1, 2], 2) + tp.build_block_scatter([0, 5], 2) == [1, 0, 5]
    assert tp.build_block_scatter(tp.build_scatter([2, 3], [3, 1]), 2) == [3, 3, 3]
    assert tp.build_block_scatter([0, 0, 0, 0], 2) == [0, 0]
    assert tp.build_block_scatter(list(range(5)), 2) == [1, 1, 0]
    assert tp.build_block_scatter([], 4) == []
    with pytest.raises(tp.TprimsError):
        tp.build_block_scatter([0, 1], 0)  # the reference panics on a zero block size
    with pytest.raises(tp.TprimsError):
        tp.build_scatter([-1], [1])


def test_operand_contiguous_axis_leads_its_role_when_the_operand_is_large():
    """Two M axes, A contiguous along one and D along the other: with 4 K >=
    N (A at least a quarter of D's size) A's contiguous axis comes first in
    M, so a tile's rows read runs of A (TCCG ajbdc-ckbad-jk-like); with D
    much the larger tensor D's contiguous axis keeps the lead. N likewise
    with B; and where D runs along the other role the operand leads at any
    size."""
    import tprims_amd as tp
    from harness import T, make_plan
    # D[a, c, n] = sum_k A[c, k, a] B[k, n], a and c M axes (A contiguous in
    # c, D in a); K = 200 (up to K = 128 the rows are blocked instead)
    ea, ec, ek = 6, 5, 200
    a = T([ec, ek, ea], [1, ec, ec * ek], [2, 3, 1])
    for en, leads_with_c in ((4, True), (800, True), (1000, False)):
        b = T([ek, en], [1, ek], [3, 4])
        d = T([ea, ec, en], [1, ea, ea * ec], [1, 2, 4])
        sc = make_plan(tp.DType.F64, a, b, d).scatters()
        if leads_with_c:  # m = c + ec * a
            assert sc["am"][:ec + 1] == list(range(ec)) + [ec * ek], sc["am"][:8]
            assert sc["dm"][:2] == [0, ea]
        else:  # m = a + ea * c
            assert sc["dm"][:ea + 1] == list(range(ea)) + [ea], sc["dm"][:8]
            assert sc["am"][:2] == [0, ec * ek]
    # N: D[m, j, l] = sum_k A[m, k] B[l, k, j], B contiguous in l, D in j
    # (D's contiguous axis an N axis: the 4 K >= M rule decides)
    em, ej, el, ek = 4, 3, 5, 8
    b = T([el, ek, ej], [1, el, el * ek], [6, 3, 5])
    for em, leads_with_l in ((4, True), (32, True), (64, False)):
        a = T([em, ek], [1, em], [1, 3])
        d = T([em, ej, el], [ej, 1, em * ej], [1, 5, 6])
        sc = make_plan(tp.DType.F64, a, b, d).scatters()
        assert (sc["bn"][1] == 1) == leads_with_l, sc["bn"][:4]
    # D contiguous in m (an M axis): its writes run along M whatever the N
    # order, so B's contiguous axis leads N at any size (TCCG CCSD(T)
    # abcijk-ikmc-mjab, A likewise with D along N: f32 1.21x)
    for em in (4, 64, 4096):
        a = T([em, ek], [1, em], [1, 3])
        d = T([em, ej, el], [1, em, em * ej], [1, 5, 6])
        sc = make_plan(tp.DType.F64, a, b, d).scatters()
        assert sc["bn"][1] == 1, (em, sc["bn"][:4])


def test_rows_are_blocked_when_a_and_d_run_along_different_m_axes_and_k_is_short():
    """D[a, c, n] = sum_k A[c, k, a] B[k, n] with extents 24 (TCCG InTensLi
    ajbdc-ckbad-jk-like: A contiguous in c, D in a): with K <= 128 the M
    order is a in blocks of 8, c in blocks of 8, then the rest of a and of c,
    so a tile's rows write runs of 8 of D and read runs of 8 of A; the short
    K also takes the 64x64 tile. K = 200 keeps one axis leading."""
    import tprims_amd as tp
    from harness import T, make_plan
    e = 24
    for ek, blocked in ((24, True), (128, True), (200, False)):
        a = T([e, ek, e], [1, e, e * ek], [2, 3, 1])
        b = T([ek, e], [1, ek], [3, 4])
        d = T([e, e, e], [1, e, e * e], [1, 2, 4])
        plan = make_plan(tp.DType.F64, a, b, d)
        sc = plan.scatters()
        if blocked:
            assert sc["dm"][:9] == list(range(8)) + [e], sc["dm"][:9]
            assert sc["am"][:2] == [0, e * ek] and sc["am"][8] == 1, sc["am"][:9]
            # rows 64..: the second block of a (D offsets 8..15)
            assert sc["dm"][64] == 8 and sc["am"][64] == 8 * e * ek
        else:  # 4 K >= N: A's axis c leads (m = c + 24 a)
            assert sc["am"][:e] == list(range(e)) and sc["dm"][:2] == [0, e], sc["am"][:4]
    # the short-K tile rule: 24^3 rows x 24 columns of f32, K = 24
    a = T([e, e, e, e], [1, e, e * e, e ** 3], [2, 3, 1, 5])
    b = T([e, e], [1, e], [3, 4])
    d = T([e, e, e, e], [1, e, e * e, e ** 3], [1, 2, 4, 5])
    rep = make_plan(tp.DType.F32, a, b, d).report()
    assert rep["tile"] == 1, rep


def test_plan_scatters_are_the_reduced_matrix_view():
    """The plan's twelve scatter vectors: a column-major matrix product, and
    the f64 24^6 contraction (no fold across a and b, K ordered with B's unit
    stride first), whose M/N runs of 24 are what lets it gather 16-byte runs."""
    import tprims_amd as tp
    from harness import T, make_plan
    p = make_plan(tp.DType.F64, T([4, 6], [1, 4], [0, 2]), T([6, 5], [1, 6], [2, 1]), T([4, 5], [1, 4], [0, 1]))
    sc = p.scatters()
    assert sc["am"] == [0, 1, 2, 3] and sc["ak"] == [0, 4, 8, 12, 16, 20]
    assert sc["bk"] == [0, 1, 2, 3, 4, 5] and sc["bn"] == [0, 6, 12, 18, 24]
    assert sc["dm"] == sc["cm"] == [0, 1, 2, 3] and sc["dn"] == sc["cn"] == [0, 4, 8, 12, 16]
    assert sc["ha"] == sc["hb"] == sc["hc"] == sc["hd"] == [0]
    e = 24
    st = [1, e, e * e, e ** 3]
    p = make_plan(tp.DType.F64, T([e] * 4, st, [0, 4, 1, 5]), T([e] * 4, st, [5, 2, 4, 3]), T([e] * 4, st, [0, 1, 2, 3]))
    sc = p.scatters()
    assert sc["am"] == tp.build_scatter([e, e], [1, e * e])
    assert sc["dm"] == tp.build_scatter([e, e], [1, e])
    assert sc["bn"] == tp.build_scatter([e, e], [e, e ** 3])
    assert sc["dn"] == tp.build_scatter([e, e], [e * e, e ** 3])
    assert sc["bk"] == tp.build_scatter([e, e], [1, e * e])
    assert sc["ak"] == tp.build_scatter([e, e], [e ** 3, e])
    # runs of 24 along A's M and B's K: regular with stride 1 at the f64 run of 2
    assert set(tp.build_block_scatter(sc["am"], 2)) == {1} and set(tp.build_block_scatter(sc["bk"], 2)) == {1}
    assert ".regs16" in p.report()["kernel"]  # 16-byte run loads
    # one odd offset breaks the runs: A with a leading dimension of 25 (odd row starts)
    q = make_plan(tp.DType.F64, T([6, 8], [1, 7], [2, 0]), T([6, 4], [1, 6], [2, 1]), T([8, 4], [1, 8], [0, 1]))
    assert q.report()["a_k_major"] == 1 and ".dma16" not in q.report()["kernel"]


def _build_c_consumer(tmp_path):
    """Compiles tests/c/consumer.c as C99 against include/ and libtprims."""
    from tprims_amd import _capi
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.dirname(_capi.LIB_PATH)
    exe = str(tmp_path / "consumer")
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                           os.path.join(root, "tests", "c", "consumer.c"), "-L", libdir, "-ltprims",
                           "-Wl,-rpath," + libdir, "-o", exe])
    return exe


def build_fixed_cost(out_dir):
    """Compiles tests/c/fixed_cost.c (C99, HIP runtime API for the device
    buffers) against include/ and libtprims; returns the executable."""
    from tprims_amd import _capi
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.dirname(_capi.LIB_PATH)
    exe = os.path.join(str(out_dir), "fixed_cost")
    subprocess.check_call(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                           "-I", os.path.join(root, "include"), "-I", "/opt/rocm/include",
                           os.path.join(root, "tests", "c", "fixed_cost.c"), "-L", libdir, "-ltprims",
                           "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib",
                           "-o", exe])
    return exe


def test_fixed_cost_program_compiles(tmp_path):
    """The C timing program of one small TAPP_execute_product builds as C99
    (bench.py runs it on the GPU)."""
    assert os.path.exists(build_fixed_cost(tmp_path))


def test_c_consumer_compiles_against_the_headers(tmp_path):
    """A C host that uses only the TAPP C ABI and the tprims extensions
    compiles as strict C99 and links against libtprims (it runs on a GPU:
    tests/test_gpu_c_consumer.py)."""
    assert os.path.exists(_build_c_consumer(tmp_path))


def test_batched_null_item_pointer_is_refused_before_any_work():
    """TAPP_execute_batched_product validates every item before running any
    (crates/tprims-capi/src/execute.rs:235): a null D in the second item is an
    invalid-argument error, raised without touching a device."""
    import tprims_amd as tp
    from harness import T, make_plan
    p = make_plan(tp.DType.F64, T([4, 3], [1, 4], [0, 2]), T([3, 5], [1, 3], [2, 1]), T([4, 5], [1, 4], [0, 1]))
    with pytest.raises(tp.TprimsError) as e:
        p.execute_batched_raw(1.0, [4096, 8192], [4096, 8192], 0.0, None, [8192, None])
    assert e.value.status == 1
