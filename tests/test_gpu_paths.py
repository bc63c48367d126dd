"""GPU parity of the paths the library ships beyond the default executor:
the f64 24^6 benchmark contraction through the plain TAPP C ABI (default
split-K plan, one K range, host buffers); executors on caller streams
(asynchronous, as the benchmark runs them) and TAPP_create_executor
executors, two of them interleaved; the fused batched launch and split K on
an asynchronous executor; the reference conformance corners that need their
own driver (repeated execution of one plan, tensor-info setters before
planning, alpha = 0 with NaN operands); DLPack operands in ROCm memory; the
f32 adjoint against central finite differences. Every result is compared with
the FP64 label oracle (oracle/)."""
import ctypes

import numpy as np
import pytest

import oracle
import tprims_amd as tp
from cases import CASES, scalar
from harness import DTYPES, NP, T, check, make_plan, run, tol, values
from tprims_amd import _capi
from tprims_amd._capi import i64_array, intptr, lib

A_, E_, B_, F_, C_, D_ = 1, 2, 3, 4, 5, 6


# ---- the f64 24^6 contraction through the TAPP C ABI ---------------------------

def _tapp_product(dtype, tensors, ops=(0, 0, 0, 0), cfg=None):
    """TAPP_create_tensor_product (or the _ex form with a configuration) on
    four (T) tensors A, B, C, D, as a C caller writes it."""
    h = intptr(0)
    assert lib().TAPP_create_handle(ctypes.byref(h)) == 0
    infos, labels = [], []
    for t in tensors:
        i = intptr(0)
        assert lib().TAPP_create_tensor_info(ctypes.byref(i), int(dtype), len(t.extents), i64_array(t.extents),
                                             i64_array(t.strides)) == 0
        infos.append(i.value)
        labels.append(i64_array(t.labels))
    p = intptr(0)
    args = []
    for k in range(4):
        args += [ops[k], infos[k], labels[k]]
    if cfg is None:
        st = lib().TAPP_create_tensor_product(ctypes.byref(p), h.value, *args, _capi.TAPP_DEFAULT_PREC)
    else:
        st = lib().tprims_create_tensor_product_ex(ctypes.byref(p), h.value, *args, _capi.TAPP_DEFAULT_PREC,
                                                   ctypes.byref(cfg.to_c()))
    assert st == 0, lib().tprims_last_error()
    for i in infos:
        assert lib().TAPP_destroy_tensor_info(i) == 0
    return p.value, h.value


def _report(plan):
    r = _capi.PlanReportC()
    assert lib().tprims_product_report(plan, ctypes.byref(r)) == 0
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["default", "one_k_range", "split_k3", "host_buffers"])
def test_f64_24_contraction_through_the_tapp_abi(gpu, variant):
    """D[a,b,c,d] = sum_{e,f} A[a,e,b,f] B[f,c,e,d], every extent 24,
    column-major, U[-1,1] seeded, alpha 1, beta 0: the whole D against the
    oracle at the reference's f64 budget (driver/tests/correctness.rs:77-88).
    The planner's choice (the K-group 32x32 tile on 16-byte runs), one K range
    on the 64x64 tile, and three K ranges with the split-K reduction."""
    import torch

    e = 24
    a = T([e] * 4, tp.col_major([e] * 4), [A_, E_, B_, F_])
    b = T([e] * 4, tp.col_major([e] * 4), [F_, C_, E_, D_])
    d = T([e] * 4, tp.col_major([e] * 4), [A_, B_, C_, D_])
    cfg = {"one_k_range": tp.PlanConfig(split_k=1), "split_k3": tp.PlanConfig(split_k=3)}.get(variant)
    plan, handle = _tapp_product(tp.DType.F64, [a, b, d, d], cfg=cfg)
    kern = _report(plan).kernel.decode()
    if variant == "one_k_range":
        assert ".splitk" not in kern and ".kgroup" not in kern, kern
    elif variant == "split_k3":
        assert ".dma16.splitk3" in kern, kern
    else:
        assert "32x48x16.regs16.kgroup8" in kern, kern  # the planner's choice: 16-byte runs, eight K groups
    rng = np.random.default_rng(24)
    av, bv = rng.uniform(-1, 1, e ** 4), rng.uniform(-1, 1, e ** 4)
    dv = np.full(e ** 4, np.nan)  # beta = 0: D is written, never read
    alpha, beta = ctypes.c_double(1.0), ctypes.c_double(0.0)
    st = intptr(0)
    if variant == "host_buffers":
        res = lib().TAPP_execute_product(plan, 0, ctypes.byref(st), ctypes.byref(alpha), av.ctypes.data,
                                         bv.ctypes.data, ctypes.byref(beta), None, dv.ctypes.data)
        got = dv
    else:
        ta, tb = torch.from_numpy(av).to(gpu), torch.from_numpy(bv).to(gpu)
        td = torch.from_numpy(dv).to(gpu)
        torch.cuda.synchronize()
        res = lib().TAPP_execute_product(plan, 0, ctypes.byref(st), ctypes.byref(alpha), ta.data_ptr(),
                                         tb.data_ptr(), ctypes.byref(beta), None, td.data_ptr())
        torch.cuda.synchronize()
        got = td.cpu().numpy()
    assert res == 0, lib().tprims_last_error()
    want = oracle.contract(1.0, a.oracle_op(av), b.oracle_op(bv), 0.0, None, d.oracle_op(np.zeros(e ** 4)))
    err = oracle.max_rel_err(got, want)
    assert err <= 1e-11, "max_rel_err %.3e (%s)" % (err, kern)
    assert lib().TAPP_destroy_tensor_product(plan) == 0
    assert lib().TAPP_destroy_handle(handle) == 0


# ---- executors ---------------------------------------------------------------------

def _executor(kind):
    import torch

    if kind == "caller_stream_async":  # the benchmark's executor
        s = torch.cuda.Stream()
        ex = tp.Executor(0, s.cuda_stream, asynchronous=True)
        ex._keep = s
        return ex
    if kind == "own_stream_async":
        return tp.Executor(0, None, asynchronous=True)
    # TAPP_create_executor: a synchronous executor with a stream of its own
    ex = tp.Executor.__new__(tp.Executor)
    h = intptr(0)
    assert lib().TAPP_create_executor(ctypes.byref(h)) == 0
    ex.handle = h.value
    return ex


EXECUTORS = ["caller_stream_async", "own_stream_async", "tapp_create_executor"]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", EXECUTORS)
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_conformance_case_on_executor(gpu, name, dtype, kind):
    kw = CASES[name](dtype)
    ex = _executor(kind)
    try:
        got, want, plan = run(dtype, executor=ex, **kw)
    finally:
        ex.close()
    exact = name == "alpha_zero" or name.startswith("empty_k")
    check(dtype, got, want, exact=exact)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_two_executors_interleaved(gpu, dtype):
    """Two executors (a caller-stream asynchronous one and a TAPP one) take
    turns over several plans without synchronizing in between; each result
    matches the oracle."""
    import torch

    exs = [_executor("caller_stream_async"), _executor("tapp_create_executor")]
    names = ["large_k", "padded_view", "general_strides", "conj_mask_13_deep_k", "reduction_in_b", "alpha_beta"]
    dev = torch.device("cuda:0")
    jobs = []
    for i, nm in enumerate(names):
        kw = CASES[nm](dtype)
        a, b, d = kw["a"], kw["b"], kw["d"]
        c, cv = kw.get("c"), kw.get("cv")
        plan = make_plan(dtype, a, b, d, c)
        es = dtype.size
        ta, tb = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (kw["av"], kw["bv"]))
        tds = [torch.from_numpy(np.ascontiguousarray(kw["dv"])).to(dev) for _ in range(2)]  # one D per round
        tc = torch.from_numpy(cv).to(dev) if cv is not None else None
        jobs.append((nm, kw, plan, ta, tb, tc, tds))
    torch.cuda.synchronize()
    for rnd in range(2):
        for i, (nm, kw, plan, ta, tb, tc, tds) in enumerate(jobs):
            a, b, d, c = kw["a"], kw["b"], kw["d"], kw.get("c")
            es = dtype.size
            pc = tc.data_ptr() + c.origin() * es if tc is not None else None
            plan.execute_raw(kw.get("alpha", 1.0), ta.data_ptr() + a.origin() * es, tb.data_ptr() + b.origin() * es,
                             kw.get("beta", 0.0), pc, tds[rnd].data_ptr() + d.origin() * es, exs[(i + rnd) % 2])
    for ex in exs:
        ex.synchronize()
    torch.cuda.synchronize()
    for nm, kw, plan, ta, tb, tc, tds in jobs:
        a, b, d, c = kw["a"], kw["b"], kw["d"], kw.get("c")
        want = oracle.contract(kw.get("alpha", 1.0), a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]),
                               kw.get("beta", 0.0), c.oracle_op(kw["cv"]) if c is not None else None,
                               d.oracle_op(kw["dv"]), cplx=dtype.is_complex)
        for td in tds:
            check(dtype, td.cpu().numpy(), want)
    for ex in exs:
        ex.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_split_k_and_batched_on_an_asynchronous_executor(gpu, dtype):
    """Split-K plans of growing size (the executor's workspace grows between
    them) and a fused batched call, queued back to back on one asynchronous
    executor and synchronized once."""
    import torch

    ex = _executor("caller_stream_async")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)

    def vals(t):
        v = rng.uniform(-1, 1, t.storage()) + (1j * rng.uniform(-1, 1, t.storage()) if dtype.is_complex else 0)
        return v.astype(NP[dtype])

    pending = []
    for M, K, N in ((40, 1500, 24), (96, 2600, 80), (130, 3000, 70)):
        a, b = T([M, K], [1, M], [0, 2], conj=dtype.is_complex), T([K, N], [1, K], [2, 1])
        d = T([M, N], [1, M], [0, 1])
        plan = make_plan(dtype, a, b, d, d)
        assert ".splitk" in plan.report()["kernel"]
        av, bv, cv = vals(a), vals(b), vals(d)
        ta, tb, tc = (torch.from_numpy(x).to(dev) for x in (av, bv, cv))
        td = torch.zeros_like(tc)
        torch.cuda.synchronize()
        plan.execute_raw(0.75, ta.data_ptr(), tb.data_ptr(), -0.5, tc.data_ptr(), td.data_ptr(), ex)
        want = oracle.contract(0.75, a.oracle_op(av), b.oracle_op(bv), -0.5, d.oracle_op(cv),
                               d.oracle_op(np.zeros(d.storage(), NP[dtype])), cplx=dtype.is_complex)
        pending.append((td, want, (ta, tb, tc)))
    # The fused batched launch: 6 items of one plan, operands in one allocation each.
    a, b, d = T([33, 70], [1, 33], [0, 2]), T([70, 21], [1, 70], [2, 1]), T([33, 21], [1, 33], [0, 1])
    plan = make_plan(dtype, a, b, d, d)
    n = 6
    av, bv = vals(T([a.storage() * n], [1], [0])), vals(T([b.storage() * n], [1], [0]))
    ta, tb = torch.from_numpy(av).to(dev), torch.from_numpy(bv).to(dev)
    td = torch.zeros(d.storage() * n, dtype=ta.dtype, device=dev)
    es = dtype.size
    torch.cuda.synchronize()
    plan.execute_batched_raw(1.25, [ta.data_ptr() + i * a.storage() * es for i in range(n)],
                             [tb.data_ptr() + i * b.storage() * es for i in range(n)], 0.0, None,
                             [td.data_ptr() + i * d.storage() * es for i in range(n)], ex)
    ex.synchronize()
    torch.cuda.synchronize()
    for tdi, want, _ in pending:
        check(dtype, tdi.cpu().numpy(), want)
    got = td.cpu().numpy()
    for i in range(n):
        want = oracle.contract(1.25, a.oracle_op(av[i * a.storage():(i + 1) * a.storage()]),
                               b.oracle_op(bv[i * b.storage():(i + 1) * b.storage()]), 0.0, None,
                               d.oracle_op(np.zeros(d.storage(), NP[dtype])), cplx=dtype.is_complex)
        check(dtype, got[i * d.storage():(i + 1) * d.storage()], want)
    ex.close()


# ---- conformance corners with their own driver --------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("where", ["gpu", "host"])
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_one_plan_executes_repeatedly(gpu, dtype, where):
    """tapp_conformance.rs:774-813: one plan run three times over fresh output
    buffers gives the same result, bit for bit, each time."""
    kw = CASES["large_k"](dtype)
    plan = make_plan(dtype, kw["a"], kw["b"], kw["d"])
    outs = []
    for rep in range(3):
        k = dict(kw)
        k["dv"] = np.zeros_like(kw["dv"])
        got, want, _ = run(dtype, plan=plan, where=where, **k)
        check(dtype, got, want)
        outs.append(got)
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_tensor_info_setters_take_effect_before_planning(gpu, dtype):
    """tapp_conformance.rs:848-900: infos created with the wrong extents and
    strides, corrected with TAPP_set_extents / TAPP_set_strides, then planned
    and executed; the result is the oracle's."""
    import torch

    I, J, K = ord("i"), ord("j"), ord("k")
    a, b, d = T([5, 3], [1, 5], [I, K]), T([3, 4], [4, 1], [K, J]), T([5, 4], [1, 5], [I, J])
    h = intptr(0)
    assert lib().TAPP_create_handle(ctypes.byref(h)) == 0
    infos = []
    for t in (a, b, d, d):
        i = intptr(0)
        assert lib().TAPP_create_tensor_info(ctypes.byref(i), int(dtype), 2, i64_array([2, 2]),
                                             i64_array([2, 1])) == 0
        assert lib().TAPP_set_extents(i.value, i64_array(t.extents)) == 0
        assert lib().TAPP_set_strides(i.value, i64_array(t.strides)) == 0
        infos.append(i.value)
    p = intptr(0)
    assert lib().TAPP_create_tensor_product(ctypes.byref(p), h.value, 0, infos[0], i64_array(a.labels), 0, infos[1],
                                            i64_array(b.labels), 0, infos[2], i64_array(d.labels), 0, infos[3],
                                            i64_array(d.labels), _capi.TAPP_DEFAULT_PREC) == 0
    av, bv = values(dtype, a, 281), values(dtype, b, 282)
    dev = torch.device("cuda:0")
    ta, tb = torch.from_numpy(av).to(dev), torch.from_numpy(bv).to(dev)
    td = torch.zeros(d.storage(), dtype=ta.dtype, device=dev)
    pa, _k1 = tp._scalar(1.0, dtype)
    pb, _k2 = tp._scalar(0.0, dtype)
    torch.cuda.synchronize()
    assert lib().TAPP_execute_product(p.value, 0, None, pa, ta.data_ptr(), tb.data_ptr(), pb, None,
                                      td.data_ptr()) == 0
    torch.cuda.synchronize()
    want = oracle.contract(1.0, a.oracle_op(av), b.oracle_op(bv), 0.0, None,
                           d.oracle_op(np.zeros(d.storage(), NP[dtype])), cplx=dtype.is_complex)
    check(dtype, td.cpu().numpy(), want)
    for i in infos:
        lib().TAPP_destroy_tensor_info(i)
    lib().TAPP_destroy_tensor_product(p.value)
    lib().TAPP_destroy_handle(h.value)


@pytest.mark.gpu
@pytest.mark.parametrize("method", [tp.ComplexMethod.NATIVE, tp.ComplexMethod.ONE_M, tp.ComplexMethod.THREE_M])
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_alpha_zero_never_reads_a_or_b(gpu, dtype, method):
    """tapp_conformance.rs:446 with A and B poisoned by NaN: alpha = 0 gives
    D = beta * C exactly, so A and B were never read (0 * NaN would be NaN)."""
    if method != tp.ComplexMethod.NATIVE and not dtype.is_complex:
        pytest.skip("complex methods apply to complex types")
    kw = CASES["large_k"](dtype)
    nan = np.nan + (1j * np.nan if dtype.is_complex else 0)
    kw["av"] = np.full_like(kw["av"], nan)
    kw["bv"] = np.full_like(kw["bv"], nan)
    d = kw["d"]
    cv = values(dtype, d, 7)
    beta = scalar(dtype, 0.5, -0.25)
    kw.update(alpha=0.0, beta=beta, c=d, cv=cv)
    got, _, plan = run(dtype, config=tp.PlanConfig(complex_method=method), **kw)
    want = (np.asarray(beta, dtype=NP[dtype]) * cv).astype(NP[dtype])
    np.testing.assert_array_equal(got, want)


# ---- DLPack operands in ROCm memory ------------------------------------------------------

class _DevView:
    """A kDLROCM DLTensor over a device buffer: base pointer, byte_offset of
    the logical origin, the case's (possibly negative) strides."""

    def __init__(self, buf, t, dtype, device_id=0):
        self.buf = buf
        self.shape = i64_array(t.extents)
        self.strides = i64_array(t.strides)
        v = _capi.DLTensor()
        v.data = buf.data_ptr()
        v.device.device_type, v.device.device_id = 10, device_id
        v.ndim = len(t.extents)
        v.dtype.code, v.dtype.bits, v.dtype.lanes = (5 if dtype.is_complex else 2), 8 * dtype.size, 1
        v.shape = ctypes.cast(self.shape, _capi.i64p)
        v.strides = ctypes.cast(self.strides, _capi.i64p)
        v.byte_offset = t.origin() * dtype.size
        self.v = v

    def borrow(self):
        return lib().tprims_tensor_borrow_raw(ctypes.byref(self.v), 0)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", ["large_k", "negative_strides", "padded_view", "output_diagonal", "hadamard",
                                  "full_reduction", "conj_mask_15", "c_layout_differs", "empty_k"])
def test_execute_dlpack_on_rocm_views(gpu, dtype, name):
    """tprims_execute_product_dlpack on kDLROCM views (byte offsets to the
    logical origin, negative strides, diagonals, rank 0, empty K) against the
    oracle; a view on another device id is refused before any work."""
    import torch

    kw = CASES[name](dtype)
    a, b, d = kw["a"], kw["b"], kw["d"]
    c, cv = kw.get("c"), kw.get("cv")
    plan = make_plan(dtype, a, b, d, c)
    dev = torch.device("cuda:0")
    ta, tb, td = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (kw["av"], kw["bv"], kw["dv"]))
    tc = torch.from_numpy(cv).to(dev) if cv is not None else None
    va, vb, vd = _DevView(ta, a, dtype), _DevView(tb, b, dtype), _DevView(td, d, dtype)
    vc = _DevView(tc, c, dtype) if tc is not None else None
    alpha, beta = kw.get("alpha", 1.0), kw.get("beta", 0.0)
    pa, _k1 = tp._scalar(alpha, dtype)
    pb, _k2 = tp._scalar(beta, dtype)
    null = lib().tprims_tensor_borrow_raw(None, 0)
    torch.cuda.synchronize()
    other = _DevView(td, d, dtype, device_id=1)
    assert lib().tprims_execute_product_dlpack(plan.handle, 0, pa, va.borrow(), vb.borrow(), pb,
                                               vc.borrow() if vc else null, other.borrow()) == \
        _capi.TPRIMS_ERR_DEVICE
    st = lib().tprims_execute_product_dlpack(plan.handle, 0, pa, va.borrow(), vb.borrow(), pb,
                                             vc.borrow() if vc else null, vd.borrow())
    assert st == 0, lib().tprims_last_error()
    torch.cuda.synchronize()
    want = oracle.contract(alpha, a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]), beta,
                           c.oracle_op(cv) if c is not None else None, d.oracle_op(kw["dv"]), cplx=dtype.is_complex)
    check(dtype, td.cpu().numpy(), want)


# ---- f32 adjoint against central finite differences ---------------------------------------

@pytest.mark.gpu
def test_f32_vjp_matches_central_finite_differences(gpu):
    """dA and dB of an f32 contraction with a diagonal, a reduced index and a
    batch index against central differences (h = 1e-3) of the same
    contraction evaluated in FP64 (an f64 plan of the same problem), along
    random directions."""
    import torch

    from test_gpu_vjp import diag_problem, rand

    rng = np.random.default_rng(9)
    a, b, d = diag_problem(5, 4, 3, 2, 6)
    f32, f64 = tp.DType.F32, tp.DType.F64
    av, bv, W = rand(f32, a.storage(), rng), rand(f32, b.storage(), rng), rand(f32, d.storage(), rng)
    plan32, plan64 = make_plan(f32, a, b, d), make_plan(f64, a, b, d)
    ta, tb, tw = (torch.from_numpy(x).to(gpu) for x in (av, bv, W))
    da, db = torch.empty_like(ta), torch.empty_like(tb)
    plan32.vjp(0.5, ta, tb, tw, da, db)
    torch.cuda.synchronize()
    gA, gB = da.cpu().numpy().astype(np.float64), db.cpu().numpy().astype(np.float64)

    def loss(Av, Bv):
        tA, tB = torch.from_numpy(Av).to(gpu), torch.from_numpy(Bv).to(gpu)
        out = torch.zeros(d.storage(), dtype=torch.float64, device=gpu)
        plan64.execute_raw(0.5, tA.data_ptr(), tB.data_ptr(), 0.0, None, out.data_ptr())
        torch.cuda.synchronize()
        return float(np.dot(W.astype(np.float64), out.cpu().numpy()))

    h = 1e-3
    A64, B64 = av.astype(np.float64), bv.astype(np.float64)
    for trial in range(3):
        V = rng.standard_normal(a.storage())
        fd = (loss(A64 + h * V, B64) - loss(A64 - h * V, B64)) / (2 * h)
        an = float(np.dot(gA, V))
        assert abs(fd - an) <= 2e-5 * max(1.0, np.linalg.norm(gA) * np.linalg.norm(V)), (trial, fd, an)
        U = rng.standard_normal(b.storage())
        fd = (loss(A64, B64 + h * U) - loss(A64, B64 - h * U)) / (2 * h)
        an = float(np.dot(gB, U))
        assert abs(fd - an) <= 2e-5 * max(1.0, np.linalg.norm(gB) * np.linalg.norm(U)), (trial, fd, an)


@pytest.mark.gpu
def test_destroying_an_executor_with_queued_work_is_busy(gpu):
    """An asynchronous call returns with its kernel still queued: destroying
    the executor then reports TPRIMS_BUSY (the reference's in-flight rule,
    crates/tprims-capi/src/executor.rs:285-305) and succeeds once synchronized."""
    import torch

    n = 2048
    a, b, d = T([n, n], [1, n], [0, 2]), T([n, n], [1, n], [2, 1]), T([n, n], [1, n], [0, 1])
    plan = make_plan(tp.DType.F64, a, b, d)
    dev = torch.device("cuda:0")
    ta = torch.randn(n * n, dtype=torch.float64, device=dev)
    tb = torch.randn(n * n, dtype=torch.float64, device=dev)
    td = torch.zeros(n * n, dtype=torch.float64, device=dev)
    ex = tp.Executor(0, None, asynchronous=True)
    torch.cuda.synchronize()
    for _ in range(4):  # about a millisecond of GPU work
        plan.execute_raw(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr(), ex)
    assert lib().TAPP_destroy_executor(ex.handle) == _capi.TPRIMS_BUSY
    ex.synchronize()
    assert lib().TAPP_destroy_executor(ex.handle) == 0
    ex.handle = 0
