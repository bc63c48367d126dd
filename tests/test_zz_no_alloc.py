"""GPU tests of tprims_execute_product_noalloc, the entry point that
allocates nothing (its device scratch is one static array per GPU): through
the C ABI, its output equals TAPP_execute_product's on the same plan and
inputs (bit for bit: the same kernels in the same order) and the oracle's,
over two calls in a row whose second is smaller, on a synchronous executor
and queued back to back on an asynchronous one; every conformance case; and
its refusals (host operands, split-BF16, a scratch need beyond the array)."""
import ctypes

import numpy as np
import pytest

import oracle
import tprims_amd as tp
from cases import CASES
from harness import DTYPES, NP, T, check, make_plan, values
from tprims_amd import _capi
from tprims_amd._capi import intptr, lib

SCRATCH_BYTES = 64 << 20  # TPRIMS_NOALLOC_SCRATCH_BYTES


def _call(fn, plan, ex, alpha, a, b, beta, c, d, dtype):
    pa, _k1 = tp._scalar(alpha, dtype)
    pb, _k2 = tp._scalar(beta, dtype)
    st = intptr(0)
    return fn(plan.handle, ex.handle if ex is not None else 0, ctypes.byref(st), pa, a, b, pb, c, d)


def _big_case():
    """The f64 24^6 contraction D[a,b,c,d] = sum_{e,f} A[a,e,b,f] B[f,c,e,d]
    with three K ranges asked for (split K: scatter vectors and partial tiles
    both in the scratch)."""
    e = 24
    a = T([e] * 4, tp.col_major([e] * 4), [1, 2, 3, 4])
    b = T([e] * 4, tp.col_major([e] * 4), [4, 5, 2, 6])
    d = T([e] * 4, tp.col_major([e] * 4), [1, 3, 5, 6])
    return tp.DType.F64, a, b, d, 1.0, 0.0


def _small_case():
    """A smaller c64 product with conj(A), alpha, beta and a separate C."""
    a = T([37, 50], [1, 37], [0, 2], conj=True)
    b = T([50, 29], [1, 50], [2, 1])
    d = T([37, 29], [1, 37], [0, 1])
    return tp.DType.C64, a, b, d, complex(0.75, -0.5), complex(-0.5, 0.25)


def _device_operands(dtype, a, b, d, seed):
    import torch

    dev = torch.device("cuda:0")
    av, bv, cv = values(dtype, a, seed), values(dtype, b, seed + 1), values(dtype, d, seed + 2)
    return av, bv, cv, [torch.from_numpy(x).to(dev) for x in (av, bv, cv)]


@pytest.mark.gpu
@pytest.mark.parametrize("asynchronous", [False, True], ids=["sync", "async"])
def test_noalloc_matches_execute_product_on_two_calls(gpu, asynchronous):
    import torch

    ex = tp.Executor(0, None, asynchronous=True) if asynchronous else None
    jobs = []
    for k, case in enumerate((_big_case, _small_case)):  # the second call is smaller
        dtype, a, b, d, alpha, beta = case()
        plan = make_plan(dtype, a, b, d, d, tp.PlanConfig(split_k=3) if k == 0 else None)
        if k == 0:
            assert ".splitk" in plan.report()["kernel"]
        av, bv, cv, (ta, tb, tc) = _device_operands(dtype, a, b, d, 10 * k + 1)
        d_ref = torch.full_like(tc, float("nan"))
        d_new = torch.full_like(tc, float("nan"))
        torch.cuda.synchronize()
        assert _call(lib().TAPP_execute_product, plan, None, alpha, ta.data_ptr(), tb.data_ptr(), beta,
                     tc.data_ptr(), d_ref.data_ptr(), dtype) == 0, lib().tprims_last_error()
        torch.cuda.synchronize()
        jobs.append((dtype, a, b, d, alpha, beta, plan, av, bv, cv, ta, tb, tc, d_ref, d_new))
    # the no-allocation calls, back to back (queued without a wait when asynchronous)
    for (dtype, a, b, d, alpha, beta, plan, av, bv, cv, ta, tb, tc, d_ref, d_new) in jobs:
        assert _call(lib().tprims_execute_product_noalloc, plan, ex, alpha, ta.data_ptr(), tb.data_ptr(), beta,
                     tc.data_ptr(), d_new.data_ptr(), dtype) == 0, lib().tprims_last_error()
    if ex is not None:
        ex.synchronize()
    torch.cuda.synchronize()
    for (dtype, a, b, d, alpha, beta, plan, av, bv, cv, ta, tb, tc, d_ref, d_new) in jobs:
        got, ref = d_new.cpu().numpy(), d_ref.cpu().numpy()
        np.testing.assert_array_equal(got, ref)
        want = oracle.contract(alpha, a.oracle_op(av), b.oracle_op(bv), beta, d.oracle_op(cv),
                               d.oracle_op(np.zeros(d.storage(), NP[dtype])), cplx=dtype.is_complex)
        check(dtype, got, want)
    if ex is not None:
        ex.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_noalloc_conformance_case(gpu, name, dtype):
    """Every conformance case through the no-allocation entry point against
    the oracle (device operands)."""
    import torch

    kw = CASES[name](dtype)
    a, b, d = kw["a"], kw["b"], kw["d"]
    c, cv = kw.get("c"), kw.get("cv")
    in_place = kw.get("in_place", False)
    plan = make_plan(dtype, a, b, d, c)
    dev = torch.device("cuda:0")
    es = dtype.size
    ta, tb, td = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (kw["av"], kw["bv"], kw["dv"]))
    tc = torch.from_numpy(cv).to(dev) if cv is not None else None
    pd = td.data_ptr() + d.origin() * es
    pc = pd if in_place else (tc.data_ptr() + c.origin() * es if tc is not None else None)
    alpha, beta = kw.get("alpha", 1.0), kw.get("beta", 0.0)
    torch.cuda.synchronize()
    st = _call(lib().tprims_execute_product_noalloc, plan, None, alpha, ta.data_ptr() + a.origin() * es,
               tb.data_ptr() + b.origin() * es, beta, pc, pd, dtype)
    assert st == 0, lib().tprims_last_error()
    torch.cuda.synchronize()
    cop = (d if in_place else c).oracle_op(kw["dv"] if in_place else cv) if (in_place or cv is not None) else None
    want = oracle.contract(alpha, a.oracle_op(kw["av"]), b.oracle_op(kw["bv"]), beta, cop, d.oracle_op(kw["dv"]),
                           cplx=dtype.is_complex)
    check(dtype, td.cpu().numpy(), want, exact=name == "alpha_zero" or name.startswith("empty_k"))


@pytest.mark.gpu
def test_noalloc_refusals(gpu):
    """Host operands (TPRIMS_ERR_DEVICE), split-BF16 (TPRIMS_ERR_UNSUPPORTED)
    and a split-K partial area beyond the array (TPRIMS_ERR_UNSUPPORTED) are
    refused before any work; D is left untouched."""
    import torch

    dtype, a, b, d, alpha, beta = _small_case()
    plan = make_plan(dtype, a, b, d, d)
    av, bv, cv, (ta, tb, tc) = _device_operands(dtype, a, b, d, 5)
    host_d = np.full(d.storage(), 7.0, dtype=NP[dtype])
    assert _call(lib().tprims_execute_product_noalloc, plan, None, alpha, ta.data_ptr(), tb.data_ptr(), beta,
                 tc.data_ptr(), host_d.ctypes.data, dtype) == _capi.TPRIMS_ERR_DEVICE
    assert np.all(host_d == 7.0)
    f32 = make_plan(tp.DType.F32, a, b, d, d, tp.PlanConfig(precision=tp.Precision.SPLIT_BF16_3))
    xa, xb, xd = (torch.zeros(4096, dtype=torch.float32, device="cuda") for _ in range(3))
    assert _call(lib().tprims_execute_product_noalloc, f32, None, 1.0, xa.data_ptr(), xb.data_ptr(), 0.0,
                 None, xd.data_ptr(), tp.DType.F32) == _capi.TPRIMS_ERR_UNSUPPORTED
    n = 2048  # 4 K ranges of 2048^2 f64 partials: 128 MiB of partials
    big = make_plan(tp.DType.F64, T([n, n], [1, n], [0, 2]), T([n, n], [1, n], [2, 1]), T([n, n], [1, n], [0, 1]),
                    config=tp.PlanConfig(split_k=4))
    assert 4 * n * n * 8 > SCRATCH_BYTES
    y = torch.zeros(3 * n * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert _call(lib().tprims_execute_product_noalloc, big, None, 1.0, y.data_ptr(), y.data_ptr() + 8 * n * n, 0.0,
                 None, y.data_ptr() + 16 * n * n, tp.DType.F64) == _capi.TPRIMS_ERR_UNSUPPORTED
    assert "bytes of device scratch" in lib().tprims_last_error().decode()
    torch.cuda.synchronize()
    assert float(y.abs().max()) == 0.0
