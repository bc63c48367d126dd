"""GPU parity of TAPP_execute_batched_product (crates/tprims-capi/src/execute.rs:235):
one plan, many items, each item checked against the FP64 oracle. Items whose
operands are device memory run as one launch (their origins become extra batch
entries); items that write memory another item touches, or that live in host
memory, run one by one, in order, as TAPP runs them."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from cases import scalar
from harness import DTYPES, NP, T, check, make_plan, values


def _want(dtype, a, av, b, bv, d, dv, alpha, beta, cv=None):
    return oracle.contract(alpha, a.oracle_op(av), b.oracle_op(bv), beta,
                           d.oracle_op(cv) if cv is not None else None, d.oracle_op(dv), cplx=dtype.is_complex)


def _items(torch, dtype, a, b, d, n, seed, gap=3):
    """n items carved out of one device buffer per operand, `gap` elements
    apart (so their origins differ by arbitrary element counts). Returns, per
    item, the origin addresses and readers of the item's storage."""
    dev = torch.device("cuda:0")
    es = dtype.size
    bufs = {}
    for name, t, s in (("a", a, 1), ("b", b, 2), ("c", d, 3), ("d", d, 4)):
        per = t.storage() + gap
        flat = np.zeros(per * n, NP[dtype])
        if name != "d":
            for i in range(n):
                flat[i * per: i * per + t.storage()] = values(dtype, t, seed + 10 * i + s)
        bufs[name] = (torch.from_numpy(flat).to(dev), per, t)
    items = []
    for i in range(n):
        item = {}
        for name, (buf, per, t) in bufs.items():
            item[name] = buf.data_ptr() + (i * per + t.origin()) * es
            item[name + "_np"] = lambda buf=buf, i=i, per=per, t=t: buf.cpu().numpy()[i * per: i * per + t.storage()]
        items.append(item)
    return items


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("shape", ["gemm", "batched_strided", "split_k", "long_k", "elementwise", "negative_strides"])
def test_batched_items_match_the_oracle(gpu, dtype, shape):
    """Items of one plan in one fused launch against the oracle: a GEMM, a
    padded batched layout with conj, split K (asked for), the planner's own
    choice for a small output with a long K (the register-direct K-group
    tile for f64), the elementwise strategy and negative strides."""
    import torch

    if shape == "gemm":
        a, b, d = T([37, 29], [1, 37], [0, 2]), T([29, 41], [1, 29], [2, 1]), T([37, 41], [1, 37], [0, 1])
        n = 7
    elif shape == "batched_strided":
        a = T([20, 33, 3], [2, 40, 40 * 33], [0, 2, 3], conj=dtype.is_complex)  # padded, conj
        b = T([33, 18, 3], [18, 1, 33 * 18], [2, 1, 3])
        d = T([20, 18, 3], [1, 20, 360], [0, 1, 3])
        n = 5
    elif shape == "negative_strides":  # 64-bit gathers (register-staged kernel)
        a = T([20, 30], [-1, -20], [0, 2], conj=dtype.is_complex)
        b = T([30, 24], [24, 1], [2, 1])
        d = T([20, 24], [-24, 1], [0, 1])
        n = 6
    elif shape in ("split_k", "long_k"):
        a, b, d = T([40, 900], [1, 40], [0, 2]), T([900, 24], [1, 900], [2, 1]), T([40, 24], [1, 40], [0, 1])
        n = 3
    else:
        H = 1100
        a, b, d = T([2, 3, H], [1, 2, 6], [0, 2, 3]), T([3, 2, H], [1, 3, 6], [2, 1, 3]), T([2, 2, H], [1, 2, 4], [0, 1, 3])
        n = 4
    plan = make_plan(dtype, a, b, d, d, tp.PlanConfig(split_k=3) if shape == "split_k" else None)
    if shape == "split_k":
        assert ".splitk" in plan.report()["kernel"]
    if shape == "long_k" and dtype == tp.DType.F64:
        assert ".regs16.kgroup8" in plan.report()["kernel"], plan.report()
    if shape == "elementwise":
        assert plan.report()["strategy"] == 3
    alpha, beta = scalar(dtype, 0.75, -0.25), scalar(dtype, -0.5, 0.5)
    items = _items(torch, dtype, a, b, d, n, seed=50)
    plan.execute_batched_raw(alpha, [it["a"] for it in items], [it["b"] for it in items], beta,
                             [it["c"] for it in items], [it["d"] for it in items])
    torch.cuda.synchronize()
    for it in items:
        want = _want(dtype, a, it["a_np"](), b, it["b_np"](), d, np.zeros(d.storage(), NP[dtype]), alpha, beta,
                     it["c_np"]())
        check(dtype, it["d_np"](), want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.F64, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("gap", [3, 16], ids=["origins_misaligned", "origins_aligned"])
def test_batched_items_in_16_byte_runs(gpu, dtype, gap):
    """A layout that gathers in 16-byte runs: the fused launch keeps the run
    gathers only when every item's origin is 16-byte aligned (gap 16), and
    falls back to 4-byte gathers otherwise (gap 3); both match the oracle."""
    import torch

    a, b, d = T([64, 48], [1, 64], [0, 2]), T([48, 32], [1, 48], [2, 1]), T([64, 32], [1, 64], [0, 1])
    plan = make_plan(dtype, a, b, d, d)
    assert ".dma16" in plan.report()["kernel"], plan.report()
    alpha, beta = 1.25, -0.5
    items = _items(torch, dtype, a, b, d, 5, seed=90, gap=gap)
    plan.execute_batched_raw(alpha, [it["a"] for it in items], [it["b"] for it in items], beta,
                             [it["c"] for it in items], [it["d"] for it in items])
    torch.cuda.synchronize()
    for it in items:
        want = _want(dtype, a, it["a_np"](), b, it["b_np"](), d, np.zeros(d.storage(), NP[dtype]), alpha, beta,
                     it["c_np"]())
        check(dtype, it["d_np"](), want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F64, tp.DType.C32], ids=lambda d: d.name)
def test_batched_items_sharing_one_output_accumulate_in_order(gpu, dtype):
    """Every item accumulates into the same D (C = D, beta = 1): the items
    write what the others read, so they run one by one and the result is
    D0 + sum_i alpha A_i B_i."""
    import torch

    a, b, d = T([24, 16], [1, 24], [0, 2]), T([16, 20], [1, 16], [2, 1]), T([24, 20], [1, 24], [0, 1])
    n = 4
    dev = torch.device("cuda:0")
    avs = [values(dtype, a, 70 + i) for i in range(n)]
    bvs = [values(dtype, b, 80 + i) for i in range(n)]
    d0 = values(dtype, d, 90)
    ta = [torch.from_numpy(x).to(dev) for x in avs]
    tb = [torch.from_numpy(x).to(dev) for x in bvs]
    td = torch.from_numpy(d0.copy()).to(dev)
    plan = make_plan(dtype, a, b, d, d)
    alpha = scalar(dtype, 0.5, 0.25)
    plan.execute_batched_raw(alpha, [x.data_ptr() for x in ta], [x.data_ptr() for x in tb], 1.0,
                             [td.data_ptr()] * n, [td.data_ptr()] * n)
    torch.cuda.synchronize()
    want = d0.astype(np.complex128 if dtype.is_complex else np.float64)
    for i in range(n):
        want = _want(dtype, a, avs[i], b, bvs[i], d, np.zeros(d.storage(), NP[dtype]), alpha, 1.0, want)
    check(dtype, td.cpu().numpy(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C64], ids=lambda d: d.name)
def test_batched_host_items_and_output_pass(gpu, dtype):
    """Host (pageable) items are staged one by one; alpha = 0 makes every item
    the output pass D_i = beta C_i."""
    a, b, d = T([9, 7], [1, 9], [0, 2]), T([7, 5], [1, 7], [2, 1]), T([9, 5], [1, 9], [0, 1])
    n = 3
    plan = make_plan(dtype, a, b, d, d)
    for alpha in (scalar(dtype, 1.5, -1.0), 0.0):
        avs = [values(dtype, a, 100 + i) for i in range(n)]
        bvs = [values(dtype, b, 110 + i) for i in range(n)]
        cvs = [values(dtype, d, 120 + i) for i in range(n)]
        dvs = [np.zeros(d.storage(), NP[dtype]) for _ in range(n)]
        plan.execute_batched_raw(alpha, [x.ctypes.data for x in avs], [x.ctypes.data for x in bvs], 0.5,
                                 [x.ctypes.data for x in cvs], [x.ctypes.data for x in dvs])
        for i in range(n):
            check(dtype, dvs[i], _want(dtype, a, avs[i], b, bvs[i], d, np.zeros(d.storage(), NP[dtype]), alpha, 0.5,
                                       cvs[i]))


@pytest.mark.gpu
def test_batched_null_c_needs_beta_zero(gpu):
    import torch

    a, b, d = T([4, 3], [1, 4], [0, 2]), T([3, 5], [1, 3], [2, 1]), T([4, 5], [1, 4], [0, 1])
    plan = make_plan(tp.DType.F64, a, b, d, d)
    dev = torch.device("cuda:0")
    A = torch.randn(12, dtype=torch.float64, device=dev)
    B = torch.randn(15, dtype=torch.float64, device=dev)
    D = torch.zeros(2, 20, dtype=torch.float64, device=dev)
    with pytest.raises(tp.TprimsError):
        plan.execute_batched_raw(1.0, [A.data_ptr()] * 2, [B.data_ptr()] * 2, 1.0, None,
                                 [D[0].data_ptr(), D[1].data_ptr()])
    plan.execute_batched_raw(1.0, [A.data_ptr()] * 2, [B.data_ptr()] * 2, 0.0, None, [D[0].data_ptr(), D[1].data_ptr()])
    torch.cuda.synchronize()
    want = (A.cpu().numpy().reshape(3, 4).T @ B.cpu().numpy().reshape(5, 3).T).T.reshape(-1)
    np.testing.assert_allclose(D.cpu().numpy()[0], want, rtol=1e-12)
    np.testing.assert_allclose(D.cpu().numpy()[1], want, rtol=1e-12)
