"""GPU tests of grouped execution (one persistent kernel over many uneven
contractions, optionally captured in a hipGraph) against the FP64 oracle and
against one launch per contraction."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from harness import NP, T, make_plan, tol


def mps_items(dtype, n, seed, dims=(16, 32, 48), phys=(2, 4)):
    """E[a,s,b] = sum_{c,t} L[a,c] * M[c,t,s,b] with uneven bond dims."""
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(n):
        a, b, c = (int(rng.choice(dims)) for _ in range(3))
        s, t = (int(rng.choice(phys)) for _ in range(2))
        L = T([a, c], tp.col_major([a, c]), [1, 3])
        Mt = T([c, t, s, b], tp.col_major([c, t, s, b]), [3, 4, 2, 5])
        E = T([a, s, b], tp.col_major([a, s, b]), [1, 2, 5])
        def draw(k):
            x = rng.standard_normal(k) + (1j * rng.standard_normal(k) if dtype.is_complex else 0)
            return x.astype(NP[dtype])
        items.append((L, draw(a * c), Mt, draw(c * t * s * b), E, draw(a * s * b)))
    return items


def build(dtype, items, gpu, beta=0.0):
    import torch

    plans, A, B, C, D, want = [], [], [], [], [], []
    alphas, betas = [], []
    for k, (L, lv, Mt, mv, E, cv) in enumerate(items):
        plans.append(make_plan(dtype, L, Mt, E))
        A.append(torch.from_numpy(lv).to(gpu))
        B.append(torch.from_numpy(mv).to(gpu))
        C.append(torch.from_numpy(cv).to(gpu))
        D.append(torch.zeros(len(cv), dtype=A[-1].dtype, device=gpu))
        al = complex(1.0 + 0.1 * k, -0.2) if dtype.is_complex else 1.0 + 0.1 * k
        alphas.append(al)
        betas.append(beta)
        want.append(oracle.contract(al, L.oracle_op(lv), Mt.oracle_op(mv), beta, E.oracle_op(cv),
                                    E.oracle_op(np.zeros(len(cv), NP[dtype])), cplx=dtype.is_complex))
    return plans, alphas, betas, A, B, C, D, want


def check_all(dtype, D, want, label):
    for k, (d, w) in enumerate(zip(D, want)):
        err = oracle.max_rel_err(d.cpu().numpy().astype(np.complex128), w)
        assert err <= tol(dtype), "%s item %d: %.3e" % (label, k, err)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.C64, tp.DType.F64, tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
def test_group_matches_oracle_and_per_item_launches(gpu, dtype):
    import torch

    items = mps_items(dtype, 24, 11)
    plans, al, be, A, B, C, D, want = build(dtype, items, gpu, beta=0.5)
    g = tp.Group(plans, al, A, B, be, C, D)
    g.execute()
    torch.cuda.synchronize()
    check_all(dtype, D, want, "grouped")
    grouped = [d.clone() for d in D]
    for d in D:
        d.zero_()
    g.execute_each()
    torch.cuda.synchronize()
    check_all(dtype, D, want, "per-item")
    info = g.info()
    assert info["tiles"] > 0 and info["flops"] == sum(p.report()["flops"] for p in plans)
    # K is never split: the same tile shape gives the same bits.
    for k, p in enumerate(plans):
        if p.report()["tile"] == info["tile"]:
            assert torch.equal(grouped[k], D[k])


@pytest.mark.gpu
def test_group_graph_replays_match_eager(gpu):
    import torch

    dtype = tp.DType.C64
    items = mps_items(dtype, 16, 5)
    plans, al, be, A, B, C, D, want = build(dtype, items, gpu)
    g = tp.Group(plans, al, A, B, be, C, D)
    g.execute()
    torch.cuda.synchronize()
    eager = [d.clone() for d in D]
    g.use_graph(True)
    for rep in range(3):
        for d in D:
            d.fill_(float("nan"))
        torch.cuda.synchronize()
        g.execute()
        torch.cuda.synchronize()
        for k in range(len(D)):
            assert torch.equal(D[k], eager[k]), "replay %d item %d" % (rep, k)
        # eager work between replays
        _ = (A[0] * 2).sum().item()
    check_all(dtype, D, want, "graph")


@pytest.mark.gpu
def test_group_autotune_keeps_results(gpu):
    import torch

    dtype = tp.DType.C64
    plans, al, be, A, B, C, D, want = build(dtype, mps_items(dtype, 12, 9), gpu)
    g = tp.Group(plans, al, A, B, be, C, D)
    tile, ms = g.autotune(reps=2)
    assert tile in (1, 2) and ms > 0  # c64 groups are not offered the 128x128 tile
    for d in D:
        d.zero_()
    g.execute()
    torch.cuda.synchronize()
    check_all(dtype, D, want, "autotuned")
    for t in (1, 2):
        g.set_tile(t)
        g.execute()
        torch.cuda.synchronize()
        check_all(dtype, D, want, "tile %d" % t)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.C64, tp.DType.F64, tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("method", [tp.ComplexMethod.NATIVE, tp.ComplexMethod.ONE_M, tp.ComplexMethod.THREE_M],
                         ids=lambda m: m.name)
def test_group_on_every_dma_tile(gpu, dtype, method):
    """The grouped kernel on the LDS-DMA tiles (64x64, 128x64 where offered,
    128x128 where offered: not for c64 or the c32 1m schedule, which set_tile
    refuses) for every complex schedule, against the oracle."""
    import torch

    if not dtype.is_complex and method != tp.ComplexMethod.NATIVE:
        pytest.skip("complex schedules apply to complex dtypes")
    items = mps_items(dtype, 10, 21, dims=(16, 48, 80))
    plans, al, be, A, B, C, D, want = build(dtype, items, gpu, beta=0.25)
    if method != tp.ComplexMethod.NATIVE:
        plans = [tp.Plan(p.problem, tp.PlanConfig(complex_method=method)) for p in plans]
    g = tp.Group(plans, al, A, B, be, C, D)
    for t in (1, 3, 4):
        if t == 3 and dtype == tp.DType.C64:
            continue
        if t == 4 and (dtype == tp.DType.C64 or (dtype == tp.DType.C32 and method == tp.ComplexMethod.ONE_M)):
            with pytest.raises(tp.TprimsError):
                g.set_tile(t)
            continue
        g.set_tile(t)
        for d in D:
            d.fill_(float("nan"))
        g.execute()
        torch.cuda.synchronize()
        for k, (d, w) in enumerate(zip(D, want)):
            err = oracle.max_rel_err(d.cpu().numpy().astype(np.complex128), w)
            assert err <= tol(dtype, method), "tile %d item %d: %.3e" % (t, k, err)


def test_group_rejects_mixed_dtypes_without_a_gpu_call():
    a = T([4, 3], [1, 4], [0, 2])
    b = T([3, 5], [1, 3], [2, 1])
    d = T([4, 5], [1, 4], [0, 1])
    p1 = make_plan(tp.DType.F64, a, b, d)
    p2 = make_plan(tp.DType.F32, a, b, d)
    import ctypes
    bufs = [np.zeros(32) for _ in range(4)]
    with pytest.raises(tp.TprimsError) as e:
        tp.Group([p1, p2], [1.0, 1.0], [bufs[0], bufs[0]], [bufs[1], bufs[1]], [0.0, 0.0], None,
                 [bufs[2], bufs[3]])
    assert e.value.status in (3, 4)


def _gemm_plan(dtype=tp.DType.F64):
    return make_plan(dtype, T([4, 3], [1, 4], [0, 2]), T([3, 5], [1, 3], [2, 1]), T([4, 5], [1, 4], [0, 1]))


def test_group_rejects_overlapping_items_without_a_gpu_call():
    """Items run at the same time: an item whose output meets another item's
    operand or output is refused (TPRIMS_ERR_ALIASED) before the GPU runtime
    is touched; disjoint items and an item reading its own output pass that
    check (and then need a device, absent here)."""
    p = _gemm_plan()
    buf = np.zeros(256)
    base = buf.ctypes.data
    A0, B0, D0 = base, base + 8 * 16, base + 8 * 40
    A1, B1, D1 = base + 8 * 80, base + 8 * 100, base + 8 * 130
    cases = {
        "D1 on A0": ([A0, A1], [B0, B1], [D0, A0 + 8], None, 0.0),
        "D1 on D0": ([A0, A1], [B0, B1], [D0, D0 + 8 * 5], None, 0.0),
        "D0 on C1": ([A0, A1], [B0, B1], [D0, D1], [D1, D0], 1.0),
    }
    for label, (A, B, D, C, beta) in cases.items():
        with pytest.raises(tp.TprimsError) as e:
            tp.Group([p, p], [1.0, 1.0], A, B, [beta, beta], C, D)
        assert e.value.status == 6, label  # TPRIMS_ERR_ALIASED
    # disjoint items, the second accumulating into itself: past the check
    for C, betas in ((None, [0.0, 0.0]), ([D0, D1], [0.0, 1.0])):
        try:
            tp.Group([p, p], [1.0, 1.0], [A0, A1], [B0, B1], betas, C, [D0, D1])
        except tp.TprimsError as e:
            assert e.status == 4  # no HIP device in this container
        else:  # a GPU machine: host buffers are not grouped operands
            pytest.fail("host operands were accepted")


@pytest.mark.gpu
def test_group_autotune_with_in_place_items_runs_them_once(gpu):
    """Items with C = D and beta = 1 accumulate into their output: after
    autotune (many timing launches) every output holds one execution's
    result, D0 + alpha * L M, checked against the oracle."""
    import torch

    dtype = tp.DType.C64
    items = mps_items(dtype, 12, 13)
    plans, al, A, B, D, want = [], [], [], [], [], []
    for k, (L, lv, Mt, mv, E, cv) in enumerate(items):
        plans.append(make_plan(dtype, L, Mt, E))
        A.append(torch.from_numpy(lv).to(gpu))
        B.append(torch.from_numpy(mv).to(gpu))
        D.append(torch.from_numpy(cv.copy()).to(gpu))
        a = complex(0.5 + 0.1 * k, 0.3)
        al.append(a)
        want.append(oracle.contract(a, L.oracle_op(lv), Mt.oracle_op(mv), 1.0, E.oracle_op(cv),
                                    E.oracle_op(np.zeros(len(cv), NP[dtype])), cplx=True))
    g = tp.Group(plans, al, A, B, [1.0] * len(plans), D, D)
    tile, ms = g.autotune(reps=3)
    torch.cuda.synchronize()
    check_all(dtype, D, want, "after autotune (tile %d)" % tile)


@pytest.mark.gpu
def test_group_keeps_its_plans_alive(gpu):
    """Plans destroyed after the group was built stay valid for it (the
    group holds a reference to each)."""
    import torch

    dtype = tp.DType.F64
    plans, al, be, A, B, C, D, want = build(dtype, mps_items(dtype, 8, 17), gpu, beta=0.5)
    g = tp.Group(plans, al, A, B, be, C, D)
    for p in plans:
        p.close()
    plans.clear()
    import gc
    gc.collect()
    g.execute()
    torch.cuda.synchronize()
    check_all(dtype, D, want, "plans released")
    g.close()
