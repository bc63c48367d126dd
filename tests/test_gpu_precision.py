"""GPU tests of the split-BF16 precision paths (f32 and c32 storage): every
conformance case at the reference's f32 budget, and the accuracy of each
scheme on the precision benchmark's input distribution (N(0,1) scaled by
10^U(-3,3) per row) against the FP64 oracle."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from cases import CASES
from harness import NP, T, check, run

SPLITS = [tp.Precision.SPLIT_BF16_2, tp.Precision.SPLIT_BF16_3]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", SPLITS, ids=lambda p: p.name)
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_split_conformance_case(gpu, name, dtype, precision):
    kw = CASES[name](dtype)
    got, want, plan = run(dtype, config=tp.PlanConfig(precision=precision), **kw)
    if plan.report()["strategy"] == 0:
        assert "split_bf16" in plan.report()["kernel"]
    check(dtype, got, want, precision=precision)


def _transform_inputs(dtype, p, t, s, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((p * p * p, t))
    U = rng.standard_normal((t, s))
    if dtype.is_complex:
        A = A + 1j * rng.standard_normal(A.shape)
        U = U + 1j * rng.standard_normal(U.shape)
    A *= 10.0 ** rng.uniform(-3, 3, size=(p * p * p, 1))
    U *= 10.0 ** rng.uniform(-3, 3, size=(t, 1))
    return A.astype(NP[dtype]), U.astype(NP[dtype])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
def test_transform_step_accuracy_vs_fp64(gpu, dtype):
    """D[p,q,r,s] = sum_t A[p,q,r,t] U[t,s] with rows spanning 10^+-3."""
    p, t, s = 12, 96, 80
    A, U = _transform_inputs(dtype, p, t, s, 5)
    # A stored (p,q,r,t) column-major: element (x, t) at x + t * p^3.
    a = T([p, p, p, t], tp.col_major([p, p, p, t]), [0, 1, 2, 3])
    u = T([t, s], [1, t], [3, 4])
    d = T([p, p, p, s], tp.col_major([p, p, p, s]), [0, 1, 2, 4])
    av = A.reshape(-1, order="F")
    uv = U.reshape(-1, order="F")
    dv = np.zeros(p ** 3 * s, NP[dtype])
    errs, pct = {}, {}
    for prec in [tp.Precision.STORAGE] + SPLITS:
        got, want, plan = run(dtype, a, av, u, uv, d, dv, config=tp.PlanConfig(precision=prec))
        errs[prec] = oracle.max_rel_err(got.astype(np.complex128), want)
        pct[prec] = elementwise_rel_err(got, want)
        if dtype == tp.DType.F32 and prec != tp.Precision.STORAGE:
            assert plan.report()["kernel"].endswith(".a_in_gemm")
    print({k.name: "normwise %.2e p50 %.2e p99 %.2e" % ((v,) + pct[k]) for k, v in errs.items()})
    # elementwise: the 3-term scheme is as accurate as FP32 storage on every row scale
    assert pct[tp.Precision.SPLIT_BF16_3][1] <= 4 * max(pct[tp.Precision.STORAGE][1], 1e-7)
    assert pct[tp.Precision.SPLIT_BF16_3][0] <= 1e-6
    assert errs[tp.Precision.SPLIT_BF16_3] <= 1e-6
    assert errs[tp.Precision.STORAGE] <= 1e-6
    assert errs[tp.Precision.SPLIT_BF16_2] <= 1e-4
    # the 2-term scheme is measurably coarser than the 3-term one
    assert errs[tp.Precision.SPLIT_BF16_2] > errs[tp.Precision.SPLIT_BF16_3]


def elementwise_rel_err(got, want):
    """p50 and p99 of |got - want| / |want| over the nonzero outputs."""
    got, want = np.asarray(got).astype(np.complex128), np.asarray(want).astype(np.complex128)
    nz = np.abs(want) > 0
    r = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    return float(np.percentile(r, 50)), float(np.percentile(r, 99))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", SPLITS, ids=lambda p: p.name)
@pytest.mark.parametrize("tile", [11, 12, 13])
@pytest.mark.parametrize("shape", [(132, 100, 70, 1), (260, 64, 130, 2), (1000, 333, 257, 1), (64, 31, 16, 3),
                                   (4, 40, 9, 1)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("b_major", ["k", "n"])
def test_split_a_inside_the_gemm(gpu, precision, tile, shape, b_major):
    """An f32 A gathered along M in 16-byte runs is split inside the GEMM
    (A's f32 tile DMA'd to LDS, each lane splitting the values it feeds the
    MFMA; no A panels): every split tile, ragged M, N and K, batch entries,
    both B majors, alpha and beta, against the oracle."""
    M, K, N, H = shape
    lda = M + 4  # a padded leading dimension that keeps 16-byte runs
    a = T([M, K, H], [1, lda, lda * K], [0, 2, 3])
    b = T([K, N, H], [1, K, K * N], [2, 1, 3]) if b_major == "k" else T([N, K, H], [1, N, N * K], [1, 2, 3])
    d = T([M, N, H], [1, M, M * N], [0, 1, 3])
    from harness import values
    av, bv, cv = values(tp.DType.F32, a, 71), values(tp.DType.F32, b, 72), values(tp.DType.F32, d, 73)
    got, want, plan = run(tp.DType.F32, a, av, b, bv, d, np.zeros(d.storage(), np.float32), alpha=1.25, beta=-0.5,
                          c=d, cv=cv, config=tp.PlanConfig(precision=precision, tile=tile))
    rep = plan.report()
    assert rep["a_k_major"] == 0 and rep["kernel"].endswith(".a_in_gemm"), rep["kernel"]
    check(tp.DType.F32, got, want, precision=precision)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [11, 12])
def test_split_tiles(gpu, tile):
    kw = CASES["large_k"](tp.DType.F32)
    got, want, plan = run(tp.DType.F32, config=tp.PlanConfig(precision=tp.Precision.SPLIT_BF16_3, tile=tile), **kw)
    assert plan.report()["tile"] == tile
    check(tp.DType.F32, got, want, precision=tp.Precision.SPLIT_BF16_3)


def test_split_precision_is_refused_for_double():
    d = T([4, 4], [1, 4], [0, 1])
    from harness import make_plan
    with pytest.raises(tp.TprimsError) as e:
        make_plan(tp.DType.F64, d, T([4, 4], [1, 4], [1, 2]), T([4, 4], [1, 4], [0, 2]),
                  config=tp.PlanConfig(precision=tp.Precision.SPLIT_BF16_3))
    assert e.value.status == 15


def k_runs_layout(L, I, Kx, J):
    """D[i,k] = sum_{j,l} A[l,i,j] B[l,k,j], column-major: both operands are
    contiguous along l only, so their K offsets run in groups of L
    consecutive elements with jumps between groups (16-byte K vectors for
    the packers when L is a multiple of 8; L = 4 keeps the gather path)."""
    a = T([L, I, J], tp.col_major([L, I, J]), [0, 1, 2])
    b = T([L, Kx, J], tp.col_major([L, Kx, J]), [0, 3, 2])
    d = T([I, Kx], tp.col_major([I, Kx]), [1, 3])
    return a, b, d


@pytest.mark.gpu
@pytest.mark.parametrize("precision", SPLITS, ids=lambda p: p.name)
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("L", [4, 8, 16, 24])
def test_split_packer_k_runs(gpu, L, dtype, precision):
    """Split-BF16 packing of operands whose K offsets are contiguous only
    within runs of L (aligned groups of 8 k load as 16-byte vectors through
    the K table), against the oracle."""
    I, Kx, J = 72, 40, 11
    a, b, d = k_runs_layout(L, I, Kx, J)
    rng = np.random.default_rng(L)
    def vals(n):
        v = rng.standard_normal(n)
        if dtype.is_complex:
            v = v + 1j * rng.standard_normal(n)
        return v.astype(NP[dtype])
    av, bv = vals(a.storage()), vals(b.storage())
    dv = np.zeros(d.storage(), NP[dtype])
    got, want, plan = run(dtype, a, av, b, bv, d, dv, config=tp.PlanConfig(precision=precision))
    assert "split_bf16" in plan.report()["kernel"]
    check(dtype, got, want, precision=precision)
