"""GPU parity: every conformance case, every dtype and complex method, through
the C ABI (TAPP_execute_product) on the HIP kernels, against the FP64 label
oracle over the whole output buffer."""
import numpy as np
import pytest

import tprims_amd as tp
from cases import CASES, scalar
from harness import DTYPES, NP, T, check, run, values

METHODS = [tp.ComplexMethod.NATIVE, tp.ComplexMethod.ONE_M, tp.ComplexMethod.THREE_M]


def _variants():
    out = []
    for dt in DTYPES:
        ms = METHODS if dt.is_complex else [tp.ComplexMethod.NATIVE]
        for m in ms:
            out.append((dt, m))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,method", _variants(), ids=lambda v: str(getattr(v, "name", v)))
@pytest.mark.parametrize("name", sorted(CASES))
def test_conformance_case(gpu, name, dtype, method):
    kw = CASES[name](dtype)
    got, want, plan = run(dtype, config=tp.PlanConfig(complex_method=method), **kw)
    exact = (name == "alpha_zero" or name.startswith("empty_k")) and method != tp.ComplexMethod.THREE_M
    check(dtype, got, want, method, exact=exact)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_single_complex_product_is_hand_checkable(gpu, dtype):
    # tapp_conformance.rs:88-111: (1+2i)(3+4i) = -5+10i; 3 for real types.
    a = T([1, 1], [1, 1], [0, 2])
    b = T([1, 1], [1, 1], [2, 1])
    d = T([1, 1], [1, 1], [0, 1])
    av = np.array([scalar(dtype, 1, 2)], dtype=NP[dtype])
    bv = np.array([scalar(dtype, 3, 4)], dtype=NP[dtype])
    got, _, _ = run(dtype, a, av, b, bv, d, np.zeros(1, NP[dtype]))
    want = complex(-5, 10) if dtype.is_complex else 3.0
    assert got[0] == want


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_small_real_product_is_hand_checkable(gpu, dtype):
    # tapp_conformance.rs:114-140: A B = [[19,22],[43,50]] column-major.
    a = T([2, 2], [1, 2], [0, 2])
    b = T([2, 2], [1, 2], [2, 1])
    d = T([2, 2], [1, 2], [0, 1])
    av = np.array([1, 3, 2, 4], dtype=NP[dtype])
    bv = np.array([5, 7, 6, 8], dtype=NP[dtype])
    got, _, _ = run(dtype, a, av, b, bv, d, np.zeros(4, NP[dtype]))
    np.testing.assert_array_equal(got, np.array([19, 43, 22, 50], dtype=NP[dtype]))


@pytest.mark.gpu
def test_readme_c_example_values(gpu):
    # crates/tprims-capi/tests/c/standard_consumer.c: D = [22, 28, 49, 64]
    a = T([2, 3], [1, 2], [ord("i"), ord("k")])
    b = T([3, 2], [1, 3], [ord("k"), ord("j")])
    d = T([2, 2], [1, 2], [ord("i"), ord("j")])
    av = np.arange(1, 7, dtype=np.float64)
    got, _, _ = run(tp.DType.F64, a, av, b, av.copy(), d, np.full(4, -1.0))
    np.testing.assert_array_equal(got, [22, 28, 49, 64])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("name", ["general_strides", "padded_view", "output_diagonal", "negative_strides",
                                  "accumulate_in_place", "alpha_zero"])
def test_host_operands_are_staged(gpu, name, dtype):
    """Host (pageable) buffers go through the executor's device scratch."""
    kw = CASES[name](dtype)
    got, want, _ = run(dtype, where="host", **kw)
    check(dtype, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [1, 2, 3, 4])
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_every_tile(gpu, tile, dtype):
    if tile == 3 and dtype == tp.DType.C64:
        pytest.skip("tile 3 is not offered for c64")
    kw = CASES["large_k"](dtype)
    got, want, plan = run(dtype, config=tp.PlanConfig(tile=tile), **kw)
    assert plan.report()["tile"] == tile
    check(dtype, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("layout", ["affine", "two_batch_axes", "hadamard", "hadamard_dense"])
def test_elementwise_strategy(gpu, dtype, layout):
    """The all-batch strategy (tiny M x N and K, many batch entries) against the
    oracle: conj(A), beta * op_C(C), one folded batch axis (affine offsets) or
    two unfoldable ones (scatter vectors), and a pure Hadamard product."""
    from harness import values
    if layout == "hadamard":
        n = 5000
        a, b, d = T([n], [1], [0], conj=dtype.is_complex), T([n], [2], [0]), T([n], [1], [0])
    elif layout == "hadamard_dense":  # unit strides: 16-byte vectors and a ragged end
        n = 5003
        a, b, d = T([n], [1], [0], conj=dtype.is_complex), T([n], [1], [0]), T([n], [1], [0])
    elif layout == "affine":
        H = 3000
        a = T([2, 3, H], [1, 2, 6], [0, 2, 3], conj=dtype.is_complex)
        b = T([3, 2, H], [1, 3, 6], [2, 1, 3])
        d = T([2, 2, H], [1, 2, 4], [0, 1, 3])
    else:
        H1, H2 = 60, 50
        a = T([2, 3, H1, H2], [1, 2, 6 * H2, 6], [0, 2, 3, 4], conj=dtype.is_complex)
        b = T([3, 2, H1, H2], [1, 3, 6, 6 * H1], [2, 1, 3, 4])
        d = T([2, 2, H1, H2], [1, 2, 4, 4 * H1], [0, 1, 3, 4])
    av, bv = values(dtype, a, 31), values(dtype, b, 32)
    dv = np.zeros(d.storage(), NP[dtype])
    cv = values(dtype, d, 33)
    got, want, plan = run(dtype, a, av, b, bv, d, dv, alpha=1.25, beta=-0.5, c=d, cv=cv)
    assert plan.report()["strategy"] == 3
    check(dtype, got, want)


def _dma_variants():
    real = [(tp.DType.F32, tp.ComplexMethod.NATIVE), (tp.DType.F64, tp.ComplexMethod.NATIVE)]
    return real + [(dt, m) for dt in (tp.DType.C32, tp.DType.C64) for m in METHODS]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,method", _dma_variants(), ids=lambda v: str(getattr(v, "name", v)))
@pytest.mark.parametrize("name", sorted(CASES))
def test_128x128_dma_tile_every_case(gpu, name, dtype, method):
    """The 128x128 tile on the LDS-DMA kernel (the register-staged 64x64
    kernel when an operand needs 64-bit offsets) on every conformance case,
    with each complex schedule: native, 1m (the real GEMM of the expanded A)
    and 3m (three real products)."""
    kw = CASES[name](dtype)
    got, want, plan = run(dtype, config=tp.PlanConfig(tile=4, complex_method=method), **kw)
    rep = plan.report()
    if rep["strategy"] == 0:
        assert rep["tile"] == 4
    exact = (name == "alpha_zero" or name.startswith("empty_k")) and method != tp.ComplexMethod.THREE_M
    check(dtype, got, want, method, exact=exact)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,method", _dma_variants()[2:], ids=lambda v: str(getattr(v, "name", v)))
@pytest.mark.parametrize("tile", [1, 3, 4])
def test_dma_complex_schedules_on_large_k(gpu, dtype, method, tile):
    """Every complex schedule on the DMA tiles (64x64, 128x64, 128x128) over a
    long K with conj(A), alpha, beta and a padded (non-unit stride) operand."""
    if tile == 3 and dtype == tp.DType.C64:
        pytest.skip("tile 3 is not offered for c64")
    M, K, N = 150, 333, 140
    a = T([K, M], [2, 2 * K], [2, 0], conj=True)  # every 2nd element, K-major
    b = T([N, K], [1, N], [1, 2])                  # N-major
    d = T([M, N], [1, M], [0, 1])
    av, bv, cv = values(dtype, a, 41), values(dtype, b, 42), values(dtype, d, 43)
    got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), alpha=complex(0.75, -0.5),
                          beta=complex(-0.5, 0.25), c=d, cv=cv,
                          config=tp.PlanConfig(tile=tile, complex_method=method, split_k=1))
    assert plan.report()["kernel"].endswith(".dma"), plan.report()["kernel"]
    check(dtype, got, want, method)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("shape", [(40, 3000, 24, 1), (100, 700, 90, 2), (576, 576, 576, 1)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("d_major", ["m", "n"])
def test_split_k(gpu, dtype, shape, d_major):
    """Few output tiles and a long K: the planner splits K (fixed-order
    reduction of the partial tiles, then the write-back with alpha, beta and
    conj) and the result matches the oracle, with D (and C) contiguous along
    M or along N."""
    from harness import values
    M, K, N, H = shape
    a = T([M, K, H], [1, M, M * K], [0, 2, 3], conj=dtype.is_complex)
    b = T([K, N, H], [1, K, K * N], [2, 1, 3])
    d = T([M, N, H], [1, M, M * N] if d_major == "m" else [N, 1, M * N], [0, 1, 3])
    rng = np.random.default_rng(3)
    def vals(t):
        v = rng.uniform(-1, 1, t.storage())
        if dtype.is_complex:
            v = v + 1j * rng.uniform(-1, 1, t.storage())
        return v.astype(NP[dtype])
    # K up to 1024: the planner takes a K-group 32x32 tile where its gathers
    # allow (f64 at any tile count; c32 with 324 tiles), so three K ranges
    # are asked for.
    cfg = tp.PlanConfig(split_k=3) if K <= 1024 else None
    got, want, plan = run(dtype, a, vals(a), b, vals(b), d, np.zeros(d.storage(), NP[dtype]),
                          alpha=1.5, beta=-0.5, c=d, cv=vals(d), config=cfg)
    assert ".splitk" in plan.report()["kernel"]
    check(dtype, got, want)


_RUN_VARIANTS = [(tp.DType.F32, tp.ComplexMethod.NATIVE), (tp.DType.F64, tp.ComplexMethod.NATIVE),
                 (tp.DType.C32, tp.ComplexMethod.NATIVE), (tp.DType.C32, tp.ComplexMethod.ONE_M),
                 (tp.DType.C32, tp.ComplexMethod.THREE_M)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,method", _RUN_VARIANTS, ids=lambda v: v.name)
@pytest.mark.parametrize("tile", [1, 3, 4])
@pytest.mark.parametrize("majors", ["kk", "kn", "mk", "mn"])
@pytest.mark.parametrize("split", [1, 3])
def test_dma_16_byte_runs(gpu, dtype, method, tile, majors, split):
    """Operands whose gathers are 16-byte runs (4 f32, 2 f64 or 2 c32
    consecutive along the major index, every offset a multiple of 16 bytes)
    load whole runs per lane: every tile, both majors of A and B, a batch
    index, edges that do not fill a tile, K split or not, and for c32 every
    complex schedule with conj(A)."""
    M, K, N, H = 148, 340, 136, 2
    cj = dtype.is_complex
    a = T([K, M, H], [1, K, K * M], [2, 0, 3], conj=cj) if majors[0] == "k" else \
        T([M, K, H], [1, M, M * K], [0, 2, 3], conj=cj)
    b = T([K, N, H], [1, K, K * N], [2, 1, 3]) if majors[1] == "k" else T([N, K, H], [1, N, N * K], [1, 2, 3])
    d = T([M, N, H], [1, M, M * N], [0, 1, 3])
    av, bv, cv = values(dtype, a, 61), values(dtype, b, 62), values(dtype, d, 63)
    alpha = complex(0.75, -0.5) if cj else 0.75
    beta = complex(-0.5, 0.25) if cj else -0.5
    got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), alpha=alpha, beta=beta, c=d,
                          cv=cv, config=tp.PlanConfig(tile=tile, split_k=split, complex_method=method))
    rep = plan.report()
    assert ".dma16" in rep["kernel"] and rep["a_k_major"] == (majors[0] == "k") and \
        rep["b_k_major"] == (majors[1] == "k"), rep
    check(dtype, got, want, method)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.F64], ids=lambda d: d.name)
@pytest.mark.parametrize("case", ["k_not_multiple", "odd_row_offsets", "short_runs"])
def test_dma_without_16_byte_runs(gpu, dtype, case):
    """Layouts that do not qualify for 16-byte runs keep the 4-byte gathers
    and still match the oracle."""
    M, K, N = 132, 203, 68
    if case == "k_not_multiple":  # K-major A with K odd
        a, b = T([K, M], [1, K], [2, 0]), T([K, N], [1, K], [2, 1])
    elif case == "odd_row_offsets":  # K-major A whose rows start at odd elements (leading dimension K + 1)
        K = 204
        a, b = T([K, M], [1, K + 1], [2, 0]), T([K, N], [1, K], [2, 1])
    else:  # M-major A in runs of 2 (stride-2 pairs): runs too short for f32
        a = T([2, M // 2, K], [1, 4, 2 * M], [0, 4, 2])
        b = T([K, N], [1, K], [2, 1])
    d = T([M, N], [1, M], [0, 1]) if case != "short_runs" else T([2, M // 2, N], [1, 2, M], [0, 4, 1])
    av, bv = values(dtype, a, 71), values(dtype, b, 72)
    got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), alpha=1.25,
                          config=tp.PlanConfig(split_k=1))
    kern = plan.report()["kernel"]
    assert ".dma16" not in kern or (case == "short_runs" and dtype == tp.DType.F64), kern
    check(dtype, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_split_k_repeats_bitwise(gpu, dtype):
    """Split K reduces its partial tiles in a fixed order: repeated
    executions give bitwise identical, oracle-checked results."""
    import torch

    M, K, N = 96, 2000, 80
    a = T([M, K], [1, M], [0, 2], conj=dtype.is_complex)
    b = T([K, N], [1, K], [2, 1])
    d = T([M, N], [1, M], [0, 1])
    av, bv, cv = values(dtype, a, 61), values(dtype, b, 62), values(dtype, d, 63)
    got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), alpha=1.25, beta=0.5, c=d, cv=cv)
    assert ".splitk" in plan.report()["kernel"]
    check(dtype, got, want)
    dev = torch.device("cuda:0")
    ta, tb, tc = (torch.from_numpy(x).to(dev) for x in (av, bv, cv))
    outs = []
    for rep in range(4):
        td = torch.zeros(d.storage(), dtype=ta.dtype, device=dev)
        plan.execute_raw(1.25, ta.data_ptr(), tb.data_ptr(), 0.5, tc.data_ptr(), td.data_ptr())
        torch.cuda.synchronize()
        outs.append(td.cpu().numpy())
    for o in outs:
        np.testing.assert_array_equal(o, got)



def _random_run_case(rng, dtype):
    """A random contraction whose A and B gather in 16-byte runs: extents
    multiples of the run (down to one run), column-major operands with
    leading dimensions that keep every row start aligned, optional batch."""
    q = 16 // dtype.size
    M = int(q * rng.integers(1, 50))
    N = int(q * rng.integers(1, 40))
    K = int(q * rng.integers(1, 90))
    H = int(rng.integers(1, 4))
    pad = int(q * rng.integers(0, 3))  # extra leading-dimension elements (a multiple of the run)
    if rng.random() < 0.5:   # A K-major
        lda = K + pad
        a = T([K, M, H], [1, lda, lda * M], [2, 0, 3], conj=dtype.is_complex and rng.random() < 0.5)
    else:
        lda = M + pad
        a = T([M, K, H], [1, lda, lda * K], [0, 2, 3], conj=dtype.is_complex and rng.random() < 0.5)
    if rng.random() < 0.5:   # B K-major
        ldb = K + pad
        b = T([K, N, H], [1, ldb, ldb * N], [2, 1, 3])
    else:
        ldb = N + pad
        b = T([N, K, H], [1, ldb, ldb * K], [1, 2, 3])
    d = T([M, N, H], [1, M, M * N], [0, 1, 3]) if rng.random() < 0.5 else T([M, N, H], [N, 1, M * N], [0, 1, 3])
    return a, b, d


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.F64, tp.DType.C32], ids=lambda d: d.name)
def test_dma_16_byte_runs_random_sweep(gpu, dtype):
    """Seeded random sweep of run layouts (24 cases per dtype): random tile,
    split and complex schedule, alpha and beta; every case matches the
    oracle, and most take the run path (the planner may orient an operand
    whose unit-stride extent is a single short run along its other index)."""
    rng = np.random.default_rng(1234 + dtype.size)
    runs = 0
    methods = [tp.ComplexMethod.NATIVE, tp.ComplexMethod.ONE_M, tp.ComplexMethod.THREE_M]
    for case in range(24):
        a, b, d = _random_run_case(rng, dtype)
        tile = int(rng.choice([1, 3, 4]))
        split = int(rng.choice([0, 1, 2, 5]))
        method = methods[int(rng.integers(0, 3))] if dtype.is_complex else tp.ComplexMethod.NATIVE
        av, bv, cv = values(dtype, a, 300 + case), values(dtype, b, 400 + case), values(dtype, d, 500 + case)
        alpha = complex(0.5, 0.25) if dtype.is_complex else 1.5
        beta = complex(-0.25, 0.5) if dtype.is_complex else -0.75
        got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), alpha=alpha, beta=beta, c=d,
                              cv=cv, config=tp.PlanConfig(tile=tile, split_k=split, complex_method=method))
        kern = plan.report()["kernel"]
        runs += ".dma16" in kern
        print("case", case, kern, a.extents, a.strides, b.extents, b.strides, flush=True)
        check(dtype, got, want, method)
    assert runs >= 18, runs
