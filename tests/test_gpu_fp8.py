"""GPU tests of the FP8 precision path (f32 and c32 storage; OCP e4m3 with one
power-of-two scale per 128-k panel of every row, scaled fp8 MFMA):

- exact results when every input is exactly representable in e4m3: products and
  partial sums are then exact in fp32 and power-of-two scaling loses nothing,
  so any mistake in packing, scale handling, panel indexing, conjugation or
  the write-back shows as a mismatch. Ragged shapes cross 128-k panels and
  tile edges, with batches, both gather orientations and beta != 0;
- the error on Gaussian inputs against the FP64 oracle sits at the e4m3 level
  (above the split-BF16 schemes, far below O(1));
- every conformance case at an FP8-level budget (the semantics of diagonals,
  broadcasts, beta = 0 with a NaN C, in-place accumulation, ...)."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from cases import CASES
from harness import NP, T, check, run

FP8 = tp.PlanConfig(precision=tp.Precision.FP8_E4M3)
GRID = np.array([-2.0, -1.5, -1.0, -0.5, 0.5, 1.0, 1.5, 2.0])
FP8_TOL = 0.25  # normwise; e4m3 keeps 3 mantissa bits


def grid_values(dtype, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.choice(GRID, size=n)
    if dtype.is_complex:
        v = v + 1j * rng.choice(GRID, size=n)
    return v.astype(NP[dtype])


def batched_gemm(M, K, N, H, a_kmajor, conj_a=False, conj_b=False, b_kmajor=True):
    a_str = [K, 1, M * K] if a_kmajor else [1, M, M * K]
    b_str = [1, K, K * N] if b_kmajor else [N, 1, K * N]
    a = T([M, K, H], a_str, [0, 2, 3], conj=conj_a)
    b = T([K, N, H], b_str, [2, 1, 3], conj=conj_b)
    d = T([M, N, H], [1, M, M * N], [0, 1, 3])
    return a, b, d


# (M, K, N, H, A K-major, B K-major, conj A, conj B, beta). Operands gathered
# along their rows (not K-major) take the row packers: four rows per lane
# when the rows come in 16-byte runs (M or N a multiple of 4 here), one row
# per lane otherwise.
SHAPES = [
    (40, 300, 24, 3, True, True, False, False, 0.0),
    (40, 300, 24, 3, False, True, True, False, -0.5),
    (300, 130, 200, 1, True, True, False, True, 0.0),
    (17, 1000, 9, 2, True, True, True, True, -0.5),
    (257, 128, 129, 1, False, True, False, False, 0.0),
    (5, 7, 3, 4, True, True, False, False, -0.5),
    (40, 300, 24, 3, False, False, False, True, 0.0),
    (52, 260, 129, 2, False, False, True, False, -0.5),
    (257, 384, 68, 1, True, False, False, False, 0.0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(str(v) for v in s[:4]) + ("K" if s[4] else "M")
                         + ("K" if s[5] else "N"))
def test_fp8_exact_on_representable_inputs(gpu, dtype, shape):
    M, K, N, H, akm, bkm, ca, cb, beta = shape
    if not dtype.is_complex and (ca or cb):
        ca = cb = False
    a, b, d = batched_gemm(M, K, N, H, akm, ca, cb, b_kmajor=bkm)
    av = grid_values(dtype, a.storage(), 1)
    bv = grid_values(dtype, b.storage(), 2)
    dv = np.zeros(d.storage(), NP[dtype])
    kw = {}
    if beta != 0.0:
        kw = dict(c=d, cv=grid_values(dtype, d.storage(), 3))
    got, want, plan = run(dtype, a, av, b, bv, d, dv, alpha=1.5, beta=beta, config=FP8, **kw)
    assert "fp8_e4m3" in plan.report()["kernel"]
    check(dtype, got, want, exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
def test_fp8_gaussian_error_vs_fp64(gpu, dtype):
    M, K, N, H = 192, 640, 160, 2
    a, b, d = batched_gemm(M, K, N, H, True, conj_a=dtype.is_complex)
    rng = np.random.default_rng(7)

    def gauss(n):
        v = rng.standard_normal(n)
        if dtype.is_complex:
            v = v + 1j * rng.standard_normal(n)
        return v.astype(NP[dtype])

    av, bv = gauss(a.storage()), gauss(b.storage())
    dv = np.zeros(d.storage(), NP[dtype])
    errs = {}
    for prec in [tp.Precision.FP8_E4M3, tp.Precision.SPLIT_BF16_2]:
        got, want, _ = run(dtype, a, av, b, bv, d, dv, config=tp.PlanConfig(precision=prec))
        errs[prec] = oracle.max_rel_err(got.astype(np.complex128), want)
    print({k.name: "%.2e" % v for k, v in errs.items()})
    assert 1e-4 < errs[tp.Precision.FP8_E4M3] < 0.1
    assert errs[tp.Precision.FP8_E4M3] > 10 * errs[tp.Precision.SPLIT_BF16_2]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("name", sorted(CASES))
def test_fp8_conformance_case(gpu, name, dtype):
    kw = CASES[name](dtype)
    got, want, plan = run(dtype, config=FP8, **kw)
    if plan.report()["strategy"] == 0:
        assert "fp8_e4m3" in plan.report()["kernel"]
    err = oracle.max_rel_err(got.astype(np.complex128 if dtype.is_complex else np.float64), want)
    assert np.all(np.isfinite(got.astype(np.complex128)))
    assert err <= FP8_TOL, "max_rel_err %.3e" % err


def test_fp8_is_refused_for_double():
    from harness import make_plan
    d = T([4, 4], [1, 4], [0, 1])
    with pytest.raises(tp.TprimsError) as e:
        make_plan(tp.DType.C64, d, T([4, 4], [1, 4], [1, 2]), T([4, 4], [1, 4], [0, 2]), config=FP8)
    assert e.value.status == 15


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F32, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("L", [4, 8, 16, 24])
def test_fp8_packer_k_runs(gpu, L, dtype):
    """FP8 packing of operands whose K offsets are contiguous only within runs
    of L, D[i,k] = sum_{j,l} A[l,i,j] B[l,k,j] (aligned groups of 4 k load as
    16-byte vectors through the K table when L is a multiple of 8): exact on
    representable inputs."""
    I, Kx, J = 72, 40, 11
    a = T([L, I, J], tp.col_major([L, I, J]), [0, 1, 2])
    b = T([L, Kx, J], tp.col_major([L, Kx, J]), [0, 3, 2])
    d = T([I, Kx], tp.col_major([I, Kx]), [1, 3])
    av = grid_values(dtype, a.storage(), 4)
    bv = grid_values(dtype, b.storage(), 5)
    dv = np.zeros(d.storage(), NP[dtype])
    got, want, plan = run(dtype, a, av, b, bv, d, dv, config=FP8)
    assert "fp8_e4m3" in plan.report()["kernel"]
    check(dtype, got, want, exact=True)
