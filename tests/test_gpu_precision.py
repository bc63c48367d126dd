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
    errs = {}
    for prec in [tp.Precision.STORAGE] + SPLITS:
        got, want, plan = run(dtype, a, av, u, uv, d, dv, config=tp.PlanConfig(precision=prec))
        errs[prec] = oracle.max_rel_err(got.astype(np.complex128), want)
    print({k.name: "%.2e" % v for k, v in errs.items()})
    assert errs[tp.Precision.SPLIT_BF16_3] <= 1e-6
    assert errs[tp.Precision.STORAGE] <= 1e-6
    assert errs[tp.Precision.SPLIT_BF16_2] <= 1e-4
    # the 2-term scheme is measurably coarser than the 3-term one
    assert errs[tp.Precision.SPLIT_BF16_2] > errs[tp.Precision.SPLIT_BF16_3]


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
