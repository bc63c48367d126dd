"""GPU tests of tprims_product_autotune and the tuning cache: after
autotuning, D holds one execution's result (checked against the FP64 label
oracle, also when D accumulates into itself), the plan runs the chosen tile
(report source 3) and keeps giving oracle results, a split-BF16 plan tunes over
its own tiles, and a cache saved to a file drives a fresh plan of the same
problem after a clear and a load."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from harness import DTYPES, NP, T, check, make_plan, values

I, K, J = 1, 2, 3


@pytest.fixture
def clean_cache():
    tp.tuning_cache_clear()
    yield
    tp.tuning_cache_clear()


def _upload(*arrs):
    import torch

    dev = torch.device("cuda:0")
    out = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in arrs]
    torch.cuda.synchronize()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("in_place", [False, True], ids=["separate_c", "c_is_d"])
def test_autotune_leaves_one_execution_and_freezes_the_tile(gpu, clean_cache, dtype, in_place):
    import torch

    a = T([200, 330], [1, 200], [I, K], conj=dtype.is_complex)
    b = T([330, 150], [1, 330], [K, J])
    d = T([200, 150], [1, 200], [I, J])
    av, bv, cv = values(dtype, a, 81), values(dtype, b, 82), values(dtype, d, 83)
    alpha = 0.75 if not dtype.is_complex else complex(0.75, -0.25)
    beta = -0.5 if not dtype.is_complex else complex(-0.5, 0.25)
    plan = make_plan(dtype, a, b, d, d)
    ta, tb, tc = _upload(av, bv, cv)
    td = tc.clone() if in_place else torch.full_like(tc, float("nan"))
    pc = td.data_ptr() if in_place else tc.data_ptr()
    torch.cuda.synchronize()
    tile, ksplit, ms = plan.autotune(alpha, ta.data_ptr(), tb.data_ptr(), beta, pc, td.data_ptr(), reps=2)
    torch.cuda.synchronize()
    assert ms > 0
    r = plan.report()
    assert r["source"] == 3 and r["tile"] == tile, r
    assert tp.tuning_cache_size() == 1
    want = oracle.contract(alpha, a.oracle_op(av), b.oracle_op(bv), beta, d.oracle_op(cv),
                           d.oracle_op(np.zeros(d.storage(), NP[dtype])), cplx=dtype.is_complex)
    check(dtype, td.cpu().numpy(), want)
    # The frozen plan keeps giving oracle results.
    td2 = tc.clone() if in_place else torch.full_like(tc, float("nan"))
    pc2 = td2.data_ptr() if in_place else tc.data_ptr()
    plan.execute_raw(alpha, ta.data_ptr(), tb.data_ptr(), beta, pc2, td2.data_ptr())
    torch.cuda.synchronize()
    check(dtype, td2.cpu().numpy(), want)


@pytest.mark.gpu
def test_autotune_split_bf16_over_its_own_tiles(gpu, clean_cache):
    import torch

    dtype = tp.DType.F32
    a = T([512, 384], [1, 512], [I, K])
    b = T([384, 256], [1, 384], [K, J])
    d = T([512, 256], [1, 512], [I, J])
    av, bv = values(dtype, a, 91), values(dtype, b, 92)
    plan = make_plan(dtype, a, b, d, config=tp.PlanConfig(precision=tp.Precision.SPLIT_BF16_3))
    ta, tb, td = _upload(av, bv, np.full(d.storage(), np.nan, np.float32))
    tile, ksplit, ms = plan.autotune(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr(), reps=2)
    torch.cuda.synchronize()
    assert tile in (11, 12, 13) and ksplit == 1, (tile, ksplit)
    assert plan.report()["kernel"].startswith("f32.split_bf16x3")
    want = oracle.contract(1.0, a.oracle_op(av), b.oracle_op(bv), 0.0, None, d.oracle_op(np.zeros(d.storage())))
    check(dtype, td.cpu().numpy(), want, precision=tp.Precision.SPLIT_BF16_3)


@pytest.mark.gpu
def test_saved_cache_drives_a_fresh_plan(gpu, clean_cache, tmp_path):
    """The f64 24^6 contraction: autotune, save, clear, load; a new plan of the
    same problem takes the cached configuration and matches the oracle."""
    import torch

    e = 24
    cm = tp.col_major([e] * 4)
    a, b, d = T([e] * 4, cm, [1, 2, 3, 4]), T([e] * 4, cm, [4, 5, 2, 6]), T([e] * 4, cm, [1, 3, 5, 6])
    rng = np.random.default_rng(24)
    av, bv = rng.uniform(-1, 1, e ** 4), rng.uniform(-1, 1, e ** 4)
    ta, tb, td = _upload(av, bv, np.full(e ** 4, np.nan))
    plan = make_plan(tp.DType.F64, a, b, d)
    tile, ksplit, _ = plan.autotune(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr(), reps=3)
    kern = plan.report()["kernel"]
    path = tmp_path / "tuning.txt"
    tp.tuning_cache_save(str(path))
    tp.tuning_cache_clear()
    assert make_plan(tp.DType.F64, a, b, d).report()["source"] == 0
    tp.tuning_cache_load(str(path))
    fresh = make_plan(tp.DType.F64, a, b, d)
    r = fresh.report()
    assert r["source"] == 2 and r["tile"] == tile and r["kernel"] == kern, (r, kern)
    td2 = torch.full_like(td, float("nan"))
    fresh.execute_raw(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td2.data_ptr())
    torch.cuda.synchronize()
    want = oracle.contract(1.0, a.oracle_op(av), b.oracle_op(bv), 0.0, None, d.oracle_op(np.zeros(e ** 4)))
    for got in (td.cpu().numpy(), td2.cpu().numpy()):
        assert oracle.max_rel_err(got, want) <= 1e-11
