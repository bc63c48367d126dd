"""GPU parity over the TCCG corpus (tests/tccg.py; the reference's tcbench
corpus, benchmarks/benchmarks/tcbench/corpus.rs:262): every one of the 49
cases in every dtype at a small nominal size (64 KiB per tensor), compared
with the FP64 oracle over the whole output, and at the 64 MiB benchmark size
(the reference's published corpus size, docs/archive/tensorprimitives/results.md:84)
on a fixed sample of outputs evaluated in FP64 from the definition."""
import numpy as np
import pytest

import oracle
import tccg
import tprims_amd as tp
from harness import DTYPES, NP, T, check, run, tol
from test_tccg import plan_of

CASES = tccg.corpus()


def tensors(s):
    lab = lambda x: [ord(ch) for ch in x]
    return tuple(T(*s.layout(x), lab(x)) for x in (s.a, s.b, s.c))


def rand(dtype, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if dtype.is_complex else 0)
    return v.astype(NP[dtype])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_tccg_case_small_vs_oracle(gpu, case, dtype):
    s = tccg.Sized(case, 64 * 1024)
    a, b, d = tensors(s)
    av, bv = rand(dtype, a.storage(), 1), rand(dtype, b.storage(), 2)
    got, want, plan = run(dtype, a, av, b, bv, d, np.zeros(d.storage(), NP[dtype]), plan=plan_of(s, dtype))
    check(dtype, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F64, tp.DType.C32], ids=lambda d: d.name)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_tccg_case_benchmark_size_sampled(gpu, case, dtype):
    import torch

    s = tccg.Sized(case, 64 * 1024 * 1024)
    plan = plan_of(s, dtype)
    dev = torch.device("cuda:0")
    tdt = {tp.DType.F64: torch.float64, tp.DType.C32: torch.complex64}[dtype]
    g = torch.Generator(device=dev).manual_seed(3)
    ta = torch.randn(s.storage(s.a), dtype=tdt, device=dev, generator=g)
    tb = torch.randn(s.storage(s.b), dtype=tdt, device=dev, generator=g)
    td = torch.zeros(s.storage(s.c), dtype=tdt, device=dev)
    plan.execute_raw(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr())
    torch.cuda.synchronize()
    cd = np.complex128 if dtype.is_complex else np.float64
    A, B, D = (t.cpu().numpy().astype(cd) for t in (ta, tb, td))
    sa, sb, sc = (dict(zip(x, s.layout(x)[1])) for x in (s.a, s.b, s.c))
    kl = [ch for ch in s.labels if ch not in s.c]  # contracted labels
    kd = [s.extent[ch] for ch in kl]
    grid = np.indices(kd).reshape(len(kd), -1)
    ak = sum(grid[p] * sa[ch] for p, ch in enumerate(kl))
    bk = sum(grid[p] * sb[ch] for p, ch in enumerate(kl))
    rng = np.random.default_rng(4)
    got, want = [], []
    for _ in range(32):
        idx = {ch: int(rng.integers(0, s.extent[ch])) for ch in s.c}
        aoff = sum(idx[ch] * sa[ch] for ch in s.a if ch in idx)
        boff = sum(idx[ch] * sb[ch] for ch in s.b if ch in idx)
        want.append(np.dot(A[aoff + ak], B[boff + bk]))
        got.append(D[sum(idx[ch] * sc[ch] for ch in s.c)])
    got, want = np.array(got), np.array(want)
    err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
    # sums of len(ak) N(0,1) products: allow the budget to grow with sqrt(K) past K = 1024
    assert err <= tol(dtype) * max(1.0, np.sqrt(len(ak)) / 32), "%s: sampled rel err %.3e" % (case[0], err)
