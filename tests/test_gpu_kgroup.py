"""GPU parity of the K-group 32x32 tiles. Tile id 5 (LDS-DMA): four waves
each sum every fourth K slab of one 32 x 32 output tile and add their partial
tiles in LDS in a fixed order. Tile id 6 (f64 in 16-byte runs): eight waves,
each over every eighth slab, with operands loaded straight into registers
(K-major pairs feed two MFMA k steps, M/N-major pairs the two row blocks).
The planner takes them instead of split K for real products with a small
output and K up to 1024. Checked against the FP64 label oracle over every
dtype and complex method, both gather orientations of A and B, 16-byte-run
and 4-byte gathers, ragged M, N and K, conj, alpha and beta * C, batch
entries, K shorter than one round of the groups, repeated calls (bitwise
equal: the sum order is fixed), and the planner's choice of the tile."""
import numpy as np
import pytest

import tprims_amd as tp
from harness import DTYPES, NP, T, check, make_plan, run, values

I, K, J, H = 1, 2, 3, 4
METHODS = [tp.ComplexMethod.NATIVE, tp.ComplexMethod.ONE_M, tp.ComplexMethod.THREE_M]


def _operands(M, Kx, N, a_km, b_km, conj):
    """A[m,k] and B[k,n] gathered along K (km) or along M / N, D column-major."""
    a = T([M, Kx], [Kx, 1] if a_km else [1, M], [I, K], conj=conj)
    b = T([Kx, N], [1, Kx] if b_km else [N, 1], [K, J])
    d = T([M, N], [1, M], [I, J])
    return a, b, d


def _variants():
    out = []
    for dt in DTYPES:
        for m in (METHODS if dt.is_complex else [tp.ComplexMethod.NATIVE]):
            for majors in ("kk", "km", "mk", "mm"):
                for runs in (True, False):
                    out.append((dt, m, majors, runs))
    return out


def _supported(dtype, a_km, b_km, runs):
    """Where the K-group kernel runs (else the plan runs the register-staged
    32x32 tile): K-major A and B, or 16-byte runs of f64 / c32 in any
    orientation, or of f32 along K."""
    if runs and dtype != tp.DType.C64:
        return (a_km and b_km) or dtype in (tp.DType.F64, tp.DType.C32)
    return a_km and b_km


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,method,majors,runs", _variants(),
                         ids=lambda v: getattr(v, "name", str(v)))
def test_kgroup_tile_against_the_oracle(gpu, dtype, method, majors, runs):
    a_km, b_km = majors[0] == "k", majors[1] == "k"
    # Extents: multiples of 8 give 16-byte runs for every dtype; odd extents
    # give none. Neither is a multiple of the 32 x 32 tile or of 64 k.
    M, Kx, N = (136, 600, 104) if runs else (131, 597, 99)
    a, b, d = _operands(M, Kx, N, a_km, b_km, conj=dtype.is_complex)
    c = T([M, N], [N, 1], [I, J])  # C row-major while D is column-major
    cfg = tp.PlanConfig(complex_method=method, tile=5)
    av, bv, cv = values(dtype, a, 51), values(dtype, b, 52), values(dtype, c, 53)
    dv = np.full(d.storage(), np.nan, NP[dtype])
    alpha = 0.75 if not dtype.is_complex else complex(0.75, -0.25)
    got, want, plan = run(dtype, a, av, b, bv, d, dv, alpha=alpha, beta=-0.5, c=c, cv=cv, config=cfg)
    kern = plan.report()["kernel"]
    assert plan.report()["tile"] == 5
    assert (".kgroup4" in kern) == _supported(dtype, a_km, b_km, runs), kern
    assert ".splitk" not in kern, kern
    check(dtype, got, want, method)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_kgroup_tile_with_batch_entries_and_short_k(gpu, dtype):
    """Batch entries (tiles of every entry in one launch) with K shorter than
    one iteration of the four groups (groups past K add zeros), and a K of
    exactly one slab per group."""
    for Kx in (5, 64):
        a = T([40, Kx, 3], [Kx, 1, 40 * Kx], [I, K, H], conj=dtype.is_complex)
        b = T([Kx, 36, 3], [1, Kx, 36 * Kx], [K, J, H])
        d = T([40, 36, 3], [1, 40, 40 * 36], [I, J, H])
        av, bv = values(dtype, a, 61 + Kx), values(dtype, b, 62 + Kx)
        dv = np.zeros(d.storage(), NP[dtype])
        got, want, plan = run(dtype, a, av, b, bv, d, dv, alpha=1.5, config=tp.PlanConfig(tile=5))
        assert ".kgroup4" in plan.report()["kernel"]
        check(dtype, got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_planner_takes_the_kgroup_tile_for_small_outputs_with_long_k(gpu, dtype):
    """A 520 x 520 output (81 tiles of 64 x 64, 289 of 32 x 32) over K = 1000
    gathered along K: the planner takes the K-group tile instead of split K
    for f32 and f64 (split K for c32 and c64); two calls give bitwise equal
    results."""
    import torch

    M, Kx, N = 520, 1000, 520
    a, b, d = _operands(M, Kx, N, True, True, conj=dtype.is_complex)
    plan = make_plan(dtype, a, b, d)
    kern = plan.report()["kernel"]
    if dtype.is_complex:
        assert ".splitk" in kern, kern
    elif dtype == tp.DType.F64:
        assert ".regs16.kgroup8" in kern and ".splitk" not in kern, kern
    else:
        assert ".kgroup4" in kern and ".splitk" not in kern, kern
    av, bv = values(dtype, a, 71), values(dtype, b, 72)
    dv = np.zeros(d.storage(), NP[dtype])
    got, want, _ = run(dtype, a, av, b, bv, d, dv, plan=plan)
    check(dtype, got, want)
    dev = torch.device("cuda:0")
    ta, tb = torch.from_numpy(av).to(dev), torch.from_numpy(bv).to(dev)
    outs = []
    for _ in range(2):
        td = torch.zeros(d.storage(), dtype=ta.dtype, device=dev)
        plan.execute_raw(1.0, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr())
        torch.cuda.synchronize()
        outs.append(td.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], got)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [6, 7])
@pytest.mark.parametrize("majors", ["kk", "km", "mk", "mm"])
@pytest.mark.parametrize("runs", [True, False], ids=["runs", "no_runs"])
def test_register_direct_tile_against_the_oracle(gpu, tile, majors, runs):
    """Tiles 6 (32x32) and 7 (32x48; 32x32 where B gathers along N) on f64:
    every gather orientation, ragged M, N (not multiples of the tile) and K
    (not a multiple of 16), C row-major while D is column-major, alpha and
    beta; without 16-byte runs (odd extents) the plan runs as the K-group or
    register-staged 32x32 tile."""
    a_km, b_km = majors[0] == "k", majors[1] == "k"
    M, Kx, N = (202, 590, 130) if runs else (201, 589, 131)
    a, b, d = _operands(M, Kx, N, a_km, b_km, conj=False)
    c = T([M, N], [N, 1], [I, J])
    av, bv, cv = values(tp.DType.F64, a, 81), values(tp.DType.F64, b, 82), values(tp.DType.F64, c, 83)
    dv = np.full(d.storage(), np.nan)
    got, want, plan = run(tp.DType.F64, a, av, b, bv, d, dv, alpha=0.75, beta=-0.5, c=c, cv=cv,
                          config=tp.PlanConfig(tile=tile))
    kern = plan.report()["kernel"]
    assert plan.report()["tile"] == tile
    assert (".regs16.kgroup8" in kern) == runs, kern
    assert ("32x48" in kern) == (runs and tile == 7 and b_km), kern
    check(tp.DType.F64, got, want)


@pytest.mark.gpu
def test_register_direct_tile_with_batch_entries_short_k_and_repeats(gpu):
    """Batch entries in one launch, K from two pairs (one slab, partly past K)
    to one slab past a round of the eight groups (groups past K load zeros
    from the sentinel), and two calls bitwise equal."""
    import torch

    for tile, Kx in ((6, 4), (6, 16), (6, 94), (6, 144), (7, 4), (7, 94), (7, 144)):
        a = T([40, Kx, 3], [1, 40, 40 * Kx], [I, K, H])
        b = T([Kx, 36, 3], [1, Kx, 36 * Kx], [K, J, H])
        d = T([40, 36, 3], [1, 40, 40 * 36], [I, J, H])
        av, bv = values(tp.DType.F64, a, 91 + Kx), values(tp.DType.F64, b, 92 + Kx)
        dv = np.full(d.storage(), np.nan)
        plan = make_plan(tp.DType.F64, a, b, d, config=tp.PlanConfig(tile=tile))
        assert ".regs16.kgroup8" in plan.report()["kernel"], plan.report()
        got, want, _ = run(tp.DType.F64, a, av, b, bv, d, dv, alpha=1.5, plan=plan)
        check(tp.DType.F64, got, want)
        dev = torch.device("cuda:0")
        ta, tb = torch.from_numpy(av).to(dev), torch.from_numpy(bv).to(dev)
        td = torch.full((d.storage(),), float("nan"), dtype=torch.float64, device=dev)
        plan.execute_raw(1.5, ta.data_ptr(), tb.data_ptr(), 0.0, None, td.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(td.cpu().numpy(), got)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_kgroup_tile_hands_partial_sums_on_in_order(gpu, dtype):
    """The K groups add their partial tiles through LDS in turn; each group's
    stores must be complete before the next group reads them. M = 72, N =
    140608, K = 72 (two slabs per group, every call's groups finishing close
    together) over 300 calls: every result bitwise equal to the first and to
    a float64 product within the storage tolerance. Without the wait before
    the hand-off barrier the f32 tile gave wrong tiles in about 1 call in 100
    here."""
    import torch

    M, Kx, N = 72, 72, 140608
    a, b, d = _operands(M, Kx, N, True, True, conj=False)
    plan = make_plan(dtype, a, b, d, config=tp.PlanConfig(tile=5))
    assert ".kgroup4" in plan.report()["kernel"], plan.report()
    dev = torch.device("cuda:0")
    tdt = {tp.DType.F32: torch.float32, tp.DType.F64: torch.float64,
           tp.DType.C32: torch.complex64, tp.DType.C64: torch.complex128}[dtype]
    g = torch.Generator(device=dev).manual_seed(7)
    A = torch.randn(a.storage(), dtype=tdt, device=dev, generator=g)
    B = torch.randn(b.storage(), dtype=tdt, device=dev, generator=g)
    # A[k, m] at k + K m, B[k, n] at k + K n, D[m, n] at m + M n.
    want = (A.view(M, Kx).to(torch.complex128 if dtype.is_complex else torch.float64)
            @ B.view(N, Kx).to(torch.complex128 if dtype.is_complex else torch.float64).T).T.reshape(-1)
    D = torch.empty(d.storage(), dtype=tdt, device=dev)
    first = None
    for i in range(300):
        D.fill_(float("nan"))
        plan.execute_raw(1.0, A.data_ptr(), B.data_ptr(), 0.0, None, D.data_ptr())
        torch.cuda.synchronize()
        if first is None:
            first = D.clone()
            err = ((first.to(want.dtype) - want).abs().max() / want.abs().max()).item()
            assert err < (1e-5 if dtype in (tp.DType.F32, tp.DType.C32) else 1e-13), err
        else:
            assert torch.equal(D, first), "call %d differs from the first" % i
