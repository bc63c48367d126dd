"""GPU tests of the hand-written adjoints (tprims_product_vjp): dA and dB
against the FP64 oracle applied to the adjoint problem (diagonal output,
broadcast over the reduced index), against central finite differences of the
forward, and through torch.autograd (gradcheck over the DLPack forward)."""
import numpy as np
import pytest

import oracle
import tprims_amd as tp
from harness import DTYPES, NP, T, make_plan, tol


def diag_problem(i=6, j=5, r=3, b=2, k=7):
    """D[i,k,b] = sum_{j,r} A[i,i2=i,j,r,b] * B[j,k,b] (A diagonal over i,i2)."""
    I, I2, J, R, Bl, K = 1, 1, 2, 3, 4, 5
    a = T([i, i, j, r, b], tp.col_major([i, i, j, r, b]), [I, I2, J, R, Bl])
    bb = T([j, k, b], tp.col_major([j, k, b]), [J, K, Bl])
    d = T([i, k, b], tp.col_major([i, k, b]), [I, K, Bl])
    return a, bb, d


def rand(dtype, n, rng):
    x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if dtype.is_complex else 0)
    return x.astype(NP[dtype])


def adjoint_want(dtype, alpha, a, av, b, bv, d, gv, conj_a, conj_b, conj_d):
    """The oracle on the adjoint problems, written from the VJP formulas."""
    ca = complex(alpha).conjugate() if dtype.is_complex else alpha
    g_op = oracle.Operand(gv, d.extents, d.strides, d.labels, conj=conj_d)
    b_op = oracle.Operand(bv, b.extents, b.strides, b.labels, conj=not conj_b)
    a_op = oracle.Operand(av, a.extents, a.strides, a.labels, conj=not conj_a)
    zA = oracle.Operand(np.zeros(a.storage(), NP[dtype]), a.extents, a.strides, a.labels, conj=conj_a)
    zB = oracle.Operand(np.zeros(b.storage(), NP[dtype]), b.extents, b.strides, b.labels, conj=conj_b)
    cplx = dtype.is_complex
    wa = oracle.contract(ca, g_op, b_op, 0.0, None, zA, cplx=cplx)
    wb = oracle.contract(ca, a_op, g_op, 0.0, None, zB, cplx=cplx)
    return wa, wb


@pytest.mark.gpu
@pytest.mark.parametrize("conj", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 0)])
@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: d.name)
def test_vjp_matches_adjoint_oracle(gpu, dtype, conj):
    import torch

    if not dtype.is_complex and conj != (0, 0, 0):
        pytest.skip("conjugation is the identity for real types")
    rng = np.random.default_rng(3)
    a, b, d = diag_problem()
    a.conj, b.conj, d.conj = bool(conj[0]), bool(conj[1]), bool(conj[2])
    av, bv, gv = rand(dtype, a.storage(), rng), rand(dtype, b.storage(), rng), rand(dtype, d.storage(), rng)
    alpha = complex(0.75, -0.5) if dtype.is_complex else 0.75
    plan = make_plan(dtype, a, b, d)
    ta, tb, tg = (torch.from_numpy(x).to(gpu) for x in (av, bv, gv))
    da = torch.full_like(ta, float("nan"))
    db = torch.full_like(tb, float("nan"))
    plan.vjp(alpha, ta, tb, tg, da, db)
    torch.cuda.synchronize()
    wa, wb = adjoint_want(dtype, alpha, a, av, b, bv, d, gv, *[bool(c) for c in conj])
    for got, want, name in ((da, wa, "dA"), (db, wb, "dB")):
        g = got.cpu().numpy()
        assert np.all(np.isfinite(g.view(np.float32 if g.dtype in (np.float32, np.complex64) else np.float64))), name
        err = oracle.max_rel_err(g.astype(np.complex128), want)
        assert err <= tol(dtype), "%s %.3e" % (name, err)
    # off-diagonal entries of dA are exactly zero
    dA = da.cpu().numpy().reshape(a.extents, order="F")
    off = ~np.eye(a.extents[0], dtype=bool)
    assert np.all(dA[off] == 0)
    rep = plan.vjp_report(0)
    assert rep["strategy"] == 0 and rep["m"] == a.extents[0]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F64, tp.DType.C64], ids=lambda d: d.name)
def test_vjp_accumulates_into_the_diagonal(gpu, dtype):
    import torch

    rng = np.random.default_rng(4)
    a, b, d = diag_problem(4, 3, 2, 2, 5)
    av, bv, gv = rand(dtype, a.storage(), rng), rand(dtype, b.storage(), rng), rand(dtype, d.storage(), rng)
    plan = make_plan(dtype, a, b, d)
    ta, tb, tg = (torch.from_numpy(x).to(gpu) for x in (av, bv, gv))
    base = rand(dtype, a.storage(), rng)
    da = torch.from_numpy(base.copy()).to(gpu)
    plan.vjp(1.0, ta, tb, tg, da, None, accumulate=True)
    torch.cuda.synchronize()
    wa, _ = adjoint_want(dtype, 1.0, a, av, b, bv, d, gv, False, False, False)
    got = da.cpu().numpy().astype(np.complex128)
    err = oracle.max_rel_err(got, base.astype(np.complex128) + wa)
    assert err <= tol(dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tp.DType.F64, tp.DType.C64], ids=lambda d: d.name)
def test_vjp_matches_central_finite_differences(gpu, dtype):
    """Directional derivatives of L = Re sum conj(W) D, h = 1e-3 in FP64."""
    import torch

    rng = np.random.default_rng(8)
    a, b, d = diag_problem(4, 3, 2, 2, 5)
    av, bv, W = rand(dtype, a.storage(), rng), rand(dtype, b.storage(), rng), rand(dtype, d.storage(), rng)
    alpha = complex(0.5, 0.25) if dtype.is_complex else 0.5
    plan = make_plan(dtype, a, b, d)
    ta, tb, tw = (torch.from_numpy(x).to(gpu) for x in (av, bv, W))
    da, db = torch.empty_like(ta), torch.empty_like(tb)
    plan.vjp(alpha, ta, tb, tw, da, db)
    torch.cuda.synchronize()

    def loss(Av, Bv):
        tA, tB = torch.from_numpy(Av).to(gpu), torch.from_numpy(Bv).to(gpu)
        out = torch.zeros(d.storage(), dtype=tA.dtype, device=gpu)
        plan.execute_raw(alpha, tA.data_ptr(), tB.data_ptr(), 0.0, None, out.data_ptr())
        torch.cuda.synchronize()
        return float(np.real(np.vdot(W, out.cpu().numpy())))

    h = 1e-3
    for trial in range(3):
        V = rand(dtype, a.storage(), rng)
        fd = (loss(av + h * V, bv) - loss(av - h * V, bv)) / (2 * h)
        an = float(np.real(np.vdot(da.cpu().numpy(), V)))
        assert abs(fd - an) <= 1e-8 * max(1.0, abs(fd)), (trial, fd, an)
        U = rand(dtype, b.storage(), rng)
        fd = (loss(av, bv + h * U) - loss(av, bv - h * U)) / (2 * h)
        an = float(np.real(np.vdot(db.cpu().numpy(), U)))
        assert abs(fd - an) <= 1e-8 * max(1.0, abs(fd)), (trial, fd, an)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "complex128"])
def test_autograd_function_gradcheck(gpu, dtype):
    import torch

    from tprims_amd.autograd import contract

    dt = getattr(torch, dtype)
    g = torch.Generator(device=gpu).manual_seed(1)
    # D[i,k,b] = sum_{j,r} A[i,i,j,r,b] B[j,k,b]
    A = torch.randn(3, 3, 2, 2, 2, dtype=dt, device=gpu, generator=g).requires_grad_()
    B = torch.randn(2, 4, 2, dtype=dt, device=gpu, generator=g).requires_grad_()
    f = lambda x, y: contract("iijrb", x, "jkb", y, "ikb", alpha=0.5)
    assert torch.autograd.gradcheck(f, (A, B), eps=1e-6, atol=1e-7)
    # forward agrees with torch's own einsum
    ref = 0.5 * torch.einsum("iijrb,jkb->ikb", A.detach().cpu(), B.detach().cpu())
    assert torch.allclose(f(A, B).detach().cpu(), ref, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "complex128"])
def test_autograd_on_a_side_stream(gpu, dtype):
    """Forward and backward under torch.cuda.stream(s): the library's kernels
    run on s, ordered with the torch work that produces their inputs and
    consumes their outputs on s (no host synchronization in between)."""
    import torch

    from tprims_amd.autograd import contract

    dt = getattr(torch, dtype)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        g = torch.Generator(device=gpu).manual_seed(2)
        A0 = torch.randn(64, 64, 48, 3, 4, dtype=dt, device=gpu, generator=g)
        B0 = torch.randn(48, 56, 4, dtype=dt, device=gpu, generator=g)
        A = (A0 * 1.5).requires_grad_()  # produced on s just before the forward reads it
        B = (B0 - 0.25).requires_grad_()
        D = contract("iijrb", A, "jkb", B, "ikb", alpha=0.5)
        W = torch.randn(D.shape, dtype=dt, device=gpu, generator=g)
        (D * W.conj()).real.sum().backward()
        out = (D.detach().clone(), A.grad.clone(), B.grad.clone())
    s.synchronize()
    a64 = A.detach().cpu().to(torch.complex128 if A.is_complex() else torch.float64)
    b64 = B.detach().cpu().to(a64.dtype)
    w64 = W.cpu().to(a64.dtype)
    ref = 0.5 * torch.einsum("iijrb,jkb->ikb", a64, b64)
    tol_ = 1e-4 if dtype == "float32" else 1e-12
    scale = ref.abs().max()
    assert (out[0].cpu().to(a64.dtype) - ref).abs().max() <= tol_ * scale
    # dB = conj(alpha) sum_{i,r} conj(A[i,i,j,r,b]) W[i,k,b]
    diag = torch.diagonal(a64, dim1=0, dim2=1).movedim(-1, 0)  # [i, j, r, b]
    want_db = 0.5 * torch.einsum("ijrb,ikb->jkb", diag.conj(), w64)
    assert (out[2].cpu().to(a64.dtype) - want_db).abs().max() <= tol_ * want_db.abs().max()
    want_da_diag = 0.5 * torch.einsum("ikb,jkb->ijb", w64, b64.conj())
    got_da = out[1].cpu().to(a64.dtype)
    got_diag = torch.diagonal(got_da, dim1=0, dim2=1).movedim(-1, 0)  # [i, j, r, b]
    for r in range(got_diag.shape[2]):
        assert (got_diag[:, :, r, :] - want_da_diag).abs().max() <= tol_ * want_da_diag.abs().max()
