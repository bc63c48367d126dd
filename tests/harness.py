"""Shared test harness: TAPP-style cases run through the C ABI on the GPU (or
with host buffers, staged by the library) and checked against the FP64 label
oracle over the WHOLE output buffer, padding included, as the reference's
conformance suite does (crates/tprims-capi/tests/common/mod.rs:449-460).
"""
import numpy as np

import oracle
import tprims_amd as tp

NP = {tp.DType.F32: np.float32, tp.DType.F64: np.float64, tp.DType.C32: np.complex64, tp.DType.C64: np.complex128}
DTYPES = [tp.DType.F32, tp.DType.F64, tp.DType.C32, tp.DType.C64]


def tol(dtype, method=tp.ComplexMethod.NATIVE, precision=tp.Precision.STORAGE):
    """The reference's budgets (crates/tprims-capi/tests/common/mod.rs:57-63,
    crates/tprims-contract/src/driver/tests/correctness.rs:77-88): f32 2e-4,
    f64 1e-11, x100 for 3m. Split-BF16 3-term is held to 1e-5 (fp32-class)."""
    base = 2e-4 if dtype in (tp.DType.F32, tp.DType.C32) else 1e-11
    if method == tp.ComplexMethod.THREE_M and dtype.is_complex:
        base *= 100
    if precision == tp.Precision.SPLIT_BF16_3:
        base = 1e-5
    if precision == tp.Precision.SPLIT_BF16_2:
        base = 2e-4
    return base


class T:
    """A tensor as a C caller describes it: extents, strides, labels, op."""

    def __init__(self, extents, strides, labels, conj=False):
        self.extents = [int(e) for e in extents]
        self.strides = [int(s) for s in strides]
        self.labels = [int(l) for l in labels]
        self.conj = conj

    def c(self):
        return T(self.extents, self.strides, self.labels, True)

    def span(self):
        if any(e == 0 for e in self.extents):
            return 0, -1
        lo = sum(min(0, (e - 1) * s) for e, s in zip(self.extents, self.strides))
        hi = sum(max(0, (e - 1) * s) for e, s in zip(self.extents, self.strides))
        return lo, hi

    def storage(self):
        lo, hi = self.span()
        return max(0, hi - lo + 1)

    def origin(self):
        """Index of the logical origin inside the storage."""
        return -self.span()[0]

    def layout(self):
        return tp.LayoutSpec(self.extents, self.strides)

    def spec(self):
        return tp.OperandSpec(self.layout(), tp.Op.CONJUGATE if self.conj else tp.Op.IDENTITY)

    def oracle_op(self, data):
        return oracle.Operand(data, self.extents, self.strides, self.labels, offset=self.origin(), conj=self.conj)


def values(dtype, t, seed):
    return oracle.seq(max(1, t.storage()), seed, dtype.is_complex)[: max(1, t.storage())].astype(NP[dtype])


def make_plan(dtype, a, b, d, c=None, config=None):
    cdesc = c if c is not None else d
    prob = tp.Problem.from_labels(dtype, a.spec(), b.spec(), tp.CSpec.separate(cdesc.spec()), d.spec(),
                                  tp.Labels(a.labels, b.labels, d.labels, cdesc.labels))
    return tp.Plan(prob, config)


def run(dtype, a, av, b, bv, d, dv, alpha=1.0, beta=0.0, c=None, cv=None, in_place=False, config=None,
        where="gpu", executor=None, plan=None):
    """Runs one case (on `plan` when given, else a plan made from the
    tensors); returns (got_storage, want_storage, plan). With an executor the
    call runs on its stream (asynchronously when it is set so): the inputs are
    uploaded and synchronized before, the device synchronized after."""
    import torch

    if plan is None:
        plan = make_plan(dtype, a, b, d, c, config)
    es = dtype.size
    npdt = NP[dtype]
    av, bv, dv = (np.asarray(x, dtype=npdt).copy() for x in (av, bv, dv))
    cv = None if cv is None else np.asarray(cv, dtype=npdt).copy()

    want = oracle.contract(alpha, a.oracle_op(av), b.oracle_op(bv), beta,
                           (d if in_place else c).oracle_op(dv if in_place else cv) if (in_place or cv is not None)
                           else None, d.oracle_op(dv), cplx=dtype.is_complex)

    if where == "gpu":
        dev = torch.device("cuda:0")
        ta, tb, td = (torch.from_numpy(x).to(dev) for x in (av, bv, dv))
        tc = torch.from_numpy(cv).to(dev) if cv is not None else None
        pa = ta.data_ptr() + a.origin() * es
        pb = tb.data_ptr() + b.origin() * es
        pd = td.data_ptr() + d.origin() * es
        pc = pd if in_place else (tc.data_ptr() + c.origin() * es if tc is not None else None)
        torch.cuda.synchronize()  # the uploads, before work on another stream reads them
        plan.execute_raw(alpha, pa, pb, beta, pc, pd, executor)
        if executor is not None:
            executor.synchronize()
        torch.cuda.synchronize()
        got = td.cpu().numpy()
    else:  # host buffers, staged through the executor's device scratch
        pa = av.ctypes.data + a.origin() * es
        pb = bv.ctypes.data + b.origin() * es
        pd = dv.ctypes.data + d.origin() * es
        pc = pd if in_place else (cv.ctypes.data + c.origin() * es if cv is not None else None)
        plan.execute_raw(alpha, pa, pb, beta, pc, pd, executor)
        got = dv
    return got, want, plan


def check(dtype, got, want, method=tp.ComplexMethod.NATIVE, precision=tp.Precision.STORAGE, exact=False):
    got = np.asarray(got)
    assert np.all(np.isfinite(got.view(np.float32 if got.dtype in (np.float32, np.complex64) else np.float64))), \
        "non-finite output"
    if exact:
        np.testing.assert_array_equal(got.astype(np.complex128), np.asarray(want).astype(np.complex128))
        return 0.0
    err = oracle.max_rel_err(got.astype(np.complex128 if dtype.is_complex else np.float64), want)
    assert err <= tol(dtype, method, precision), "max_rel_err %.3e > %.1e" % (err, tol(dtype, method, precision))
    return err
