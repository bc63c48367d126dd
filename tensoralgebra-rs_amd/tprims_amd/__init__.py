"""tprims_amd: the Python host mirror of the reference's contraction API,
running on libtprims.so (hand-written HIP kernels for MI355X / gfx950).

Mirrors crates/tprims-contract/src/api (Problem::from_labels,
Problem::from_dot_general, Labels, DotGeneral, LayoutSpec, OperandSpec,
CSpec) and crates/tprims-contract/src/plan (Plan::new, execute_into,
execute_into_accum, execute_raw, report). A Plan lowers to one TAPP product of
the C ABI; execution runs on the GPU through that product.

Operands are torch tensors (device or host), numpy arrays (host, staged by the
library) or raw integer addresses (`execute_raw`). The library is required:
nothing here computes a contraction in Python.
"""
import array
import ctypes
import os
import enum

from . import _capi
from ._capi import TprimsError, check, i64_array, lib

__all__ = [
    "DType", "Op", "LayoutSpec", "OperandSpec", "CSpec", "Labels", "DotGeneral", "Problem", "PlanConfig", "Plan",
    "Executor", "TprimsError", "ComplexMethod", "Precision", "col_major", "row_major", "build_scatter",
    "build_block_scatter", "IRREGULAR", "SCATTERS", "tuning_cache_save", "tuning_cache_load",
    "tuning_cache_clear", "tuning_cache_size",
]


class DType(enum.IntEnum):
    """Storage type (crates/tprims-contract/src/api/problem.rs:17)."""
    F32 = _capi.TAPP_F32
    F64 = _capi.TAPP_F64
    C32 = _capi.TAPP_C32
    C64 = _capi.TAPP_C64

    @property
    def size(self):
        return {DType.F32: 4, DType.F64: 8, DType.C32: 8, DType.C64: 16}[self]

    @property
    def is_complex(self):
        return self in (DType.C32, DType.C64)

    @staticmethod
    def of(x):
        """DType of a torch tensor or numpy array."""
        name = str(x.dtype).replace("torch.", "")
        table = {"float32": DType.F32, "float64": DType.F64, "complex64": DType.C32, "complex128": DType.C64}
        if name not in table:
            raise TprimsError(_capi.TPRIMS_ERR_DTYPE, "unsupported dtype %s" % name)
        return table[name]


class Op(enum.IntEnum):
    IDENTITY = _capi.TAPP_IDENTITY
    CONJUGATE = _capi.TAPP_CONJUGATE


class ComplexMethod(enum.IntEnum):
    NATIVE = _capi.TPRIMS_CPLX_NATIVE
    ONE_M = _capi.TPRIMS_CPLX_1M
    THREE_M = _capi.TPRIMS_CPLX_3M


class Precision(enum.IntEnum):
    STORAGE = _capi.TPRIMS_PREC_STORAGE
    SPLIT_BF16_2 = _capi.TPRIMS_PREC_SPLIT_BF16_2
    SPLIT_BF16_3 = _capi.TPRIMS_PREC_SPLIT_BF16_3
    FP8_E4M3 = _capi.TPRIMS_PREC_FP8_E4M3


def col_major(dims):
    """Column-major element strides (fixtures.rs:44)."""
    s, acc = [], 1
    for d in dims:
        s.append(acc)
        acc *= max(int(d), 1)
    return s


def row_major(dims):
    """Row-major element strides (fixtures.rs:55)."""
    s, acc = [0] * len(dims), 1
    for k in range(len(dims) - 1, -1, -1):
        s[k] = acc
        acc *= max(int(dims[k]), 1)
    return s


class LayoutSpec:
    """Extents, signed element strides and a logical offset (problem.rs:71)."""

    def __init__(self, dims, strides, offset=0):
        if len(dims) != len(strides):
            raise TprimsError(_capi.TPRIMS_ERR_SHAPE, "extents and strides differ in length")
        self.dims = [int(d) for d in dims]
        self.strides = [int(s) for s in strides]
        self.offset = int(offset)

    @staticmethod
    def of(t):
        """The layout of a torch tensor or numpy array (origin at its first element)."""
        if hasattr(t, "stride") and callable(t.stride):
            return LayoutSpec(list(t.shape), list(t.stride()), 0)
        item = t.itemsize
        return LayoutSpec(list(t.shape), [s // item for s in t.strides], 0)

    @property
    def rank(self):
        return len(self.dims)

    def __repr__(self):
        return "LayoutSpec(%r, %r, %r)" % (self.dims, self.strides, self.offset)


class OperandSpec:
    """A layout plus an element-wise operation (problem.rs:196)."""

    def __init__(self, layout, op=Op.IDENTITY):
        self.layout = layout
        self.op = Op(op)

    def conj(self):
        return OperandSpec(self.layout, Op.CONJUGATE)


class CSpec:
    """How beta*op_C(C) is supplied (problem.rs:236): Absent, Output(op), Separate(spec)."""

    def __init__(self, kind, op=Op.IDENTITY, spec=None):
        self.kind, self.op, self.spec = kind, Op(op), spec

    @staticmethod
    def absent():
        return CSpec("absent")

    @staticmethod
    def output(op=Op.IDENTITY):
        return CSpec("output", op)

    @staticmethod
    def separate(spec):
        return CSpec("separate", spec.op, spec)


class Labels:
    """Per-operand int64 labels (crates/tprims-contract/src/api/labels.rs:19)."""

    def __init__(self, a, b, d, c=None):
        self.a = [int(x) for x in a]
        self.b = [int(x) for x in b]
        self.d = [int(x) for x in d]
        self.c = None if c is None else [int(x) for x in c]

    def with_c(self, c):
        return Labels(self.a, self.b, self.d, c)


class DotGeneral:
    """Contracted and batch axis lists (crates/tprims-contract/src/api/dot_general.rs:20).

    Output extents are [lhs_free, rhs_free, batch]; C is D.
    """

    def __init__(self, lhs_contract, rhs_contract, lhs_batch=(), rhs_batch=()):
        self.lhs_contract = [int(x) for x in lhs_contract]
        self.rhs_contract = [int(x) for x in rhs_contract]
        self.lhs_batch = [int(x) for x in lhs_batch]
        self.rhs_batch = [int(x) for x in rhs_batch]

    def validate(self, a_dims, b_dims):
        """dot_general.rs:82-141: returns (lhs_free, rhs_free, out_dims)."""
        for what, l, r in (("contracting", self.lhs_contract, self.rhs_contract),
                           ("batch", self.lhs_batch, self.rhs_batch)):
            if len(l) != len(r):
                raise TprimsError(_capi.TPRIMS_ERR_INVALID_ARGUMENT, "%s axis lists differ in length" % what)

        def free(name, rank, c, bt):
            seen = [False] * rank
            for x in list(c) + list(bt):
                if x >= rank or x < 0:
                    raise TprimsError(_capi.TPRIMS_ERR_INVALID_ARGUMENT, "axis %d out of range for %s" % (x, name))
                if seen[x]:
                    raise TprimsError(_capi.TPRIMS_ERR_INVALID_ARGUMENT, "axis %d of %s repeated" % (x, name))
                seen[x] = True
            return [x for x in range(rank) if not seen[x]]

        lf = free("A", len(a_dims), self.lhs_contract, self.lhs_batch)
        rf = free("B", len(b_dims), self.rhs_contract, self.rhs_batch)
        for l, r in list(zip(self.lhs_contract, self.rhs_contract)) + list(zip(self.lhs_batch, self.rhs_batch)):
            if a_dims[l] != b_dims[r]:
                raise TprimsError(_capi.TPRIMS_ERR_SHAPE, "paired axes %d/%d have extents %d and %d"
                                  % (l, r, a_dims[l], b_dims[r]))
        out = [a_dims[x] for x in lf] + [b_dims[x] for x in rf] + [a_dims[x] for x in self.lhs_batch]
        return lf, rf, out


class Problem:
    """D = op_D(alpha * sum_K op_A(A) op_B(B) + beta * op_C(C)) (problem.rs:1-10).

    Validation happens when the Plan lowers it through the C ABI.
    """

    def __init__(self, dtype, a, b, c, d, labels):
        self.dtype = DType(dtype)
        self.a, self.b, self.c, self.d, self.labels = a, b, c, d, labels
        if (c.kind == "separate") != (labels.c is not None):
            raise TprimsError(_capi.TPRIMS_ERR_LABELS, "a separate C needs its labels")

    @staticmethod
    def from_labels(dtype, a, b, c, d, labels):
        return Problem(dtype, a, b, c, d, labels)

    @staticmethod
    def from_dot_general(dtype, a, b, d, dot):
        """problem.rs:396-449: synthetic labels, C is D with identity op_C."""
        lf, rf, out = dot.validate(a.layout.dims, b.layout.dims)
        if out != d.layout.dims:
            raise TprimsError(_capi.TPRIMS_ERR_SHAPE, "output extents %r, expected %r" % (d.layout.dims, out))
        nb, nc = len(dot.lhs_batch), len(dot.lhs_contract)
        la, lb, ld = [0] * a.layout.rank, [0] * b.layout.rank, []
        nxt = nb + nc
        for x in lf:
            la[x] = nxt
            ld.append(nxt)
            nxt += 1
        for x in rf:
            lb[x] = nxt
            ld.append(nxt)
            nxt += 1
        for p, (l, r) in enumerate(zip(dot.lhs_batch, dot.rhs_batch)):
            la[l] = p
            lb[r] = p
            ld.append(p)
        for p, (l, r) in enumerate(zip(dot.lhs_contract, dot.rhs_contract)):
            la[l] = nb + p
            lb[r] = nb + p
        return Problem(dtype, a, b, CSpec.output(Op.IDENTITY), d, Labels(la, lb, ld))


class PlanConfig:
    """Plan configuration (crates/tprims-contract/src/plan/config.rs:101), GPU knobs."""

    def __init__(self, complex_method=ComplexMethod.NATIVE, precision=Precision.STORAGE, tile=0, split_k=0):
        self.complex_method = int(complex_method)
        self.precision = int(precision)
        self.tile = int(tile)
        self.split_k = int(split_k)  # 0 planner's choice, 1 never, n > 1 that many K ranges

    def to_c(self):
        c = _capi.PlanConfigC()
        c.complex_method, c.precision, c.tile, c.split_k = self.complex_method, self.precision, self.tile, self.split_k
        return c


class Executor:
    """A GPU executor: a device and a stream (tprims_tapp_executor_create_hip)."""

    def __init__(self, device=0, stream=None, asynchronous=False):
        h = _capi.intptr(0)
        check(lib().tprims_tapp_executor_create_hip(ctypes.byref(h), int(device), stream))
        self.handle = h.value
        if asynchronous:
            check(lib().tprims_tapp_executor_set_async(self.handle, 1))

    def synchronize(self):
        check(lib().tprims_tapp_executor_synchronize(self.handle))

    @property
    def stream(self):
        s = ctypes.c_void_p(0)
        check(lib().tprims_tapp_executor_get_stream(self.handle, ctypes.byref(s)))
        return s.value

    def close(self):
        if self.handle:
            check(lib().TAPP_destroy_executor(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_HANDLE = None


def _handle():
    global _HANDLE
    if _HANDLE is None:
        h = _capi.intptr(0)
        check(lib().TAPP_create_handle(ctypes.byref(h)))
        _HANDLE = h.value
    return _HANDLE


def _info(dtype, layout):
    h = _capi.intptr(0)
    check(lib().TAPP_create_tensor_info(ctypes.byref(h), int(dtype), layout.rank, i64_array(layout.dims),
                                        i64_array(layout.strides)))
    return h.value


def _addr(x, layout=None):
    """Address of the element at logical index zero of a tensor/array/int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data  # numpy


IRREGULAR = -(1 << 63)  # block-scatter entry of a block that is not an arithmetic run (scatter.rs:43)
# The plan's scatter vectors (the reduced matrix view, analysis.rs:210), by TPRIMS_SCATTER_* id.
SCATTERS = ("am", "ak", "bk", "bn", "cm", "cn", "dm", "dn", "ha", "hb", "hc", "hd")


def _sized(call):
    """Runs call(out, cap, len_ptr) twice: once to size, once to fill."""
    n = ctypes.c_int64(0)
    check(call(None, 0, ctypes.byref(n)))
    buf = (ctypes.c_int64 * max(1, n.value))()
    check(call(buf, n.value, ctypes.byref(n)))
    return list(buf[:n.value])


def build_scatter(extents, strides):
    """Element offsets of the mixed-radix index tuples, fastest mode first
    (crates/tprims-kernel/src/pack/scatter.rs:70); [] modes give [0]."""
    if len(extents) != len(strides):
        raise TprimsError(_capi.TPRIMS_ERR_SHAPE, "extents and strides differ in length")
    e, s = i64_array(extents), i64_array(strides)
    return _sized(lambda out, cap, n: lib().tprims_build_scatter(e, s, len(extents), out, cap, n))


def build_block_scatter(scat, blk):
    """Per block of blk entries: the run's common difference, 0 for a single
    entry, or IRREGULAR (crates/tprims-kernel/src/pack/scatter.rs:130)."""
    v = i64_array(scat)
    return _sized(lambda out, cap, n: lib().tprims_build_block_scatter(v, len(scat), int(blk), out, cap, n))


def tuning_cache_save(path):
    """Writes the process's tuning cache to `path` (tprims_tuning_cache_save)."""
    check(lib().tprims_tuning_cache_save(os.fsencode(path)))


def tuning_cache_load(path):
    """Merges a saved tuning cache into the process's (all entries or none)."""
    check(lib().tprims_tuning_cache_load(os.fsencode(path)))


def tuning_cache_clear():
    check(lib().tprims_tuning_cache_clear())


def tuning_cache_size():
    n = ctypes.c_int64(0)
    check(lib().tprims_tuning_cache_size(ctypes.byref(n)))
    return n.value


def _scalar(v, dtype):
    v = complex(v)
    if dtype == DType.F32:
        c = ctypes.c_float(v.real)
        return ctypes.byref(c), c
    if dtype == DType.F64:
        c = ctypes.c_double(v.real)
        return ctypes.byref(c), c
    if dtype == DType.C32:
        c = (ctypes.c_float * 2)(v.real, v.imag)
        return ctypes.cast(c, ctypes.c_void_p), c
    c = (ctypes.c_double * 2)(v.real, v.imag)
    return ctypes.cast(c, ctypes.c_void_p), c


class Plan:
    """A prepared contraction (crates/tprims-contract/src/plan/mod.rs:139).

    Holds no data pointer; reusable across buffers and executors.
    """

    def __init__(self, problem, config=None):
        self.problem = problem
        self.config = config or PlanConfig()
        p = problem
        la, lb, ld = p.labels.a, p.labels.b, p.labels.d
        if p.c.kind == "separate":
            c_layout, lc, c_op = p.c.spec.layout, p.labels.c, p.c.op
        else:
            c_layout, lc, c_op = p.d.layout, ld, p.c.op
        self._c_kind = p.c.kind
        infos = [_info(p.dtype, l) for l in (p.a.layout, p.b.layout, c_layout, p.d.layout)]
        try:
            h = _capi.intptr(0)
            cfg = self.config.to_c()
            check(lib().tprims_create_tensor_product_ex(
                ctypes.byref(h), _handle(), int(p.a.op), infos[0], i64_array(la), int(p.b.op), infos[1],
                i64_array(lb), int(c_op), infos[2], i64_array(lc), int(p.d.op), infos[3], i64_array(ld),
                _capi.TAPP_DEFAULT_PREC, ctypes.byref(cfg)))
            self.handle = h.value
        finally:
            for i in infos:
                lib().TAPP_destroy_tensor_info(i)

    @staticmethod
    def new(problem, config=None):
        return Plan(problem, config)

    def scatters(self):
        """The plan's twelve scatter vectors (am, ak, bk, bn, cm, cn, dm, dn,
        ha, hb, hc, hd) after role assignment, sorting and folding
        (crates/tprims-contract/src/plan/analysis.rs:210, Plan::oriented_scatters)."""
        return {name: _sized(lambda out, cap, n, i=i: lib().tprims_product_scatter(self.handle, i, out, cap, n))
                for i, name in enumerate(SCATTERS)}

    def report(self):
        """The planner's conclusions (plan/report.rs:33) as a dict."""
        r = _capi.PlanReportC()
        check(lib().tprims_product_report(self.handle, ctypes.byref(r)))
        out = {f: getattr(r, f) for f, _ in _capi.PlanReportC._fields_}
        out["kernel"] = r.kernel.decode()
        return out

    def autotune(self, alpha, a, b, beta, c, d, executor=None, reps=3):
        """tprims_product_autotune: times every tile (and K split or none) this
        plan can run on the given device operands (element addresses of index
        zero, as execute_raw takes them), freezes the fastest into the plan,
        records it in the process's tuning cache and leaves D as one execution
        does. Returns (tile id, K ranges, mean ms of the chosen configuration)."""
        dt = self.problem.dtype
        pa, _ka = _scalar(alpha, dt)
        pb, _kb = _scalar(beta, dt)
        t, s, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0)
        ex = executor.handle if executor is not None else 0
        check(lib().tprims_product_autotune(self.handle, ex, pa, a, b, pb, c, d, int(reps), ctypes.byref(t),
                                            ctypes.byref(s), ctypes.byref(ms)))
        return t.value, s.value, ms.value

    def tuning_key(self):
        """The key of this plan's reduced problem in the tuning cache."""
        n = ctypes.c_int64(0)
        check(lib().tprims_product_tuning_key(self.handle, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().tprims_product_tuning_key(self.handle, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def execute_raw(self, alpha, a, b, beta, c, d, executor=None):
        """Execute on element addresses of index zero (plan/mod.rs:452)."""
        dt = self.problem.dtype
        pa, _ka = _scalar(alpha, dt)
        pb, _kb = _scalar(beta, dt)
        st = _capi.intptr(0)
        ex = executor.handle if executor is not None else 0
        check(lib().TAPP_execute_product(self.handle, ex, ctypes.byref(st), pa, a, b, pb, c, d))

    def execute_batched_raw(self, alpha, a, b, beta, c, d, executor=None):
        """TAPP_execute_batched_product (crates/tprims-capi/src/execute.rs:235):
        the plan on len(d) items; a, b, d (and c, or None) are lists of element
        addresses of index zero. Same alpha and beta for every item; items run
        as one launch when that equals running them in order."""
        dt = self.problem.dtype
        n = len(d)
        if len(a) != n or len(b) != n or (c is not None and len(c) != n):
            raise ValueError("a, b, c and d need one address per item")
        pa, _ka = _scalar(alpha, dt)
        pb, _kb = _scalar(beta, dt)
        keep = []

        def arr(xs):
            # One uint64 buffer viewed as the void* array (no per-address
            # objects); None is a null pointer, which the library refuses.
            try:
                buf = array.array("Q", xs) if n else array.array("Q", [0])
            except TypeError:
                buf = array.array("Q", [0 if x is None else x for x in xs])
            keep.append(buf)
            return (ctypes.c_void_p * len(buf)).from_buffer(buf)

        A, B, D = arr(a), arr(b), arr(d)
        C = arr(c) if c is not None else None
        st = _capi.intptr(0)
        ex = executor.handle if executor is not None else 0
        check(lib().TAPP_execute_batched_product(self.handle, ex, ctypes.byref(st), n, pa, A, B, pb, C, D))

    def execute_into(self, alpha, a, b, d, executor=None):
        """D = op_D(alpha * op_A(A) op_B(B)) (plan/mod.rs:271); C is not read."""
        self.execute_raw(alpha, _addr(a), _addr(b), 0.0, None, _addr(d), executor)

    def execute_into_accum(self, alpha, a, b, beta, c, d, executor=None):
        """D = op_D(alpha * op_A(A) op_B(B) + beta * op_C(C)) (plan/mod.rs:309).

        For a problem whose C is D (CSpec.output), pass c=None: D is read in place.
        """
        if self._c_kind == "absent":
            beta, c = 0.0, None
        elif self._c_kind == "output":
            c = d
        self.execute_raw(alpha, _addr(a), _addr(b), beta, _addr(c), _addr(d), executor)

    def execute_dlpack(self, alpha, a, b, beta, c, d, executor=None):
        """Execute with DLPack operands (tprims_execute_product_dlpack); each of
        a, b, c (or None), d is a torch tensor exported through DLPack."""
        from .dlpack import borrow

        dt = self.problem.dtype
        pa, _ka = _scalar(alpha, dt)
        pb, _kb = _scalar(beta, dt)
        keep = []
        ta, tb, td = borrow(a, keep), borrow(b, keep), borrow(d, keep)
        tc = borrow(c, keep) if c is not None else _capi.TprimsTensor()
        ex = executor.handle if executor is not None else 0
        check(lib().tprims_execute_product_dlpack(self.handle, ex, pa, ta, tb, pb, tc, td))

    def vjp(self, alpha, a, b, dd, da=None, db=None, accumulate=False, executor=None):
        """Cotangents dA, dB of D with respect to A and B (tprims_product_vjp);
        operands are device tensors or addresses in the plan's layouts."""
        dt = self.problem.dtype
        pa, _ka = _scalar(alpha, dt)
        ex = executor.handle if executor is not None else 0
        check(lib().tprims_product_vjp(self.handle, ex, pa, _addr(a), _addr(b), _addr(dd), _addr(da), _addr(db),
                                       int(bool(accumulate))))

    def vjp_report(self, which):
        r = _capi.PlanReportC()
        check(lib().tprims_product_vjp_report(self.handle, int(which), ctypes.byref(r)))
        out = {f: getattr(r, f) for f, _ in _capi.PlanReportC._fields_}
        out["kernel"] = r.kernel.decode()
        return out

    def close(self):
        if getattr(self, "handle", 0):
            lib().TAPP_destroy_tensor_product(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def contract(labels_a, a, labels_b, b, labels_d, d, alpha=1.0, beta=0.0, c=None, labels_c=None, conj_a=False,
             conj_b=False, conj_c=False, conj_d=False, config=None, executor=None):
    """One-shot label contraction on torch tensors / numpy arrays."""
    dt = DType.of(d)
    spec = lambda x, cj: OperandSpec(LayoutSpec.of(x), Op.CONJUGATE if cj else Op.IDENTITY)
    if c is not None:
        cs = CSpec.separate(spec(c, conj_c))
        lab = Labels(labels_a, labels_b, labels_d, labels_c if labels_c is not None else labels_d)
    else:
        cs = CSpec.absent()
        lab = Labels(labels_a, labels_b, labels_d)
    if c is None:
        # TAPP needs a C description: reuse D's, unread at beta 0.
        cs = CSpec.separate(spec(d, conj_c))
        lab = lab.with_c(labels_d)
        beta = 0.0
    plan = Plan(Problem.from_labels(dt, spec(a, conj_a), spec(b, conj_b), cs, spec(d, conj_d), lab), config)
    plan.execute_raw(alpha, _addr(a), _addr(b), beta, _addr(c), _addr(d), executor)
    return plan


class Group:
    """Many contractions of one dtype run as one persistent kernel
    (tprims_group_create; the GPU form of contract_batched,
    crates/tprims-contract/src/batch.rs:125). Operands are device tensors."""

    def __init__(self, plans, alphas, a, b, betas, c, d, executor=None):
        n = len(plans)
        dt = plans[0].problem.dtype
        scal = (ctypes.c_double if dt in (DType.F64, DType.C64) else ctypes.c_float)
        w = 2 if dt.is_complex else 1

        def scalars(vs):
            flat = []
            for v in vs:
                v = complex(v)
                flat += [v.real, v.imag][:w]
            return (scal * max(1, len(flat)))(*flat)

        def ptrs(xs):
            return (ctypes.c_void_p * max(1, n))(*[_addr(x) for x in xs])

        self._keep = [scalars(alphas), scalars(betas), ptrs(a), ptrs(b), ptrs(c) if c is not None else None, ptrs(d)]
        self.plans = plans
        h = _capi.intptr(0)
        ex = executor.handle if executor is not None else 0
        hs = (_capi.intptr * max(1, n))(*[p.handle for p in plans])
        self._keep.append(hs)
        check(lib().tprims_group_create(ctypes.byref(h), ex, n, hs, ctypes.cast(self._keep[0], ctypes.c_void_p),
                                        self._keep[2], self._keep[3], ctypes.cast(self._keep[1], ctypes.c_void_p),
                                        self._keep[4], self._keep[5]))
        self.handle = h.value
        self.executor = executor

    def _ex(self, executor):
        e = executor or self.executor
        return e.handle if e is not None else 0

    def execute(self, executor=None):
        check(lib().tprims_group_execute(self.handle, self._ex(executor)))

    def execute_each(self, executor=None):
        check(lib().tprims_group_execute_each(self.handle, self._ex(executor)))

    def use_graph(self, enable=True, executor=None):
        check(lib().tprims_group_use_graph(self.handle, self._ex(executor), int(bool(enable))))

    def set_tile(self, tile):
        check(lib().tprims_group_set_tile(self.handle, int(tile)))

    def autotune(self, reps=3, executor=None):
        t, ms = ctypes.c_int32(0), ctypes.c_double(0)
        check(lib().tprims_group_autotune(self.handle, self._ex(executor), int(reps), ctypes.byref(t),
                                          ctypes.byref(ms)))
        return t.value, ms.value

    def info(self):
        tiles, flops, tile = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int32(0)
        check(lib().tprims_group_info(self.handle, ctypes.byref(tiles), ctypes.byref(flops), ctypes.byref(tile)))
        return {"tiles": tiles.value, "flops": flops.value, "tile": tile.value}

    def close(self):
        if getattr(self, "handle", 0):
            lib().tprims_group_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
