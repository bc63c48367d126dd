"""The CPU oracle of the contraction path: TEST INFRASTRUCTURE ONLY.

Only `tests/`, `__graft_entry__.smoke()` and `bench.py`'s CPU-baseline leg
import this package, and only as the checker (or, for the baseline, as the
timed CPU restatement). The HIP product path never imports it.

Contents
- `contract()`: the FP64 label oracle (oracle.c), a restatement of
  `contract_reference` (crates/tprims-testkit/src/oracle.rs:92-204).
- `max_rel_err()`: `max_rel_err` (crates/tprims-testkit/src/fixtures.rs:81-87).
- `seq()`: the conformance suite's xorshift input sequence
  (crates/tprims-capi/tests/common/mod.rs:72-93), so TAPP cases replay the
  reference suite's exact inputs.
- `einsum_reference()`: an independent numpy restatement used to cross-check
  `contract()` (numpy.einsum handles diagonals and reductions).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    """Compile oracle.c into oracle/_build (the packed CPU driver, oracle/packed.py,
    compiles itself on first use for the host CPU)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        op = [dp, ctypes.c_int, i64p, i64p, ctypes.c_int64, i64p, ctypes.c_int]
        lib.oracle_contract.argtypes = [ctypes.c_int, dp] + op + op + [dp] + op + op
        lib.oracle_contract.restype = ctypes.c_int
        lib.oracle_max_rel_err.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int64]
        lib.oracle_max_rel_err.restype = ctypes.c_double
        _lib = lib
    return _lib


def _i64(xs):
    arr = (ctypes.c_int64 * max(1, len(xs)))(*[int(x) for x in xs])
    return arr


def _f64(a, cplx):
    """A flat float64 view of storage `a` (interleaved for complex)."""
    a = np.ascontiguousarray(a)
    if cplx:
        a = a.astype(np.complex128, copy=False)
        return np.ascontiguousarray(a).view(np.float64)
    return np.ascontiguousarray(a.astype(np.float64, copy=False))


class Operand:
    """Flat storage plus the layout the oracle reads it through."""

    def __init__(self, data, dims, strides, labels, offset=0, conj=False):
        self.data = np.asarray(data).reshape(-1)
        self.dims = [int(d) for d in dims]
        self.strides = [int(s) for s in strides]
        self.labels = [int(l) for l in labels]
        self.offset = int(offset)
        self.conj = bool(conj)


def contract(alpha, a, b, beta, c, d, cplx=None):
    """D = op_D(alpha * sum op_A(A) op_B(B) + beta * op_C(C)) in FP64.

    `a`, `b`, `c` (or None), `d` are `Operand`s. Returns a new flat array:
    d.data (as complex128 or float64) with every addressed element written.
    """
    lib = _load()
    if cplx is None:
        cplx = any(np.iscomplexobj(o.data) for o in (a, b, d) if o is not None)
    w = 2 if cplx else 1
    dp = ctypes.POINTER(ctypes.c_double)
    bufs = []

    def op_args(o, writable=False):
        if o is None:
            return [None, 0, None, None, 0, None, 0]
        arr = _f64(o.data, cplx).copy() if writable else _f64(o.data, cplx)
        bufs.append(arr)
        return [arr.ctypes.data_as(dp), len(o.dims), _i64(o.dims), _i64(o.strides), o.offset, _i64(o.labels),
                int(o.conj)]

    al = np.array([complex(alpha).real, complex(alpha).imag], dtype=np.float64)
    be = np.array([complex(beta).real, complex(beta).imag], dtype=np.float64)
    aa, ba = op_args(a), op_args(b)
    ca = op_args(c)
    da = op_args(d, writable=True)
    dbuf = bufs[-1]
    rc = lib.oracle_contract(int(cplx), al.ctypes.data_as(dp), *aa, *ba, be.ctypes.data_as(dp), *ca, *da)
    if rc != 0:
        raise ValueError("oracle refused the labels (status %d)" % rc)
    out = dbuf.view(np.complex128) if cplx else dbuf
    return out.copy()


def max_rel_err(got, want):
    """Normwise relative error: max |got-want| / max(1, max |want|)."""
    got = np.asarray(got).reshape(-1)
    want = np.asarray(want).reshape(-1)
    cplx = np.iscomplexobj(got) or np.iscomplexobj(want)
    g, w = _f64(got, cplx), _f64(want, cplx)
    lib = _load()
    dp = ctypes.POINTER(ctypes.c_double)
    return float(lib.oracle_max_rel_err(int(cplx), g.ctypes.data_as(dp), w.ctypes.data_as(dp), want.size))


def seq(n, seed, cplx=False):
    """The conformance suite's deterministic values: multiples of 1/2048 in [-1, 1)."""
    mask = (1 << 64) - 1
    s = ((seed * 0x9E3779B97F4A7C15) & mask) | 1

    def nxt():
        nonlocal s
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        return (s >> 52) / 2048.0 - 1.0

    out = []
    for _ in range(n):
        re = nxt()
        im = nxt() if cplx else 0.0
        out.append(complex(re, im) if cplx else re)
    return np.array(out, dtype=np.complex128 if cplx else np.float64)


def einsum_reference(alpha, a, b, beta, c, d):
    """numpy restatement of the label semantics, for cross-checking `contract`.

    Handles repeated labels (diagonals) in inputs and output and reductions.
    """
    def view(o):
        base = o.data.astype(np.complex128)
        idx = np.zeros(o.dims, dtype=np.int64) + o.offset
        for ax, (n, s) in enumerate(zip(o.dims, o.strides)):
            shape = [1] * len(o.dims)
            shape[ax] = n
            idx = idx + (np.arange(n, dtype=np.int64) * s).reshape(shape)
        v = base[idx] if len(o.dims) else base[np.array(o.offset)]
        return np.conj(v) if o.conj else v

    letters = {}

    def sub(labels):
        return "".join(letters.setdefault(l, chr(ord("a") + len(letters))) for l in labels)

    sa, sb = sub(a.labels), sub(b.labels)
    # Output: one letter per distinct D label (first occurrence order).
    dl = []
    for l in d.labels:
        if l not in dl:
            dl.append(l)
    sd = sub(dl)
    prod = np.einsum("%s,%s->%s" % (sa, sb, sd), view(a), view(b))
    res = complex(alpha) * prod
    if c is not None and complex(beta) != 0:
        # C with possibly repeated labels: take its diagonal onto the distinct labels.
        cv = view(c)
        sc = sub(c.labels)
        res = res + complex(beta) * np.einsum("%s->%s" % (sc, sd), cv)
    if d.conj:
        res = np.conj(res)
    out = d.data.astype(np.complex128).copy()
    # Scatter onto D's addresses (a repeated D label writes the diagonal).
    pos = {l: i for i, l in enumerate(dl)}
    it = np.ndindex(*[d.dims[d.labels.index(l)] for l in dl]) if dl else [()]
    for tup in it:
        off = d.offset + sum(tup[pos[l]] * s for l, s in zip(d.labels, d.strides))
        out[off] = res[tup] if dl else res
    return out
