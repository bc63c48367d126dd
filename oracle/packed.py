"""The C++ port of the reference's packed CPU driver (packed_cpu.cpp), used
as the CPU baseline beside the GPU figure: TEST AND BASELINE INFRASTRUCTURE.

The library is compiled for the host it runs on (-O3 -march=native, OpenMP)
the first time it is needed on that host.
"""
import ctypes
import hashlib
import os
import platform
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _cpu_tag():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") or line.startswith("flags"):
                    model += line
    except OSError:
        model = platform.processor()
    return hashlib.sha1(model.encode()).hexdigest()[:10]


_lib = None


def lib():
    global _lib
    if _lib is None:
        src = os.path.join(HERE, "packed_cpu.cpp")
        with open(src, "rb") as f:
            tag = hashlib.sha1(f.read()).hexdigest()[:8]
        out = os.path.join(HERE, "_build", "packed_%s_%s.so" % (_cpu_tag(), tag))
        if not os.path.exists(out):
            os.makedirs(os.path.dirname(out), exist_ok=True)
            # One temporary name per process: parallel test workers may compile at the same time.
            tmp = "%s.tmp%x" % (out, os.getpid())
            subprocess.check_call(["g++", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                                   "-o", tmp, src])
            os.replace(tmp, out)
        l = ctypes.CDLL(out)
        i64p = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        op = [ctypes.c_void_p, ctypes.c_int, i64p, i64p, i64p, ctypes.c_int]
        l.packed_contract.argtypes = [ctypes.c_int, dp] + op + op + [dp] + op + op + [ctypes.c_int]
        l.packed_contract.restype = ctypes.c_int
        l.packed_pack_panel.argtypes = [ctypes.c_int, ctypes.c_void_p, i64p, ctypes.c_int64, i64p, ctypes.c_int64,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        l.packed_pack_panel.restype = ctypes.c_int
        _lib = l
    return _lib


DT = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


def _arr(xs):
    return (ctypes.c_int64 * max(1, len(xs)))(*[int(x) for x in xs])


def contract(alpha, a, b, beta, c, d, threads=None):
    """Packed-driver contraction on oracle.Operand-like objects whose `data`
    arrays share one dtype; the origin is data[offset]. Writes d.data in place."""
    threads = threads or (os.cpu_count() or 1)
    dt = DT[np.asarray(d.data).dtype]
    dp = ctypes.POINTER(ctypes.c_double)
    keep = []

    def op(o):
        if o is None:
            return [None, 0, None, None, None, 0]
        arr = o.data
        keep.append(arr)
        ptr = arr.ctypes.data + o.offset * arr.itemsize
        return [ptr, len(o.dims), _arr(o.dims), _arr(o.strides), _arr(o.labels), int(o.conj)]

    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    rc = lib().packed_contract(dt, al.ctypes.data_as(dp), *op(a), *op(b), be.ctypes.data_as(dp), *op(c), *op(d),
                               int(threads))
    if rc:
        raise ValueError("packed_contract refused the dtype")


def pack_panel(data, vscat, kscat, vr, conj=False):
    """The port's packed panel of `data` (a 1-D array) as reals: Real format
    for real dtypes, Planar for complex (pack.rs:57)."""
    data = np.ascontiguousarray(data)
    dt = DT[data.dtype]
    nsl = (len(vscat) + vr - 1) // vr
    real = np.float32 if data.dtype in (np.float32, np.complex64) else np.float64
    out = np.zeros(nsl * len(kscat) * vr * (2 if np.iscomplexobj(data) else 1), real)
    rc = lib().packed_pack_panel(dt, data.ctypes.data, _arr(vscat), len(vscat), _arr(kscat), len(kscat), int(vr),
                                 int(conj), out.ctypes.data)
    if rc:
        raise ValueError("packed_pack_panel refused the dtype")
    return out


class _Op:
    def __init__(self, data, dims, strides, labels, offset=0, conj=False):
        self.data, self.dims, self.strides, self.labels = data, dims, strides, labels
        self.offset, self.conj = offset, conj


def physical_cores():
    try:
        out = subprocess.check_output(["lscpu", "-p=Core,Socket"]).decode()
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        return len(cores)
    except Exception:
        return os.cpu_count() or 1


def bench_complex_sample(dtype="c64", seconds=10.0, threads=None):
    """Times the benchmark's complex contraction on one batch entry (n = 1 of
    16): i = k = 512, j = l = 64, conj(A) with A a stride-2 view, beta 0.5.
    Returns the cpu_baseline record (TFLOP/s for the whole workload's rate)."""
    threads = threads or physical_cores()
    i = k = 512
    j = l = 64
    cdt = np.complex128 if dtype == "c64" else np.complex64
    rng = np.random.default_rng(0)
    abuf = (rng.standard_normal(2 * j * i * l) + 1j * rng.standard_normal(2 * j * i * l)).astype(cdt)
    bbuf = (rng.standard_normal(l * k * j) + 1j * rng.standard_normal(l * k * j)).astype(cdt)
    cbuf = (rng.standard_normal(i * k) + 1j * rng.standard_normal(i * k)).astype(cdt)
    dbuf = np.zeros(i * k, cdt)
    J, I, L, K = 1, 2, 3, 5
    A = _Op(abuf, [j, i, l], [2, 2 * j, 2 * j * i], [J, I, L], conj=True)
    B = _Op(bbuf, [l, k, j], [1, l, l * k], [L, K, J])
    C = _Op(cbuf, [i, k], [1, i], [I, K])
    D = _Op(dbuf, [i, k], [1, i], [I, K])
    flops = 8 * i * k * j * l
    contract(1.0, A, B, 0.5, C, D, threads)  # warm-up
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        contract(1.0, A, B, 0.5, C, D, threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 50:
            break
    med = float(np.median(times))
    return {"value": round(flops / med / 1e12, 5), "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": "one batch entry (n=1 of 16) of the %s workload, %d runs, median %.1f ms" %
                      (dtype, len(times), med * 1e3)}
