"""CPU tests of the DLPack operand contract (tprims_tensor_dtype,
tprims_tensor_layout, and the checks tprims_execute_product_dlpack applies
before any device work), one per test of the reference's
crates/tprims-capi/tests/dlpack.rs:32-213, with the same tensors and the same
expected statuses; plus the ROCm device type this build accepts."""
import ctypes
import subprocess

import numpy as np
import pytest

import tprims_amd as tp
from tprims_amd import _capi
from tprims_amd._capi import DLTensor, TensorLayoutInfoC, intptr, lib

kDLCPU, kDLCUDA, kDLROCM = 1, 2, 10
kDLFloat, kDLComplex = 2, 5
READ_ONLY_FLAG = 1  # DLPACK_FLAG_BITMASK_READ_ONLY
I64_MAX = (1 << 63) - 1


class View:
    """A DLTensor over a numpy buffer, kept alive with its shape/stride arrays."""

    def __init__(self, data, shape, strides=None, byte_offset=0, code=kDLFloat, bits=64):
        self.buf = data
        self.shape = (ctypes.c_int64 * max(1, len(shape)))(*shape)
        self.strides = None if strides is None else (ctypes.c_int64 * max(1, len(strides)))(*strides)
        t = DLTensor()
        t.data = None if data is None else data.ctypes.data
        t.device.device_type = kDLCPU
        t.device.device_id = 0
        t.ndim = len(shape)
        t.dtype.code, t.dtype.bits, t.dtype.lanes = code, bits, 1
        t.shape = ctypes.cast(self.shape, _capi.i64p)
        t.strides = ctypes.cast(self.strides, _capi.i64p) if strides is not None else None
        t.byte_offset = byte_offset
        self.t = t

    def borrow(self, flags=0):
        return lib().tprims_tensor_borrow_raw(ctypes.byref(self.t), flags)


def layout(op, dtype, writable=False):
    """(status, dims, strides, info) of tprims_tensor_layout."""
    dims = (ctypes.c_int64 * 8)()
    strides = (ctypes.c_int64 * 8)()
    info = TensorLayoutInfoC()
    st = lib().tprims_tensor_layout(op, int(dtype), int(writable), dims, strides, 8, ctypes.byref(info))
    n = info.ndim if st == 0 else 0
    return st, list(dims[:n]), list(strides[:n]), info


def dtype_of(op):
    out = ctypes.c_int(-1)
    st = lib().tprims_tensor_dtype(op, ctypes.byref(out))
    return st, out.value


def element(info, strides, idx, buf):
    """Value at a logical index, read through the validated origin and strides."""
    off = sum(i * s for i, s in zip(idx, strides))
    return np.frombuffer((ctypes.c_char * 8).from_address(info.origin + 8 * off), dtype=np.float64)[0]


def test_null_strides_mean_row_major_and_views_are_zero_copy():
    # dlpack.rs:32-45
    d = np.arange(6, dtype=np.float64)
    v = View(d, [2, 3])
    op = v.borrow()
    assert dtype_of(op) == (0, tp.DType.F64)
    st, dims, strides, info = layout(op, tp.DType.F64)
    assert st == 0
    assert strides == [3, 1]
    assert element(info, strides, [1, 2], d) == 5.0
    assert info.origin == d.ctypes.data, "zero copy: same address"


def test_column_major_negative_strides_and_byte_offset():
    # dlpack.rs:47-61: reversed rows of a column-major 2x3 at offset 5 elements
    d = np.arange(12, dtype=np.float64)
    v = View(d, [2, 3], [-1, 2], byte_offset=5 * 8)
    st, dims, strides, info = layout(v.borrow(), tp.DType.F64)
    assert st == 0
    assert element(info, strides, [0, 0], d) == 5.0
    assert element(info, strides, [1, 0], d) == 4.0
    assert element(info, strides, [1, 2], d) == 8.0
    assert info.origin == d.ctypes.data + 5 * 8
    assert (info.span_lo, info.span_hi) == (-1, 4)


def test_rejections_are_typed():
    # dlpack.rs:63-125
    d = np.zeros(4, dtype=np.float64)
    v = View(d, [2, 2])
    # read-only output
    assert layout(v.borrow(READ_ONLY_FLAG), tp.DType.F64)[0] == 0
    assert layout(v.borrow(READ_ONLY_FLAG), tp.DType.F64, writable=True)[0] == _capi.TPRIMS_ERR_READ_ONLY
    # dtype mismatch, lanes, device, misaligned offset
    assert layout(v.borrow(), tp.DType.F32)[0] == _capi.TPRIMS_ERR_DTYPE
    v.t.dtype.lanes = 2
    assert dtype_of(v.borrow())[0] == _capi.TPRIMS_ERR_DTYPE
    v.t.dtype.lanes = 1
    v.t.device.device_type = kDLCUDA
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_DEVICE
    v.t.device.device_type = kDLCPU
    v.t.byte_offset = 3
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    v.t.byte_offset = 0
    # null view, null shape
    null = lib().tprims_tensor_borrow_raw(None, 0)
    assert dtype_of(null)[0] == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert layout(null, tp.DType.F64)[0] == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    v.t.shape = None
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert "null shape" in lib().tprims_last_error().decode()


def test_unsupported_types_and_negative_ndim():
    # tensor.rs:103-114 (code/bits table) and :154-155 (negative ndim)
    d = np.zeros(4, dtype=np.float64)
    for code, bits in ((kDLFloat, 16), (0, 32), (kDLComplex, 32), (4, 16)):
        v = View(d, [2, 2], code=code, bits=bits)
        assert dtype_of(v.borrow())[0] == _capi.TPRIMS_ERR_DTYPE
    for code, bits, want in ((kDLFloat, 32, tp.DType.F32), (kDLComplex, 64, tp.DType.C32),
                             (kDLComplex, 128, tp.DType.C64)):
        assert dtype_of(View(d, [2, 2], code=code, bits=bits).borrow()) == (0, want)
    v = View(d, [2, 2])
    v.t.ndim = -1
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_SHAPE


def test_overflowing_spans_are_shape_errors():
    # dlpack.rs:127-148
    d = np.zeros(1, dtype=np.float64)
    v = View(d, [3, 3], [I64_MAX // 2, 1])
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_SHAPE
    v = View(d, [-1])
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_SHAPE
    # a span that fits in elements but not in bytes
    v = View(d, [2], [I64_MAX // 8])
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_SHAPE
    # compact row-major strides that overflow
    v = View(d, [1 << 40, 1 << 40])
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_SHAPE


def test_empty_tensors_never_touch_data():
    # dlpack.rs:150-175
    v = View(None, [0, 5], [1 << 40, 7])
    st, dims, strides, info = layout(v.borrow(), tp.DType.F64)
    assert st == 0 and info.empty == 1
    assert dims == [0, 5]
    assert info.origin  # a placeholder, never dereferenced


def test_null_data_of_a_non_empty_tensor_is_rejected():
    # tensor.rs:203-208
    v = View(None, [2, 2])
    assert layout(v.borrow(), tp.DType.F64)[0] == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert "null data" in lib().tprims_last_error().decode()


def test_c_mirror_matches_the_dlpack_layout(tmp_path):
    # dlpack.rs:177-183: the struct sizes and offsets a C consumer compiles against
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stddef.h>\n#include "tprims/tprims.h"\n'
        "_Static_assert(sizeof(DLTensor) == 48, \"DLTensor\");\n"
        "_Static_assert(offsetof(DLManagedTensorVersioned, dl_tensor) == 32, \"dl_tensor\");\n"
        "_Static_assert(sizeof(tprims_tensor) == 16, \"tprims_tensor\");\n"
        "int main(void) { return 0; }\n")
    subprocess.check_call(["gcc", "-std=c99", "-c", "-I", os.path.join(root, "include"), str(src), "-o",
                           str(tmp_path / "sizes.o")])
    assert ctypes.sizeof(DLTensor) == 48 and ctypes.sizeof(_capi.TprimsTensor) == 16


@pytest.mark.parametrize("bits,dtype,off,ok", [(128, tp.DType.C64, 8, True), (64, tp.DType.C32, 4, True),
                                               (128, tp.DType.C64, 4, False), (64, tp.DType.C32, 2, False)])
def test_complex_byte_offsets_need_element_alignment_not_element_size(bits, dtype, off, ok):
    # dlpack.rs:185-213
    d = np.zeros(4, dtype=np.float64)
    v = View(d, [1, 1], byte_offset=off, code=kDLComplex, bits=bits)
    st, _, _, info = layout(v.borrow(), dtype)
    if ok:
        assert st == 0, lib().tprims_last_error()
        assert info.origin == d.ctypes.data + off
    else:
        assert st == _capi.TPRIMS_ERR_INVALID_ARGUMENT


def test_rocm_device_memory_is_accepted():
    """This build's extension: kDLROCM views pass the field checks (their
    device is compared with the executor's at execution)."""
    d = np.zeros(4, dtype=np.float64)
    v = View(d, [2, 2])
    v.t.device.device_type = kDLROCM
    assert layout(v.borrow(), tp.DType.F64)[0] == 0


def _plan_2x2():
    a = tp.OperandSpec(tp.LayoutSpec([2, 2], [2, 1]))
    prob = tp.Problem.from_labels(tp.DType.F64, a, a, tp.CSpec.separate(a), a, tp.Labels([0, 2], [2, 1], [0, 1], [0, 1]))
    return tp.Plan(prob)


def _exec(plan, A, B, C, D):
    one = ctypes.c_double(1.0)
    zero = ctypes.c_double(0.0)
    return lib().tprims_execute_product_dlpack(plan.handle, 0, ctypes.byref(one), A, B, ctypes.byref(zero), C, D)


def test_execute_dlpack_rejects_before_any_device_work():
    """Every field error, a layout different from the plan's and a read-only D
    are reported by tprims_execute_product_dlpack without touching the GPU
    runtime (this container has none)."""
    plan = _plan_2x2()
    bufs = [np.zeros(4) for _ in range(4)]
    views = [View(b, [2, 2]) for b in bufs]  # compact row-major = the plan's strides [2, 1]
    null = lib().tprims_tensor_borrow_raw(None, 0)
    A, B, D = views[0].borrow(), views[1].borrow(), views[3].borrow()
    assert _exec(plan, null, B, null, D) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    assert "A:" in lib().tprims_last_error().decode()
    assert _exec(plan, A, B, null, views[3].borrow(READ_ONLY_FLAG)) == _capi.TPRIMS_ERR_READ_ONLY
    wrong = View(bufs[1], [2, 2], [1, 2])  # column-major: not the plan's layout
    assert _exec(plan, A, wrong.borrow(), null, D) == _capi.TPRIMS_ERR_SHAPE
    f32 = View(bufs[1], [2, 2], code=kDLFloat, bits=32)
    assert _exec(plan, A, f32.borrow(), null, D) == _capi.TPRIMS_ERR_DTYPE
    noshape = View(bufs[1], [2, 2])
    noshape.t.shape = None
    assert _exec(plan, A, noshape.borrow(), null, D) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    nodata = View(None, [2, 2])
    assert _exec(plan, A, B, null, nodata.borrow()) == _capi.TPRIMS_ERR_INVALID_ARGUMENT
    # D overlapping A is refused by the alias preflight, still before any device work
    assert _exec(plan, A, B, null, views[0].borrow()) == _capi.TPRIMS_ERR_ALIASED
