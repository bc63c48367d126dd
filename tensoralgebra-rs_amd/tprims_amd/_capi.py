"""ctypes binding of libtprims.so (the HIP contraction library) and its C ABI.

The library is loaded from this directory. A missing library is an error:
there is no other implementation behind these names.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtprims.so")

# TAPP datatypes / precisions / element ops (include/tapp.h)
TAPP_F32, TAPP_F64, TAPP_C32, TAPP_C64, TAPP_F16, TAPP_BF16 = 0, 1, 2, 3, 4, 5
TAPP_DEFAULT_PREC = -1
TAPP_F32F32_ACCUM_F32 = 0
TAPP_F64F64_ACCUM_F64 = 1
TAPP_BF16BF16_ACCUM_F32 = 6
TAPP_IDENTITY, TAPP_CONJUGATE = 0, 1

# tprims statuses (include/tprims/core.h)
TPRIMS_OK = 0
TPRIMS_ERR_INVALID_ARGUMENT = 1
TPRIMS_ERR_SHAPE = 2
TPRIMS_ERR_DTYPE = 3
TPRIMS_ERR_DEVICE = 4
TPRIMS_ERR_READ_ONLY = 5
TPRIMS_ERR_ALIASED = 6
TPRIMS_ERR_WOULD_MATERIALIZE = 7
TPRIMS_ERR_NUMERICAL = 8
TPRIMS_BUSY = 9
TPRIMS_ERR_WOULD_DEADLOCK = 10
TPRIMS_ERR_PANIC = 12
TPRIMS_ERR_INTERNAL = 13
TPRIMS_ERR_LABELS = 14
TPRIMS_ERR_UNSUPPORTED = 15

# plan configuration (include/tprims/hip.h)
TPRIMS_CPLX_NATIVE, TPRIMS_CPLX_1M, TPRIMS_CPLX_3M = 0, 1, 2
TPRIMS_PREC_STORAGE, TPRIMS_PREC_SPLIT_BF16_2, TPRIMS_PREC_SPLIT_BF16_3, TPRIMS_PREC_FP8_E4M3 = 0, 1, 2, 3

intptr = ctypes.c_ssize_t
i64p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p


class PlanConfigC(ctypes.Structure):
    _fields_ = [("complex_method", ctypes.c_int32), ("precision", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("split_k", ctypes.c_int32), ("reserved", ctypes.c_int32 * 4)]


class PlanReportC(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64), ("batch", ctypes.c_int64),
                ("flops", ctypes.c_int64), ("dtype", ctypes.c_int32), ("complex_method", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("tile", ctypes.c_int32), ("tile_m", ctypes.c_int32),
                ("tile_n", ctypes.c_int32), ("tile_k", ctypes.c_int32), ("a_k_major", ctypes.c_int32),
                ("b_k_major", ctypes.c_int32), ("d_m_major", ctypes.c_int32), ("m_axes", ctypes.c_int32),
                ("n_axes", ctypes.c_int32), ("k_axes", ctypes.c_int32), ("h_axes", ctypes.c_int32),
                ("strategy", ctypes.c_int32), ("source", ctypes.c_int32), ("kernel", ctypes.c_char * 64)]


class DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class DLTensor(ctypes.Structure):
    _fields_ = [("data", vp), ("device", DLDevice), ("ndim", ctypes.c_int32), ("dtype", DLDataType),
                ("shape", i64p), ("strides", i64p), ("byte_offset", ctypes.c_uint64)]


class TprimsTensor(ctypes.Structure):
    _fields_ = [("view", ctypes.POINTER(DLTensor)), ("flags", ctypes.c_uint64)]


class TensorLayoutInfoC(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("empty", ctypes.c_int32), ("origin", vp), ("span_lo", ctypes.c_int64),
                ("span_hi", ctypes.c_int64)]


# Every symbol the public headers declare, in header order (checked by the tests).
EXPORTED = [
    "TAPP_check_success", "TAPP_explain_error", "TAPP_attr_set", "TAPP_attr_get", "TAPP_attr_clear",
    "TAPP_create_handle", "TAPP_destroy_handle", "TAPP_create_executor", "TAPP_destroy_executor",
    "TAPP_destroy_status", "TAPP_create_tensor_info", "TAPP_destroy_tensor_info", "TAPP_get_nmodes",
    "TAPP_set_nmodes", "TAPP_get_extents", "TAPP_set_extents", "TAPP_get_strides", "TAPP_set_strides",
    "TAPP_create_tensor_product", "TAPP_destroy_tensor_product", "TAPP_execute_product",
    "TAPP_execute_batched_product",
    "tprims_tensor_borrow_versioned", "tprims_tensor_borrow_raw", "tprims_tensor_dtype", "tprims_tensor_layout",
    "tprims_last_error", "tprims_abi_version",
    "tprims_has_part", "TAPP_implementation_name", "TAPP_implementation_version",
    "tprims_tapp_executor_create_rayon", "tprims_tapp_executor_set_budget", "tprims_tapp_executor_get_threads",
    "tprims_tapp_executor_create_hip", "tprims_tapp_executor_set_async", "tprims_tapp_executor_synchronize",
    "tprims_tapp_executor_get_stream", "tprims_tapp_executor_get_device", "tprims_create_tensor_product_ex",
    "tprims_product_report", "tprims_execute_product_noalloc", "tprims_execute_product_dlpack", "tprims_group_create", "tprims_group_execute",
    "tprims_group_use_graph", "tprims_group_execute_each", "tprims_group_info", "tprims_group_set_tile",
    "tprims_group_autotune", "tprims_group_destroy", "tprims_product_vjp", "tprims_product_vjp_report",
    "tprims_product_autotune", "tprims_product_tuning_key", "tprims_tuning_cache_save", "tprims_tuning_cache_load",
    "tprims_tuning_cache_clear", "tprims_tuning_cache_size",
]


def _declare(lib):
    c_int = ctypes.c_int
    sig = {
        "TAPP_check_success": (ctypes.c_bool, [c_int]),
        "TAPP_explain_error": (ctypes.c_size_t, [c_int, ctypes.c_size_t, ctypes.c_char_p]),
        "TAPP_create_handle": (c_int, [ctypes.POINTER(intptr)]),
        "TAPP_destroy_handle": (c_int, [intptr]),
        "TAPP_attr_set": (c_int, [intptr, c_int, vp]),
        "TAPP_attr_get": (c_int, [intptr, c_int, ctypes.POINTER(vp)]),
        "TAPP_attr_clear": (c_int, [intptr, c_int]),
        "TAPP_create_executor": (c_int, [ctypes.POINTER(intptr)]),
        "TAPP_destroy_executor": (c_int, [intptr]),
        "TAPP_destroy_status": (c_int, [intptr]),
        "TAPP_create_tensor_info": (c_int, [ctypes.POINTER(intptr), c_int, c_int, i64p, i64p]),
        "TAPP_destroy_tensor_info": (c_int, [intptr]),
        "TAPP_get_nmodes": (c_int, [intptr]),
        "TAPP_set_nmodes": (c_int, [intptr, c_int]),
        "TAPP_get_extents": (None, [intptr, i64p]),
        "TAPP_set_extents": (c_int, [intptr, i64p]),
        "TAPP_get_strides": (None, [intptr, i64p]),
        "TAPP_set_strides": (c_int, [intptr, i64p]),
        "TAPP_create_tensor_product": (c_int, [ctypes.POINTER(intptr), intptr] + [c_int, intptr, i64p] * 4 + [c_int]),
        "tprims_create_tensor_product_ex": (c_int, [ctypes.POINTER(intptr), intptr] + [c_int, intptr, i64p] * 4
                                            + [c_int, ctypes.POINTER(PlanConfigC)]),
        "TAPP_destroy_tensor_product": (c_int, [intptr]),
        "TAPP_execute_product": (c_int, [intptr, intptr, ctypes.POINTER(intptr), vp, vp, vp, vp, vp, vp]),
        "tprims_execute_product_noalloc": (c_int, [intptr, intptr, ctypes.POINTER(intptr), vp, vp, vp, vp, vp, vp]),
        "TAPP_execute_batched_product": (c_int, [intptr, intptr, ctypes.POINTER(intptr), c_int, vp,
                                                 ctypes.POINTER(vp), ctypes.POINTER(vp), vp, ctypes.POINTER(vp),
                                                 ctypes.POINTER(vp)]),
        "tprims_last_error": (ctypes.c_char_p, []),
        "tprims_abi_version": (ctypes.c_uint32, []),
        "tprims_has_part": (ctypes.c_int32, [ctypes.c_char_p]),
        "TAPP_implementation_name": (ctypes.c_char_p, []),
        "TAPP_implementation_version": (ctypes.c_char_p, []),
        "tprims_tapp_executor_create_rayon": (c_int, [ctypes.POINTER(intptr), ctypes.c_size_t, vp]),
        "tprims_tapp_executor_set_budget": (c_int, [intptr, ctypes.c_size_t]),
        "tprims_tapp_executor_get_threads": (c_int, [intptr, ctypes.POINTER(ctypes.c_size_t),
                                                     ctypes.POINTER(ctypes.c_size_t)]),
        "tprims_tapp_executor_create_hip": (c_int, [ctypes.POINTER(intptr), c_int, vp]),
        "tprims_tapp_executor_set_async": (c_int, [intptr, c_int]),
        "tprims_tapp_executor_synchronize": (c_int, [intptr]),
        "tprims_tapp_executor_get_stream": (c_int, [intptr, ctypes.POINTER(vp)]),
        "tprims_tapp_executor_get_device": (c_int, [intptr, ctypes.POINTER(c_int)]),
        "tprims_product_report": (c_int, [intptr, ctypes.POINTER(PlanReportC)]),
        "tprims_build_scatter": (c_int, [i64p, i64p, ctypes.c_int32, i64p, ctypes.c_int64, i64p]),
        "tprims_build_block_scatter": (c_int, [i64p, ctypes.c_int64, ctypes.c_int64, i64p, ctypes.c_int64, i64p]),
        "tprims_product_scatter": (c_int, [intptr, ctypes.c_int32, i64p, ctypes.c_int64, i64p]),
        "tprims_tensor_borrow_raw": (TprimsTensor, [ctypes.POINTER(DLTensor), ctypes.c_uint64]),
        "tprims_tensor_dtype": (c_int, [TprimsTensor, ctypes.POINTER(c_int)]),
        "tprims_tensor_layout": (c_int, [TprimsTensor, c_int, c_int, i64p, i64p, ctypes.c_int32,
                                         ctypes.POINTER(TensorLayoutInfoC)]),
        "tprims_execute_product_dlpack": (c_int, [intptr, intptr, vp, TprimsTensor, TprimsTensor, vp, TprimsTensor,
                                                  TprimsTensor]),
        "tprims_group_create": (c_int, [ctypes.POINTER(intptr), intptr, c_int, ctypes.POINTER(intptr), vp,
                                        ctypes.POINTER(vp), ctypes.POINTER(vp), vp, ctypes.POINTER(vp),
                                        ctypes.POINTER(vp)]),
        "tprims_group_execute": (c_int, [intptr, intptr]),
        "tprims_group_use_graph": (c_int, [intptr, intptr, c_int]),
        "tprims_group_execute_each": (c_int, [intptr, intptr]),
        "tprims_group_info": (c_int, [intptr, i64p, i64p, ctypes.POINTER(ctypes.c_int32)]),
        "tprims_group_destroy": (c_int, [intptr]),
        "tprims_product_vjp": (c_int, [intptr, intptr, vp, vp, vp, vp, vp, vp, c_int]),
        "tprims_product_vjp_report": (c_int, [intptr, c_int, ctypes.POINTER(PlanReportC)]),
        "tprims_group_set_tile": (c_int, [intptr, ctypes.c_int32]),
        "tprims_group_autotune": (c_int, [intptr, intptr, c_int, ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_double)]),
        "tprims_product_autotune": (c_int, [intptr, intptr, vp, vp, vp, vp, vp, vp, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_double)]),
        "tprims_product_tuning_key": (c_int, [intptr, ctypes.c_char_p, ctypes.c_int64, i64p]),
        "tprims_tuning_cache_save": (c_int, [ctypes.c_char_p]),
        "tprims_tuning_cache_load": (c_int, [ctypes.c_char_p]),
        "tprims_tuning_cache_clear": (c_int, []),
        "tprims_tuning_cache_size": (c_int, [i64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


_lib = None


def lib():
    """The loaded library (raises OSError if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libtprims.so is not built at %s (run __graft_entry__.build())" % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        _declare(l)
        _lib = l
    return _lib


class TprimsError(RuntimeError):
    """A non-zero status from the C ABI, with tprims_last_error's message."""

    def __init__(self, status, message):
        super().__init__("status %d: %s" % (status, message))
        self.status = status
        self.message = message


def check(status):
    if status != 0:
        msg = lib().tprims_last_error()
        raise TprimsError(status, msg.decode() if msg else "")
    return status


def i64_array(xs):
    xs = [int(x) for x in xs]
    return (ctypes.c_int64 * max(1, len(xs)))(*xs)
