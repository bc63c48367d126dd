"""Borrowing torch tensors as DLPack operands for the C ABI
(tprims_tensor_borrow_raw over the DLTensor inside a DLPack capsule)."""
import ctypes

from . import _capi

_get = ctypes.pythonapi.PyCapsule_GetPointer
_get.restype = ctypes.c_void_p
_get.argtypes = [ctypes.py_object, ctypes.c_char_p]
_valid = ctypes.pythonapi.PyCapsule_IsValid
_valid.restype = ctypes.c_int
_valid.argtypes = [ctypes.py_object, ctypes.c_char_p]


def borrow(t, keep):
    """A tprims_tensor viewing torch tensor `t`. The capsule is kept alive in
    `keep` for the duration of the call; its deleter runs when it is dropped
    (the tensor's storage is never freed by the library)."""
    from torch.utils.dlpack import to_dlpack

    cap = to_dlpack(t)
    keep.append(cap)
    if _valid(cap, b"dltensor"):
        ptr = _get(cap, b"dltensor")
        # DLManagedTensor starts with its DLTensor.
        view = ctypes.cast(ptr, ctypes.POINTER(_capi.DLTensor))
        return _capi.lib().tprims_tensor_borrow_raw(view, 0)
    raise TypeError("not a DLPack capsule")
