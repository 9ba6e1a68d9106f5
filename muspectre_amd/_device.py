"""Device buffers owned from Python, allocated through the C-ABI."""
import ctypes

import numpy as np

from . import _capi


class DeviceBuffer:
    """A block of device memory with numpy upload/download helpers."""

    def __init__(self, nbytes):
        lib = _capi.load()
        _capi.check(lib.mugpu_init(0))
        self.nbytes = int(nbytes)
        ptr = ctypes.c_void_p()
        _capi.check(lib.mugpu_alloc(self.nbytes, ctypes.byref(ptr)))
        self.ptr = ptr

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array of {} bytes does not fit a {}-byte buffer"
                             .format(a.nbytes, self.nbytes))
        _capi.check(_capi.load().mugpu_upload(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                               a.nbytes))

    def download(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested {} bytes from a {}-byte buffer"
                             .format(out.nbytes, self.nbytes))
        _capi.check(_capi.load().mugpu_download(out.ctypes.data_as(ctypes.c_void_p),
                                                 self.ptr, out.nbytes))
        return out

    def zero(self):
        _capi.check(_capi.load().mugpu_memset_zero(self.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _capi.load().mugpu_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
