"""Device buffers for the Python host layer: HBM allocations made through the library's own
C ABI (hipMalloc underneath), with explicit host<->device copies."""
import ctypes

import numpy as np

from . import _lib

_DTYPES = {np.dtype(np.float64): _lib.F64, np.dtype(np.float32): _lib.F32}


class DeviceArray:
    """A dense matrix in device memory.  ``order`` is "C" (row-major) or "F" (column-major); the
    leading dimension is the row length (C) or the column length (F)."""

    def __init__(self, shape, dtype=np.float64, order="C"):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.order = order
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().mcnla_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        self._owner = True

    @classmethod
    def from_numpy(cls, arr, order=None):
        arr = np.asarray(arr)
        if order is None:
            order = "F" if (arr.flags["F_CONTIGUOUS"] and not arr.flags["C_CONTIGUOUS"]) else "C"
        host = np.asarray(arr, order=order)
        out = cls(host.shape, host.dtype, order)
        host = np.require(host, requirements=[order])
        if out.nbytes:
            _lib.check(_lib.lib().mcnla_memcpy_htod(out.ptr, host.ctypes.data, out.nbytes))
        return out

    @classmethod
    def view(cls, ptr, shape, dtype=np.float64, order="C", ld=None):
        """A non-owning view of device memory (stage results, or a block of a larger matrix with
        leading dimension ``ld``)."""
        out = cls.__new__(cls)
        out.shape = tuple(int(s) for s in shape)
        out.dtype = np.dtype(dtype)
        out.order = order
        out.nbytes = int(np.prod(out.shape)) * out.dtype.itemsize
        out.ptr = ptr
        out._owner = False
        if ld is not None:
            out._ld = int(ld)
        return out

    def numpy(self):
        host = np.empty(self.shape, dtype=self.dtype, order=self.order)
        if self.nbytes:
            _lib.check(_lib.lib().mcnla_memcpy_dtoh(host.ctypes.data, self.ptr, self.nbytes))
        return host

    @property
    def ld(self):
        if getattr(self, "_ld", None) is not None:
            return self._ld
        if len(self.shape) == 1:
            return self.shape[0]
        return self.shape[1] if self.order == "C" else self.shape[0]

    @property
    def dtype_code(self):
        return _DTYPES[self.dtype]

    @property
    def order_code(self):
        return _lib.ROW_MAJOR if self.order == "C" else _lib.COL_MAJOR

    def free(self):
        if getattr(self, "_owner", False) and self.ptr:
            _lib.lib().mcnla_free(self.ptr)
        self.ptr = None
        self._owner = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _lib.check(_lib.lib().mcnla_device_synchronize())


def device_count():
    n = ctypes.c_int(0)
    st = _lib.lib().mcnla_device_count(ctypes.byref(n))
    return n.value if st == _lib.OK else 0


def mem_info():
    free = ctypes.c_size_t()
    total = ctypes.c_size_t()
    _lib.check(_lib.lib().mcnla_mem_info(ctypes.byref(free), ctypes.byref(total)))
    return free.value, total.value
