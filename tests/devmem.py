"""Device buffers for GPU tests through the HIP runtime the library itself links (ctypes), so
tests that hand device pointers to the C ABI need no other GPU framework in the process."""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("HIP runtime library not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """Device copy of a numpy array; data_ptr() for the C ABI, numpy() copies it back."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), max(self.a.nbytes, 8)) != 0:
            raise MemoryError("hipMalloc failed")
        self._copy(self.ptr, self.a.ctypes.data_as(C.c_void_p), H2D)

    def _copy(self, dst, src, kind):
        if self.a.nbytes and hip().hipMemcpy(dst, src, self.a.nbytes, kind) != 0:
            raise RuntimeError("hipMemcpy failed")

    def data_ptr(self):
        return self.ptr.value

    def numpy(self):
        hip().hipDeviceSynchronize()
        out = np.empty_like(self.a)
        self._copy(out.ctypes.data_as(C.c_void_p), self.ptr, D2H)
        return out

    def __del__(self):
        if self.ptr and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
