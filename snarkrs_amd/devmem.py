"""Device buffers through the HIP runtime's own C API (hipMalloc / hipMemcpy), for callers
that hand device pointers to the ``*_device`` entry points (resident witnesses, resident MSM
inputs) without a framework's allocator."""
import ctypes

import numpy as np

_hip = None


def _lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def synchronize() -> None:
    rc = _lib().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


class DeviceArray:
    H2D, D2H = 1, 2

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        if _lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")
        if self.nbytes and _lib().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, self.H2D) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    @property
    def addr(self) -> int:
        return self.ptr.value

    def to_host(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        if _lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, self.D2H) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        """Device-to-device copy of min(sizes) bytes (the runtime's copy kernel)."""
        n = min(self.nbytes, other.nbytes)
        if n and _lib().hipMemcpy(self.ptr, other.ptr, n, 3) != 0:
            raise RuntimeError("hipMemcpy device to device failed")
        synchronize()

    def free(self):
        if self.ptr:
            _lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()
