"""Minimal HIP runtime access (device buffers, copies, sync) for Python
callers that hand the painter an HBM-resident catalogue."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


class DeviceArray:
    """A device copy of a numpy array (freed on close / garbage collection)."""

    def __init__(self, host, device=0):
        h = hip()
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = C.c_void_p()
        if h.hipSetDevice(device) != 0:
            raise RuntimeError("hipSetDevice failed")
        if h.hipMalloc(C.byref(self.ptr), self.nbytes) != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed")
        if h.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    @property
    def address(self):
        return self.ptr.value

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    if hip().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def h2d_copy_ms(dev, host_ptr, nbytes, reps=5, device=0):
    """Median wall time (ms) of a synchronous host-to-device copy of nbytes
    from host_ptr (pinned or pageable) into the device buffer `dev`."""
    import time
    h = hip()
    if h.hipSetDevice(device) != 0:
        raise RuntimeError("hipSetDevice failed")
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        if h.hipMemcpy(C.c_void_p(dev), C.c_void_p(host_ptr), nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))
