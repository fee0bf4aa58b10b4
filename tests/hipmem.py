"""Device buffers for GPU tests, from the HIP runtime libtpgpu.so itself
uses (libamdhip64.so.7, already loaded with the library, so the same
runtime instance and the same device memory space).  torch's bundled HIP
runtime is a second runtime in the process and does not see the GPU once
libtpgpu.so has initialised it, so device buffers for the C ABI are made
here."""
import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        from test_particle_amd.gpu import lib
        lib()                                   # libamdhip64.so.7 comes in with it
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """A device allocation holding a numpy array's bytes."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc: {rc}")
        self.upload()

    @property
    def address(self):
        return self.ptr.value

    def upload(self):
        rc = hip().hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p),
                             self.host.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D: {rc}")

    def download(self):
        out = np.empty_like(self.host)
        rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
