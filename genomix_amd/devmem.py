"""Device buffers through the HIP runtime that libgenomix_gx.so links
(/opt/rocm's libamdhip64), for callers that hand the library device-resident
input.  A second HIP runtime in the process (for instance the one PyTorch
bundles) is not needed for that."""

import ctypes

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    return _hip


class DeviceBytes:
    """`data` copied into a new device allocation on the current device."""

    def __init__(self, data: bytes):
        self.ptr = ctypes.c_void_p()
        self.n = len(data)
        if hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(1, self.n))) != 0:
            raise MemoryError("hipMalloc failed")
        if self.n and hip().hipMemcpy(self.ptr, data, ctypes.c_size_t(self.n), 1) != 0:  # host -> device
            self.free()
            raise RuntimeError("hipMemcpy failed")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_host(address, n):
    """`n` bytes at device address `address`, copied to host memory."""
    buf = ctypes.create_string_buffer(max(1, n))
    if n and hip().hipMemcpy(buf, ctypes.c_void_p(address), ctypes.c_size_t(n), 2) != 0:  # device -> host
        raise RuntimeError("hipMemcpy failed")
    return buf.raw[:n]


def synchronize():
    if hip().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
