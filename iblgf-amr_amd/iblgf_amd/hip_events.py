"""HIP events on an arbitrary stream (the solver's), via the HIP runtime C API."""
import ctypes

_hip = None


def _rt():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return _hip


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        rc = _rt().hipEventCreate(ctypes.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")

    def record(self, stream):
        rc = _rt().hipEventRecord(self.h, ctypes.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def synchronize(self):
        _rt().hipEventSynchronize(self.h)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        rc = _rt().hipEventElapsedTime(ctypes.byref(ms), self.h, end.h)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return float(ms.value)

    def __del__(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.hipEventDestroy(self.h)
