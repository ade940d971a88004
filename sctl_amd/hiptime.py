"""HIP event timing through the HIP runtime (libamdhip64), for timing work
enqueued by libsctl_amd on an explicit stream.  Events are recorded on the
same stream as the kernels they bracket."""
import ctypes

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        vp = ctypes.c_void_p
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipEventDestroy.argtypes = [vp]
        L.hipDeviceSynchronize.argtypes = []
        L.hipMemsetAsync.argtypes = [vp, ctypes.c_int, ctypes.c_size_t, vp]
        _hip = L
    return _hip


class Event:
    def __init__(self):
        self.ev = ctypes.c_void_p()
        rc = hip().hipEventCreate(ctypes.byref(self.ev))
        if rc != 0:
            raise RuntimeError("hipEventCreate failed (%d)" % rc)

    def record(self, stream):
        rc = hip().hipEventRecord(self.ev, stream)
        if rc != 0:
            raise RuntimeError("hipEventRecord failed (%d)" % rc)

    def synchronize(self):
        rc = hip().hipEventSynchronize(self.ev)
        if rc != 0:
            raise RuntimeError("hipEventSynchronize failed (%d)" % rc)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        end.synchronize()
        rc = hip().hipEventElapsedTime(ctypes.byref(ms), self.ev, end.ev)
        if rc != 0:
            raise RuntimeError("hipEventElapsedTime failed (%d)" % rc)
        return ms.value

    def __del__(self):
        try:
            if self.ev:
                hip().hipEventDestroy(self.ev)
        except Exception:
            pass


def device_synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed (%d)" % rc)


def memset_async(ptr, value, nbytes, stream):
    """hipMemsetAsync on `stream` (benchmark reference only; never inside a captured graph)."""
    rc = hip().hipMemsetAsync(ptr, value, nbytes, stream)
    if rc != 0:
        raise RuntimeError("hipMemsetAsync failed (%d)" % rc)

