"""Calibration of rocprofv3's memory counters (FETCH_SIZE, WRITE_SIZE): a
device-to-device copy of a known size, run inside the profiled process, shows
how many KiB the counters report for a byte read or written by the runtime's
copy kernel (__amd_rocclr_copyBuffer)."""
from __future__ import annotations

import ctypes as C
import sys


def calib_copy(mib: int, log=None) -> None:
    """One hipMemcpy device-to-device of `mib` MiB (the runtime's copy kernel:
    reads and writes exactly that many bytes)."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    (log or (lambda m: print(m, file=sys.stderr, flush=True)))(f"calibration copy: {n} bytes device-to-device")
