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

    def check(status: int, what: str) -> None:
        if status != 0:
            raise RuntimeError(f"calibration copy: {what} failed (hipError {status})")

    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    check(hip.hipMalloc(C.byref(a), C.c_size_t(n)), "hipMalloc")
    try:
        check(hip.hipMalloc(C.byref(b), C.c_size_t(n)), "hipMalloc")
        check(hip.hipMemset(a, 1, C.c_size_t(n)), "hipMemset")
        check(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        check(hip.hipMemcpy(b, a, C.c_size_t(n), 3), "hipMemcpy")  # hipMemcpyDeviceToDevice
        check(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    finally:
        hip.hipFree(a)
        if b.value:
            hip.hipFree(b)
    (log or (lambda m: print(m, file=sys.stderr, flush=True)))(f"calibration copy: {n} bytes device-to-device")
