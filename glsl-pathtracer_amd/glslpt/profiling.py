"""Calibration helper for HBM traffic counters.

    python -m glslpt.profiling copy NBYTES

copies NBYTES device-to-device with hipMemcpy (one blit-kernel dispatch that
reads and writes NBYTES). Profiled with `rocprofv3 --pmc FETCH_SIZE` (then
WRITE_SIZE), its per-dispatch counter values give the factor that turns the
counters of the path's kernels into bytes of HBM traffic.
"""
import ctypes as C
import sys


def copy(nbytes):
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        rc = hip.hipMalloc(C.byref(p), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    print(f"copied {nbytes} bytes device to device")


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "copy":
        copy(int(sys.argv[2]))
    else:
        sys.exit("usage: python -m glslpt.profiling copy NBYTES")
