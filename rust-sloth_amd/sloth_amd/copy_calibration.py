"""Device-to-device copy of a known size, for calibrating HBM byte counters.

`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` read the copy kernel
(__amd_rocclr_copyBuffer) that this runs: a copy of N bytes moves N bytes in
and N bytes out, so the counters' reading for it gives the scale from counter
units to bytes.  Calls the HIP runtime directly through ctypes (no torch).

    python3 copy_calibration.py [MiB]    (default 256 MiB, copied 4 times)
"""
from __future__ import annotations

import ctypes as C
import sys


def main() -> int:
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    n = mib << 20
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(4):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {mib} MiB device-to-device 4 times", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
