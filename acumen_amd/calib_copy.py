"""Device-to-device copy of a known size, for calibrating rocprofv3's HBM
byte counters (FETCH_SIZE, MemWrites32B) against bytes actually moved.

    rocprofv3 --pmc FETCH_SIZE -- python -m acumen_amd.calib_copy [MiB] [reps]
"""
import ctypes
import sys


def main() -> int:
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {mib} MiB x {reps}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
