"""dem_scan_noalloc, the allocation-free variant of dem_debug_scan, through the C ABI.

The test allocates the device buffers itself (the entry point allocates nothing) and compares the
variant's output and total with dem_debug_scan's on the same values, bit for bit, on two calls in a
row where the second is smaller than the first, then an empty scan and the size limit.
"""
import ctypes

import numpy as np
import pytest

from taichi_dem_julia_port_amd import _native as N

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    # the HIP runtime libdemhip.so is linked against
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMalloc.restype = ctypes.c_int
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipFree.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    return hip


class _DeviceBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 8)) == 0

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _debug_scan(lib, a):
    out = np.zeros_like(a)
    tot = ctypes.c_uint64(0)
    rc = lib.dem_debug_scan(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(a),
                            ctypes.byref(tot))
    assert rc == 0, lib.dem_last_error(None)
    return out, tot.value


def _noalloc_scan(lib, hip, din, dout, dtot, a):
    """Upload a, scan it with dem_scan_noalloc into dout/dtot (device), download both."""
    n = len(a)
    if n:
        assert hip.hipMemcpy(din.ptr, a.ctypes.data, a.nbytes, H2D) == 0
    # a stale total shows up if the entry point does not write it
    stale = np.array([0xDEADBEEF], dtype=np.uint64)
    assert hip.hipMemcpy(dtot.ptr, stale.ctypes.data, 8, H2D) == 0
    rc = lib.dem_scan_noalloc(din.ptr, dout.ptr, n, dtot.ptr, None)
    assert rc == 0, lib.dem_last_error(None)
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(n, dtype=np.uint32)
    if n:
        assert hip.hipMemcpy(out.ctypes.data, dout.ptr, out.nbytes, D2H) == 0
    tot = np.zeros(1, dtype=np.uint64)
    assert hip.hipMemcpy(tot.ctypes.data, dtot.ptr, 8, D2H) == 0
    return out, int(tot[0])


def test_noalloc_scan_matches_debug_scan_on_successive_calls():
    lib = N.load_library()
    hip = _hip()
    rng = np.random.default_rng(2024)
    # the first call is large (257 long tiles), the second smaller (35 short tiles)
    sizes = [(1 << 21) + 3, 70001]
    cap = max(sizes)
    din, dout, dtot = _DeviceBuf(hip, 4 * cap), _DeviceBuf(hip, 4 * cap), _DeviceBuf(hip, 8)
    try:
        for n in sizes:
            print(f"[noalloc] n={n}", flush=True)
            a = rng.integers(0, 1000, size=n, dtype=np.uint32)
            ref, ref_tot = _debug_scan(lib, a)
            got, got_tot = _noalloc_scan(lib, hip, din, dout, dtot, a)
            assert np.array_equal(got, ref)
            assert got_tot == ref_tot == int(a.sum(dtype=np.uint64))
        print("[noalloc] empty scan", flush=True)
        got, got_tot = _noalloc_scan(lib, hip, din, dout, dtot, np.zeros(0, dtype=np.uint32))
        assert got_tot == 0
        # beyond the scratch's limit: rejected before any device work
        rc = lib.dem_scan_noalloc(din.ptr, dout.ptr, (1 << 30) + 1, dtot.ptr, None)
        assert rc == N.DEM_ERR_INVALID
        assert b"2^30" in lib.dem_last_error(None)
    finally:
        for b in (din, dout, dtot):
            b.close()
