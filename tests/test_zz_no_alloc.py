"""pppm_hashgrid_build_noalloc through the C ABI: the same grid as
pppm_hashgrid_build, bit for bit, from device buffers, on two calls in a row
(the second smaller, so the shared scratch array still holds the first call's
contents); sizes above its limits and host pointers are refused.

Device buffers for the test come from the HIP runtime the library itself
links (/opt/rocm), not from another runtime in the process.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import pppm_amd as P

pytestmark = pytest.mark.gpu

HIP = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
HIP.hipFree.argtypes = [C.c_void_p]
HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
H2D, D2H = 1, 2


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


class DeviceArray:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert HIP.hipMalloc(C.byref(self.p), max(nbytes, 16)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert HIP.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            assert HIP.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            HIP.hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(ctx, dpos, n, S, scale, bmin, dsorted, dstart, dend):
    b = np.ascontiguousarray(bmin, np.float64)
    return ctx.L.pppm_hashgrid_build_noalloc(ctx.h, dpos.p if dpos else None, n, S, scale, b.ctypes.data,
                                             dsorted.p if dsorted else None, dstart.p, dend.p)


def test_noalloc_grid_equals_allocating_grid(gpu_ctx):
    rng = np.random.default_rng(17)
    cap_n, cap_s = 400_000, 300_000
    dpos, dsorted = DeviceArray(12 * cap_n), DeviceArray(4 * cap_n)
    dstart, dend = DeviceArray(4 * cap_s), DeviceArray(4 * cap_s)
    try:
        for n, S in ((cap_n, cap_s), (50_000, 30_011)):  # second call smaller: stale scratch contents
            say(f"no-alloc grid n={n} S={S}")
            pos = rng.uniform(0.0, 7.0, (n, 3)).astype(np.float32)
            pos[:100] = pos[100:200]        # duplicate points
            pos[200:300] = np.floor(pos[200:300] * 8) / 8  # points on cell boundaries
            scale, bmin = 4.0, np.array([-0.25, 0.0, 0.5])
            dpos.put(pos)
            rc = noalloc(gpu_ctx, dpos, n, S, scale, bmin, dsorted, dstart, dend)
            assert rc == P.OK, gpu_ctx.L.pppm_last_error(gpu_ctx.h)
            got = (dsorted.get(np.int32, n), dstart.get(np.int32, S), dend.get(np.int32, S))
            want = gpu_ctx.hashgrid(pos, S, scale, bmin)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
            assert got[2][-1] == n and np.all(got[0] >= 0)
    finally:
        for d in (dpos, dsorted, dstart, dend):
            d.free()


def test_noalloc_grid_limits_and_host_pointers(gpu_ctx):
    say("no-alloc grid: limits and host pointers")
    d = DeviceArray(1024)
    try:
        big = P.PPPM_NOALLOC_MAX_POINTS if hasattr(P, "PPPM_NOALLOC_MAX_POINTS") else 1 << 20
        assert noalloc(gpu_ctx, d, big + 1, 16, 1.0, [0, 0, 0], d, d, d) == P.ERR_CAPACITY
        assert noalloc(gpu_ctx, d, 4, (1 << 20) + 1, 1.0, [0, 0, 0], d, d, d) == P.ERR_CAPACITY
        host = np.zeros(64, np.int32)

        class Host:
            p = C.c_void_p(host.ctypes.data)

        assert noalloc(gpu_ctx, d, 4, 16, 1.0, [0, 0, 0], d, Host, d) == P.ERR_INVALID
        # the context stays usable
        pos = np.random.default_rng(1).uniform(0, 1, (8, 3)).astype(np.float32)
        gpu_ctx.hashgrid(pos, 7, 2.0, [0, 0, 0])
    finally:
        d.free()
