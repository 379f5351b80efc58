"""nb_forces_direct_noalloc: the direct sum on caller-owned device arrays with
the library's static scratch, against nb_forces_direct on a context (the same
kernels), bit for bit, on two calls in a row, the second with fewer bodies."""
import ctypes

import numpy as np
import pytest

from helpers import uniform_cube

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs for device arrays of its own."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.lib.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.held = []

    def array(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(8, host.nbytes)) == 0
        self.held.append(p)
        if host.size:
            assert self.lib.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D) == 0
        return p

    def read(self, p, n):
        out = np.empty(n)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.int64)


@pytest.mark.parametrize("precision", ["fp32c", "fp64_fast"])
def test_noalloc_matches_forces_direct(gpu, precision):
    prec = {"fp32c": gpu.PREC_FP32C, "fp64_fast": gpu.PREC_FP64_FAST}[precision]
    lib = gpu.load()
    hip = Hip()
    stream = ctypes.c_void_p()
    assert hip.lib.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # 20,000 bodies use 51 source splits (their partial sums in the
        # scratch); the second call, with fewer bodies and splits, must not be
        # disturbed by what the first left in the shared array
        for n, seed in ((20000, 3), (3001, 4), (20000, 5)):
            print(f"noalloc {precision}: n={n}", flush=True)
            d = uniform_cube(n, seed)
            with gpu.Context(n, precision=prec) as c:
                c.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
                want = c.forces_direct()
                G, eps2 = c.params.G, c.params.eps2
            x, y, z, m = (hip.array(d[k]) for k in ("x", "y", "z", "m"))
            ax, ay, az = (hip.array(np.full(n, np.nan)) for _ in range(3))
            rc = lib.nb_forces_direct_noalloc(n, x, y, z, m, G, eps2, prec, ax, ay, az, stream)
            assert rc == gpu.NB_OK
            assert hip.lib.hipStreamSynchronize(stream) == 0
            got = [hip.read(p, n) for p in (ax, ay, az)]
            for u, v in zip(got, want):
                np.testing.assert_array_equal(bits(u), bits(v))
            hip.close()
    finally:
        hip.close()
        hip.lib.hipStreamDestroy(stream)


def test_noalloc_rejects_what_it_cannot_do(gpu):
    lib = gpu.load()
    hip = Hip()
    try:
        d = uniform_cube(8, 1)
        x, y, z, m, ax, ay, az = (hip.array(d[k]) for k in ("x", "y", "z", "m", "x", "y", "z"))
        p = gpu.default_params()
        args = (x, y, z, m, p.G, p.eps2)
        # FP64 reference arithmetic has no split kernel
        assert lib.nb_forces_direct_noalloc(8, *args, gpu.PREC_FP64, ax, ay, az,
                                            None) == gpu.NB_E_INVALID_ARG
        # more bodies than the scratch is sized for (checked before any launch)
        assert lib.nb_forces_direct_noalloc(gpu.NOALLOC_MAX_BODIES + 1, *args, gpu.PREC_FP32C,
                                            ax, ay, az, None) == gpu.NB_E_INVALID_ARG
        assert lib.nb_forces_direct_noalloc(0, *args, gpu.PREC_FP32C, ax, ay, az,
                                            None) == gpu.NB_E_INVALID_ARG
        assert lib.nb_forces_direct_noalloc(8, None, y, z, m, p.G, p.eps2, gpu.PREC_FP32C, ax,
                                            ay, az, None) == gpu.NB_E_INVALID_ARG
    finally:
        hip.close()
