"""efvm_state_sums_noalloc (the allocation-free entry point) against efvm_conservation_sums,
through the C ABI: two calls in a row on one GPU, the second on a smaller state, FP64 and FP32;
the fixed partial-sum array keeps the first call's contents, which the second must not see.
The test's own output buffer comes from the HIP runtime the library uses (not torch's)."""
import ctypes as C

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases

pytestmark = pytest.mark.gpu


def hip():
    efvm.load_library()                       # loads the HIP runtime the library links
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DeviceDoubles:
    def __init__(self, n):
        self.h, self.n = hip(), n
        self.p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(self.p), 8 * n) == 0
        host = np.full(n, np.nan)
        assert self.h.hipMemcpy(self.p, host.ctypes.data, 8 * n, 1) == 0        # host to device

    def get(self):
        out = np.zeros(self.n)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0    # device to host
        return out

    def free(self):
        self.h.hipFree(self.p)


def context(n, precision, seed):
    c = efvm.Context(n, n, hi=(10.0, 10.0), muscl=True, limiter="mc", precision=precision)
    c.set_boundaries(["periodic"] * 4)
    c.upload(cases.vortex_modes_state(n, n, seed=seed))
    c.step(2)
    c.synchronize()
    return c


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_sums_equal_the_context_sums(precision):
    out = DeviceDoubles(5)
    for n, seed in ((640, 3), (96, 4)):          # the second call is on a smaller state
        c = context(n, precision, seed)
        ref = c.sums()
        vol = (10.0 / n) ** 3
        efvm.efvm.state_sums_noalloc(c.device_planes(), precision, n * n, vol, out.p.value)
        got = out.get()
        print(precision, n, got, ref)
        assert np.array_equal(got, ref), (precision, n, got, ref)
        c.close()
    out.free()


def test_noalloc_sums_reject_bad_arguments():
    c = context(64, "fp64", 5)
    out = DeviceDoubles(5)
    for bad in (0, 2 ** 31):
        with pytest.raises(efvm.EfvmError):
            efvm.efvm.state_sums_noalloc(c.device_planes(), "fp64", bad, 1.0, out.p.value)
    out.free()
    c.close()
