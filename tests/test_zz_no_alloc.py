"""The allocation-free entry point ofx_median_device_noalloc (no counterpart in the reference; for
callers that cannot allocate device memory) against ofx_median, the entry point it varies, through
the C ABI: bit-identical medians on two calls in a row, the second with fewer values, so that the
shared scratch array must be cleared between calls. The test's own device arrays come from the HIP
runtime that libofx.so loaded (hipMalloc / hipMemcpy / hipFree through ctypes), so the test works
whatever else the process has loaded before it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def hip(eng):
    """The HIP runtime libofx.so runs on (already loaded: dlopen of its path returns it)."""
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libofx.so did not load a HIP runtime"
    rt = ctypes.CDLL(path)
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip, self.ptr = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(int(nbytes), 8)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def download(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _device_median(eng, hip, values):
    d, out = DevArray(hip, values.nbytes), DevArray(hip, 8)
    try:
        d.upload(values)
        out.upload(np.array([12345.0]))
        eng.median_device_noalloc(d.ptr.value, values.size, out.ptr.value)
        return float(out.download(np.zeros(1))[0])
    finally:
        d.free()
        out.free()


def test_noalloc_median_matches_ofx_median_on_two_calls(eng, hip):
    rng = np.random.default_rng(42)
    first = rng.lognormal(0.0, 3.0, 1_000_001)           # odd count
    first[::97] = -first[::97]                           # both signs
    second = rng.standard_normal(4096) * 1e-3            # even count, smaller, a different range
    second[7] = -0.0
    for label, v in (("first call, 1000001 values", first), ("second call, 4096 values", second)):
        print(f"noalloc median: {label}", flush=True)
        got = _device_median(eng, hip, v)
        want = eng.median(v)
        print(f"  ofx_median_device_noalloc {got!r}  ofx_median {want!r}  numpy {np.median(v)!r}", flush=True)
        assert np.float64(got).tobytes() == np.float64(want).tobytes()
        assert got == np.median(v)


def test_noalloc_median_nan_empty_and_too_many(eng, hip):
    v = np.array([3.0, np.nan, 1.0])
    print("noalloc median: a NaN among the values", flush=True)
    assert np.isnan(_device_median(eng, hip, v)) and np.isnan(eng.median(v))
    out = DevArray(hip, 8)
    try:
        out.upload(np.zeros(1))
        print("noalloc median: no values", flush=True)
        eng.median_device_noalloc(out.ptr.value, 0, out.ptr.value)
        assert np.isnan(out.download(np.zeros(1))[0])
        # refused on the host before any work: the count exceeds what the histograms hold
        with pytest.raises(ValueError, match="at most 4294967295"):
            eng.median_device_noalloc(out.ptr.value, 1 << 32, out.ptr.value)
    finally:
        out.free()
