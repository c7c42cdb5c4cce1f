"""ash3d_tr_mass_totals_noalloc: the allocation-free variant of
ash3d_tr_mass_totals (partial sums in the library's fixed device array,
result written to caller-provided device memory) gives bitwise the totals
of the allocating entry point, on two calls in a row with the second on a
smaller input, and rejects contexts beyond its bin limit."""
import ctypes
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr(pkg):
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


@pytest.fixture(scope="module")
def hip(tr):
    """The HIP runtime the library runs on (the test, not the entry point,
    allocates the output buffer)."""
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


class DeviceBuffer:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), 8 * n) == 0
        nan = np.full(n, np.nan)
        assert hip.hipMemcpy(self.ptr, nan.ctypes.data, 8 * n, 1) == 0  # host to device

    def read(self):
        h = np.zeros(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device to host
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def _totals_both(tr, hip, st, steps):
    T = tr.Transport(st)
    T.tune(line_solver="partitioned")
    for it in range(1, steps + 1):
        if it > 1:
            T.step_glue()
        T.transport_step(it, True)
    ref = np.concatenate(T.mass_totals())
    out = DeviceBuffer(hip, 3 * st.ns)
    T.mass_totals_noalloc(out.ptr.value)
    T.synchronize()
    got = out.read()
    out.free()
    T.close()
    return ref, got


def test_noalloc_mass_totals_matches(pkg, tr, hip):
    # a larger input first, then a smaller one: the second call must not see
    # the first call's partial sums
    big = pkg.workloads.random_state(40, 30, 16, 9, seed=71, latlon=True, periodic=True)
    ref1, got1 = _totals_both(tr, hip, big, 3)
    assert np.array_equal(ref1, got1), np.max(np.abs(ref1 - got1))
    assert np.all(ref1[:9] > 0)
    small = pkg.workloads.random_state(12, 9, 7, 3, seed=72, latlon=True)
    ref2, got2 = _totals_both(tr, hip, small, 2)
    assert np.array_equal(ref2, got2), np.max(np.abs(ref2 - got2))
    assert ref2.shape == (9,) and not np.array_equal(ref2, ref1[:9])


def test_noalloc_rejects_too_many_bins(pkg, tr, hip):
    st = pkg.workloads.random_state(6, 5, 4, 65, seed=73, diffusivity=False)
    T = tr.Transport(st)
    out = DeviceBuffer(hip, 3 * 65)
    with pytest.raises(tr.Ash3dError) as e:
        T.mass_totals_noalloc(out.ptr.value)
    assert e.value.code == 1 and "at most 64 bins" in str(e.value)
    assert np.isnan(out.read()).all()
    out.free()
    T.close()
