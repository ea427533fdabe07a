"""The allocation-free energy entry point, ccsdt_tengy_device_noalloc (needs a GPU).

It runs ccsd_tengy_omp's reduction on device-resident arrays with its per-tile
partials in a static device array.  Checked through the C ABI against the
synchronous drop-in ccsd_tengy_omp on the same values (both run the same
kernel arithmetic: bitwise equal), on two calls in a row, the second with a
smaller nvir (the shared array keeps the first call's partials, which the
second must not read), on a stream of the caller's, and the size limit.
The test's own device buffers come from the HIP runtime through ctypes; the
entry point allocates nothing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(cc):
    cc.init(0)
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    fn = cc.lib().ccsdt_tengy_device_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_void_p] * 15 + [ctypes.c_double, ctypes.c_void_p]
                   + [ctypes.c_int] * 3 + [ctypes.c_void_p])
    return L


class Dev:
    """Device copies of host arrays (freed on close)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


VEC = ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2")


def case(cc, nocc, nvir, seed):
    arrays = cc.make_inputs(nocc, nvir, 1, seed)
    rng = np.random.default_rng(seed)
    f = [rng.uniform(-0.05, 0.05, nvir * nvir) for _ in range(8)]
    vec = {n: arrays[n][:nvir].copy() for n in VEC}
    eaijk = arrays["eorb"][nocc] - (arrays["eorb"][0] + arrays["eorb"][1 % nocc] + arrays["eorb"][0])
    return f, vec, arrays["eorb"], eaijk


def noalloc(cc, dev, f, vec, eorb, eaijk, nocc, nvir, stream):
    fp = [dev.put(x) for x in f]
    vp = {n: dev.put(vec[n]) for n in VEC}
    ep = dev.put(eorb)
    out = dev.put(np.full(4, np.nan))
    rc = cc.lib().ccsdt_tengy_device_noalloc(*fp, vp["dintc1"], vp["dintx1"], vp["t1v1"], vp["dintc2"],
                                             vp["dintx2"], vp["t1v2"], ep, eaijk, out, 0, nocc, nvir,
                                             stream)
    assert rc == 0, cc.last_error()
    return out


def test_matches_dropin_two_calls(cc, hip):
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    dev = Dev(hip)
    try:
        # each call smaller than the one before; nvir 45 and 33 (a last tile one row
        # wide) take the odd-nvir load form, nvir = 1 single loads
        for nocc, nvir, seed in ((30, 300, 41), (5, 45, 42), (4, 33, 44), (2, 1, 45)):
            f, vec, eorb, eaijk = case(cc, nocc, nvir, seed)
            out = noalloc(cc, dev, f, vec, eorb, eaijk, nocc, nvir, stream)
            assert hip.hipStreamSynchronize(stream) == 0
            got = dev.get(out, 4)
            ref = np.array(cc.tengy(f, vec, eorb, eaijk, 0, nocc, nvir))
            assert np.all(np.isfinite(ref))
            assert np.array_equal(got, ref), (nvir, got, ref)
    finally:
        hip.hipStreamSynchronize(stream)
        dev.close()
        hip.hipStreamDestroy(stream)


def test_default_stream_and_repeat(cc, hip):
    """The null (default) stream, and the same call twice: the same bits."""
    dev = Dev(hip)
    try:
        f, vec, eorb, eaijk = case(cc, 8, 77, 43)
        outs = [noalloc(cc, dev, f, vec, eorb, eaijk, 8, 77, None) for _ in range(2)]
        assert hip.hipStreamSynchronize(None) == 0
        a, b = dev.get(outs[0], 4), dev.get(outs[1], 4)
        assert np.array_equal(a, b)
        assert np.array_equal(a, np.array(cc.tengy(f, vec, eorb, eaijk, 0, 8, 77)))
    finally:
        dev.close()


def test_size_limit_rejected(cc, hip):
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = cc.lib().ccsdt_tengy_device_noalloc(*([dummy] * 15), 0.0, dummy, 0, 4, 2049, None)
    assert rc == -1
    assert "2048" in cc.last_error()[1]
    rc = cc.lib().ccsdt_tengy_device_noalloc(*([dummy] * 15), 0.0, None, 0, 4, 100, None)
    assert rc == -1
