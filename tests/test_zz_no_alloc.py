"""The allocation-free LU entry point (capsx_lu_logdet_batch_noalloc)
against the allocating one (capsx_lu_logdet_batch), through the C ABI: the
same kernels and schedule on the same matrices give the same determinants
bit for bit, on two calls in a row, the second with a smaller batch of
smaller matrices (the fixed scratch array keeps the first call's contents,
which the second must not see).  The caller's device buffers and stream
come from the HIP runtime the library itself uses (libamdhip64 as already
loaded; torch bundles another build of it under the same soname, so it is
not used here); the entry point itself allocates nothing."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _matrices(rng, N, n):
    mats = []
    for b in range(n):
        if b % 3 == 0:                     # row interchanges in every panel
            A = rng.normal(size=(N, N))
            A[np.arange(N), np.arange(N)] = 0.0
        else:
            X = rng.normal(size=(N, N)) / math.sqrt(N)
            A = np.eye(N) - 0.3 * (X + X.T) / 2
        mats.append(A)
    return np.array(mats)


class _Hip:
    """The caller's side: device buffers and a stream from the HIP runtime
    that libcaps_amd.so is linked against."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                     "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.L, name).restype = ctypes.c_int
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p

    def to_host(self, p, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out


def test_noalloc_lu_matches_allocating_entry():
    import caps_amd
    L = caps_amd.lib()
    f = L.capsx_lu_logdet_batch_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    hip = _Hip()
    stream = ctypes.c_void_p()
    assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
    rng = np.random.default_rng(21)
    bufs = []
    try:
        for N, n in ((300, 8), (129, 3)):
            mats = _matrices(rng, N, n)
            ref = caps_amd.lu_logdet_batch(mats)                    # the allocating entry point
            dA = hip.to_device(np.transpose(mats, (0, 2, 1)))       # column-major per matrix
            dout = hip.to_device(np.full(n, np.nan))
            bufs += [dA, dout]
            rc = f(dA, N, N, n, dout, stream.value)
            assert rc == 0
            assert hip.L.hipStreamSynchronize(stream.value) == 0
            got = hip.to_host(dout, n)
            print("noalloc N=%d n=%d: max |diff| vs allocating entry %.1e" % (N, n, np.abs(got - ref).max()))
            assert np.array_equal(got, ref), (N, n, got, ref)
            for A, v in zip(mats, got):
                assert abs(v - np.linalg.slogdet(A)[1]) < 1e-10 * N
        # a batch beyond the fixed array is refused before anything is launched
        small = hip.malloc(128)
        bufs.append(small)
        assert f(small, 24000, 24000, 4, small, stream.value) == 3
        assert f(small, 10, 5, 1, small, stream.value) == 1
    finally:
        hip.L.hipStreamSynchronize(stream.value)
        for p in bufs:
            hip.L.hipFree(p)
        hip.L.hipStreamDestroy(stream.value)
