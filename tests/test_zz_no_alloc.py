"""negf_invert_batch_noalloc, the allocation-free batched inverse: called through
the C ABI on device buffers, it must give bitwise the inverses (and INFO) of the
allocating entry point negf_invert_batch on the same input, on two calls in a row
-- the second with a smaller input, which reuses the static scratch array the
first one left dirty -- in both element types, and reject a request beyond its
static workspace."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("nonequilibriumgreensfunction_amd.capi")


def _hip():
    """The HIP runtime the library itself is linked with (matched by soname once the
    library is loaded), for the test's own device buffers."""
    capi.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


def _noalloc(G, complex64, m3=True):
    """(inverses, info) of the stack G (batch, n, n) through the no-alloc entry point,
    on device buffers the test allocates."""
    h = _hip()
    batch, n, _ = G.shape
    g = np.ascontiguousarray(np.swapaxes(G, 1, 2))  # column-major matrices
    out = np.zeros_like(g)
    info = np.full(batch, -7, np.int32)
    bufs = [ctypes.c_void_p() for _ in range(3)]
    try:
        for b, nbytes in zip(bufs, (g.nbytes, g.nbytes, info.nbytes)):
            _ok(h.hipMalloc(ctypes.byref(b), nbytes))
        _ok(h.hipMemcpy(bufs[0], g.ctypes.data, g.nbytes, 1))
        _ok(h.hipMemcpy(bufs[2], info.ctypes.data, info.nbytes, 1))
        _ok(h.hipDeviceSynchronize())
        capi._check(capi.lib().negf_invert_batch_noalloc(int(complex64), int(m3), n, batch,
                                                         bufs[0], bufs[1], bufs[2], None))
        _ok(h.hipDeviceSynchronize())
        _ok(h.hipMemcpy(out.ctypes.data, bufs[1], out.nbytes, 2))
        _ok(h.hipMemcpy(info.ctypes.data, bufs[2], info.nbytes, 2))
    finally:
        for b in bufs:
            if b.value:
                h.hipFree(b)
    return np.swapaxes(out, 1, 2), info


@pytest.mark.parametrize("complex64", [False, True], ids=["c128", "c64"])
def test_noalloc_equals_allocating_inverse(complex64):
    rng = np.random.default_rng(11)
    dt = np.complex64 if complex64 else np.complex128
    for n, batch in ((1100, 3), (300, 2)):  # the second call is smaller than the first
        print(f"no-alloc inverse n={n} batch={batch} complex64={complex64}", flush=True)
        G = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(dt)
        G += (2 * np.sqrt(n) * np.eye(n)).astype(dt)
        G[-1, :, 5] = 0  # a singular matrix: INFO must report it in both entry points
        X, info = _noalloc(G, complex64)
        Xr, infor = capi.invert_batch(G, complex64=complex64, m3=True)
        assert np.array_equal(info, infor), (info, infor)
        assert info[-1] > 0 and np.all(info[:-1] == 0)
        ok = slice(0, batch - 1)
        assert np.array_equal(X[ok], Xr[ok])


def test_noalloc_rejects_oversized_requests():
    with pytest.raises(capi.NegfError) as e:
        capi._check(capi.lib().negf_invert_batch_noalloc(0, 1, 3072, 8, ctypes.c_void_p(16),
                                                         ctypes.c_void_p(32), None, None))
    assert e.value.code == capi.NEGF_ERR_MEMORY
