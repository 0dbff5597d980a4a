"""GPU: cupss_c2r_noalloc (the inverse transform with its workspace in a
static device array, no device allocation) against cupss_c2r (workspace
passed in) through the C-ABI: two calls in a row, the second on a smaller
grid, each bit-identical to the allocating entry point; sizes above the
static workspace are rejected before any device work."""
import ctypes

import numpy as np
import pytest

import cupss_amd as C

pytestmark = pytest.mark.gpu


def _spectrum(dims, seed):
    sx, sy, sz = dims
    x = np.random.default_rng(seed).standard_normal(sx * sy * sz)
    return np.fft.rfftn(x.reshape(sz, sy, sx)).ravel().astype(np.complex64)


def _run(fn_name, X, dims, mask_p, scale, extra):
    sx, sy, sz = dims
    n = sx * sy * sz
    din = C.DeviceBuffer(X.nbytes)
    din.upload(X)
    dout = C.DeviceBuffer(n * 4)
    fn = getattr(C.lib(), fn_name)
    C.check(fn(sx, sy, sz, din.ptr, mask_p, scale, dout.ptr, *extra, None))
    C.check(C.lib().cupss_device_synchronize())
    return dout.download(np.float32, n)


def _reference(X, dims, mask_p, scale):
    scratch = C.DeviceBuffer(X.nbytes)
    return _run("cupss_c2r", X, dims, mask_p, scale, [scratch.ptr])


def _noalloc(X, dims, mask_p, scale):
    return _run("cupss_c2r_noalloc", X, dims, mask_p, scale, [])


@pytest.mark.parametrize("first,second", [((1024, 1024, 1), (256, 128, 1)), ((64, 64, 64), (32, 16, 8))])
def test_noalloc_matches_c2r_two_calls(first, second):
    for i, (dims, p) in enumerate(((first, 3), (second, 2))):
        n = dims[0] * dims[1] * dims[2]
        print("c2r_noalloc call", i, dims, "mask", p, flush=True)
        X = _spectrum(dims, 100 + i)
        want = _reference(X, dims, p, 1.0 / n)
        got = _noalloc(X, dims, p, 1.0 / n)
        assert np.array_equal(got, want), (dims, float(np.max(np.abs(got - want))))
        # and the plain inverse: the round trip of the spectrum
        sx, sy, sz = dims
        x = np.fft.irfftn(X.astype(np.complex128).reshape(sz, sy, sx // 2 + 1), s=(sz, sy, sx), axes=(0, 1, 2)).ravel()
        assert np.max(np.abs(_noalloc(X, dims, 0, 1.0 / n) - x)) < 1e-5 * np.max(np.abs(x)) + 1e-6


def test_noalloc_rejects_sizes_beyond_the_static_workspace():
    dims = (4096, 1024, 1)    # 2049 * 1024 bins > 2^21
    rc = C.lib().cupss_c2r_noalloc(dims[0], dims[1], dims[2], ctypes.c_void_p(16), 0, 1.0, ctypes.c_void_p(16), None)
    assert rc == -1
    assert b"static scratch" in C.lib().cupss_last_error()
