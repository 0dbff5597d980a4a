"""GPU: the hand-written Stockham R2C/C2R transforms against numpy (float64),
through the C-ABI (cupss_r2c / cupss_c2r on device buffers)."""
import ctypes

import numpy as np
import pytest

import cupss_amd as C
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

DIMS = [(2, 1, 1), (8, 1, 1), (64, 1, 1), (4096, 1, 1), (16, 8, 1), (8, 16, 1), (64, 64, 1), (256, 128, 1),
        (1024, 1024, 1), (2048, 32, 1), (32, 4096, 1), (16, 16, 16), (32, 16, 8), (64, 64, 64), (8, 4, 256)]


def _r2c(x, dims):
    sx, sy, sz = dims
    nh = (sx // 2 + 1) * sy * sz
    din = C.DeviceBuffer(x.nbytes)
    din.upload(x)
    dout = C.DeviceBuffer(nh * 8)
    C.check(C.lib().cupss_r2c(sx, sy, sz, din.ptr, dout.ptr, None))
    C.check(C.lib().cupss_device_synchronize())
    return dout.download(np.complex64, nh)


def _c2r(X, dims, mask_p, scale):
    sx, sy, sz = dims
    n = sx * sy * sz
    din = C.DeviceBuffer(X.nbytes)
    din.upload(X.astype(np.complex64))
    scratch = C.DeviceBuffer(X.nbytes)
    dout = C.DeviceBuffer(n * 4)
    C.check(C.lib().cupss_c2r(sx, sy, sz, din.ptr, mask_p, scale, dout.ptr, scratch.ptr, None))
    C.check(C.lib().cupss_device_synchronize())
    return dout.download(np.float32, n)


@pytest.mark.parametrize("dims", DIMS)
def test_r2c_matches_numpy(dims):
    sx, sy, sz = dims
    print("r2c", dims, flush=True)
    x = np.random.default_rng(sx + 3 * sy + 7 * sz).standard_normal(sx * sy * sz).astype(np.float32)
    got = _r2c(x, dims)
    ref = np.fft.rfftn(x.astype(np.float64).reshape(sz, sy, sx)).ravel()
    assert rel_l2(got, ref) < 5e-6
    # self-conjugate bins of the x transform are exactly real
    h = got.reshape(sz, sy, sx // 2 + 1)
    assert np.all(h[:, :, 0].imag == 0) or sy * sz > 1


@pytest.mark.parametrize("dims", DIMS)
def test_c2r_roundtrip_and_mask(dims):
    sx, sy, sz = dims
    print("c2r", dims, flush=True)
    n = sx * sy * sz
    x = np.random.default_rng(11 + sx).standard_normal(n).astype(np.float32)
    X = np.fft.rfftn(x.astype(np.float64).reshape(sz, sy, sx))
    back = _c2r(X.ravel(), dims, 0, 1.0 / n)
    assert rel_l2(back, x) < 5e-6
    for p in (2, 3):
        def keep(s, idx):
            nn = np.where(idx > s // 2, idx - s, idx)
            return np.abs(nn) <= s // (p + 1)
        kz, ky, kx = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx // 2 + 1), indexing="ij")
        m = keep(sx, kx) & keep(sy, ky) & keep(sz, kz)
        ref = np.fft.irfftn(X * m, s=(sz, sy, sx), axes=(0, 1, 2)).ravel()
        got = _c2r(X.ravel(), dims, p, 1.0 / n)
        assert rel_l2(got, ref) < 5e-6, p


def test_normals_match_oracle_stream():
    from oracle import oracle as O
    n = 4099
    d = C.DeviceBuffer(n * 4)
    C.check(C.lib().cupss_normals(77, 5, 2, d.ptr, n, None))
    C.check(C.lib().cupss_device_synchronize())
    got = d.download(np.float32, n)
    ref = O.normals(77, 5, 2, n)
    assert np.max(np.abs(got - ref)) < 2e-5
