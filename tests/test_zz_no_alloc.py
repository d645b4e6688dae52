"""sr_fft_conv_noalloc (no device allocation: its scratch buffers live in a
fixed device array) against sr_fft_conv through the C-ABI: bit-identical
results for the forward and the adjoint, separable and general kernels,
host and device arrays, on two calls in a row whose second is smaller; the
size limits are rejected with SR_ERR_UNSUPPORTED before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _kernel(rng, px, pz, separable):
    if separable:
        u, w = rng.random(px) + 0.1, rng.random(pz) + 0.1
        k = np.outer(u + u[::-1], w + w[::-1])
    else:
        k = rng.random((px, pz))
        k = k + k[::-1, ::-1]
    return np.asfortranarray((k / k.sum()).astype(np.float32))


def _call(gpu, fn, img, ker, adj, device):
    from spectrecon_amd import _lib as L
    nx, nz = img.shape
    px, pz = ker.shape
    if not device:
        out = np.zeros_like(img)
        L.check(fn(out.ctypes.data, img.ctypes.data, nx, nz, ker.ctypes.data, px, pz, adj, L.SR_MEM_HOST))
        return out
    di, dk, do = gpu.DeviceArray.from_host(img), gpu.DeviceArray.from_host(ker), gpu.DeviceArray((nx, nz))
    L.check(fn(do.ptr, di.ptr, nx, nz, dk.ptr, px, pz, adj, L.SR_MEM_DEVICE))
    return do.to_host()


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("separable", [True, False])
def test_noalloc_equals_fft_conv(gpu, device, separable):
    lib = gpu.lib()
    rng = np.random.default_rng(11 + int(device) + 2 * int(separable))
    # a large plane and PSF first, then a smaller one (shapes the unit-operator tests cover)
    for nx, nz, px, pz in ((128, 80, 37, 37), (20, 14, 5, 5)):
        print(f"noalloc: {nx}x{nz}, {px}x{pz}, separable={separable}, device={device}", flush=True)
        img = np.asfortranarray(rng.standard_normal((nx, nz)).astype(np.float32))
        ker = _kernel(rng, px, pz, separable)
        for adj in (0, 1):
            ref = _call(gpu, lib.sr_fft_conv, img, ker, adj, device)
            got = _call(gpu, lib.sr_fft_conv_noalloc, img, ker, adj, device)
            np.testing.assert_array_equal(got, ref)
            assert np.abs(ref).max() > 0


def test_noalloc_limits(gpu):
    from spectrecon_amd import _lib as L
    lib = gpu.lib()
    img = np.zeros((256, 129), np.float32, order="F")  # 33024 > 32768 floats
    ker = np.ones((3, 3), np.float32, order="F") / 9
    out = np.zeros_like(img)
    rc = lib.sr_fft_conv_noalloc(out.ctypes.data, img.ctypes.data, 256, 129, ker.ctypes.data, 3, 3, 0, L.SR_MEM_HOST)
    assert rc == L.SR_ERR_UNSUPPORTED and b"scratch" in lib.sr_last_error()
    big = np.ones((65, 65), np.float32, order="F") / 65 ** 2
    img2 = np.zeros((32, 16), np.float32, order="F")
    out2 = np.zeros_like(img2)
    rc = lib.sr_fft_conv_noalloc(out2.ctypes.data, img2.ctypes.data, 32, 16, big.ctypes.data, 65, 65, 0,
                                 L.SR_MEM_HOST)
    assert rc == L.SR_ERR_UNSUPPORTED and b"scratch" in lib.sr_last_error()
