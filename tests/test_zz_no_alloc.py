"""gnd_hip_denoise_noalloc (the denoiser with a static scratch array instead
of a device allocation) against gnd_hip_denoise: identical images on two
calls in a row, the second smaller than the first, and a clear refusal of
an image larger than its scratch."""
import ctypes as C

import numpy as np
import pytest

from gianduia_amd import _abi as A
from gianduia_amd.api import DeviceBuffer, GndError, _lib

pytestmark = pytest.mark.gpu


def _inputs(rng, h, w):
    color = rng.gamma(2.0, 0.5, (h, w, 3)).astype(np.float32)
    albedo = rng.uniform(0.05, 0.9, (h, w, 3)).astype(np.float32)
    n = rng.normal(size=(h, w, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return color, albedo, (n * 0.5 + 0.5).astype(np.float32)


@pytest.mark.parametrize("guides", [False, True])
def test_noalloc_denoise_equals_denoise(device, guides):
    rng = np.random.default_rng(5)
    for h, w in ((517, 931), (61, 37)):  # second call smaller than the first
        c, a, n = _inputs(rng, h, w)
        ga, gn = (a, n) if guides else (None, None)
        ref = device.denoise(c, ga, gn)
        got = device.denoise(c, ga, gn, noalloc=True)
        assert np.array_equal(got, ref), f"{w}x{h}: the no-allocation denoiser differs"


def test_noalloc_denoise_refuses_oversize(device):
    buf = DeviceBuffer(device, 64)
    ms = C.c_double()
    rc = _lib().gnd_hip_denoise_noalloc(device.ctx, buf.ptr, None, None, 2049, 1024, buf.ptr, C.byref(ms))
    assert rc == A.GND_E_INVALID
    assert "static scratch" in device.last_error()
    with pytest.raises(GndError):
        device._check(rc)
    buf.free()
