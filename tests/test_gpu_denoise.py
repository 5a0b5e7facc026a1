"""Film denoiser on the GPU (gnd_hip_denoise, in place of the reference's
OIDN call in Denoiser::execute, src/core/denoiser.cpp:6-55).

OIDN's trained network is not available, so the GPU filter is an
edge-avoiding a-trous wavelet filter; these tests check the kernels against
the oracle's CPU restatement of that filter, its exact HDR scaling, the error
reduction it gives on rendered films, and the host plugin's denoise().  How
close it comes to OIDN's output is not measured (parity unpinned)."""
import numpy as np
import pytest

from gianduia_amd import DeviceScene, GndError, HostScene, Integrator, resolve, resolve_aovs
import oracle as O

pytestmark = pytest.mark.gpu


def _film(device, params, spp, seed_offset=0):
    hs = HostScene.generate("cbox", params)
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=spp)
    job.seed_offset = seed_offset
    out, _ = device.render(ds, [job], want_aov=True)
    g, aov = out[0]
    av = resolve_aovs(g, aov)
    return resolve(g), av["albedo"], av["normal"]


@pytest.mark.parametrize("guides", ["none", "albedo", "both"])
@pytest.mark.parametrize("size", [(64, 48), (37, 29), (33, 17), (5, 70), (1, 1)])
def test_denoise_matches_oracle(device, guides, size):
    """Also at sizes that are not multiples of the a-trous steps (1..16), so
    sub-lattices of unequal width, lattice origins near the edge and the
    LDS tiles' halo bounds are compared with the oracle."""
    w, h = size
    c, a, n = _film(device, f"width={w};height={h}", 4)
    ga = a if guides in ("albedo", "both") else None
    gn = n if guides == "both" else None
    print(f"denoise {w}x{h} guides={guides}", flush=True)
    got = device.denoise(c, ga, gn)
    want = O.denoise(c, ga, gn)
    # same float operations in the same order; exp differs in the last bits
    assert np.allclose(got, want, rtol=2e-5, atol=1e-7)


def test_denoise_exact_hdr_scaling(device):
    c, a, n = _film(device, "width=40;height=40", 2)
    d = device.denoise(c, a, n)
    assert np.array_equal(device.denoise(c * 8.0, a, n), d * 8.0)
    assert np.array_equal(device.denoise(c * 0.25, a, n), d * 0.25)


def test_denoise_reduces_error(device):
    """A 4-spp Cornell box denoised is closer to a 1024-spp render than the
    4-spp film itself by more than 2x in relative MSE."""
    ref, _, _ = _film(device, "width=96;height=96", 1024, seed_offset=1 << 40)
    c, a, n = _film(device, "width=96;height=96", 4)
    d = device.denoise(c, a, n)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))

    print(f"relMSE noisy {relmse(c):.5f} denoised {relmse(d):.5f}", flush=True)
    assert relmse(d) < 0.5 * relmse(c)


def test_integrator_denoise_equals_device_denoise(device):
    hs = HostScene.generate("cbox", "width=48;height=40;spp=4")
    it = Integrator(device, "path", max_depth=8, passes_per_callback=4)
    it.set_aovs(True)
    it.render(hs)
    before, _ = it.film()
    av = it.aovs()
    it.denoise()
    after, _ = it.film()
    assert np.array_equal(after, device.denoise(before, av["albedo"], av["normal"]))
    # without AOVs the radiance alone is filtered
    it2 = Integrator(device, "path", max_depth=8, passes_per_callback=4)
    it2.render(hs)
    it2.denoise()
    assert np.array_equal(it2.film()[0], device.denoise(before))


def test_denoise_rejects_bad_sizes(device):
    with pytest.raises(GndError):
        device.denoise(np.zeros((0, 4, 3), np.float32))
