"""The Integrator plugin on the GPU (gnd_host_integrator_*): render(scene)
equals gnd_hip_render of the same job, the callback fires once per batch of
passes with the last pass index (integrator.h:147), cancel() from the
callback stops at the next batch, the scene's own integrator is used, and the
film is saved."""
import numpy as np
import pytest

from gianduia_amd import DeviceScene, GndError, HostScene, Integrator, resolve, resolve_aovs

pytestmark = pytest.mark.gpu


def test_integrator_equals_render(device):
    hs = HostScene.generate("cbox", "width=64;height=48;spp=4")
    it = Integrator(device, "path", max_depth=8, passes_per_callback=4)
    it.render(hs)
    img, passes = it.film()
    assert passes == 4
    ds = DeviceScene(device, hs)
    films, _ = device.render(ds, [hs.default_job(spp=4)])
    assert np.array_equal(img, resolve(films[0]))


def test_callbacks_per_batch_and_cancel(device):
    hs = HostScene.generate("cbox", "width=32;height=32;spp=5")
    it = Integrator(device, "path_hip", max_depth=8, passes_per_callback=2)
    seen = []
    it.set_callback(lambda idx, rgb: seen.append((idx, rgb.shape, float(rgb.mean()))))
    it.render(hs)
    assert [s[0] for s in seen] == [1, 3, 4]
    assert all(s[1] == (32, 32, 3) for s in seen)
    assert it.film()[1] == 5
    # batches of 2 + 2 + 1 passes accumulate on the device: the bits of one 5-pass render
    ds = DeviceScene(device, hs)
    assert np.array_equal(it.film()[0], resolve(device.render(ds, [hs.default_job(spp=5)])[0][0]))
    # cancel from inside the callback: the render stops after the first batch
    it2 = Integrator(device, "path", max_depth=8, passes_per_callback=2)
    it2.set_callback(lambda idx, rgb: it2.cancel())
    it2.render(hs)
    assert it2.film()[1] == 2


def test_scene_integrator_and_save(device, tmp_path):
    hs = HostScene.generate("volumetric", "grid=16;width=40;height=32;spp=2;depth=12")
    it = Integrator(device, scene=hs)          # <integrator type="volpath">
    it.render(hs)
    img, passes = it.film()
    ds = DeviceScene(device, hs)
    films, _ = device.render(ds, [hs.default_job()])
    assert passes == 2 and np.array_equal(img, resolve(films[0]))
    it.save(str(tmp_path / "out.png"))
    it.save(str(tmp_path / "out.pfm"))
    assert (tmp_path / "out.png").read_bytes()[:4] == b"\x89PNG"
    assert (tmp_path / "out.pfm").read_bytes()[:2] == b"PF"
    with pytest.raises(GndError):
        it.save(str(tmp_path / "out.bmp"))


def test_unknown_integrator(device):
    with pytest.raises(GndError):
        Integrator(device, "photonmapper")


@pytest.mark.parametrize("name,params", [("direct", "integrator=direct"), ("ao", "integrator=ao;aolength=0.7")])
def test_direct_and_ao_plugins_equal_render(device, name, params):
    """The direct and ao integrators by name and from the scene (its AO
    length travels with the scene) render what gnd_hip_render renders."""
    hs = HostScene.generate("cbox", f"width=40;height=32;spp=3;{params}")
    ds = DeviceScene(device, hs)
    films, _ = device.render(ds, [hs.default_job()])
    # one batch of 3 passes, as the device render sums them (per-pass batches
    # would sum on the host in double and round differently)
    for it in (Integrator(device, name, passes_per_callback=3), Integrator(device, scene=hs, passes_per_callback=3)):
        it.render(hs)
        img, passes = it.film()
        assert passes == 3 and np.array_equal(img, resolve(films[0]))


def test_integrator_aovs_equal_device_render(device):
    """Film AOV images of the plugin (one batch) resolve the device's AOV sums
    exactly like resolve_aovs; without set_aovs the film has none."""
    hs = HostScene.generate("scaled", "level=3;instances=4;width=48;height=32;spp=3")
    it = Integrator(device, "path", max_depth=6, passes_per_callback=3)
    it.set_aovs(True)
    it.render(hs)
    got = it.aovs()
    ds = DeviceScene(device, hs)
    j = hs.default_job(spp=3)
    j.max_depth = 6
    out, _ = device.render(ds, [j], want_aov=True)

    g, gaov = out[0]
    want = resolve_aovs(g, gaov)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    it2 = Integrator(device, "path", max_depth=6)
    it2.render(hs)
    with pytest.raises(GndError):
        it2.aovs()


def test_host_scene_outlives_its_context():
    """A host scene whose device copy was made on a context that is shut down
    first: the shutdown frees that copy, so freeing the host scene afterwards
    does not touch the dead context (it used to synchronise its stream)."""
    from gianduia_amd import Device
    dev = Device(0)
    hs = HostScene.generate("cbox", "width=16;height=16;spp=1")
    it = Integrator(dev, "path", max_depth=2)
    it.render(hs)  # uploads the scene on dev's context and caches the copy
    it.close()
    dev.close()
    hs.free()


def test_integrator_saves_exr_and_aov_images(device, tmp_path):
    """Film::saveEXR and the AOV images (Image2D::saveEXR, "Y" for the scalar
    ones) read back equal to the film's resolved images."""
    from imgio import read_exr, read_pfm
    hs = HostScene.generate("cbox", "width=24;height=20;spp=2")
    it = Integrator(device, "path", max_depth=4, passes_per_callback=2)
    it.set_aovs(True)
    it.render(hs)
    img, _ = it.film()
    av = it.aovs()
    it.save(str(tmp_path / "beauty.exr"))
    _, ch = read_exr((tmp_path / "beauty.exr").read_bytes())
    assert np.array_equal(np.stack([ch["R"], ch["G"], ch["B"]], -1), img)
    for kind in ("albedo", "normal", "depth", "roughness", "metallic"):
        it.save_aov(kind, str(tmp_path / f"{kind}.exr"))
        it.save_aov(kind, str(tmp_path / f"{kind}.pfm"))
        _, ch = read_exr((tmp_path / f"{kind}.exr").read_bytes())
        want = av[kind] if av[kind].ndim == 3 else av[kind][..., None]
        got = np.stack([ch["R"], ch["G"], ch["B"]], -1) if "R" in ch else ch["Y"][..., None]
        assert np.array_equal(got, want), kind
        assert np.array_equal(read_pfm((tmp_path / f"{kind}.pfm").read_bytes()), want), kind
    with pytest.raises(GndError):
        it.save_aov("emission", str(tmp_path / "x.exr"))



def test_cancel_keeps_the_first_passes(device):
    """A render cancelled after k passes (from the callback, and from another
    thread) holds exactly the film of the first k passes: the same bits as a
    render of k passes."""
    import threading
    import time
    hs = HostScene.generate("cbox", "width=48;height=40;spp=64")
    ds = DeviceScene(device, hs)
    it = Integrator(device, "path", max_depth=8, passes_per_callback=1)
    it.set_callback(lambda idx, rgb: it.cancel() if idx == 2 else None)
    it.render(hs)
    img, k = it.film()
    assert k == 3
    films, _ = device.render(ds, [hs.default_job(spp=k)])
    assert np.array_equal(img, resolve(films[0]))
    # from another thread, at whatever pass the render has reached
    big = HostScene.generate("cbox", "width=256;height=256;spp=100000")
    it2 = Integrator(device, "path", max_depth=8, passes_per_callback=1)
    th = threading.Thread(target=lambda: it2.render(big))
    th.start()
    time.sleep(0.5)
    it2.cancel()
    th.join(timeout=120)
    assert not th.is_alive()
    img2, k2 = it2.film()
    assert 0 < k2 < 100000
    dsb = DeviceScene(device, big)
    films, _ = device.render(dsb, [big.default_job(spp=k2)])
    assert np.array_equal(img2, resolve(films[0]))


def test_render_saves_exr_and_png_at_the_end(device, tmp_path):
    """With an output set, render() ends by writing <output>.exr and
    <output>.png, as SamplerIntegrator::render does (integrator.h:152-153)."""
    import imgio
    hs = HostScene.generate("cbox", "width=32;height=24;spp=2")
    it = Integrator(device, "path", max_depth=8)
    it.set_output(str(tmp_path / "cbox"))
    it.render(hs)
    assert (tmp_path / "cbox.png").read_bytes()[:4] == b"\x89PNG"
    _, ch = imgio.read_exr((tmp_path / "cbox.exr").read_bytes())
    img, _ = it.film()
    assert np.array_equal(np.stack([ch["R"], ch["G"], ch["B"]], -1), img)


def test_nan_and_inf_samples_are_zeroed(device):
    """SamplerIntegrator::render's guard (integrator.h:137-141): a sample whose
    radiance has a NaN or an Inf counts once and adds nothing.  A small sphere
    emitting +inf sits in view beside a finite quad light: the film stays
    finite, the guard fires, and the GPU film equals the oracle's at the same
    seeds."""
    from oracle import RefScene
    from scenes import SceneBuilder, quad, trs
    b = SceneBuilder()
    wall = b.matte((0.6, 0.6, 0.6))
    b.prim(b.mesh(quad(-3, 3, -3, 3, -3.0, facing=+1)), material=wall)
    b.prim(b.mesh(quad(-1, 1, -1, 1, 2.5, facing=-1)), material=wall, radiance=(4.0, 4.0, 4.0))
    M, Mi = trs(trans=(0.6, 0.3, -2.0))
    b.prim(-1, M, Mi, material=b.matte(), sphere_radius=0.25, radiance=(np.inf, np.inf, np.inf))
    d = b.desc(width=48, height=48)
    d.max_depth = 4
    hs = HostScene.from_desc(d)
    ds = DeviceScene(device, hs)
    films, st = device.render(ds, [hs.default_job(spp=2)])
    g = films[0]
    assert np.all(np.isfinite(g)) and st["nan_samples"] > 0
    r, _ = RefScene(hs.desc()).render(threads=8, spp=2)
    assert np.all(np.isfinite(r))
    same = np.all(np.isclose(g[..., :3], r[..., :3], rtol=1e-4, atol=1e-6), axis=-1).mean()
    assert same >= 0.99, same
