"""GPU path and volumetric-path integrators vs the CPU oracle.

The GPU follows the reference's per-path random-number order, so most
samples reproduce the CPU samples at the same seeds; transcendental
functions (ROCm device libm vs glibc) differ in the last bits, so a few
paths diverge after a bounce.  Agreement is checked two ways:
  * single-sample replay (same seeds): at 1 spp with the box filter every
    film pixel is one path's value, and the large majority must equal the
    oracle's;
  * statistics (independent seeds: the GPU renders other sample indices than
    the CPU, so the two images share no path): per-pixel Welch t-test over
    the per-pass squares, with the fraction of failing pixel-channels inside
    the binomial 99.9% bound of 1%, and relative MSE against a 16384-spp CPU
    render (test_equal_image_error).
"""
import numpy as np
import pytest
from scipy import stats as sst

from gianduia_amd import DeviceScene, GndError, HostScene, Job, camera_lookat, resolve, resolve_aovs
from gianduia_amd import _abi as A
from oracle import RefScene

pytestmark = pytest.mark.gpu


# sample-index offset of the GPU side of the statistical comparisons: the CPU
# renders sample indices [0, spp*W*H), the GPU [GPU_STREAM, GPU_STREAM + spp*W*H)
GPU_STREAM = 1 << 36


def _render_pair(device, name, params, spp, squares=True):
    """GPU and oracle renders of one generated scene; with squares (the
    statistical comparisons) the GPU uses independent random streams."""
    hs = HostScene.generate(name, params)
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    cam = hs.camera()
    job = hs.default_job(spp=spp, seed_offset=GPU_STREAM if squares else 0)
    out, st = device.render(ds, [job], want_sq=squares)
    if squares:
        g, gsq = out[0]
        r, rsq, _ = ref.render(threads=16, spp=spp, squares=True, width=cam.width, height=cam.height)
        return g, gsq, r, rsq, st
    r, _ = ref.render(threads=16, spp=spp, width=cam.width, height=cam.height)
    return out[0], None, r, None, st


def _pixel_moments(f, sq, n):
    """Per pixel: the mean of the n per-pass values (the resolved value,
    sum(w L) / sum(w); exact for the box filter, the ratio estimate for the
    Gaussian one) and their sample variance from the per-pass squares (the
    film's squares are of each pass's resolved value)."""
    m = f[..., :3] / np.maximum(f[..., 3:4], 1e-30)
    v = np.maximum(sq / n - m * m, 0) * n / (n - 1)
    return m, v


def _welch(g, gsq, r, rsq, n):
    gm, gv = _pixel_moments(g, gsq, n)
    rm, rv = _pixel_moments(r, rsq, n)
    with np.errstate(all="ignore"):
        _, p = sst.ttest_ind_from_stats(gm, np.sqrt(gv), n, rm, np.sqrt(rv), n, equal_var=False)
    # no variance on either side (e.g. pixels that see only black): equal
    # means pass, unequal ones fail
    p = np.where(np.isnan(p), np.where(np.isclose(gm, rm, rtol=1e-5, atol=1e-8), 1.0, 0.0), p)
    frac = float((p < 0.01).mean())
    # independent tests: every pixel-channel for the box filter; for the
    # Gaussian filter (radius 2) a sample reaches a 4 x 4 footprint of pixels,
    # so neighbouring pixels are correlated: one test per footprint
    box = np.allclose(g[..., 3], n) and np.allclose(r[..., 3], n)
    tests = p.size if box else p.size / 16
    bound = 0.01 + 3.1 * np.sqrt(0.01 * 0.99 / tests)
    return frac, bound


def _check_stats(g, gsq, r, rsq, n, label, mean_tol=0.03):
    # filter weights: exact for the box filter (n per pixel), statistical for
    # the Gaussian one (its weights depend on the jitter of each sample)
    assert abs(g[..., 3].mean() - r[..., 3].mean()) < 0.01 * r[..., 3].mean(), f"{label}: filter weights differ"
    frac, bound = _welch(g, gsq, r, rsq, n)
    print(f"{label}: welch fail fraction {frac:.4f} (bound {bound:.4f})", flush=True)
    assert frac <= bound, label
    # image means: within mean_tol, or within 5 standard errors of their
    # difference (from the per-pixel variances; x4 for the Gaussian filter,
    # whose neighbouring pixels share samples)
    img_g, img_r = resolve(g), resolve(r)
    _, gv = _pixel_moments(g, gsq, n)
    _, rv = _pixel_moments(r, rsq, n)
    shared = 1.0 if np.allclose(r[..., 3], n) else 4.0
    se = shared * float(np.sqrt((gv.sum() + rv.sum()) / n)) / gv.size
    diff = abs(float(img_g.mean()) - float(img_r.mean()))
    rel = diff / float(img_r.mean())
    print(f"{label}: mean {img_g.mean():.5f} vs {img_r.mean():.5f} (rel {rel:.4f}, {diff / max(se, 1e-30):.2f} s.e.)",
          flush=True)
    assert rel < mean_tol or diff < 5 * se, label


def _replay_fraction(g, r):
    same = np.all(np.isclose(g[..., :3], r[..., :3], rtol=1e-4, atol=1e-6), axis=-1)
    return float(same.mean())


# ---------------------------------------------------------------- Cornell box (diffuse, path)
def test_cbox_matches_oracle(device):
    print("render cbox 128x128 16spp", flush=True)
    g, gsq, r, rsq, _ = _render_pair(device, "cbox", "width=128;height=128", 16)
    _check_stats(g, gsq, r, rsq, 16, "cbox")


@pytest.mark.parametrize("name,params,min_frac", [
    ("cbox", "width=96;height=96", 0.999),
    # GGX conductor / glass / plastic + thin lens
    ("scaled", "level=4;instances=12;width=96;height=54", 0.99),
    # homogeneous medium in a glass sphere, no grid (fused shade kernel)
    ("volumetric", "grid=0;width=96;height=64", 0.995),
    # + the index-material box holding a 32^3 density grid (split kernels)
    ("volumetric", "grid=32;width=96;height=64", 0.99),
    # the full 256^3 grid of the benchmark workload: 8^3 voxels per macro cell
    ("volumetric", "grid=256;width=80;height=45", 0.99),
    # the full scaled-geometry workload's scene (1.31M-tri icosphere + 64 instances)
    ("scaled", "width=96;height=54", 0.99),
])
@pytest.mark.parametrize("tail", [0, -1])
def test_single_sample_replay(device, name, params, min_frac, tail):
    """At 1 spp each pixel holds one path: most must equal the oracle's path,
    with the tail kernel finishing the wave (tail 0, the default) and with
    the bounce kernels throughout (tail -1)."""
    print(f"replay {name} {params} tail {tail}", flush=True)
    old = device.tuning()
    device.set_tuning(tail_paths=tail)
    try:
        g, _, r, _, _ = _render_pair(device, name, params, 1, squares=False)
    finally:
        device.set_tuning(**old)
    frac = _replay_fraction(g, r)
    print(f"{name} {params}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert frac >= min_frac


@pytest.mark.parametrize("name,params", [("cbox", "width=48;height=48"),
                                         ("volumetric", "grid=32;width=40;height=32")])
def test_equal_image_error(device, name, params):
    """Equal image error at equal spp: the relative MSE of 16-spp GPU renders
    against a 16384-spp CPU render is within 2% of that of 16-spp CPU renders.
    All three use disjoint sample indices (independent streams); each side
    averages 32 independent renders so that the comparison is not noise."""
    n, trials, ref_spp = 16, 32, 16384
    hs = HostScene.generate(name, params)
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    cam = hs.camera()
    W, H = cam.width, cam.height
    span = n * W * H  # sample indices one render uses
    print(f"equal image error {name}: CPU reference {ref_spp} spp", flush=True)
    hi, _ = ref.render(threads=16, spp=ref_spp, seed_offset=1 << 40, width=W, height=H)
    r = resolve(hi)

    def relmse(x):
        img = resolve(x)
        return float(np.mean((img - r) ** 2 / (r * r + 1e-2)))

    jobs = [hs.default_job(spp=n, seed_offset=(1 << 38) + k * span) for k in range(trials)]
    gpu, _ = device.render(ds, jobs)
    e_gpu = float(np.mean([relmse(f) for f in gpu]))
    e_cpu = float(np.mean([relmse(ref.render(threads=16, spp=n, seed_offset=(1 << 39) + k * span, width=W,
                                             height=H)[0]) for k in range(trials)]))
    print(f"{name}: mean relMSE over {trials} renders: GPU {e_gpu:.6f} CPU {e_cpu:.6f}", flush=True)
    assert e_gpu <= 1.02 * e_cpu


# ---------------------------------------------------------------- AOVs (film.h:13-69)
@pytest.mark.parametrize("front,rough,integrator,depth", [
    ("conductor", 0.0, A.GND_INTEGRATOR_PATH, 4), ("conductor", 0.3, A.GND_INTEGRATOR_PATH, 4),
    ("conductor", 0.0, A.GND_INTEGRATOR_VOLPATH, 1), ("glass", 0.2, A.GND_INTEGRATOR_VOLPATH, 4),
    (None, 0.0, A.GND_INTEGRATOR_PATH, 4)])
def test_aov_mirror_scenes_equal_oracle(device, front, rough, integrator, depth):
    """The AOV film of the closed-form mirror scenes (specular chain, first-hit
    values, last-bounce rule, rough glass, misses) equals the oracle's bit for
    bit at 1 spp, and so does the radiance weight."""
    from scenes import mirror_scene
    d = mirror_scene(front, rough, integrator, max_depth=depth)
    hs = HostScene.from_desc(d)
    ds = DeviceScene(device, hs)
    print(f"aov mirror {front} {rough} integrator {integrator} depth {depth}", flush=True)
    out, _ = device.render(ds, [hs.default_job(spp=1)], want_aov=True)

    g, gaov = out[0]
    r, _, raov = RefScene(d).render(threads=4, spp=1, width=16, height=16, aovs=True)
    assert np.array_equal(g[..., 3], r[..., 3])
    assert np.array_equal(gaov, raov)


@pytest.mark.parametrize("name,params,spp,min_frac", [
    ("cbox", "width=64;height=48", 1, 0.999),
    ("cbox", "width=48;height=40;filter=gaussian", 3, 0.99),
    ("scaled", "level=3;instances=6;width=64;height=40", 1, 0.99),
    ("volumetric", "grid=16;width=48;height=32", 2, 0.99),
])
def test_aov_films_match_oracle(device, name, params, spp, min_frac):
    """AOV films of generated scenes against the oracle's on the same seeds:
    the same paths give the same records, so (as for radiance) nearly every
    pixel's nine sums agree; camera rays through a thin lens and GGX sampling
    use device libm, so a few may differ in the last bits."""
    hs = HostScene.generate(name, params)
    ds = DeviceScene(device, hs)
    cam = hs.camera()
    print(f"aov {name} {params} {spp} spp", flush=True)
    out, _ = device.render(ds, [hs.default_job(spp=spp)], want_aov=True)

    g, gaov = out[0]
    r, _, raov = RefScene(hs.desc()).render(threads=16, spp=spp, width=cam.width, height=cam.height, aovs=True)
    assert np.allclose(g[..., 3], r[..., 3], rtol=1e-5)
    same = np.all(np.isclose(gaov, raov, rtol=1e-4, atol=1e-6), axis=-1)
    print(f"{name}: {same.mean():.4f} of pixels match the oracle's AOVs", flush=True)
    assert same.mean() >= min_frac
    # the radiance film is unchanged by asking for AOVs
    plain, _ = device.render(ds, [hs.default_job(spp=spp)])
    assert np.array_equal(plain[0], g)
    av = resolve_aovs(g, gaov)
    if spp == 1 and "gaussian" not in params:
        # one sample per pixel: a hit in front of the camera or the miss value
        assert np.all((av["depth"] > 0) | (av["depth"] == -1.0))
    assert np.all(av["metallic"] >= 0) and np.all(av["metallic"] <= 1)


# ---------------------------------------------------------------- materials and thin lens (path)
def test_scaled_materials_match_oracle(device):
    print("render scaled level=4 instances=12 96x54 32spp", flush=True)
    g, gsq, r, rsq, st = _render_pair(device, "scaled", "level=4;instances=12;width=96;height=54", 32)
    assert st["nan_samples"] == 0
    _check_stats(g, gsq, r, rsq, 32, "scaled")


def test_gaussian_filter_matches_oracle(device):
    print("render cbox gaussian filter 64x64 32spp", flush=True)
    g, gsq, r, rsq, _ = _render_pair(device, "cbox", "width=64;height=64;filter=gaussian", 32)
    # Gaussian splat: every pixel gathers from up to 5x5 samples per pass
    assert np.all(g[..., 3] > 0)
    _check_stats(g, gsq, r, rsq, 32, "gaussian")


# ---------------------------------------------------------------- point / spot / directional / environment lights
LIGHT_SCENES = [
    # all four kinds beside the area light; env seen through the open front
    ("cbox", "width=64;height=64;lights=point,spot,sun,env;envres=32"),
    # outdoor: sun + sky over GGX conductor / glass / plastic, thin lens
    ("scaled", "level=4;instances=12;width=96;height=54;lights=sun,env;envres=64"),
    # volpath: NEE from inside the media (phase-function MIS) to every kind
    ("volumetric", "grid=32;width=80;height=48;lights=point,spot,env;envres=16"),
]


@pytest.mark.parametrize("name,params", LIGHT_SCENES)
def test_lights_single_sample_replay(device, name, params):
    """Delta-light NEE (MIS weight 1), environment importance sampling and
    the environment on escape (MIS against the BSDF pdf) follow the oracle's
    paths: at 1 spp nearly every pixel equals the oracle's."""
    print(f"replay lights {name} {params}", flush=True)
    g, _, r, _, st = _render_pair(device, name, params, 1, squares=False)
    assert st["nan_samples"] == 0
    frac = _replay_fraction(g, r)
    print(f"{name} {params}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert frac >= 0.99


@pytest.mark.parametrize("name,params", LIGHT_SCENES)
def test_lights_match_oracle(device, name, params):
    print(f"render lights {name} {params} 16spp", flush=True)
    g, gsq, r, rsq, st = _render_pair(device, name, params, 16)
    assert st["nan_samples"] == 0 and st["rays_shadow"] > 0
    _check_stats(g, gsq, r, rsq, 16, f"lights {name}")


def test_environment_miss_albedo_aov(device):
    """A camera ray that escapes records the environment's value as its
    albedo AOV (the miss G-buffer write, path.cpp:44-51); the oracle's AOV
    film is matched pixel for pixel."""
    hs = HostScene.generate("cbox", "width=48;height=48;lights=env;envres=16")
    ds = DeviceScene(device, hs)
    cam = hs.camera()
    out, _ = device.render(ds, [hs.default_job(spp=1)], want_aov=True)
    g, gaov = out[0]
    r, _, raov = RefScene(hs.desc()).render(threads=16, spp=1, width=cam.width, height=cam.height, aovs=True)
    same = np.all(np.isclose(gaov, raov, rtol=1e-4, atol=1e-6), axis=-1)
    print(f"env miss AOVs: {same.mean():.4f} of pixels match", flush=True)
    assert same.mean() >= 0.999
    av = resolve_aovs(g, gaov)
    miss = av["depth"] == -1.0
    assert miss.any() and np.all(av["albedo"][miss].max(axis=-1) > 0)


# ---------------------------------------------------------------- direct-lighting and AO integrators
INTEGRATOR_SCENES = [
    ("cbox", "width=64;height=64;integrator=direct"),
    # every emitter kind; the environment only through the light sample
    ("cbox", "width=64;height=64;integrator=direct;lights=point,spot,sun,env;envres=32"),
    ("scaled", "level=4;instances=12;width=96;height=54;integrator=direct"),
    # null-BSDF box: it occludes shadow rays (general shadow kernel, opaque rays)
    ("volumetric", "grid=32;width=80;height=48;integrator=direct"),
    ("cbox", "width=64;height=64;integrator=ao;aolength=0.5"),
    ("scaled", "level=4;instances=12;width=96;height=54;integrator=ao;aolength=1.5"),
    ("volumetric", "grid=32;width=80;height=48;integrator=ao"),
]


@pytest.mark.parametrize("name,params", INTEGRATOR_SCENES)
def test_direct_ao_single_sample_replay(device, name, params):
    """DirectLight (direct.cpp:10-76) and AmbientOcclusion (ao.cpp:14-40) on
    the wavefront loop: at 1 spp nearly every pixel equals the oracle's."""
    print(f"replay {name} {params}", flush=True)
    g, _, r, _, st = _render_pair(device, name, params, 1, squares=False)
    assert st["nan_samples"] == 0
    frac = _replay_fraction(g, r)
    print(f"{name} {params}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert frac >= 0.99


@pytest.mark.parametrize("name,params", INTEGRATOR_SCENES)
def test_direct_ao_match_oracle(device, name, params):
    print(f"render {name} {params} 16spp", flush=True)
    g, gsq, r, rsq, st = _render_pair(device, name, params, 16)
    assert st["nan_samples"] == 0
    _check_stats(g, gsq, r, rsq, 16, f"{name} {params}")


def test_direct_ao_leave_aov_defaults(device):
    """The direct and AO integrators write no G-buffer, so their AOV films
    resolve to the AOVRecord defaults (film.h:13-19)."""
    for integ in ("direct", "ao"):
        hs = HostScene.generate("cbox", f"width=32;height=24;integrator={integ}")
        ds = DeviceScene(device, hs)
        out, _ = device.render(ds, [hs.default_job(spp=2)], want_aov=True)
        g, gaov = out[0]
        av = resolve_aovs(g, gaov)
        assert np.all(av["depth"] == -1.0) and not av["albedo"].any() and not av["normal"].any()
        assert not av["roughness"].any() and not av["metallic"].any()


# ---------------------------------------------------------------- media (volpath)
@pytest.mark.parametrize("params", ["grid=0;width=80;height=48", "grid=32;width=80;height=48"])
def test_volumetric_matches_oracle(device, params):
    print(f"render volumetric {params} 32spp", flush=True)
    g, gsq, r, rsq, st = _render_pair(device, "volumetric", params, 32)
    assert st["rays_shadow"] > 0
    _check_stats(g, gsq, r, rsq, 32, f"volumetric {params}")


# ---------------------------------------------------------------- batching
def test_batch_equals_solo(device):
    """Jobs rendered together in one call give the same films, bit for bit,
    as each job rendered alone (mixed integrators, poses, sizes, spp, seeds)."""
    hs = HostScene.generate("volumetric", "grid=16;width=48;height=32")
    ds = DeviceScene(device, hs)
    cam = hs.camera()
    c2 = camera_lookat(cam, (0.4, 0.2, 3.5), (0.0, -0.2, 0.0))
    c2.width, c2.height = 40, 24
    c3 = camera_lookat(cam, (-0.5, 0.0, 3.0), (0.1, 0.0, 0.0))
    c3.width, c3.height = 17, 13
    jobs = [Job(cam, 3, 12, A.GND_INTEGRATOR_VOLPATH, 0),
            Job(c2, 5, 8, A.GND_INTEGRATOR_PATH, 7),
            Job(c3, 2, 20, A.GND_INTEGRATOR_VOLPATH, 3),
            Job(c2, 3, 0, A.GND_INTEGRATOR_DIRECT, 5),
            Job(c3, 4, 0, A.GND_INTEGRATOR_AO, 1, 0.7)]
    together, st = device.render(ds, jobs)
    assert st["samples"] == 48 * 32 * 3 + 40 * 24 * 5 + 17 * 13 * 2 + 40 * 24 * 3 + 17 * 13 * 4
    for i, j in enumerate(jobs):
        alone, _ = device.render(ds, [j])
        assert np.array_equal(alone[0], together[i]), f"job {i} differs when batched"


@pytest.mark.parametrize("scene,params,spp", [
    ("scaled", "level=4;instances=12;width=96;height=54", 4),
    ("volumetric", "grid=32;width=64;height=40", 4),
])
def test_lean_and_general_shading_kernels_equal(device, scene, params, spp):
    """A path / volpath job without AOVs on a scene lit by area emitters only
    runs the lean shading, scatter and tail kernels; the same job with AOVs
    runs the general ones.  The radiance films are the same bits, with the
    tail kernel on (the default) and off."""
    hs = HostScene.generate(scene, params)
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=spp)
    old = device.tuning()
    try:
        for tail in (old["tail_paths"], -1):
            device.set_tuning(tail_paths=tail)
            print(f"lean vs general {scene} {params} tail {tail}", flush=True)
            lean, _ = device.render(ds, [job])
            general, _ = device.render(ds, [job], want_aov=True)
            assert np.array_equal(lean[0], general[0][0]), f"{scene}: lean and general kernels differ (tail {tail})"
    finally:
        device.set_tuning(tail_paths=old["tail_paths"])


def test_batch_wave_split_equals_one_wave(device):
    """A queue too small for the whole batch splits it into several waves
    (whole passes each); the films are unchanged."""
    hs = HostScene.generate("cbox", "width=64;height=48")
    ds = DeviceScene(device, hs)
    jobs = [hs.default_job(spp=4), hs.default_job(spp=3, seed_offset=11)]
    a, _ = device.render(ds, jobs)
    old = device.tuning()
    device.set_tuning(queue_capacity=64 * 48)
    try:
        b, st = device.render(ds, jobs)
    finally:
        device.set_tuning(queue_capacity=old["queue_capacity"])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_seed_offset_changes_samples(device):
    hs = HostScene.generate("cbox", "width=32;height=32")
    ds = DeviceScene(device, hs)
    a, _ = device.render(ds, [hs.default_job(spp=2)])
    b, _ = device.render(ds, [hs.default_job(spp=2, seed_offset=2)])
    assert not np.array_equal(a[0], b[0])
    # the sampler index is pixelIdx + pass * W * H + seed offset, on both sides
    ref = RefScene(hs.desc())
    r, _ = ref.render(threads=4, spp=2, seed_offset=2, width=32, height=32)
    assert _replay_fraction(b[0] / 2, r / 2) > 0.95


# ---------------------------------------------------------------- launch shape, autotune, cancel
@pytest.mark.parametrize("params", ["level=4;instances=12;width=64;height=40",
                                    "level=3;instances=6;width=48;height=32;lights=sun,env;integrator=direct"])
def test_tunings_give_identical_films(device, params):
    """Block size, LDS stack depth, wave capacity and refill thresholds change
    scheduling only."""
    hs = HostScene.generate("scaled", params)
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=2)
    old = device.tuning()
    films = []
    try:
        for b, l, q, r, sr, tp in [(64, 8, 1 << 20, 0, 0, -1), (128, 16, 64 * 40, 1, 64, 0),
                                   (256, 24, 1 << 20, 64, 1, 500), (128, 32, 3000, 40, 24, -1)]:
            device.set_tuning(block_size=b, stack_lds=l, queue_capacity=q, refill_lanes=r, shadow_refill_lanes=sr,
                              tail_paths=tp)
            films.append(device.render(ds, [job])[0][0])
    finally:
        device.set_tuning(**old)
    for f in films[1:]:
        assert np.array_equal(films[0], f)


@pytest.mark.parametrize("name,params,spp,aov", [
    ("cbox", "width=64;height=48", 3, True),
    ("volumetric", "grid=16;width=48;height=32", 3, True),              # split mode: grid media, Tr shadows
    ("volumetric", "grid=0;width=48;height=32", 2, False),              # homogeneous medium, fused shading
    ("scaled", "level=4;instances=12;width=64;height=40", 2, True),     # GGX, instances, thin lens
    ("cbox", "width=48;height=40;integrator=direct;lights=point,spot,sun,env;envres=16", 3, False),
    ("volumetric", "grid=16;width=40;height=32;integrator=ao", 2, False),
    ("cbox", "width=48;height=40;cammove=0.3;camturn=10", 2, False),    # camera motion
    ("scaled", "level=3;instances=6;width=48;height=32;instmove=0.3", 2, False),  # moving primitives
])
def test_tail_kernel_identical_films(device, name, params, spp, aov):
    """The tail kernel (the last bounces of a wave in one persistent launch)
    gives the bounce kernels' bits: no tail, a tail from a few hundred paths
    left, and the tail for everything after the camera rays."""
    hs = HostScene.generate(name, params)
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=spp)
    old = device.tuning()
    out = []
    try:
        for tp in (-1, 300, 1 << 30):
            device.set_tuning(tail_paths=tp)
            films, st = device.render(ds, [job], want_aov=aov)
            out.append((films[0], st))
    finally:
        device.set_tuning(**old)
    assert out[0][1]["ms_tail"] == 0 and out[2][1]["ms_tail"] > 0
    assert out[0][1]["rays_closest"] == out[2][1]["rays_closest"]
    assert out[0][1]["rays_shadow"] == out[2][1]["rays_shadow"]
    for f, _ in out[1:]:
        if aov:
            assert np.array_equal(out[0][0][0], f[0]) and np.array_equal(out[0][0][1], f[1])
        else:
            assert np.array_equal(out[0][0], f)


@pytest.mark.parametrize("name,params,spp", [("cbox", "width=64;height=48", 3),
                                             ("volumetric", "grid=16;width=40;height=32", 2),
                                             ("scaled", "level=3;instances=4;width=48;height=32", 2),
                                             ("cbox", "width=48;height=40;integrator=direct;lights=point,env", 2),
                                             ("volumetric", "grid=16;width=40;height=32;integrator=ao", 2)])
def test_graph_bounce_chain_identical(device, name, params, spp):
    """The bounce chain captured in a hipGraph (no host sync between bounces)
    renders the same films as the synchronised loop, replay after replay."""
    hs = HostScene.generate(name, params)
    ds = DeviceScene(device, hs)
    jobs = [hs.default_job(spp=spp), hs.default_job(spp=1, seed_offset=5)]
    ref, st0 = device.render(ds, jobs)
    old = device.tuning()
    try:
        # the default tail (everything after the camera rays), a tail after a
        # few bounces, no tail: the first call of each learns the chain, the
        # next two replay it
        for tail in (0, 200, -1):
            device.set_tuning(use_graph=True, tail_paths=tail)
            for _ in range(3):
                got, st = device.render(ds, jobs)
                for a, b in zip(ref, got):
                    assert np.array_equal(a, b)
                assert st["rays_closest"] == st0["rays_closest"] and st["rays_shadow"] == st0["rays_shadow"]
    finally:
        device.set_tuning(**old)


def test_autotune_picks_valid_tuning(device):
    hs = HostScene.generate("cbox", "width=128;height=96")
    ds = DeviceScene(device, hs)
    old = device.tuning()
    try:
        t = device.autotune(ds, hs.default_job(spp=4))
        assert t["block_size"] in (64, 128, 256) and t["stack_lds"] in (8, 16, 24, 32)
        # 0 keeps the built-in default; otherwise one of the searched thresholds
        assert t["refill_lanes"] in (0, 12, 16, 24, 32, 40, 48, 56)
        assert t["shadow_refill_lanes"] in (0, 16, 24, 32, 40, 48, 56, 64)
        a, _ = device.render(ds, [hs.default_job(spp=2)])
    finally:
        device.set_tuning(**old)
    b, _ = device.render(ds, [hs.default_job(spp=2)])
    assert np.array_equal(a[0], b[0])


def test_cancel_stops_render(device):
    import threading
    import time
    hs = HostScene.generate("cbox", "width=1024;height=1024")
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=1024)  # seconds of work
    errors = []

    def run():
        try:
            device.render(ds, [job])
        except GndError as e:
            errors.append(e)

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.3)
    device.cancel()
    th.join(timeout=60)
    assert not th.is_alive()
    assert errors and errors[0].code == A.GND_E_CANCELLED
    # the context stays usable
    films, st = device.render(ds, [hs.default_job(spp=1)])
    assert st["samples"] == 1024 * 1024


# ---------------------------------------------------------------- determinism, stats, errors
def test_render_deterministic(device):
    hs = HostScene.generate("cbox", "width=64;height=64")
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=4)
    a, _ = device.render(ds, [job])
    b, _ = device.render(ds, [job])
    assert np.array_equal(a[0], b[0])


def test_render_stats(device):
    hs = HostScene.generate("cbox", "width=64;height=64")
    ds = DeviceScene(device, hs)
    films, st = device.render(ds, [hs.default_job(spp=2)])
    assert st["samples"] == 64 * 64 * 2
    assert st["rays_closest"] >= st["samples"]
    assert st["ms_trace"] + st["ms_tail"] > 0  # a small frame runs in the tail kernel after the camera rays


def test_render_rejects_bad_jobs(device):
    hs = HostScene.generate("cbox", "width=16;height=16")
    ds = DeviceScene(device, hs)
    j = hs.default_job(spp=1)
    j.camera.width = 0
    with pytest.raises(GndError):
        device.render(ds, [j])


# ---------------------------------------------------------------- camera motion blur, rolling shutter
CAMERA_MOTION = [
    ("cbox", "width=64;height=64;cammove=0.3;camturn=8"),                       # pinhole, slide + turn
    ("scaled", "level=3;instances=6;width=64;height=40;cammove=0.2;camturn=-5;rolling=1"),  # thin lens
    ("volumetric", "grid=16;width=48;height=32;camturn=12"),
]


@pytest.mark.parametrize("name,params", CAMERA_MOTION)
def test_camera_motion_matches_oracle(device, name, params):
    """A moving camera (AnimatedTransform: decomposed keyframes, mix / slerp
    per sample time) and the thin lens's rolling shutter: single-sample
    replay against the oracle and per-pixel statistics at 16 spp."""
    print(f"camera motion {name} {params}", flush=True)
    g, _, r, _, st = _render_pair(device, name, params, 1, squares=False)
    frac = _replay_fraction(g, r)
    print(f"{name} {params}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert st["nan_samples"] == 0 and frac >= 0.98
    g, gsq, r, rsq, _ = _render_pair(device, name, params, 16)
    _check_stats(g, gsq, r, rsq, 16, f"motion {name}")


PRIM_MOTION = [
    ("cbox", "width=64;height=64;movebox=0.3"),                                   # a sliding, turning box
    ("scaled", "level=3;instances=6;width=64;height=40;instmove=0.3"),           # moving instances, thin lens
    ("cbox", "width=48;height=48;movebox=0.3;integrator=direct;lights=point,env"),
    ("cbox", "width=48;height=48;movebox=0.3;integrator=ao;aolength=0.5"),
    ("volumetric", "grid=16;width=48;height=32;movebox=0.25"),
]


@pytest.mark.parametrize("name,params", PRIM_MOTION)
def test_moving_primitives_match_oracle(device, name, params):
    """Primitives with toWorldEnd: instances intersected at the ray's time
    (the interpolated transform's inverse per instance visit), motion-bounds
    TLAS, time-dependent shading frames and area-light samples, and the
    reference's time rules (time 0 for point-light, medium-NEE, direct BSDF
    and AO rays and for the emission MIS pdf)."""
    print(f"moving primitives {name} {params}", flush=True)
    g, _, r, _, st = _render_pair(device, name, params, 1, squares=False)
    frac = _replay_fraction(g, r)
    print(f"{name} {params}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert st["nan_samples"] == 0 and frac >= 0.98
    g, gsq, r, rsq, _ = _render_pair(device, name, params, 16)
    _check_stats(g, gsq, r, rsq, 16, f"moving {name}")


# ---------------------------------------------------------------- XML scene with every emitter kind
XML_LIGHTS = """<scene>
  <camera type="perspective">
    <transform name="toWorld"><lookat origin="0, 1, 4" target="0, 0, 0" up="0, 1, 0"/></transform>
    <float name="fov" value="50"/>
    <integer name="width" value="64"/><integer name="height" value="48"/>
    <filter type="box"/>
  </camera>
  <sampler type="independent"><integer name="spp" value="1"/></sampler>
  <integrator type="{integrator}"><integer name="maxDepth" value="6"/></integrator>
  <emitter type="point"><point name="position" value="0.5, 1.5, 0.5"/><color name="power" value="8, 7, 6"/></emitter>
  <primitive>
    <transform name="toWorld"><scale x="6" y="1" z="6"/><translate x="0" y="-0.5" z="0"/></transform>
    <shape type="mesh"><string name="filename" value="plane.obj"/></shape>
    <material type="matte"><color name="albedo" value="0.6 0.6 0.6"/></material>
  </primitive>
  <primitive>
    <shape type="sphere"><float name="radius" value="0.5"/></shape>
    <material type="conductor"><float name="roughness" value="0.2"/></material>
  </primitive>
  <primitive>
    <transform name="toWorld"><translate x="-1.2" y="0" z="0.3"/></transform>
    <shape type="sphere"><float name="radius" value="0.4"/></shape>
    <material type="plastic"><float name="roughness" value="0.3"/></material>
    <emitter type="area"><color name="radiance" value="2, 1, 0.5"/></emitter>
  </primitive>
  <emitter type="spot"><point name="position" value="-1, 2, 1"/><vector name="direction" value="0.5, -1, -0.5"/>
    <float name="spot_size" value="40"/></emitter>
  <emitter type="directional"><vector name="direction" value="-0.3, -1, -0.2"/><color name="irradiance" value="0.8, 0.8, 0.7"/></emitter>
  <emitter type="environment"><string name="filename" value="sky.exr"/><float name="strength" value="0.7"/></emitter>
</scene>
"""


@pytest.mark.parametrize("integrator", ["path", "direct"])
def test_xml_scene_with_all_emitters_matches_oracle(device, tmp_path, integrator):
    """A scene file with point, area, spot, directional and environment
    emitters (the map a ZIP-compressed half-float OpenEXR file) loads through
    the XML loader and renders on the GPU as the oracle renders it."""
    import imgio
    rng = np.random.default_rng(11)
    H, W = 16, 32
    sky = {c: (0.3 + rng.random((H, W))).astype(np.float32) for c in "RGB"}
    sky["R"][4:6, 10:13] = 40.0                       # a hot spot for the importance sampling
    imgio.write_exr(tmp_path / "sky.exr", sky, ptype=1, compression=3)
    (tmp_path / "plane.obj").write_text("v -0.5 0 -0.5\nv 0.5 0 -0.5\nv 0.5 0 0.5\nv -0.5 0 0.5\n"
                                        "vn 0 1 0\nf 1//1 3//1 2//1\nf 1//1 4//1 3//1\n")
    (tmp_path / "scene.xml").write_text(XML_LIGHTS.replace("{integrator}", integrator))
    hs = HostScene.load_xml(str(tmp_path / "scene.xml"))
    assert hs.desc().contents.nlights == 4
    ds = DeviceScene(device, hs)
    out, st = device.render(ds, [hs.default_job(spp=1)])
    r, _ = RefScene(hs.desc()).render(threads=16, spp=1, width=64, height=48)
    frac = _replay_fraction(out[0], r)
    print(f"xml {integrator}: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert st["nan_samples"] == 0 and frac >= 0.99
    out16, _ = device.render(ds, [hs.default_job(spp=16)], want_sq=True)
    g, gsq = out16[0]
    r16, rsq, _ = RefScene(hs.desc()).render(threads=16, spp=16, squares=True, width=64, height=48)
    _check_stats(g, gsq, r16, rsq, 16, f"xml {integrator}")
