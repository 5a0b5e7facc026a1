"""GPU traversal parity: gnd_hip_trace_closest / gnd_hip_trace_any against the
CPU oracle on seeded ray batches — bit-identical primitive ids, hit distances,
barycentrics and occlusion, plus reference-order visit counts."""
import numpy as np
import pytest

from gianduia_amd import DeviceScene, HostScene
from oracle import RefScene
from scenes import random_rays, soup_scene

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _compare_closest(g, r, label):
    assert np.array_equal(g["inst"], r["inst"]), f"{label}: instance ids differ at {np.nonzero(g['inst'] != r['inst'])[0][:10]}"
    hit = r["inst"] >= 0
    assert np.array_equal(g["prim"][hit], r["prim"][hit]), f"{label}: primitive ids differ"
    assert np.array_equal(_bits(g["t"][hit]), _bits(r["t"][hit])), f"{label}: hit distances differ"
    assert np.array_equal(_bits(g["u"][hit]), _bits(r["u"][hit])), f"{label}: u differs"
    assert np.array_equal(_bits(g["v"][hit]), _bits(r["v"][hit])), f"{label}: v differs"
    assert np.all(np.isinf(g["t"][~hit]))


@pytest.mark.parametrize("seed,instances,sphere", [(0, 1, False), (1, 4, True), (2, 16, False)])
def test_trace_closest_soup(device, seed, instances, sphere):
    b = soup_scene(seed=seed, n=3000, instances=instances, sphere=sphere)
    desc = b.desc()
    hs = HostScene.from_desc(desc)
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rays = random_rays(np.random.default_rng(100 + seed), 4096)
    print(f"trace_closest seed={seed} instances={instances} sphere={sphere}", flush=True)
    g, st = device.trace_closest(ds, rays)
    r = ref.trace(rays)
    _compare_closest(g, r, f"seed {seed}")
    assert (r["inst"] >= 0).sum() > 100


@pytest.mark.parametrize("block,lds", [(64, 8), (128, 16), (256, 32)])
def test_trace_closest_tunings(device, block, lds):
    hs = HostScene.from_desc(soup_scene(seed=5, n=5000, instances=3).desc())
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rays = random_rays(np.random.default_rng(7), 2048)
    old = device.tuning()
    device.set_tuning(block_size=block, stack_lds=lds)
    try:
        g, _ = device.trace_closest(ds, rays)
    finally:
        device.set_tuning(block_size=old["block_size"], stack_lds=old["stack_lds"])
    _compare_closest(g, ref.trace(rays), f"block {block} lds {lds}")


def test_trace_any_soup(device):
    b = soup_scene(seed=3, n=3000, instances=5, sphere=True)
    hs = HostScene.from_desc(b.desc())
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rays = random_rays(np.random.default_rng(9), 4096)
    rays["tmax"] = np.where(np.isinf(rays["tmax"]), np.float32(2.0), rays["tmax"]).astype(np.float32)
    occ, _ = device.trace_any(ds, rays)
    r = ref.trace(rays, any_hit=True)
    assert np.array_equal(occ, r)
    assert 0 < occ.sum() < len(occ)


def test_visit_counts_match_oracle(device):
    hs = HostScene.from_desc(soup_scene(seed=11, n=4000, instances=3).desc())
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rays = random_rays(np.random.default_rng(12), 1024)
    device.set_counters(True)
    try:
        g, st = device.trace_closest(ds, rays)
    finally:
        device.set_counters(False)
    r, cnt = ref.trace(rays, counts=True)
    _compare_closest(g, r, "counters")
    assert st["node_visits"] == int(cnt[:, 0].sum())
    assert st["pack_tests"] == int(cnt[:, 1].sum())
    assert st["inst_tests"] == int(cnt[:, 2].sum())


def test_trace_cbox_camera_rays(device):
    hs = HostScene.generate("cbox", "")
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rng = np.random.default_rng(4)
    n = 8192
    o = np.tile(np.array([0, 0, 3.9], np.float32), (n, 1))
    d = np.stack([rng.uniform(-0.15, 0.15, n), rng.uniform(-0.15, 0.15, n), -np.ones(n)], 1).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(), "dx": d[:, 0].copy(),
            "dy": d[:, 1].copy(), "dz": d[:, 2].copy(), "tmin": np.full(n, 1e-4, np.float32),
            "tmax": np.full(n, np.inf, np.float32)}
    g, _ = device.trace_closest(ds, rays)
    r = ref.trace(rays)
    _compare_closest(g, r, "cbox")
    assert (r["inst"] >= 0).all()


def test_trace_empty_batch(device):
    hs = HostScene.from_desc(soup_scene(seed=1, n=10).desc())
    ds = DeviceScene(device, hs)
    rays = {k: np.zeros(0, np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")}
    g, st = device.trace_closest(ds, rays)
    assert len(g["t"]) == 0


def _scene_rays(rng, n, radius=6.0, target=2.0):
    """Rays from a sphere of the given radius around the origin towards random
    points of [-target, target]^3 (most of them cross the scene)."""
    o = rng.normal(size=(n, 3))
    o = (o / np.linalg.norm(o, axis=1, keepdims=True) * radius).astype(np.float32)
    d = rng.uniform(-target, target, size=(n, 3)).astype(np.float32) - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(), "dx": d[:, 0].copy(),
            "dy": d[:, 1].copy(), "dz": d[:, 2].copy(), "tmin": np.full(n, 1e-4, np.float32),
            "tmax": np.full(n, np.inf, np.float32)}


def test_trace_full_scaled_scene(device):
    """The benchmark's own scene at full size (1,310,720-triangle icosphere at
    subdivision level 8 + 64 instances of a 10k-triangle sphere): 65,536
    seeded rays, bit-identical closest hits, reference-order visit counts and
    any-hit occlusion against the oracle."""
    hs = HostScene.generate("scaled", "")
    ds = DeviceScene(device, hs)
    ref = RefScene(hs.desc())
    rays = _scene_rays(np.random.default_rng(31), 65536)
    print("trace_closest on the full scaled scene", flush=True)
    device.set_counters(True)
    try:
        g, st = device.trace_closest(ds, rays)
    finally:
        device.set_counters(False)
    r, cnt = ref.trace(rays, counts=True)
    _compare_closest(g, r, "scaled full")
    assert (r["inst"] >= 0).sum() > 30000
    assert st["node_visits"] == int(cnt[:, 0].sum())
    assert st["pack_tests"] == int(cnt[:, 1].sum())
    assert st["inst_tests"] == int(cnt[:, 2].sum())
    short = dict(rays)
    short["tmax"] = np.random.default_rng(32).uniform(0.5, 8.0, 65536).astype(np.float32)
    occ, _ = device.trace_any(ds, short)
    assert np.array_equal(occ, ref.trace(short, any_hit=True))
    assert 0 < occ.sum() < len(occ)
