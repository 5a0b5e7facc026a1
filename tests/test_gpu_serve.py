"""Continuous-batching server (gnd_hip_server_*): jobs admitted between
bounces into free path slots give films bit-identical to rendering each job
alone with gnd_hip_render; requests that arrive while a long job renders are
admitted and finish first, on its scene and on others; cancel, wait time-outs
and release; a traversal stack overflow fails the served job as it fails the
solo render; scenes are evicted and re-uploaded under a small HBM budget;
oversize jobs are refused."""
import time

import numpy as np
import pytest

from gianduia_amd import DeviceScene, GndError, HostScene
from gianduia_amd import _abi as A
from gianduia_amd.serve import Server, poisson_load, random_job

pytestmark = pytest.mark.gpu


def _jobs(hs, n, seed):
    rng = np.random.default_rng(seed)
    return [random_job(rng, hs, res=(32, 96), spp=(1, 6)) for _ in range(n)]


def test_server_films_equal_solo_render(device):
    a = HostScene.generate("cbox", "width=48;height=40")
    b = HostScene.generate("volumetric", "grid=16;width=40;height=32")
    c = HostScene.generate("scaled", "level=3;instances=6;width=40;height=32")
    srv = Server(device, path_slots=20000, max_jobs=8)   # small pool: passes wait for slots
    try:
        ids = [srv.add_scene(h) for h in (a, b, c)]
        work = []
        for sid, hs, seed in ((ids[0], a, 1), (ids[1], b, 2), (ids[2], c, 3), (ids[0], a, 4)):
            for j in _jobs(hs, 4, seed):
                work.append((sid, hs, j))
        tickets = [(hs, j, srv.submit(sid, j)) for sid, hs, j in work]
        for hs, j, tk in tickets:
            st = srv.wait(tk, 120)
            assert st["state"] == 2 and st["passes_done"] == j.spp
        for hs, j, tk in tickets:
            ds = DeviceScene(device, hs)
            solo, _ = device.render(ds, [j])
            assert np.array_equal(srv.film(tk), solo[0]), "served film differs from the solo render"
            ds.free()
        st = srv.stats()
        assert st["samples"] == sum(j.camera.width * j.camera.height * j.spp for _, j, _ in tickets)
    finally:
        srv.close()


def test_server_aovs_equal_solo_render(device):
    """AOV films of served jobs equal the solo render's (continuous batching
    keeps per-path records and gathers them in pass order like radiance)."""
    a = HostScene.generate("cbox", "width=48;height=40")
    b = HostScene.generate("volumetric", "grid=16;width=40;height=32")
    srv = Server(device, path_slots=20000, max_jobs=8)
    try:
        ids = [srv.add_scene(h) for h in (a, b)]
        work = [(ids[0], a, j) for j in _jobs(a, 3, 7)] + [(ids[1], b, j) for j in _jobs(b, 3, 8)]
        tickets = [(hs, j, srv.submit(sid, j, want_aov=(k % 2 == 0))) for k, (sid, hs, j) in enumerate(work)]
        for hs, j, tk in tickets:
            assert srv.wait(tk, 120)["state"] == 2
        for hs, j, tk in tickets:
            ds = DeviceScene(device, hs)
            out, _ = device.render(ds, [j], want_aov=True)
            assert np.array_equal(srv.film(tk), out[0][0])
            if tk.aov is not None:
                assert np.array_equal(srv.aov(tk), out[0][1]), "served AOVs differ from the solo render"
            ds.free()
    finally:
        srv.close()


def test_server_direct_ao_and_lights_equal_solo_render(device):
    """Served jobs of every integrator, over a scene with point, spot, sun
    and sky lights, equal their solo renders bit for bit."""
    hs = HostScene.generate("cbox", "width=48;height=40;lights=point,spot,sun,env;envres=16")
    srv = Server(device, path_slots=12000, max_jobs=8)
    try:
        sid = srv.add_scene(hs)
        jobs = _jobs(hs, 8, 9)
        for k, j in enumerate(jobs):
            j.integrator = (A.GND_INTEGRATOR_PATH, A.GND_INTEGRATOR_DIRECT, A.GND_INTEGRATOR_AO)[k % 3]
            j.ao_length = 0.6
        tickets = [(j, srv.submit(sid, j)) for j in jobs]
        for j, tk in tickets:
            assert srv.wait(tk, 120)["state"] == 2
        ds = DeviceScene(device, hs)
        for j, tk in tickets:
            solo, _ = device.render(ds, [j])
            assert np.array_equal(srv.film(tk), solo[0]), f"integrator {j.integrator}: served film differs"
        ds.free()
    finally:
        srv.close()


def test_server_evicts_under_budget(device):
    scenes = [HostScene.generate("scaled", f"level=3;instances=4;seed={i};width=32;height=32") for i in range(3)]
    probe = DeviceScene(device, scenes[0])
    one = probe.nbytes
    probe.free()
    slots = 16384  # a pass of the largest job (96 x 96) fits
    # path state and one lane's queues (300 B per slot) plus room for one scene but not two
    srv = Server(device, path_slots=slots, max_jobs=4, hbm_budget=300 * slots + int(1.5 * one), lanes=1)
    try:
        ids = [srv.add_scene(h) for h in scenes]
        tickets = []
        for rnd in range(2):
            for sid, hs in zip(ids, scenes):
                j = _jobs(hs, 1, 10 * rnd + sid)[0]
                tickets.append((hs, j, srv.submit(sid, j)))
                srv.wait(tickets[-1][2], 60)
                assert srv.stats()["resident_scenes"] == 1
        for hs, j, tk in tickets:
            ds = DeviceScene(device, hs)
            assert np.array_equal(srv.film(tk), device.render(ds, [j])[0][0])
            ds.free()
    finally:
        srv.close()


def test_server_rejects_oversize_and_bad_scene(device):
    hs = HostScene.generate("cbox", "width=64;height=64")
    srv = Server(device, path_slots=2048, max_jobs=2)
    try:
        sid = srv.add_scene(hs)
        with pytest.raises(GndError) as e:
            srv.submit(sid, hs.default_job(spp=1))   # 4096 paths per pass > 2048 slots
        assert e.value.code == A.GND_E_NOMEM
        with pytest.raises(GndError):
            srv.submit(sid + 5, hs.default_job(spp=1))
    finally:
        srv.close()


def test_poisson_load_completes(device):
    pool = [HostScene.generate("procedural", f"tris={t};seed={i}") for i, t in enumerate((2e4, 2e5))]
    srv = Server(device, path_slots=2_000_000, max_jobs=64)
    try:
        ids = [srv.add_scene(h) for h in pool]
        r = poisson_load(srv, ids, qps=50, n_jobs=20, seed=3, res=(64, 256), spp=(1, 8))
        assert r["jobs"] == 20 and r["p99_ms"] >= r["p50_ms"] > 0
    finally:
        srv.close()


def _wait_until(pred, timeout_s=60.0):
    t0 = time.perf_counter()
    while not pred():
        assert time.perf_counter() - t0 < timeout_s, "condition not reached"
        time.sleep(0.001)


def test_server_overlaps_requests_with_a_running_job(device):
    """Continuous batching: while a long job A renders (its passes fill the
    slot pool), a job on another scene (B) and a job on A's scene (C) are
    admitted between bounces and finish while A is still running.  Every
    film equals its solo render bit for bit."""
    a = HostScene.generate("cbox", "width=64;height=64")
    b = HostScene.generate("scaled", "level=3;instances=6;width=48;height=40")
    srv = Server(device, path_slots=64 * 64 * 8, max_jobs=8)
    try:
        ia, ib = srv.add_scene(a), srv.add_scene(b)
        job_a = a.default_job(spp=4000)
        ta = srv.submit(ia, job_a)
        _wait_until(lambda: srv.poll(ta)["passes_done"] >= 8)
        job_b = _jobs(b, 1, 21)[0]
        job_c = a.default_job(spp=3)
        job_c.seed_offset = 7
        tb, tc = srv.submit(ib, job_b), srv.submit(ia, job_c)
        sb, sc = srv.wait(tb, 120), srv.wait(tc, 120)
        sa_mid = srv.poll(ta)
        assert sa_mid["state"] == 1, "the long job finished before the later requests: no overlap shown"
        assert sa_mid["passes_done"] < job_a.spp
        sa = srv.wait(ta, 300)
        for st in (sb, sc):
            assert st["t_submit_ms"] > sa["t_admit_ms"]
            assert st["t_admit_ms"] < sa["t_done_ms"] and st["t_done_ms"] < sa["t_done_ms"]
        for hs, job, tk in ((a, job_a, ta), (b, job_b, tb), (a, job_c, tc)):
            ds = DeviceScene(device, hs)
            solo, _ = device.render(ds, [job])
            assert np.array_equal(srv.film(tk), solo[0]), "served film differs from the solo render"
            ds.free()
    finally:
        srv.close()


def test_server_cancel_timeout_and_release(device):
    """wait() raises TimeoutError while a job runs (its film stays owned by
    the server); cancel() ends it with GND_E_CANCELLED; release() forgets an
    ended job and refuses a running one; the server then renders the next
    job correctly."""
    hs = HostScene.generate("cbox", "width=64;height=64")
    srv = Server(device, path_slots=64 * 64 * 8, max_jobs=4)
    try:
        sid = srv.add_scene(hs)
        long = srv.submit(sid, hs.default_job(spp=100000))
        with pytest.raises(TimeoutError):
            srv.wait(long, 0.05)
        with pytest.raises(GndError) as e:
            srv.release(long)
        assert e.value.code == A.GND_E_BUSY
        srv.cancel(long)
        with pytest.raises(GndError) as e:
            srv.wait(long, 60)
        assert e.value.code == A.GND_E_CANCELLED
        srv.release(long)
        with pytest.raises(GndError):
            srv.poll(long)
        j = hs.default_job(spp=5)
        tk = srv.submit(sid, j)
        assert srv.wait(tk, 60)["state"] == 2
        ds = DeviceScene(device, hs)
        assert np.array_equal(srv.film(tk), device.render(ds, [j])[0][0])
        ds.free()
    finally:
        srv.close()


def test_server_stack_overflow_fails_the_job(device):
    """A traversal stack overflow (forced with a 2-entry stack limit) fails
    the served job with GND_E_STACK, as it fails gnd_hip_render; with the
    whole stack the same job renders."""
    hs = HostScene.generate("scaled", "level=4;instances=6;width=32;height=32")  # BLAS deeper than 2 entries
    ds = DeviceScene(device, hs)
    srv = Server(device, path_slots=32 * 32 * 4, max_jobs=4)
    old = device.tuning()
    try:
        sid = srv.add_scene(hs)
        j = hs.default_job(spp=4)
        device.set_tuning(stack_limit=2)
        with pytest.raises(GndError) as e:
            device.render(ds, [j])
        assert e.value.code == A.GND_E_STACK
        tk = srv.submit(sid, j)
        with pytest.raises(GndError) as e:
            srv.wait(tk, 60)
        assert e.value.code == A.GND_E_STACK
        device.set_tuning(stack_limit=0)
        tk = srv.submit(sid, j)
        assert srv.wait(tk, 60)["state"] == 2
        assert np.array_equal(srv.film(tk), device.render(ds, [j])[0][0])
    finally:
        device.set_tuning(**old)
        srv.close()
        ds.free()


def test_serving_pool_at_full_scale(device):
    """The serving workload's own pool (8 procedural scenes of 10k..5M
    triangles, log-uniform) under a Poisson stream at 40 requests/s: every job
    completes, requests on different scenes overlap on the device (a job is
    admitted while another is still running), and served films equal solo
    renders for a job of the smallest and of the largest scene."""
    from gianduia_amd.serve import scene_pool
    pool = scene_pool()
    srv = Server(device, path_slots=8_000_000, max_jobs=64)
    try:
        ids = [srv.add_scene(h) for h in pool]
        rng = np.random.default_rng(17)
        work = []
        t = 0.0
        import time as _t
        t0 = _t.perf_counter()
        for k in range(24):
            t += float(rng.exponential(1 / 40.0))
            sid = ids[k % len(ids)]
            job = random_job(rng, pool[sid], res=(128, 640), spp=(2, 8))
            delay = t - (_t.perf_counter() - t0)
            if delay > 0:
                _t.sleep(delay)
            work.append((sid, job, srv.submit(sid, job)))
        sts = [srv.wait(tk, 300) for _, _, tk in work]
        assert all(s["state"] == 2 for s in sts)
        spans = sorted((s["t_admit_ms"], s["t_done_ms"]) for s in sts)
        overlapped = sum(1 for (a0, d0), (a1, _) in zip(spans, spans[1:]) if a1 < d0)
        print(f"serving pool: {overlapped} of {len(spans) - 1} consecutive admissions overlap a running job",
              flush=True)
        assert overlapped > 0
        tris = []
        for hs in pool:
            d = hs.desc().contents
            tris.append(sum(d.meshes[d.prims[i].mesh].ntris for i in range(d.nprims) if d.prims[i].shape_type == 0))
        for target in (int(np.argmin(tris)), int(np.argmax(tris))):
            sid, job, tk = next(w for w in work if w[0] == ids[target])
            ds = DeviceScene(device, pool[target])
            assert np.array_equal(srv.film(tk), device.render(ds, [job])[0][0]), f"scene {target}: film differs"
            ds.free()
    finally:
        srv.close()
