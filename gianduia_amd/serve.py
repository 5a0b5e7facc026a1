"""Serving driver: the continuous-batching server of libgnd_hip.so
(gnd_hip_server_*, include/gnd_hip.h) and a seeded Poisson load generator.

The reference has no serving layer; a job here is what its SamplerIntegrator
renders (include/gianduia/core/integrator.h:49-155): one camera, spp passes,
one film.  Jobs of a scene pool share the GPU by continuous batching and a
job's film is bit-identical to rendering it alone.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .api import Device, DeviceBuffer, GndError, HostScene, Job, _lib, camera_lookat


@dataclass
class Ticket:
    id: int
    job: Job
    film: DeviceBuffer
    t_submit: float
    aov: DeviceBuffer | None = None


class Server:
    """A continuous-batching server on one device.

    The server keeps each submitted job's device buffers alive until the job
    has ended (done, failed or cancelled), so a dropped Ticket or a wait that
    timed out can never free a film the device still writes."""

    def __init__(self, dev: Device, path_slots: int = 8_000_000, max_jobs: int = 256, hbm_budget: int = 0,
                 lanes: int = 0):
        self.dev = dev
        cfg = A.gnd_hip_server_config(hbm_budget, path_slots, max_jobs, lanes)
        h = C.c_void_p()
        dev._check(_lib().gnd_hip_server_create(dev.ctx, C.byref(cfg), C.byref(h)))
        self.handle = h
        self.scenes: list[HostScene] = []  # keeps the borrowed host arrays alive
        self._owned: dict[int, tuple] = {}  # ticket -> buffers of a job that has not ended

    def add_scene(self, hs: HostScene) -> int:
        arrays = hs.build()
        sid = C.c_int32()
        self.dev._check(_lib().gnd_hip_server_add_scene(self.handle, arrays, C.byref(sid)))
        self.scenes.append(hs)
        return sid.value

    def submit(self, scene_id: int, job: Job, want_aov: bool = False, film: DeviceBuffer | None = None) -> Ticket:
        """Queues `job`; `film` (optional) is a caller's buffer of at least
        W*H*16 bytes to render into instead of a new one."""
        w, h = job.camera.width, job.camera.height
        if film is None:
            film = self.dev.alloc(w * h * 16)
        elif film.nbytes < w * h * 16:
            raise ValueError("film buffer too small")
        aov = self.dev.alloc(w * h * 36) if want_aov else None
        cj = A.gnd_hip_job()
        cj.camera = job.camera
        cj.spp = job.spp
        cj.max_depth = job.max_depth
        cj.integrator = job.integrator
        cj.ao_length = job.ao_length
        cj.seed_offset = job.seed_offset
        cj.film = film.addr
        cj.film_sq = None
        cj.film_aov = aov.addr if aov is not None else None
        t = C.c_uint64()
        self.dev._check(_lib().gnd_hip_server_submit(self.handle, scene_id, C.byref(cj), C.byref(t)))
        self._owned[t.value] = (film, aov)
        return Ticket(t.value, job, film, time.perf_counter(), aov)

    def _ended(self, tid: int, d: dict) -> dict:
        if d["state"] == 2 or d["state"] < 0:
            self._owned.pop(tid, None)
        return d

    def poll(self, t: Ticket) -> dict:
        st = A.gnd_hip_job_status()
        self.dev._check(_lib().gnd_hip_server_poll(self.handle, t.id, C.byref(st)))
        return self._ended(t.id, st.as_dict())

    def wait(self, t: Ticket, timeout_s: float = 600.0) -> dict:
        """Waits for the job to end.  Raises TimeoutError while it is still
        queued or running, GndError if it failed."""
        st = A.gnd_hip_job_status()
        rc = _lib().gnd_hip_server_wait(self.handle, t.id, timeout_s, C.byref(st))
        if rc == A.GND_E_TIMEOUT:
            raise TimeoutError(f"job {t.id} not finished after {timeout_s} s (state {st.state})")
        self.dev._check(rc)
        d = self._ended(t.id, st.as_dict())
        if d["state"] < 0:
            raise GndError(d["state"], f"job {t.id} failed: {self.dev.last_error()}")
        return d

    def cancel(self, t: Ticket) -> None:
        """Stops the job (gnd_hip_server_cancel); wait() then reports GND_E_CANCELLED."""
        self.dev._check(_lib().gnd_hip_server_cancel(self.handle, t.id))

    def release(self, t: Ticket) -> None:
        """Forgets an ended job on the server side (its status can no longer be read)."""
        self.dev._check(_lib().gnd_hip_server_release(self.handle, t.id))

    def film(self, t: Ticket) -> np.ndarray:
        w, h = t.job.camera.width, t.job.camera.height
        return self.dev.download(t.film, np.float32, w * h * 4).reshape(h, w, 4)

    def aov(self, t: Ticket) -> np.ndarray:
        """The job's (H, W, 9) weighted AOV sums (submitted with want_aov)."""
        w, h = t.job.camera.width, t.job.camera.height
        return self.dev.download(t.aov, np.float32, w * h * 9).reshape(h, w, 9)

    def stats(self) -> dict:
        st = A.gnd_hip_stats()
        n = C.c_int32()
        b = C.c_size_t()
        self.dev._check(_lib().gnd_hip_server_stats(self.handle, C.byref(st), C.byref(n), C.byref(b)))
        d = st.as_dict()
        d["resident_scenes"] = n.value
        d["resident_bytes"] = b.value
        return d

    def close(self) -> None:
        """Stops the worker; buffers of jobs that had not ended are freed only
        after that (no kernel runs any more)."""
        if self.handle:
            _lib().gnd_hip_server_destroy(self.handle)
            self.handle = None
        self._owned.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_jobs(hs: HostScene, n: int = 64, seed: int = 0) -> list[Job]:
    """The batching workload: n jobs over the shared scene `hs`, each with a
    random pose (looking at the scene centre from a random direction above
    the floor, the scene's own camera model), a square resolution
    log-uniform in 64^2..1024^2 and spp log-uniform in 1..256 (seeded)."""
    rng = np.random.default_rng(seed)
    base = hs.camera()
    d = hs.desc().contents
    jobs = []
    for _ in range(n):
        side = int(round(float(np.exp(rng.uniform(np.log(64), np.log(1024))))))
        spp = int(round(float(np.exp(rng.uniform(np.log(1), np.log(256))))))
        theta = rng.uniform(0.15, 0.45) * np.pi
        phi = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(4.5, 6.5)
        eye = (r * np.sin(theta) * np.cos(phi), r * np.cos(theta), r * np.sin(theta) * np.sin(phi))
        cam = camera_lookat(base, eye, (0.0, 0.0, 0.0))
        cam.width = cam.height = side
        jobs.append(Job(cam, spp, d.max_depth, d.integrator, 0))
    return jobs


def scene_pool(n: int = 8, lo: float = 1e4, hi: float = 5e6, seed: int = 0) -> list[HostScene]:
    """n procedural scenes with log-uniform triangle counts in [lo, hi]."""
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(n):
        tris = float(np.exp(rng.uniform(math.log(lo), math.log(hi))))
        pool.append(HostScene.generate("procedural", f"tris={tris:.0f};seed={seed * 100 + i}"))
    return pool


def random_job(rng: np.random.Generator, hs: HostScene, res=(128, 1920), spp=(4, 64)) -> Job:
    """A job on `hs`: random pose around the scene centre, width log-uniform in
    `res` (height width * U[9/16, 1], at least 64), spp log-uniform in `spp`."""
    d = hs.desc().contents
    w = int(round(float(np.exp(rng.uniform(math.log(res[0]), math.log(res[1]))))))
    h = max(64, min(1080, int(round(w * rng.uniform(9 / 16, 1.0)))))
    n = int(round(float(np.exp(rng.uniform(math.log(spp[0]), math.log(spp[1]))))))
    theta = rng.uniform(0.15, 0.45) * math.pi
    phi = rng.uniform(0, 2 * math.pi)
    r = rng.uniform(4.5, 6.5)
    eye = (r * math.sin(theta) * math.cos(phi), r * math.cos(theta), r * math.sin(theta) * math.sin(phi))
    cam = camera_lookat(hs.camera(), eye, (0.0, 0.0, 0.0))
    cam.width, cam.height = w, h
    return Job(cam, n, d.max_depth, d.integrator, 0)


def poisson_load(server: Server, scene_ids: list[int], qps: float, n_jobs: int, seed: int = 0,
                 timeout_s: float = 600.0, **job_kw) -> dict:
    """Submits n_jobs with exponential inter-arrival times (rate qps) on
    scenes drawn uniformly from the pool and returns latency percentiles (ms,
    from each request's scheduled arrival to its film being complete) and
    throughput.  Films come from a pool allocated before the clock starts, as
    a serving front end would keep them; a submission that falls behind its
    arrival time is charged the delay."""
    rng = np.random.default_rng(seed)
    hosts = {sid: server.scenes[sid] for sid in scene_ids}
    plan = []
    t = 0.0
    for _ in range(n_jobs):
        t += float(rng.exponential(1.0 / qps))
        sid = int(rng.choice(scene_ids))
        plan.append((t, sid, random_job(rng, hosts[sid], **job_kw)))
    films = [server.dev.alloc(job.camera.width * job.camera.height * 16) for _, _, job in plan]
    tickets = []
    late = []
    t0 = time.perf_counter()
    for (at, sid, job), film in zip(plan, films):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        late.append(max(0.0, (time.perf_counter() - t0) - at) * 1e3)
        tickets.append(server.submit(sid, job, film=film))
    lat, queue, samples = [], [], 0
    for tk, lt in zip(tickets, late):
        st = server.wait(tk, timeout_s)
        lat.append(lt + st["latency_ms"])
        queue.append(st["queue_ms"])
        samples += tk.job.camera.width * tk.job.camera.height * tk.job.spp
        server.release(tk)
    wall = time.perf_counter() - t0
    for f in films:  # hipFree waits for the device: only after the run
        f.free()
    lat = np.array(lat)
    return {"qps_offered": qps, "jobs": n_jobs, "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "mean_queue_ms": round(float(np.mean(queue)), 2),
            "max_submit_lag_ms": round(float(max(late)), 2),
            "msamples_per_s": round(samples / wall / 1e6, 2), "wall_s": round(wall, 2),
            "qps_achieved": round(n_jobs / wall, 2)}
