#!/usr/bin/env python3
"""Benchmark of the MI355X ray-tracing hot path.

Workload (one step): one progressive render of the scaled-geometry scene —
a 1,310,720-triangle displaced icosphere plus 64 instances of a
10,000-triangle displaced sphere, conductor / glass / plastic GGX materials,
one quad light, thin-lens camera (aperture 0.05), 1920x1080, 16 spp, max
depth 16, box filter, path integrator, FP32 — rendered through the C-ABI
(gnd_hip_render) with the scene and film resident in HBM.

Prints one JSON line: Msamples/s (whole job), Mrays/s, the roofline of the
dominant kernel (closest-hit BVH4 traversal), the CPU baseline (the C++
restatement of the reference CPU integrator on the host cores, on a bounded
sample of the same frame) and the 64-job batching workload.

--scene cbox|volumetric measures the Cornell-box (512x512, 64 spp, depth 8)
and volumetric (1280x720, 32 spp, depth 32) workloads instead; --aov also
accumulates the film AOVs; --denoise times the film denoiser on the frame;
--serve runs the serving workload (Poisson arrivals over a pool of 8 scenes)
and prints its own JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak (8 TB/s)
RAY_BYTES, HIT_BYTES, NODE_BYTES, PACK_BYTES, INST_BYTES = 32, 20, 144, 160, 64
INTEGRATOR_NAMES = ("path", "volpath", "direct", "ao")
METRIC = "Msamples/s & Mrays/s (1080p, 1M tris) at equal MSE; p50/p99 job latency @QPS"


def denoise_workload(dev, sess, with_aovs: bool) -> dict:
    """Device time of gnd_hip_denoise on the frame just rendered (resolved
    radiance, plus the albedo / normal AOV guides when they were rendered):
    best of 3 calls, and the algorithmic bytes of one a-trous pass (each
    pixel's 48-byte record read once, its 16-byte result written once)."""
    import numpy as np
    from gianduia_amd import resolve, resolve_aovs
    g = sess.film(0)
    c = resolve(g)
    al = nm = None
    if with_aovs:
        av = resolve_aovs(g, sess.film_aov(0))
        al, nm = av["albedo"], av["normal"]
    best = min(dev.denoise(c, al, nm, timed=True)[1] for _ in range(3))
    H, W = c.shape[:2]
    return {"ms": round(best, 3), "width": W, "height": H, "guides": bool(with_aovs), "passes": 5,
            "pass_algorithmic_bytes": 64 * W * H}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(rays, nodes, packs, insts):
    return (RAY_BYTES + HIT_BYTES) * rays + NODE_BYTES * nodes + PACK_BYTES * packs + INST_BYTES * insts


WORKLOADS = {
    "scaled": "scaled-geometry path tracing: 1,310,720-tri displaced icosphere + 64 instanced 10k-tri spheres, "
              "GGX conductor/glass/plastic, quad light, thin lens 0.05",
    "cbox": "Cornell box path tracing: 36 triangles, diffuse, quad light, pinhole fov 40",
    "volumetric": "volumetric path tracing: Cornell box + glass sphere with a homogeneous medium (sigma_t 2, "
                  "albedo 0.8, g 0.3) + 256^3 fBm density grid (sigma_t max 8) in an index box",
}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kTracePaths_scaled1080p_hbm_traffic.md")


def measured_traffic(rays_per_launch):
    """HBM bytes per closest-hit launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by the copy-kernel calibration): the profiled
    bytes per ray times this run's rays per launch (the autotuner may cut the
    frame into other launches than the profiled run), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_ray:"):
                    return float(line.split(":", 1)[1]) * rays_per_launch
    except OSError:
        pass
    return None


def host_cores():
    """Threads for the CPU baseline: every core this process may run on
    (sched_getaffinity, what nproc reports), with what lscpu and the cgroup CPU
    quota say about the host beside it."""
    import subprocess
    n = len(os.sched_getaffinity(0))
    host = {"affinity_cores": n, "os_cpu_count": os.cpu_count()}
    try:
        host["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout)
    except Exception:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("CPU(s)", "Model name", "Thread(s) per core", "Socket(s)"):
                host["lscpu " + k.strip()] = v.strip()
    except Exception:
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            host["cgroup_cpu_max"] = q if q == "max" else round(int(q) / int(period), 2)
    except Exception:
        pass
    return n, host


def cpu_baseline(hs, target_s=12.0):
    """Times the CPU restatement (oracle) of the reference integrator on the host
    cores: a centred band of rows of pass 0 grown until it takes target_s, or
    whole passes 0, 1, ... when a full pass is faster than that."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import numpy as np
    from oracle import RefScene, lib
    threads, host = host_cores()
    t0 = time.time()
    ref = RefScene(hs.desc())
    log(f"cpu baseline: oracle scene built in {time.time() - t0:.1f}s, {threads} threads")
    cam = hs.camera()
    W, H = cam.width, cam.height
    rgbw = np.zeros((H, W, 4), np.float32)

    def run(p0, p1, r0, r1):
        stats = (C.c_double * 5)()
        rc = lib().gnd_ref_render(ref.h, threads, p0, p1, r0, r1, 0, rgbw.ctypes.data_as(C.POINTER(C.c_float)),
                                  None, stats, None)
        if rc != 0:
            raise RuntimeError("oracle render failed")
        return stats[0], (r1 - r0) * W * (p1 - p0)

    rows = 8
    while True:
        r0 = max(0, H // 2 - rows // 2)
        r1 = min(H, r0 + rows)
        secs, samples = run(0, 1, r0, r1)
        if secs >= target_s or r1 - r0 >= H:
            break
        rows = min(H, max(rows * 2, int(rows * target_s / max(secs, 1e-3) * 1.1)))
    desc = f"rows {r0}-{r1} of pass 0"
    passes = 1
    if r1 - r0 >= H and secs < target_s:
        more = max(1, int(np.ceil((target_s - secs) / max(secs, 1e-3))))
        s2, n2 = run(1, 1 + more, 0, H)
        secs += s2
        samples += n2
        passes += more
        desc = f"passes 0-{passes - 1} (full frames)"
    integ = INTEGRATOR_NAMES[hs.desc().contents.integrator]
    return {"value": samples / secs / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port", "host": host,
            "sample": f"{desc} of the same {W}x{H} frame ({samples} camera samples, {secs:.1f} s, {integ} "
                      f"integrator, C++ restatement of the reference CPU integrator, {threads} threads)"}


def batch_workload(dev, ds, hs):
    """64-job mix rendered four ways, films resident in HBM: `plain` = one job
    at a time with one pass per wave (the reference's pass loop,
    integrator.h:79-150), `solo` = one job at a time with whole-frame waves,
    `batched` = all jobs in one call sharing every bounce launch, and
    `batched_graph` = the same with the bounce chain captured in a hipGraph.
    Films are identical in all four."""
    from gianduia_amd.api import RenderSession
    from gianduia_amd.serve import batch_jobs
    jobs = batch_jobs(hs)
    samples = sum(j.camera.width * j.camera.height * j.spp for j in jobs)
    old = dev.tuning()
    solo = [RenderSession(dev, ds, [j]) for j in jobs]
    together = RenderSession(dev, ds, jobs)
    res = {"jobs": len(jobs), "samples": samples}
    try:
        together.run()  # warm-up: wave buffers at full size
        dev.set_tuning(use_graph=True)
        together.run()  # warm-up: the bounce chains of each wave shape, captured once and reused
        dev.set_tuning(**old)
        for mode in ("plain", "solo", "batched", "batched_graph"):
            t0 = time.perf_counter()
            if mode.startswith("batched"):
                dev.set_tuning(use_graph=mode == "batched_graph")
                together.run()
            else:
                for j, sess in zip(jobs, solo):
                    if mode == "plain":
                        dev.set_tuning(queue_capacity=max(1024, j.camera.width * j.camera.height))
                    sess.run()
            secs = time.perf_counter() - t0
            dev.set_tuning(**old)
            res[f"{mode}_s"] = round(secs, 3)
            res[f"{mode}_msamples_per_s"] = round(samples / secs / 1e6, 2)
            log(f"batch workload {mode}: {secs:.2f} s")
    finally:
        dev.set_tuning(**old)
    best = min(res["batched_s"], res["batched_graph_s"])
    res["batched_vs_plain"] = round(res["plain_s"] / best, 2)
    res["batched_vs_solo"] = round(res["solo_s"] / best, 2)
    return res


# The server's slot pool: a pass holds its W*H slots until its last path ends,
# so most slots of a deep pass sit idle in its late bounces; a large pool keeps
# the bounce launches full (saturation on the serving mix, 4 lanes: 16M slots
# 320, 40M 372, 120M 430 Msamples/s).  120M slots and 4 lanes' queues are
# 94 GB of the 288 GB.
SERVE_SLOTS = 120_000_000
SERVE_LANES = 4


def serve_vs_serial(dev, srv, ids, pool, n_jobs=48, seed=1000):
    """The same n_jobs random jobs over the scene pool rendered three ways,
    films resident in HBM: `served` = all submitted to the continuous-batching
    server at once (saturation), `plain` = one job at a time through
    gnd_hip_render with one pass per wave (the reference's pass loop,
    integrator.h:79-150), `solo` = one job at a time with whole-frame waves.
    Samples per second of each, and the served latency percentiles."""
    import numpy as np
    from gianduia_amd import DeviceScene
    from gianduia_amd.api import RenderSession
    from gianduia_amd.serve import random_job
    rng = np.random.default_rng(seed)
    plan = []
    for _ in range(n_jobs):
        sid = int(rng.choice(ids))
        plan.append((sid, random_job(rng, pool[sid])))
    samples = sum(j.camera.width * j.camera.height * j.spp for _, j in plan)
    res = {"jobs": n_jobs, "samples": samples}
    films = [dev.alloc(j.camera.width * j.camera.height * 16) for _, j in plan]
    t0 = time.perf_counter()
    tickets = [srv.submit(sid, j, film=f) for (sid, j), f in zip(plan, films)]
    lat = []
    for tk in tickets:
        lat.append(srv.wait(tk)["latency_ms"])
        srv.release(tk)
    served_s = time.perf_counter() - t0
    for f in films:
        f.free()
    res["served_s"] = round(served_s, 3)
    res["served_p50_ms"] = round(float(np.percentile(lat, 50)), 2)
    res["served_p99_ms"] = round(float(np.percentile(lat, 99)), 2)
    scenes = {sid: DeviceScene(dev, pool[sid]) for sid in sorted({sid for sid, _ in plan})}
    sessions = [RenderSession(dev, scenes[sid], [j]) for sid, j in plan]
    old = dev.tuning()
    try:
        for sess in sessions:  # warm-up: buffers and code objects
            sess.run()
        for mode in ("plain", "solo"):
            t0 = time.perf_counter()
            for (_, j), sess in zip(plan, sessions):
                if mode == "plain":
                    dev.set_tuning(queue_capacity=max(1024, j.camera.width * j.camera.height))
                sess.run()
            res[f"{mode}_s"] = round(time.perf_counter() - t0, 3)
            dev.set_tuning(**old)
    finally:
        dev.set_tuning(**old)
        del sessions
        for sc in scenes.values():
            sc.free()
    for mode in ("served", "plain", "solo"):
        res[f"{mode}_msamples_per_s"] = round(samples / res[f"{mode}_s"] / 1e6, 2)
    res["served_vs_plain"] = round(res["plain_s"] / res["served_s"], 2)
    res["served_vs_solo"] = round(res["solo_s"] / res["served_s"], 2)
    return res


def serve_workload(qps_levels=(5, 10, 20, 50, 100, 200), jobs_per_level=100):
    """Serving workload: a pool of 8 procedural scenes (1e4..5e6 triangles,
    log-uniform, seeded), Poisson arrivals at each rate, jobs with random pose,
    width log-uniform in 128..1920 (height 9/16..1 of it), spp log-uniform in
    4..64, rendered by the continuous-batching server.  Latency is submission
    to film complete.  Then the saturation throughput of the server against
    the serial plain and solo kernels on one job mix (serve_vs_serial)."""
    from gianduia_amd import Device
    from gianduia_amd.serve import Server, poisson_load, scene_pool
    dev = Device(0)
    t0 = time.time()
    pool = scene_pool()
    build_ms = []
    for hs in pool:
        hs.build()
        build_ms.append(round(hs.build_ms, 1))
    log(f"serve: scene pool built in {time.time() - t0:.1f} s ({build_ms} ms BVH builds)")
    srv = Server(dev, path_slots=SERVE_SLOTS, max_jobs=256, lanes=SERVE_LANES)
    ids = [srv.add_scene(h) for h in pool]
    tris = []
    for hs in pool:
        d = hs.desc().contents
        tris.append(sum(d.meshes[d.prims[i].mesh].ntris for i in range(d.nprims) if d.prims[i].shape_type == 0))
    # first use uploads each scene: one small job per scene, timed apart
    t1 = time.perf_counter()
    poisson_load(srv, ids, qps=1000.0, n_jobs=len(ids) * 2, seed=99, res=(64, 65), spp=(1, 2))
    upload_s = time.perf_counter() - t1
    levels = []
    for k, qps in enumerate(qps_levels):
        r = poisson_load(srv, ids, qps=qps, n_jobs=jobs_per_level, seed=k)
        log(f"serve: {json.dumps(r)}")
        levels.append(r)
    vs = serve_vs_serial(dev, srv, ids, pool)
    log(f"serve vs serial: {json.dumps(vs)}")
    st = srv.stats()
    srv.close()
    dev.close()
    return {"metric": "p50/p99 job latency @QPS", "unit": "ms", "mode": "serve", "n_gpus": 1,
            "data": "synthetic (procedural scenes and jobs, seeded)",
            "config": {"scenes": len(pool), "instanced_triangles": tris, "path_slots": SERVE_SLOTS,
                       "lanes": SERVE_LANES,
                       "jobs_per_level": jobs_per_level, "width": [128, 1920], "spp": [4, 64]},
            "scene_build_ms": build_ms, "scene_upload_s": round(upload_s, 2), "levels": levels,
            "vs_serial": vs,
            "rays_closest": st["rays_closest"], "rays_shadow": st["rays_shadow"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-counters", action="store_true", help="skip the untimed counting pass")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job batching workload")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tuning instead of running gnd_hip_autotune (untimed) on the workload")
    ap.add_argument("--tuning", default="",
                    help="explicit tunables, key=value;... (gnd_hip_tuning field names); skips the autotuner")
    ap.add_argument("--serve", action="store_true",
                    help="only run the serving workload (Poisson arrivals over a scene pool) and print its JSON")
    ap.add_argument("--scene", default="scaled", choices=sorted(WORKLOADS),
                    help="workload scene: scaled (default, 1080p 1M tris), cbox (512^2 Cornell box), "
                         "volumetric (720p, homogeneous + 256^3 grid media)")
    ap.add_argument("--params", default="", help="scene generator overrides, key=value;...")
    ap.add_argument("--aov", action="store_true",
                    help="also accumulate the film AOVs (albedo, normal, depth, roughness, metallic)")
    ap.add_argument("--denoise", action="store_true",
                    help="also time the film denoiser (gnd_hip_denoise) on the rendered frame, with the AOV "
                         "guides when --aov is given")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_probe:
        from gianduia_amd import Device
        dev = Device(0)
        ms = dev.copy_probe(a.copy_probe << 20, repeats=4)
        print(json.dumps({"copy_probe_mib": a.copy_probe, "ms_per_copy": ms,
                          "gbs_read_plus_write": 2 * (a.copy_probe << 20) / (ms / 1e3) / 1e9}))
        dev.close()
        return
    if a.serve:
        print(json.dumps(serve_workload()), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    from gianduia_amd import Device, DeviceScene, HostScene
    from gianduia_amd import _abi as A
    from gianduia_amd.api import RenderSession

    dev = Device(0)
    info = dev.info()
    log(f"device {info}")
    t0 = time.time()
    hs = HostScene.generate(a.scene, a.params)
    gen_s = time.time() - t0
    hs.build()
    ds = DeviceScene(dev, hs)
    d = hs.desc().contents
    ntris = sum(d.meshes[i].ntris for i in range(d.nmeshes))
    inst_tris = 0
    for i in range(d.nprims):
        p = d.prims[i]
        if p.shape_type == 0:
            inst_tris += d.meshes[p.mesh].ntris
    log(f"scene: generated {gen_s:.1f}s, BVH build {hs.build_ms:.0f} ms, {inst_tris} instanced triangles, "
        f"{ds.nbytes / 2**20:.0f} MiB in HBM")
    job = hs.default_job()
    t_tune = time.perf_counter()
    if a.tuning:
        dev.set_tuning(**{k.strip(): int(v) for k, v in (kv.split("=") for kv in a.tuning.split(";") if kv.strip())})
        tuning, how = dev.tuning(), "(given)"
    elif a.no_autotune:
        tuning, how = dev.tuning(), "(default)"
    else:
        tuning, how = dev.autotune(ds, job), "(autotuned)"
    log(f"tuning {how}: {tuning} in {time.perf_counter() - t_tune:.1f} s")
    sess = RenderSession(dev, ds, [job], want_aov=a.aov)
    for i in range(a.warmup):
        st = sess.run()
        log(f"warmup {i}: {st['ms_total']:.1f} ms")
    per_step = []
    t_start = time.perf_counter()
    for i in range(a.steps):
        st = sess.run()
        per_step.append(st)
        log(f"step {i}: {st['ms_total']:.1f} ms, trace {st['ms_trace']:.1f} ms over {st['trace_launches']} launches")
    t_end = time.perf_counter()
    wall_ms = (t_end - t_start) * 1000.0
    ms_per_step = wall_ms / a.steps
    samples = per_step[0]["samples"]
    rays = per_step[0]["rays_closest"] + per_step[0]["rays_shadow"]
    msps = samples * a.steps / (wall_ms / 1000.0) / 1e6
    mrays = rays * a.steps / (wall_ms / 1000.0) / 1e6

    roof = None
    if not a.no_counters:
        trace_steps, note = per_step, None
        if per_step[0]["trace_launches"] == 0:
            # the tuning runs every path in the tail kernel: the traversal
            # kernel is timed on the same frame with the tail kernel off
            tuned = dev.tuning()
            dev.set_tuning(**{**tuned, "tail_paths": -1})
            sess.run()
            trace_steps = [sess.run() for _ in range(a.steps)]
            note = "the tuned frame runs every path in the tail kernel; the traversal kernel is timed on the " \
                   "same frame with the tail kernel off (tail_paths=-1)"
        dev.set_counters(True)
        cst = sess.run()  # identical rays (deterministic seeds); untimed
        dev.set_counters(False)
        if note:
            dev.set_tuning(**tuned)
        # the traversal kernel's own queries: the tail kernel's closest hits
        # are in rays_closest but neither in the visit counts nor in ms_trace
        trays = cst["rays_closest"] - cst["rays_tail"]
        byts = algorithmic_bytes(trays, cst["node_visits"], cst["pack_tests"], cst["inst_tests"])
        trace_ms = sum(s["ms_trace"] for s in trace_steps) / len(trace_steps)
        launches = trace_steps[0]["trace_launches"]
        achieved = byts / (trace_ms / 1000.0) / 1e9
        traffic = measured_traffic(trays / launches) if not a.params and a.scene == "scaled" else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": (os.path.relpath(TRAFFIC_PROFILE, ROOT) + ": measured bytes per ray x this run's "
                                   "rays per launch") if traffic else None,
                "kernel": "kTracePaths (closest-hit BVH4 traversal)",
                "bytes_per_launch": byts / launches, "avg_launch_ms": trace_ms / launches,
                "max_nodes_one_ray": cst["max_ray_nodes"], "nonfinite_rays": cst["bad_rays"],
                # lane steps / lane slots of the traversal loop (node, pack and
                # instance steps; an instance costs a start and a finish step)
                "simd_utilization": round(cst["lane_steps"] / max(1, cst["lane_iterations"]), 4),
                "rays_per_launch": trays / launches,
                "bytes_note": "achieved counts every reference-layout node, pack and instance read of the "
                              "reference-order traversal; most are served by L1, L2 and the Infinity Cache, so "
                              "frac can pass 1; traffic is the measured HBM bytes per launch",
                "visits_per_ray": {"nodes": cst["node_visits"] / trays,
                                   "packs": cst["pack_tests"] / trays,
                                   "insts": cst["inst_tests"] / trays},
                **({"note": note} if note else {})}
        log(f"counting pass: {json.dumps(roof)}")

    den = None
    if a.denoise:
        den = denoise_workload(dev, sess, a.aov)
        log(f"denoise: {json.dumps(den)}")

    batch = None
    if not a.no_batch and a.scene == "scaled":
        batch = batch_workload(dev, ds, hs)
        log(f"batch workload: {json.dumps(batch)}")

    cpu = None
    if not a.no_cpu_baseline:
        try:
            cpu = cpu_baseline(hs)
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {e}")

    cam = job.camera
    out = {
        "metric": METRIC, "value": round(msps, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scene built in code, seeded)",
        "config": {"workload": WORKLOADS[a.scene],
                   "width": cam.width, "height": cam.height, "spp": job.spp, "max_depth": job.max_depth,
                   "instanced_triangles": inst_tris,
                   "integrator": INTEGRATOR_NAMES[job.integrator],
                   "aovs": bool(a.aov), **({"params": a.params} if a.params else {})},
        "mrays_per_s": round(mrays, 2),
        "scene_build_ms": round(hs.build_ms, 1),
        "trace_ms_per_step": round(sum(s["ms_trace"] for s in per_step) / a.steps, 3),
        "shadow_ms_per_step": round(sum(s["ms_shadow"] for s in per_step) / a.steps, 3),
        "shade_ms_per_step": round(sum(s["ms_shade"] for s in per_step) / a.steps, 3),
        "tuning": tuning,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batch": batch,
    }
    if den:
        out["denoise"] = den
    if cpu:
        out["vs_cpu_baseline"] = round(msps / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
