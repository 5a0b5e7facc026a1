"""Python driver of libgnd_hip.so: device context, scenes, traversal and rendering.

Device memory is owned by the library (gnd_hip_malloc); arrays cross the
boundary as numpy buffers copied to and from HBM.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi as A


class GndError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _lib():
    return A.load_library()


def host_error() -> str:
    return (_lib().gnd_host_last_error() or b"").decode()


class DeviceBuffer:
    def __init__(self, dev: "Device", nbytes: int):
        self.dev = dev
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        dev._check(_lib().gnd_hip_malloc(dev.ctx, max(1, self.nbytes), C.byref(self.ptr)))

    @property
    def addr(self) -> int:
        return int(self.ptr.value or 0)

    def free(self) -> None:
        if self.ptr.value:
            _lib().gnd_hip_free(self.dev.ctx, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            if self.dev.ctx is not None:
                self.free()
        except Exception:
            pass


class Device:
    """One HIP context (gnd_hip_init) on one GPU."""

    def __init__(self, device: int = 0):
        self.ctx = None
        ctx = C.c_void_p()
        rc = _lib().gnd_hip_init(device, C.byref(ctx))
        if rc != 0:
            raise GndError(rc, (_lib().gnd_hip_last_error(None) or b"").decode())
        self.ctx = ctx

    def close(self) -> None:
        if self.ctx is not None:
            _lib().gnd_hip_shutdown(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise GndError(rc, self.last_error())

    def last_error(self) -> str:
        return (_lib().gnd_hip_last_error(self.ctx) or b"").decode()

    # ---- info / tuning
    def info(self) -> dict:
        name = C.create_string_buffer(128)
        cus = C.c_int()
        hbm = C.c_size_t()
        self._check(_lib().gnd_hip_device_info(self.ctx, name, 128, C.byref(cus), C.byref(hbm)))
        return {"arch": name.value.decode(), "cus": cus.value, "hbm_bytes": hbm.value}

    def tuning(self) -> dict:
        t = A.gnd_hip_tuning()
        self._check(_lib().gnd_hip_get_tuning(self.ctx, C.byref(t)))
        return {"block_size": t.block_size, "stack_lds": t.stack_lds, "queue_capacity": t.queue_capacity,
                "use_graph": t.use_graph, "refill_lanes": t.refill_lanes,
                "shadow_refill_lanes": t.shadow_refill_lanes, "stack_limit": t.stack_limit,
                "tail_paths": t.tail_paths}

    def set_tuning(self, block_size=None, stack_lds=None, queue_capacity=None, use_graph=None,
                   refill_lanes=None, shadow_refill_lanes=None, stack_limit=None, tail_paths=None) -> None:
        """Changes the given tunables; refill_lanes / shadow_refill_lanes 0
        means the per-scene default, stack_limit 0 the whole traversal stack,
        tail_paths 0 the per-device default and -1 no tail kernel."""
        cur = self.tuning()
        t = A.gnd_hip_tuning(block_size or cur["block_size"], stack_lds or cur["stack_lds"],
                             queue_capacity or cur["queue_capacity"],
                             cur["use_graph"] if use_graph is None else int(use_graph),
                             cur["refill_lanes"] if refill_lanes is None else int(refill_lanes),
                             cur["shadow_refill_lanes"] if shadow_refill_lanes is None else int(shadow_refill_lanes),
                             cur["stack_limit"] if stack_limit is None else int(stack_limit),
                             cur["tail_paths"] if tail_paths is None else int(tail_paths))
        self._check(_lib().gnd_hip_set_tuning(self.ctx, C.byref(t)))

    def autotune(self, scene: "DeviceScene", job: "Job") -> dict:
        """gnd_hip_autotune: times a probe of `job` (whole passes, up to 40 M
        paths) per block size x LDS stack depth, then per closest-hit and
        any-hit refill threshold, tail-kernel threshold and wave capacity,
        keeping the fastest of each (every choice renders the same bits);
        caches the result per scene and job shape; returns the chosen tuning."""
        w, h = job.camera.width, job.camera.height
        film = self.alloc(w * h * 16)
        cj = A.gnd_hip_job()
        cj.camera = job.camera
        cj.spp = job.spp
        cj.max_depth = job.max_depth
        cj.integrator = job.integrator
        cj.ao_length = job.ao_length
        cj.seed_offset = job.seed_offset
        cj.film = film.addr
        cj.film_sq = None
        t = A.gnd_hip_tuning()
        self._check(_lib().gnd_hip_autotune(self.ctx, scene.handle, C.byref(cj), C.byref(t)))
        return self.tuning()

    def set_counters(self, on: bool) -> None:
        self._check(_lib().gnd_hip_set_counters(self.ctx, int(bool(on))))

    def cancel(self) -> None:
        """Integrator::cancel: stops a render running on this device from
        another thread at its next bounce; that render raises GND_E_CANCELLED."""
        self._check(_lib().gnd_hip_cancel(self.ctx))

    def denoise(self, color: np.ndarray, albedo: np.ndarray | None = None, normal: np.ndarray | None = None,
                timed: bool = False, noalloc: bool = False):
        """gnd_hip_denoise of (H, W, 3) float32 images (radiance; albedo and
        the AOV normal n*0.5+0.5 as optional guides).  Returns the denoised
        image, and the device milliseconds with timed=True.  noalloc=True
        calls gnd_hip_denoise_noalloc (static scratch, at most 2,097,152
        pixels)."""
        color = np.ascontiguousarray(color, np.float32)
        H, W = color.shape[:2]
        bufs = []
        try:
            for a in (color, albedo, normal):
                if a is None:
                    bufs.append(None)
                    continue
                a = np.ascontiguousarray(a, np.float32)
                assert a.shape == color.shape
                b = DeviceBuffer(self, a.nbytes)
                self._check(_lib().gnd_hip_memcpy_h2d(self.ctx, b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
                bufs.append(b)
            ms = C.c_double()
            ptr = [b.ptr if b is not None else None for b in bufs]
            fn = _lib().gnd_hip_denoise_noalloc if noalloc else _lib().gnd_hip_denoise
            self._check(fn(self.ctx, ptr[0], ptr[1], ptr[2], W, H, ptr[0], C.byref(ms)))
            out = np.empty_like(color)
            self._check(_lib().gnd_hip_memcpy_d2h(self.ctx, out.ctypes.data_as(C.c_void_p), bufs[0].ptr, out.nbytes))
        finally:
            for b in bufs:
                if b is not None:
                    b.free()
        return (out, ms.value) if timed else out

    def copy_probe(self, nbytes: int, repeats: int = 10) -> float:
        """Average ms of a plain device-to-device streaming copy of nbytes."""
        ms = C.c_double()
        self._check(_lib().gnd_hip_copy_probe(self.ctx, nbytes, repeats, C.byref(ms)))
        return ms.value

    # ---- memory
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        if arr.nbytes:
            self._check(_lib().gnd_hip_memcpy_h2d(self.ctx, b.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        return b

    def download(self, buf: DeviceBuffer, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            self._check(_lib().gnd_hip_memcpy_d2h(self.ctx, out.ctypes.data_as(C.c_void_p), buf.ptr, out.nbytes))
        return out

    # ---- traversal
    def _ray_soa(self, rays: dict, keep: list) -> A.gnd_ray_soa:
        soa = A.gnd_ray_soa()
        for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"):
            b = self.upload(np.asarray(rays[k], dtype=np.float32))
            keep.append(b)
            setattr(soa, k, b.addr)
        return soa

    def trace_closest(self, scene: "DeviceScene", rays: dict):
        n = len(rays["ox"])
        keep: list = []
        soa = self._ray_soa(rays, keep)
        hs = A.gnd_hit_soa()
        bufs = {}
        for k, dt in (("t", np.float32), ("inst", np.int32), ("prim", np.int32), ("u", np.float32),
                      ("v", np.float32)):
            bufs[k] = (self.alloc(4 * max(n, 1)), dt)
            setattr(hs, k, bufs[k][0].addr)
        st = A.gnd_hip_stats()
        self._check(_lib().gnd_hip_trace_closest(self.ctx, scene.handle, C.byref(soa), n, C.byref(hs), C.byref(st)))
        out = {k: self.download(b, dt, n) for k, (b, dt) in bufs.items()}
        return out, st.as_dict()

    def trace_any(self, scene: "DeviceScene", rays: dict):
        n = len(rays["ox"])
        keep: list = []
        soa = self._ray_soa(rays, keep)
        occ = self.alloc(4 * max(n, 1))
        st = A.gnd_hip_stats()
        self._check(_lib().gnd_hip_trace_any(self.ctx, scene.handle, C.byref(soa), n, occ.ptr, C.byref(st)))
        return self.download(occ, np.int32, n), st.as_dict()

    # ---- rendering
    def render(self, scene: "DeviceScene", jobs: list["Job"], want_sq: bool = False, want_aov: bool = False):
        """Renders the jobs in one call; returns (films, stats).  Each film is
        the (H,W,4) rgbw array, or a tuple of it with the (H,W,3) per-pass
        squares (want_sq) and the (H,W,9) weighted AOV sums (want_aov)."""
        cj = (A.gnd_hip_job * len(jobs))()
        films = []
        for i, j in enumerate(jobs):
            w, h = j.camera.width, j.camera.height
            f = self.alloc(w * h * 16)
            sq = self.alloc(w * h * 12) if want_sq else None
            av = self.alloc(w * h * 36) if want_aov else None
            films.append((f, sq, av, w, h))
            cj[i].film_aov = av.addr if av is not None else None
            cj[i].camera = j.camera
            cj[i].spp = j.spp
            cj[i].max_depth = j.max_depth
            cj[i].integrator = j.integrator
            cj[i].ao_length = j.ao_length
            cj[i].seed_offset = j.seed_offset
            cj[i].film = f.addr
            cj[i].film_sq = sq.addr if sq is not None else None
        st = A.gnd_hip_stats()
        self._check(_lib().gnd_hip_render(self.ctx, scene.handle, cj, len(jobs), C.byref(st)))
        out = []
        for f, sq, av, w, h in films:
            item = [self.download(f, np.float32, w * h * 4).reshape(h, w, 4)]
            if want_sq:
                item.append(self.download(sq, np.float32, w * h * 3).reshape(h, w, 3))
            if want_aov:
                item.append(self.download(av, np.float32, w * h * 9).reshape(h, w, 9))
            out.append(item[0] if len(item) == 1 else tuple(item))
        return out, st.as_dict()


class RenderSession:
    """Jobs with device-resident films, rendered repeatedly without host copies."""

    def __init__(self, dev: Device, scene: "DeviceScene", jobs: list["Job"], want_sq: bool = False,
                 want_aov: bool = False):
        self.dev, self.scene, self.jobs = dev, scene, jobs
        self.cj = (A.gnd_hip_job * len(jobs))()
        self.films = []
        self.aovs = []
        for i, j in enumerate(jobs):
            w, h = j.camera.width, j.camera.height
            f = dev.alloc(w * h * 16)
            sq = dev.alloc(w * h * 12) if want_sq else None
            av = dev.alloc(w * h * 36) if want_aov else None
            self.films.append((f, sq, w, h))
            self.aovs.append(av)
            self.cj[i].film_aov = av.addr if av is not None else None
            self.cj[i].camera = j.camera
            self.cj[i].spp = j.spp
            self.cj[i].max_depth = j.max_depth
            self.cj[i].integrator = j.integrator
            self.cj[i].ao_length = j.ao_length
            self.cj[i].seed_offset = j.seed_offset
            self.cj[i].film = f.addr
            self.cj[i].film_sq = sq.addr if sq is not None else None

    def run(self) -> dict:
        st = A.gnd_hip_stats()
        self.dev._check(_lib().gnd_hip_render(self.dev.ctx, self.scene.handle, self.cj, len(self.jobs), C.byref(st)))
        return st.as_dict()

    def film(self, i: int) -> np.ndarray:
        f, _, w, h = self.films[i]
        return self.dev.download(f, np.float32, w * h * 4).reshape(h, w, 4)

    def film_sq(self, i: int) -> np.ndarray:
        _, sq, w, h = self.films[i]
        return self.dev.download(sq, np.float32, w * h * 3).reshape(h, w, 3)

    def film_aov(self, i: int) -> np.ndarray:
        _, _, w, h = self.films[i]
        return self.dev.download(self.aovs[i], np.float32, w * h * 9).reshape(h, w, 9)


@dataclass
class Job:
    camera: A.gnd_camera_desc
    spp: int
    max_depth: int
    integrator: int = A.GND_INTEGRATOR_PATH
    seed_offset: int = 0
    ao_length: float = 5.0  # ao integrator: occlusion ray length (ao.cpp:9)


_PROP_KINDS = {"boolean": A.GND_PROP_BOOLEAN, "integer": A.GND_PROP_INTEGER, "float": A.GND_PROP_FLOAT,
               "string": A.GND_PROP_STRING, "point": A.GND_PROP_POINT, "vector2": A.GND_PROP_VECTOR2,
               "vector": A.GND_PROP_VECTOR, "color": A.GND_PROP_COLOR, "transform": A.GND_PROP_TRANSFORM}


class SceneObjects:
    """The reference's object graph handed over object by object
    (gnd_host_builder_*, the objects src/core/parser.cpp:41-235 creates):
    add() each object with its class tag, type name, name and properties,
    add_child() each parent -> child link in document order, finish() the
    scene.  A property is (name, kind, value) with kind one of boolean,
    integer, float, string, point, vector2, vector, color, transform; a
    transform's value is (matrix, inverse), each 16 floats column-major."""

    def __init__(self, base_dir: str = ""):
        self.handle = C.c_void_p()
        rc = _lib().gnd_host_builder_create(base_dir.encode(), C.byref(self.handle))
        if rc != 0:
            raise GndError(rc, host_error())

    def add(self, tag: str, type: str | None = None, name: str | None = None, props=()) -> int:
        props = list(props)
        arr = (A.gnd_prop * max(1, len(props)))()
        keep = []
        for p, (nm, kind, val) in zip(arr, props):
            p.name = nm.encode()
            p.kind = _PROP_KINDS[kind]
            if kind in ("boolean", "integer"):
                p.ivalue = int(val)
            elif kind == "float":
                p.v[0] = float(val)
            elif kind == "string":
                p.svalue = str(val).encode()
            elif kind == "transform":
                m = (C.c_float * 32)(*[float(x) for x in list(val[0]) + list(val[1])])
                keep.append(m)
                p.matrix = C.cast(m, A.f32p)
            else:
                for i, x in enumerate(val):
                    p.v[i] = float(x)
        oid = C.c_int32()
        rc = _lib().gnd_host_builder_object(self.handle, tag.encode(), (type or tag).encode(),
                                            name.encode() if name else None, arr, len(props), C.byref(oid))
        if rc != 0:
            raise GndError(rc, host_error())
        return oid.value

    def add_child(self, parent: int, child: int) -> None:
        rc = _lib().gnd_host_builder_add_child(self.handle, parent, child)
        if rc != 0:
            raise GndError(rc, host_error())

    def finish(self, root: int) -> "HostScene":
        h = C.c_void_p()
        rc = _lib().gnd_host_builder_finish(self.handle, root, C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        return HostScene(h)

    def close(self) -> None:
        if self.handle:
            _lib().gnd_host_builder_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostScene:
    """Host scene shell (XML / procedural / caller-built) with its BVH build."""

    def __init__(self, handle):
        self.handle = handle
        self._arrays = None
        self.build_ms = None

    @classmethod
    def generate(cls, name: str, params: str = "") -> "HostScene":
        h = C.c_void_p()
        rc = _lib().gnd_host_scene_generate(name.encode(), params.encode(), C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        return cls(h)

    @classmethod
    def load_xml(cls, path: str) -> "HostScene":
        h = C.c_void_p()
        rc = _lib().gnd_host_scene_load_xml(path.encode(), C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        return cls(h)

    @classmethod
    def from_desc(cls, desc: A.gnd_scene_desc) -> "HostScene":
        h = C.c_void_p()
        rc = _lib().gnd_host_scene_create(C.byref(desc), C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        return cls(h)

    def desc(self):
        return _lib().gnd_host_scene_desc(self.handle)

    def build(self, threads: int = 0):
        if self._arrays is None:
            out = C.POINTER(A.gnd_hip_scene_arrays)()
            ms = C.c_double()
            rc = _lib().gnd_host_scene_build(self.handle, threads, C.byref(out), C.byref(ms))
            if rc != 0:
                raise GndError(rc, host_error())
            self._arrays = out
            self.build_ms = ms.value
        return self._arrays

    def camera(self) -> A.gnd_camera_desc:
        c = A.gnd_camera_desc()
        C.memmove(C.byref(c), C.byref(self.desc().contents.camera), C.sizeof(c))
        return c

    def default_job(self, spp: int | None = None, seed_offset: int = 0) -> Job:
        d = self.desc().contents
        return Job(self.camera(), spp if spp is not None else d.spp, d.max_depth, d.integrator, seed_offset,
                   d.ao_length if d.ao_length > 0 else 5.0)

    def free(self) -> None:
        if self.handle:
            _lib().gnd_host_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceScene:
    """A scene resident in HBM (gnd_hip_scene_upload)."""

    def __init__(self, dev: Device, host: HostScene, threads: int = 0):
        self.dev = dev
        arrays = host.build(threads)
        h = C.c_void_p()
        dev._check(_lib().gnd_hip_scene_upload(dev.ctx, arrays, C.byref(h)))
        self.handle = h
        self.host = host

    @property
    def nbytes(self) -> int:
        return int(_lib().gnd_hip_scene_bytes(self.handle))

    def free(self) -> None:
        if self.handle:
            _lib().gnd_hip_scene_free(self.dev.ctx, self.handle)
            self.handle = None


class Integrator:
    """Integrator plugin (gnd_host_integrator_*, include/gianduia/core/integrator.h:18-40):
    render(scene) into the integrator's film in batches of passes, with a
    progress callback per batch and cancel() from any thread."""

    def __init__(self, dev: Device, name: str | None = None, scene: "HostScene | None" = None,
                 max_depth: int = -1, passes_per_callback: int = 1):
        self.dev = dev
        h = C.c_void_p()
        if name is None:
            if scene is None:
                raise ValueError("an integrator name or a scene is needed")
            rc = _lib().gnd_host_integrator_from_scene(dev.ctx, scene.handle, passes_per_callback, C.byref(h))
        else:
            rc = _lib().gnd_host_integrator_create(dev.ctx, name.encode(), max_depth, passes_per_callback, C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        self.handle = h
        self._cb = None
        self.callback_error = None

    def set_callback(self, fn) -> None:
        """fn(sample_index, rgb) with rgb an (H, W, 3) float32 copy of the resolved film."""
        if fn is None:
            self._cb = None
            _lib().gnd_host_integrator_set_callback(self.handle, None, None)
            return

        def tramp(idx, rgb, w, h, _user):
            try:
                img = np.ctypeslib.as_array(rgb, shape=(h * w * 3,)).reshape(h, w, 3).copy()
                fn(idx, img)
            except Exception as e:  # cannot unwind through C
                self.callback_error = e

        self._cb = A.PROGRESS_FN(tramp)
        _lib().gnd_host_integrator_set_callback(self.handle, C.cast(self._cb, C.c_void_p), None)

    def render(self, scene: "HostScene") -> None:
        rc = _lib().gnd_host_integrator_render(self.handle, scene.handle)
        if rc != 0:
            raise GndError(rc, host_error())
        if self.callback_error is not None:
            raise self.callback_error

    def cancel(self) -> None:
        _lib().gnd_host_integrator_cancel(self.handle)

    def film(self):
        """(resolved (H, W, 3) float32 image, passes rendered)."""
        w, h, n = C.c_int(), C.c_int(), C.c_int()
        _lib().gnd_host_integrator_film(self.handle, None, C.byref(w), C.byref(h), C.byref(n))
        out = np.zeros((h.value, w.value, 3), np.float32)
        _lib().gnd_host_integrator_film(self.handle, out.ctypes.data_as(A.f32p), None, None, None)
        return out, n.value

    def set_aovs(self, enabled: bool = True) -> None:
        """Film AOVs (albedo, normal, depth, roughness, metallic) on the next render."""
        _lib().gnd_host_integrator_set_aovs(self.handle, int(bool(enabled)))

    def aovs(self) -> dict:
        """The resolved AOV images of the last render (set_aovs first)."""
        w, h = C.c_int(), C.c_int()
        rc = _lib().gnd_host_integrator_aovs(self.handle, None, C.byref(w), C.byref(h))
        if rc != 0:
            raise GndError(rc, host_error())
        out = np.zeros((h.value, w.value, 9), np.float32)
        _lib().gnd_host_integrator_aovs(self.handle, out.ctypes.data_as(A.f32p), None, None)
        return {"albedo": out[..., 0:3], "normal": out[..., 3:6], "depth": out[..., 6], "roughness": out[..., 7],
                "metallic": out[..., 8]}

    def save_aov(self, aov: str, path: str) -> None:
        """Writes one AOV image ("albedo", "normal", "depth", "roughness",
        "metallic") as .exr or .pfm."""
        rc = _lib().gnd_host_integrator_save_aov(self.handle, aov.encode(), path.encode())
        if rc != 0:
            raise GndError(rc, host_error())

    def denoise(self) -> None:
        """Denoiser::execute on the film: the resolved radiance is replaced
        by its denoised version (guided by the AOVs when they are on)."""
        rc = _lib().gnd_host_integrator_denoise(self.handle)
        if rc != 0:
            raise GndError(rc, host_error())

    def set_output(self, path_stem: str | None) -> None:
        """Every render then ends by writing <path_stem>.exr and .png (the
        reference's film->saveEXR(); film->savePNG()); None turns it off."""
        rc = _lib().gnd_host_integrator_set_output(self.handle, path_stem.encode() if path_stem else None)
        if rc != 0:
            raise GndError(rc, host_error())

    def save(self, path: str) -> None:
        rc = _lib().gnd_host_integrator_save(self.handle, path.encode())
        if rc != 0:
            raise GndError(rc, host_error())

    def close(self) -> None:
        if self.handle:
            _lib().gnd_host_integrator_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def integrator_names() -> list[str]:
    n = _lib().gnd_host_integrator_names(None, 0)
    buf = C.create_string_buffer(n + 1)
    _lib().gnd_host_integrator_names(buf, n + 1)
    return buf.value.decode().split(",") if n else []


def sampler_floats(sample_index: int, n: int) -> np.ndarray:
    """IndependentSampler: the first n floats of a sample index."""
    out = np.zeros(n, np.float32)
    _lib().gnd_host_sampler_floats(sample_index, n, out.ctypes.data_as(A.f32p))
    return out


def load_exr(path: str) -> np.ndarray:
    """An OpenEXR image as (H, W, 3) float32 R, G, B, read like Film's EXR
    constructor (film.cpp:32-48): a channel the file lacks reads as 0."""
    w, h = C.c_int(), C.c_int()
    rc = _lib().gnd_host_image_load(str(path).encode(), C.byref(w), C.byref(h), None, 0)
    if rc != 0:
        raise GndError(rc, host_error())
    out = np.zeros((h.value, w.value, 3), np.float32)
    rc = _lib().gnd_host_image_load(str(path).encode(), C.byref(w), C.byref(h), out.ctypes.data_as(A.f32p), out.size)
    if rc != 0:
        raise GndError(rc, host_error())
    return out


def camera_lookat(cam: A.gnd_camera_desc, origin, target, up=(0.0, 1.0, 0.0)) -> A.gnd_camera_desc:
    """A copy of `cam` posed with the reference's lookAt (transform.h:141-144)."""
    c = A.gnd_camera_desc()
    C.memmove(C.byref(c), C.byref(cam), C.sizeof(c))
    f3 = C.c_float * 3
    rc = _lib().gnd_host_camera_lookat(C.byref(c), f3(*origin), f3(*target), f3(*up))
    if rc != 0:
        raise GndError(rc, host_error())
    return c


def resolve_aovs(rgbw: np.ndarray, aov: np.ndarray) -> dict:
    """Film::resolve for the AOV images (film.cpp:50-95): sums * (1/weight);
    where the weight is 0: albedo and normal 0, depth -1, roughness and
    metallic 0."""
    w = rgbw[..., 3:4]
    has = w > 0
    inv = np.where(has, np.float32(1.0) / np.where(has, w, np.float32(1.0)), np.float32(0.0))
    v = (aov * inv).astype(np.float32)
    depth = np.where(has[..., 0], v[..., 6], np.float32(-1.0)).astype(np.float32)
    return {"albedo": v[..., 0:3], "normal": v[..., 3:6], "depth": depth, "roughness": v[..., 7],
            "metallic": v[..., 8]}


def resolve(rgbw: np.ndarray) -> np.ndarray:
    """Film::resolve: rgb = sum * (1/weight), 0 where the weight is 0."""
    w = rgbw[..., 3:4]
    inv = np.where(w > 0, np.float32(1.0) / np.where(w > 0, w, np.float32(1.0)), np.float32(0.0))
    return (rgbw[..., :3] * inv).astype(np.float32)
