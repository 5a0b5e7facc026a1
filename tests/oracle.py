"""ctypes access to the CPU oracle (oracle/_build/libgnd_ref.so) for the tests.

The oracle is the checker: tests call it on the same inputs as the HIP
library and compare.  It is built by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from gianduia_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libgnd_ref.so")

f32p = C.POINTER(C.c_float)

_API = {
    "gnd_ref_last_error": (C.c_char_p, []),
    "gnd_ref_scene_create": (C.c_void_p, [C.POINTER(A.gnd_scene_desc)]),
    "gnd_ref_scene_free": (None, [C.c_void_p]),
    "gnd_ref_scene_counts": (None, [C.c_void_p, C.POINTER(C.c_int64)]),
    "gnd_ref_mesh_nodes": (C.c_int64, [C.c_void_p, C.c_int, C.c_void_p]),
    "gnd_ref_mesh_packs": (C.c_int64, [C.c_void_p, C.c_int, C.c_void_p]),
    "gnd_ref_mesh_cdf": (C.c_int64, [C.c_void_p, C.c_int, f32p, f32p]),
    "gnd_ref_tlas_nodes": (C.c_int64, [C.c_void_p, C.c_void_p]),
    "gnd_ref_tlas_prims": (C.c_int64, [C.c_void_p, C.POINTER(C.c_int32)]),
    "gnd_ref_sizeof_node4": (C.c_int, []),
    "gnd_ref_sizeof_pack": (C.c_int, []),
    "gnd_ref_trace": (C.c_int, [C.c_void_p, C.POINTER(A.gnd_ray_soa), C.c_int64, C.POINTER(A.gnd_hit_soa), C.c_int,
                                C.c_void_p]),
    "gnd_ref_render": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, f32p, f32p,
                                 C.POINTER(C.c_double), f32p]),
    "gnd_ref_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_uint32)]),
    "gnd_ref_splitmix64": (C.c_uint64, [C.c_uint64]),
    "gnd_ref_sampler_floats": (None, [C.c_uint64, C.c_int, f32p]),
    "gnd_ref_denoise": (None, [f32p, f32p, f32p, C.c_int, C.c_int, f32p]),
    "gnd_ref_area_sample": (C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p]),
    "gnd_ref_fr_dielectric": (C.c_float, [C.c_float, C.c_float, C.c_float]),
    "gnd_ref_refract": (C.c_int, [f32p, f32p, C.c_float, f32p]),
    "gnd_ref_fr_conductor": (None, [C.c_float, f32p, f32p, f32p, f32p]),
    "gnd_ref_roughness_to_alpha": (C.c_float, [C.c_float]),
    "gnd_ref_ggx": (C.c_float, [C.c_float, C.c_int, f32p, f32p]),
    "gnd_ref_ggx_sample": (None, [C.c_float, f32p, f32p, f32p]),
    "gnd_ref_lobe_f": (None, [C.POINTER(A.gnd_material_desc), f32p, f32p, f32p]),
    "gnd_ref_lobe_pdf": (C.c_float, [C.POINTER(A.gnd_material_desc), f32p, f32p]),
    "gnd_ref_lobe_sample": (C.c_int, [C.POINTER(A.gnd_material_desc), f32p, f32p, C.c_float, f32p, f32p, f32p]),
    "gnd_ref_bsdf_sample_weight": (C.c_int, [C.POINTER(A.gnd_material_desc), f32p, f32p, C.c_float, f32p, f32p,
                                             f32p]),
    "gnd_ref_warp": (None, [C.c_int, f32p, f32p]),
    "gnd_ref_cosine_hemisphere_pdf": (C.c_float, [f32p]),
    "gnd_ref_frame": (None, [f32p, f32p]),
    "gnd_ref_hg_pdf": (C.c_float, [C.c_float, f32p, f32p]),
    "gnd_ref_hg_sample": (C.c_float, [C.c_float, f32p, f32p, f32p]),
    "gnd_ref_luminance": (C.c_float, [C.c_float, C.c_float, C.c_float]),
    "gnd_ref_bounds_intersect": (C.c_int, [f32p, f32p, f32p, f32p, C.c_float, f32p]),
    "gnd_ref_camera_ray": (None, [C.c_void_p, f32p, f32p, C.c_float, C.c_float, f32p]),
    "gnd_ref_motion_matrix": (None, [f32p, f32p, C.c_float, f32p]),
    "gnd_ref_decompose": (None, [f32p, f32p, f32p, f32p]),
    "gnd_ref_surface": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "gnd_ref_medium": (C.c_int, [C.c_void_p, C.c_int, C.c_int, f32p, f32p, C.c_float, C.c_uint64, f32p]),
    "gnd_ref_grid_sample": (C.c_float, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]),
    "gnd_ref_path_sample": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, f32p, f32p]),
    "gnd_ref_light_sample": (None, [C.c_void_p, C.c_int, f32p, f32p, f32p]),
    "gnd_ref_env_pdf": (C.c_float, [C.c_void_p, f32p, f32p]),
}

_lib = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        L = C.CDLL(ORACLE_LIB)
        for name, (res, args) in _API.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def fp(a: np.ndarray):
    return a.ctypes.data_as(f32p)


def vec(*v) -> np.ndarray:
    return np.array(v, dtype=np.float32)


def motion_matrix(start: np.ndarray, end: np.ndarray, time: float) -> np.ndarray:
    """AnimatedTransform(start, 0, end, 1).interpolate(time) of column-major
    4x4 float32 matrices (returned column-major, flat)."""
    out = np.zeros(16, np.float32)
    lib().gnd_ref_motion_matrix(fp(np.ascontiguousarray(start, np.float32).ravel()),
                                fp(np.ascontiguousarray(end, np.float32).ravel()), time, fp(out))
    return out


def decompose(m: np.ndarray):
    """glm::decompose of a column-major affine matrix: T, quaternion (x, y, z, w), S."""
    T, Q, S = np.zeros(3, np.float32), np.zeros(4, np.float32), np.zeros(3, np.float32)
    lib().gnd_ref_decompose(fp(np.ascontiguousarray(m, np.float32).ravel()), fp(T), fp(Q), fp(S))
    return T, Q, S


def denoise(color: np.ndarray, albedo: np.ndarray | None = None, normal: np.ndarray | None = None) -> np.ndarray:
    """The film denoiser's a-trous filter on the CPU ((H, W, 3) float32)."""
    color = np.ascontiguousarray(color, np.float32)
    H, W = color.shape[:2]
    a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    n = None if normal is None else np.ascontiguousarray(normal, np.float32)
    out = np.empty_like(color)
    lib().gnd_ref_denoise(fp(color), None if a is None else fp(a), None if n is None else fp(n), W, H, fp(out))
    return out


class RefScene:
    def __init__(self, desc_ptr):
        d = desc_ptr.contents if hasattr(desc_ptr, "contents") else desc_ptr
        self.width, self.height = int(d.camera.width), int(d.camera.height)  # the film gnd_ref_render fills
        self.h = lib().gnd_ref_scene_create(desc_ptr)
        if not self.h:
            raise RuntimeError(lib().gnd_ref_last_error().decode())

    def __del__(self):
        try:
            if self.h:
                lib().gnd_ref_scene_free(self.h)
        except Exception:
            pass

    def counts(self):
        c = (C.c_int64 * 4)()
        lib().gnd_ref_scene_counts(self.h, c)
        return list(c)

    def mesh_nodes(self, m: int) -> np.ndarray:
        n = lib().gnd_ref_mesh_nodes(self.h, m, None)
        buf = np.zeros(n * 144, dtype=np.uint8)
        lib().gnd_ref_mesh_nodes(self.h, m, buf.ctypes.data_as(C.c_void_p))
        return buf.reshape(n, 144)

    def mesh_packs(self, m: int) -> np.ndarray:
        n = lib().gnd_ref_mesh_packs(self.h, m, None)
        buf = np.zeros(n * 160, dtype=np.uint8)
        lib().gnd_ref_mesh_packs(self.h, m, buf.ctypes.data_as(C.c_void_p))
        return buf.reshape(n, 160)

    def mesh_cdf(self, m: int):
        n = lib().gnd_ref_mesh_cdf(self.h, m, None, None)
        out = np.zeros(n, dtype=np.float32)
        ta = C.c_float()
        lib().gnd_ref_mesh_cdf(self.h, m, fp(out), C.byref(ta))
        return out, ta.value

    def tlas_nodes(self) -> np.ndarray:
        n = lib().gnd_ref_tlas_nodes(self.h, None)
        buf = np.zeros(n * 144, dtype=np.uint8)
        lib().gnd_ref_tlas_nodes(self.h, buf.ctypes.data_as(C.c_void_p))
        return buf.reshape(n, 144)

    def tlas_prims(self) -> np.ndarray:
        n = lib().gnd_ref_tlas_prims(self.h, None)
        out = np.zeros(n, dtype=np.int32)
        lib().gnd_ref_tlas_prims(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def trace(self, rays: dict, any_hit: bool = False, counts: bool = False):
        n = len(rays["ox"])
        arrs = {k: np.ascontiguousarray(rays[k], dtype=np.float32) for k in
                ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")}
        soa = A.gnd_ray_soa(*[arrs[k].ctypes.data for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")])
        out = {"t": np.zeros(n, np.float32), "inst": np.zeros(n, np.int32), "prim": np.zeros(n, np.int32),
               "u": np.zeros(n, np.float32), "v": np.zeros(n, np.float32)}
        hs = A.gnd_hit_soa(*[out[k].ctypes.data for k in ("t", "inst", "prim", "u", "v")])
        cnt = np.zeros(3 * max(n, 1), np.int64) if counts else None
        lib().gnd_ref_trace(self.h, C.byref(soa), n, C.byref(hs), int(any_hit),
                            cnt.ctypes.data_as(C.c_void_p) if counts else None)
        if any_hit:
            out = out["inst"].copy()
        return (out, cnt.reshape(-1, 3)[:n]) if counts else out

    def render(self, threads=1, spp=None, spp_begin=0, row_begin=0, row_end=-1, seed_offset=0, squares=False,
               width=None, height=None, aovs=False):
        """Returns (rgbw, stats), (rgbw, sq, stats) with squares, and the AOV
        sums (H, W, 9) appended with aovs."""
        W = self.width if width is None else width
        H = self.height if height is None else height
        if (W, H) != (self.width, self.height):  # the oracle writes its camera's whole film
            raise ValueError(f"film {W}x{H} != the scene camera's {self.width}x{self.height}")
        rgbw = np.zeros((H, W, 4), np.float32)
        sq = np.zeros((H, W, 3), np.float32) if squares else None
        aov = np.zeros((H, W, 9), np.float32) if aovs else None
        stats = (C.c_double * 5)()
        rc = lib().gnd_ref_render(self.h, threads, spp_begin, -1 if spp is None else spp, row_begin, row_end,
                                  seed_offset, fp(rgbw), fp(sq) if squares else None, stats,
                                  fp(aov) if aovs else None)
        if rc != 0:
            raise RuntimeError(lib().gnd_ref_last_error().decode())
        out = (rgbw, sq, list(stats)) if squares else (rgbw, list(stats))
        return out + (aov,) if aovs else out

    def camera_ray(self, pfilm, plens=(0.5, 0.5), lam=0.0, time=0.0, with_time=False):
        """Camera ray (origin, direction, weight[, ray time]) for a film
        position, lens sample, chromatic offset and time draw."""
        out = np.zeros(8, np.float32)
        lib().gnd_ref_camera_ray(self.h, fp(vec(*pfilm)), fp(vec(*plens)), lam, time, fp(out))
        if with_time:
            return out[:3], out[3:6], float(out[6]), float(out[7])
        return out[:3], out[3:6], float(out[6])

    def area_sample(self, prim, ref_p, u):
        """Area light of primitive `prim` sampled from ref_p: (p, n, pdf), pdf 0 when unset."""
        out = np.zeros(7, np.float32)
        lib().gnd_ref_area_sample(self.h, prim, fp(vec(*ref_p)), fp(vec(*u)), fp(out))
        return out[0:3], out[3:6], float(out[6])

    def light_sample(self, k, ref_p, u):
        """Light k's NEE sample from ref_p: dict(p, pdf, Li, dir, tmax)."""
        out = np.zeros(11, np.float32)
        lib().gnd_ref_light_sample(self.h, k, fp(vec(*ref_p)), fp(vec(*u)), fp(out))
        return {"p": out[0:3], "pdf": float(out[3]), "Li": out[4:7], "dir": out[7:10], "tmax": float(out[10])}

    def env_pdf(self, d):
        """(solid-angle pdf of the environment light towards d, its radiance along d)."""
        le = np.zeros(3, np.float32)
        pdf = lib().gnd_ref_env_pdf(self.h, fp(vec(*d)), fp(le))
        return pdf, le

    def path_sample(self, x, y, p, seed_offset=0, aov=False):
        """Radiance of one (pixel, pass) path; with aov also its AOV record
        (albedo rgb, normal xyz, depth, roughness, metallic)."""
        out = np.zeros(3, np.float32)
        rec = np.zeros(9, np.float32)
        lib().gnd_ref_path_sample(self.h, x, y, p, seed_offset, fp(out), fp(rec) if aov else None)
        return (out, rec) if aov else out
