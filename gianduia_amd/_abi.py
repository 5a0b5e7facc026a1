"""ctypes mirror of include/gnd_scene.h, include/gnd_hip.h and include/gnd_host.h."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgnd_hip.so")

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)

GND_SHAPE_MESH, GND_SHAPE_SPHERE = 0, 1
GND_MAT_MATTE, GND_MAT_CONDUCTOR, GND_MAT_GLASS, GND_MAT_PLASTIC, GND_MAT_INDEX = range(5)
GND_MEDIUM_HOMOGENEOUS, GND_MEDIUM_GRID = 0, 1
GND_CAMERA_PERSPECTIVE, GND_CAMERA_THINLENS = 0, 1
GND_FILTER_GAUSSIAN, GND_FILTER_BOX = 0, 1
GND_INTEGRATOR_PATH, GND_INTEGRATOR_VOLPATH, GND_INTEGRATOR_DIRECT, GND_INTEGRATOR_AO = 0, 1, 2, 3

GND_OK, GND_E_INVALID, GND_E_DEVICE, GND_E_NOMEM, GND_E_CANCELLED, GND_E_STACK = 0, -1, -2, -3, -4, -5
GND_E_TIMEOUT, GND_E_BUSY = -6, -7


class gnd_mesh_desc(C.Structure):
    _fields_ = [("ntris", C.c_uint32), ("positions", f32p), ("normals", f32p), ("uvs", f32p),
                ("vertex_ids", u32p)]


class gnd_material_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo_is_checker", C.c_int32), ("albedo", C.c_float * 3),
                ("R", C.c_float * 3), ("T", C.c_float * 3), ("eta_rgb", C.c_float * 3), ("k_rgb", C.c_float * 3),
                ("Kd", C.c_float * 3), ("Ks", C.c_float * 3), ("eta", C.c_float), ("roughness", C.c_float),
                ("medium_inside", C.c_int32), ("medium_outside", C.c_int32)]


class gnd_medium_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("g", C.c_float),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("density", f32p),
                ("density_scale", C.c_float), ("macro_res", C.c_int32), ("to_local", C.c_float * 16),
                ("to_local_inv", C.c_float * 16)]


class gnd_primitive_desc(C.Structure):
    _fields_ = [("shape_type", C.c_int32), ("mesh", C.c_int32), ("radius", C.c_float), ("material", C.c_int32),
                ("has_emitter", C.c_int32), ("radiance", C.c_float * 3), ("to_world", C.c_float * 16),
                ("to_world_inv", C.c_float * 16), ("has_motion", C.c_int32), ("to_world_end", C.c_float * 16),
                ("to_world_end_inv", C.c_float * 16)]


class gnd_camera_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("fov", C.c_float),
                ("to_world", C.c_float * 16), ("to_world_inv", C.c_float * 16), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float), ("k1", C.c_float), ("ca_axial", C.c_float),
                ("ca_lateral", C.c_float), ("vignetting", C.c_int32), ("blades", C.c_int32),
                ("filter", C.c_int32), ("filter_radius", C.c_float * 2), ("filter_alpha", C.c_float),
                ("has_motion", C.c_int32), ("to_world_end", C.c_float * 16), ("rolling_shutter", C.c_int32),
                ("scan_duration", C.c_float), ("exposure_duration", C.c_float)]


GND_LIGHT_POINT, GND_LIGHT_SPOT, GND_LIGHT_DIRECTIONAL, GND_LIGHT_ENVIRONMENT = range(4)


class gnd_light_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("emitter_index", C.c_int32), ("position", C.c_float * 3),
                ("direction", C.c_float * 3), ("power", C.c_float * 3), ("irradiance", C.c_float * 3),
                ("spot_size", C.c_float), ("blend", C.c_float), ("strength", C.c_float),
                ("env_width", C.c_int32), ("env_height", C.c_int32), ("env_rgb", f32p)]


class gnd_scene_desc(C.Structure):
    _fields_ = [("nmeshes", C.c_uint32), ("meshes", C.POINTER(gnd_mesh_desc)),
                ("nprims", C.c_uint32), ("prims", C.POINTER(gnd_primitive_desc)),
                ("nmaterials", C.c_uint32), ("materials", C.POINTER(gnd_material_desc)),
                ("nmedia", C.c_uint32), ("media", C.POINTER(gnd_medium_desc)),
                ("camera", gnd_camera_desc), ("integrator", C.c_int32), ("max_depth", C.c_int32),
                ("spp", C.c_uint32), ("nlights", C.c_uint32), ("lights", C.POINTER(gnd_light_desc)),
                ("ao_length", C.c_float)]


class gnd_ray_soa(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class gnd_hit_soa(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "inst", "prim", "u", "v")]


class gnd_bvh4_node(C.Structure):
    _fields_ = [("min_x", C.c_float * 4), ("min_y", C.c_float * 4), ("min_z", C.c_float * 4),
                ("max_x", C.c_float * 4), ("max_y", C.c_float * 4), ("max_z", C.c_float * 4),
                ("offset", C.c_uint32 * 4), ("count", C.c_uint32 * 4), ("child_type", C.c_uint8 * 4),
                ("pad", C.c_uint8 * 12)]


class gnd_tri_pack4(C.Structure):
    _fields_ = [(n, C.c_float * 4) for n in ("v0x", "v0y", "v0z", "v1x", "v1y", "v1z", "v2x", "v2y", "v2z")] + \
               [("prim", C.c_uint32 * 4)]


class gnd_mesh_info(C.Structure):
    _fields_ = [("node_base", C.c_uint32), ("n_nodes", C.c_uint32), ("pack_base", C.c_uint32),
                ("n_packs", C.c_uint32), ("tri_base", C.c_uint32), ("n_tris", C.c_uint32),
                ("cdf_base", C.c_uint32), ("has_uv", C.c_uint32), ("total_area", C.c_float),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3)]


class gnd_hip_scene_arrays(C.Structure):
    _fields_ = [("desc", C.POINTER(gnd_scene_desc)),
                ("n_tlas_nodes", C.c_uint32), ("tlas_nodes", C.POINTER(gnd_bvh4_node)),
                ("n_tlas_prims", C.c_uint32), ("tlas_prims", i32p),
                ("n_meshes", C.c_uint32), ("meshes", C.POINTER(gnd_mesh_info)),
                ("n_blas_nodes", C.c_uint32), ("blas_nodes", C.POINTER(gnd_bvh4_node)),
                ("n_packs", C.c_uint32), ("packs", C.POINTER(gnd_tri_pack4)),
                ("n_tris", C.c_uint32), ("shading", f32p),
                ("n_cdf", C.c_uint32), ("cdf", f32p)]


class gnd_hip_job(C.Structure):
    _fields_ = [("camera", gnd_camera_desc), ("spp", C.c_uint32), ("max_depth", C.c_int32),
                ("integrator", C.c_int32), ("seed_offset", C.c_uint64), ("film", C.c_void_p),
                ("film_sq", C.c_void_p), ("film_aov", C.c_void_p), ("ao_length", C.c_float),
                ("accumulate", C.c_int32)]


class gnd_hip_stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_trace", C.c_double), ("ms_shadow", C.c_double),
                ("ms_shade", C.c_double), ("samples", C.c_uint64), ("rays_closest", C.c_uint64),
                ("rays_shadow", C.c_uint64), ("node_visits", C.c_uint64), ("pack_tests", C.c_uint64),
                ("inst_tests", C.c_uint64), ("trace_launches", C.c_uint64), ("nan_samples", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_pack_tests", C.c_uint64),
                ("shadow_inst_tests", C.c_uint64), ("max_ray_nodes", C.c_uint64), ("bad_rays", C.c_uint64),
                ("lane_iterations", C.c_uint64), ("shadow_lane_iterations", C.c_uint64), ("ms_tail", C.c_double),
                ("rays_tail", C.c_uint64), ("lane_steps", C.c_uint64), ("shadow_lane_steps", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class gnd_hip_tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("stack_lds", C.c_int32), ("queue_capacity", C.c_int64),
                ("use_graph", C.c_int32), ("refill_lanes", C.c_int32), ("shadow_refill_lanes", C.c_int32),
                ("stack_limit", C.c_int32), ("tail_paths", C.c_int32)]


class gnd_hip_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("path_slots", C.c_uint32), ("max_jobs", C.c_uint32),
                ("lanes", C.c_uint32)]


class gnd_hip_job_status(C.Structure):
    _fields_ = [("state", C.c_int32), ("passes_done", C.c_uint32), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


GND_PROP_BOOLEAN, GND_PROP_INTEGER, GND_PROP_FLOAT, GND_PROP_STRING, GND_PROP_POINT, GND_PROP_VECTOR2, \
    GND_PROP_VECTOR, GND_PROP_COLOR, GND_PROP_TRANSFORM = range(9)


class gnd_prop(C.Structure):
    _fields_ = [("name", C.c_char_p), ("kind", C.c_int32), ("ivalue", C.c_int32), ("v", C.c_float * 3),
                ("svalue", C.c_char_p), ("matrix", f32p)]


# exported symbols with their signatures: name -> (restype, argtypes)
HIP_API = {
    "gnd_hip_init": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gnd_hip_shutdown": (None, [C.c_void_p]),
    "gnd_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "gnd_hip_device_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "gnd_hip_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_tuning)]),
    "gnd_hip_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_tuning)]),
    "gnd_hip_set_counters": (C.c_int, [C.c_void_p, C.c_int]),
    "gnd_hip_malloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "gnd_hip_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gnd_hip_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "gnd_hip_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "gnd_hip_memset": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    "gnd_hip_synchronize": (C.c_int, [C.c_void_p]),
    "gnd_hip_scene_upload": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_scene_arrays), C.POINTER(C.c_void_p)]),
    "gnd_hip_scene_free": (None, [C.c_void_p, C.c_void_p]),
    "gnd_hip_scene_bytes": (C.c_size_t, [C.c_void_p]),
    "gnd_hip_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gnd_ray_soa), C.c_int64,
                                        C.POINTER(gnd_hit_soa), C.POINTER(gnd_hip_stats)]),
    "gnd_hip_trace_any": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gnd_ray_soa), C.c_int64, C.c_void_p,
                                    C.POINTER(gnd_hip_stats)]),
    "gnd_hip_render": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gnd_hip_job), C.c_int,
                                 C.POINTER(gnd_hip_stats)]),
    "gnd_hip_cancel": (C.c_int, [C.c_void_p]),
    "gnd_hip_copy_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "gnd_hip_denoise": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                  C.POINTER(C.c_double)]),
    "gnd_hip_denoise_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                          C.c_void_p, C.POINTER(C.c_double)]),
    "gnd_hip_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(gnd_hip_job), C.POINTER(gnd_hip_tuning)]),
    "gnd_hip_server_create": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_server_config), C.POINTER(C.c_void_p)]),
    "gnd_hip_server_destroy": (None, [C.c_void_p]),
    "gnd_hip_server_add_scene": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_scene_arrays), C.POINTER(C.c_int32)]),
    "gnd_hip_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(gnd_hip_job), C.POINTER(C.c_uint64)]),
    "gnd_hip_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gnd_hip_job_status)]),
    "gnd_hip_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(gnd_hip_job_status)]),
    "gnd_hip_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gnd_hip_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gnd_hip_server_stats": (C.c_int, [C.c_void_p, C.POINTER(gnd_hip_stats), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_size_t)]),
}

HOST_API = {
    "gnd_host_last_error": (C.c_char_p, []),
    "gnd_host_scene_create": (C.c_int, [C.POINTER(gnd_scene_desc), C.POINTER(C.c_void_p)]),
    "gnd_host_scene_load_xml": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "gnd_host_scene_generate": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "gnd_host_scene_free": (None, [C.c_void_p]),
    "gnd_host_builder_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "gnd_host_builder_object": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(gnd_prop),
                                          C.c_int, C.POINTER(C.c_int32)]),
    "gnd_host_builder_add_child": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "gnd_host_builder_finish": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "gnd_host_builder_free": (None, [C.c_void_p]),
    "gnd_host_scene_desc": (C.POINTER(gnd_scene_desc), [C.c_void_p]),
    "gnd_host_scene_build": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.POINTER(gnd_hip_scene_arrays)),
                                       C.POINTER(C.c_double)]),
    "gnd_host_camera_lookat": (C.c_int, [C.POINTER(gnd_camera_desc), f32p, f32p, f32p]),
    "gnd_host_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, f32p, C.POINTER(gnd_hip_stats)]),
    "gnd_host_integrator_create": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "gnd_host_integrator_from_scene": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "gnd_host_integrator_set_callback": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gnd_host_integrator_render": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gnd_host_integrator_cancel": (C.c_int, [C.c_void_p]),
    "gnd_host_integrator_film": (C.c_int, [C.c_void_p, f32p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
    "gnd_host_integrator_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "gnd_host_integrator_set_output": (C.c_int, [C.c_void_p, C.c_char_p]),
    "gnd_host_integrator_set_aovs": (C.c_int, [C.c_void_p, C.c_int]),
    "gnd_host_integrator_save_aov": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "gnd_host_integrator_aovs": (C.c_int, [C.c_void_p, f32p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gnd_host_integrator_free": (None, [C.c_void_p]),
    "gnd_host_integrator_names": (C.c_int, [C.c_char_p, C.c_int]),
    "gnd_host_sampler_floats": (C.c_int, [C.c_uint64, C.c_int, f32p]),
    "gnd_host_image_save": (C.c_int, [C.c_char_p, f32p, C.c_int, C.c_int]),
    "gnd_host_integrator_denoise": (C.c_int, [C.c_void_p]),
    "gnd_host_image_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), f32p, C.c_size_t]),
}

# gnd_host_progress_fn
PROGRESS_FN = C.CFUNCTYPE(None, C.c_int, f32p, C.c_int, C.c_int, C.c_void_p)

_lib = None


def load_library(path: str | None = None) -> C.CDLL:
    """Loads libgnd_hip.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built: run `python -m gianduia_amd.build`")
    lib = C.CDLL(p)
    for table in (HIP_API, HOST_API):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
