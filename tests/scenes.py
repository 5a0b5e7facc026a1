"""Scene descriptions built in the tests (random triangle soups, instancing)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from gianduia_amd import _abi as A

IDENT = np.eye(4, dtype=np.float32)


def col_major(m: np.ndarray):
    """4x4 row-major numpy matrix -> column-major float[16] (glm memory order)."""
    return (C.c_float * 16)(*np.asarray(m, dtype=np.float32).T.reshape(-1).tolist())


def trs(scale=(1, 1, 1), rot_deg=0.0, axis=(0, 1, 0), trans=(0, 0, 0)) -> tuple[np.ndarray, np.ndarray]:
    """Returns (M, Minv) = T*R*S and S^-1*R^T*T^-1 in float32."""
    ax = np.asarray(axis, np.float64)
    ax = ax / np.linalg.norm(ax)
    a = np.radians(rot_deg)
    c, s = np.cos(a), np.sin(a)
    x, y, z = ax
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s, 0],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s, 0],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c), 0],
                  [0, 0, 0, 1]])
    S = np.diag([scale[0], scale[1], scale[2], 1.0])
    T = np.eye(4)
    T[:3, 3] = trans
    M = T @ R @ S
    return M.astype(np.float32), np.linalg.inv(M).astype(np.float32)


class SceneBuilder:
    def __init__(self):
        self.meshes = []      # (positions (n,9) float32, normals or None)
        self.prims = []
        self.materials = []
        self.keep = []

    def mesh(self, tris: np.ndarray, normals: np.ndarray | None = None) -> int:
        self.meshes.append((np.ascontiguousarray(tris, np.float32).reshape(-1, 9),
                            None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)))
        return len(self.meshes) - 1

    def matte(self, rgb=(0.5, 0.5, 0.5)) -> int:
        m = A.gnd_material_desc()
        m.type = A.GND_MAT_MATTE
        m.albedo[:] = rgb
        m.medium_inside = m.medium_outside = -1
        self.materials.append(m)
        return len(self.materials) - 1

    def light(self, kind, position=(0, 0, 0), direction=None, power=(10, 10, 10), irradiance=(1, 1, 1),
              spot_size=45.0, blend=0.15, strength=1.0, env=None) -> int:
        """Appends a light after every emitter added so far (document order)."""
        if not hasattr(self, "lights"):
            self.lights = []
        l = A.gnd_light_desc()
        l.type = {"point": A.GND_LIGHT_POINT, "spot": A.GND_LIGHT_SPOT, "directional": A.GND_LIGHT_DIRECTIONAL,
                  "environment": A.GND_LIGHT_ENVIRONMENT}[kind]
        l.emitter_index = sum(1 for p in self.prims if p.has_emitter) + len(self.lights)
        l.position[:] = position
        l.direction[:] = direction if direction is not None else ((0, 0, -1) if kind == "spot" else (0, -1, 0))
        l.power[:] = power
        l.irradiance[:] = irradiance
        l.spot_size, l.blend, l.strength = spot_size, blend, strength
        if env is not None:
            env = np.ascontiguousarray(env, np.float32)
            self.keep.append(env)
            l.env_height, l.env_width = env.shape[0], env.shape[1]
            l.env_rgb = env.ctypes.data_as(C.POINTER(C.c_float))
        self.lights.append(l)
        return len(self.lights) - 1

    def conductor(self, R=(0.9, 0.8, 0.7), roughness=0.0, eta=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14)) -> int:
        m = A.gnd_material_desc()
        m.type = A.GND_MAT_CONDUCTOR
        m.R[:] = R
        m.eta_rgb[:] = eta
        m.k_rgb[:] = k
        m.roughness = roughness
        m.medium_inside = m.medium_outside = -1
        self.materials.append(m)
        return len(self.materials) - 1

    def glass(self, eta=1.5, roughness=0.0) -> int:
        m = A.gnd_material_desc()
        m.type = A.GND_MAT_GLASS
        m.R[:] = (1.0, 1.0, 1.0)
        m.T[:] = (1.0, 1.0, 1.0)
        m.eta = eta
        m.roughness = roughness
        m.medium_inside = m.medium_outside = -1
        self.materials.append(m)
        return len(self.materials) - 1

    def index(self, medium_inside: int, medium_outside: int = -1) -> int:
        """IndexMaterial (src/materials/index.cpp:10-16): a null BSDF that only
        bounds a medium."""
        m = A.gnd_material_desc()
        m.type = A.GND_MAT_INDEX
        m.medium_inside = medium_inside
        m.medium_outside = medium_outside
        self.materials.append(m)
        return len(self.materials) - 1

    def prim(self, mesh: int, M=None, Minv=None, material: int = 0, radiance=None, sphere_radius=None) -> int:
        p = A.gnd_primitive_desc()
        if sphere_radius is not None:
            p.shape_type = A.GND_SHAPE_SPHERE
            p.mesh = -1
            p.radius = sphere_radius
        else:
            p.shape_type = A.GND_SHAPE_MESH
            p.mesh = mesh
            p.radius = 1.0
        p.material = material
        if radiance is not None:
            p.has_emitter = 1
            p.radiance[:] = radiance
        M = IDENT if M is None else M
        Minv = IDENT if Minv is None else Minv
        p.to_world = col_major(M)
        p.to_world_inv = col_major(Minv)
        self.prims.append(p)
        return len(self.prims) - 1

    def desc(self, width=64, height=64) -> A.gnd_scene_desc:
        md = (A.gnd_mesh_desc * max(1, len(self.meshes)))()
        for i, (pos, nrm) in enumerate(self.meshes):
            self.keep += [pos, nrm]
            md[i].ntris = pos.shape[0]
            md[i].positions = pos.ctypes.data_as(C.POINTER(C.c_float))
            md[i].normals = nrm.ctypes.data_as(C.POINTER(C.c_float)) if nrm is not None else None
            md[i].uvs = None
            md[i].vertex_ids = None
        pd = (A.gnd_primitive_desc * len(self.prims))(*self.prims)
        mt = (A.gnd_material_desc * max(1, len(self.materials)))(*self.materials)
        self.keep += [md, pd, mt]
        d = A.gnd_scene_desc()
        d.nmeshes = len(self.meshes)
        d.meshes = md
        d.nprims = len(self.prims)
        d.prims = pd
        d.nmaterials = len(self.materials)
        d.materials = mt
        d.nmedia = 0
        d.media = None
        cam = A.gnd_camera_desc()
        cam.type = A.GND_CAMERA_PERSPECTIVE
        cam.width, cam.height, cam.fov = width, height, 45.0
        cam.to_world = col_major(IDENT)
        cam.to_world_inv = col_major(IDENT)
        cam.filter = A.GND_FILTER_BOX
        d.camera = cam
        d.integrator = A.GND_INTEGRATOR_PATH
        d.max_depth = 4
        d.spp = 1
        lights = getattr(self, "lights", [])
        if lights:
            la = (A.gnd_light_desc * len(lights))(*lights)
            self.keep.append(la)
            d.nlights = len(lights)
            d.lights = la
        # the desc holds raw pointers into arrays this builder owns: keep the
        # builder alive as long as the desc (`soup_scene(...).desc()` would
        # otherwise free them before the library copies the scene)
        d._builder = self
        return d


def random_soup(rng: np.random.Generator, n: int, spread=1.0, size=0.2) -> np.ndarray:
    c = rng.uniform(-spread, spread, size=(n, 1, 3))
    v = c + rng.uniform(-size, size, size=(n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def soup_scene(seed=0, n=2000, instances=1, sphere=False, light=True) -> SceneBuilder:
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    mat = b.matte()
    m = b.mesh(random_soup(rng, n))
    for k in range(instances):
        if k == 0:
            b.prim(m, material=mat)
        else:
            M, Mi = trs(scale=rng.uniform(0.3, 1.0, 3), rot_deg=rng.uniform(0, 360), axis=rng.normal(size=3),
                        trans=rng.uniform(-2, 2, 3))
            b.prim(m, M, Mi, mat)
    if sphere:
        M, Mi = trs(trans=(0.3, -0.2, 0.1))
        b.prim(-1, M, Mi, mat, sphere_radius=0.5)
    if light:
        lm = b.mesh(np.array([[-0.5, 3, -0.5, 0.5, 3, -0.5, 0.5, 3, 0.5]], np.float32))
        b.prim(lm, material=mat, radiance=(5, 5, 5))
    return b


def random_rays(rng: np.random.Generator, n: int, spread=2.5, edge_cases=True) -> dict:
    o = rng.uniform(-spread, spread, size=(n, 3)).astype(np.float32)
    tgt = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    if edge_cases and n >= 64:
        # axis-aligned directions (inf inverse components) and zero components
        d[0:8] = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                           [0.6, 0.8, 0], [0, 0.6, -0.8]], np.float32)
        o[0:8] = 0.0
    tmin = np.full(n, 1e-4, np.float32)
    tmax = np.full(n, np.inf, np.float32)
    if edge_cases and n >= 64:
        tmax[8:16] = rng.uniform(0.1, 3.0, 8).astype(np.float32)   # short rays
        tmax[16:20] = 0.0                                          # empty interval
    return {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(), "dx": d[:, 0].copy(),
            "dy": d[:, 1].copy(), "dz": d[:, 2].copy(), "tmin": tmin, "tmax": tmax}


def quad(x0, x1, y0, y1, z, facing=+1) -> np.ndarray:
    """Two triangles of the rectangle [x0,x1] x [y0,y1] at depth z, wound so
    the geometric normal is +z (facing=+1) or -z (facing=-1)."""
    a, b, c, d = (x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z)
    tris = [a + b + c, a + c + d] if facing > 0 else [a + c + b, a + d + c]
    return np.array(tris, np.float32)


def mirror_scene(front_material: str | None, roughness=0.0, integrator=A.GND_INTEGRATOR_PATH, max_depth=4):
    """Camera at the origin looking down -z (16x16, fov 45, box filter, 1 spp).
    A quad of `front_material` ('conductor', 'glass' or None) at z = -2 faces
    the camera and fills the view; a matte wall (0.2, 0.4, 0.6) at z = +1
    behind the camera faces -z; a small light sits far off to the side."""
    b = SceneBuilder()
    wall = b.matte((0.2, 0.4, 0.6))
    b.prim(b.mesh(quad(-10, 10, -10, 10, 1.0, facing=-1)), material=wall)
    if front_material == "conductor":
        b.prim(b.mesh(quad(-10, 10, -10, 10, -2.0, facing=+1)), material=b.conductor(roughness=roughness))
    elif front_material == "glass":
        b.prim(b.mesh(quad(-10, 10, -10, 10, -2.0, facing=+1)), material=b.glass(roughness=roughness))
    light = np.array([[30, 30, 30, 31, 30, 30, 31, 30, 31]], np.float32)
    b.prim(b.mesh(light), material=wall, radiance=(5, 5, 5))
    d = b.desc(width=16, height=16)
    d.integrator = integrator
    d.max_depth = max_depth
    return d

