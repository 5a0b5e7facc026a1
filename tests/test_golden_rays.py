"""Ray-batch golden fixtures (tests/golden/rays_*.npz, written by
tests/golden/make_ray_goldens.py): seeded rays through the Cornell box and a
10,000-triangle displaced sphere with their closest hits (instance,
primitive, raw float32 bits of t, u and v) and any-hit occlusion.

* CPU: the oracle still produces every golden bit; on the Cornell box (all
  primitives in world space) the goldens also equal a brute-force
  Moller-Trumbore search over all 36 triangles in the reference's arithmetic
  order, so they pin the traversal independently of the BVH.
* GPU: gnd_hip_trace_closest / gnd_hip_trace_any reproduce the goldens bit for
  bit, with nothing computed by the oracle at run time."""
import ctypes as C
import os

import numpy as np
import pytest

from gianduia_amd import DeviceScene, HostScene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["cbox", "sphere10k"]


def _load(case):
    z = np.load(os.path.join(GOLDEN, f"rays_{case}.npz"))
    rays = {k[4:]: z[k] for k in z.files if k.startswith("ray_")}
    hs = HostScene.generate(str(z["generator"]), str(z["params"]))
    return z, rays, hs


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _check(z, hit, occ):
    assert np.array_equal(hit["inst"], z["inst"]), "instance ids differ from the golden"
    h = z["inst"] >= 0
    assert np.array_equal(hit["prim"][h], z["prim"][h])
    assert np.array_equal(_bits(hit["t"])[h], z["t_bits"][h])
    assert np.array_equal(_bits(hit["u"])[h], z["u_bits"][h])
    assert np.array_equal(_bits(hit["v"])[h], z["v_bits"][h])
    assert np.array_equal(np.asarray(occ, np.int8), z["occluded"])


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_golden_rays(case):
    from oracle import RefScene
    z, rays, hs = _load(case)
    ref = RefScene(hs.desc())
    short = dict(rays)
    short["tmax"] = z["any_tmax"]
    _check(z, ref.trace(rays), ref.trace(short, any_hit=True))
    assert (z["inst"] >= 0).sum() > len(z["inst"]) // 2


def _xf(m16, v, w):
    """glm's mat4 * vec4 in float32: (c0 x + c1 y) + (c2 z + c3 w), c_i the
    columns of the column-major matrix (transform.h:43-61)."""
    c = np.asarray(m16, np.float32).reshape(4, 4)[:, :3]
    f = np.float32
    return ((c[0] * f(v[0]) + c[1] * f(v[1])) + (c[2] * f(v[2]) + c[3] * f(w))).astype(np.float32)


def test_cbox_golden_equals_brute_force():
    """Every primitive's object-space ray (world-to-object matrix of the
    primitive, Primitive::rayIntersect) against all of its triangles, the
    closest over all primitives: the goldens' instance, primitive and t."""
    from test_oracle_pins import brute_force
    z, rays, hs = _load("cbox")
    d = hs.desc().contents
    n = len(rays["ox"])
    best_t = np.full(n, np.inf, np.float32)
    best = np.full((n, 2), -1, np.int64)
    t_of = np.full((n, d.nprims), np.inf, np.float32)  # each primitive's closest t
    for i in range(d.nprims):
        p = d.prims[i]
        mesh = d.meshes[p.mesh]
        tris = np.ctypeslib.as_array(mesh.positions, shape=(mesh.ntris * 9,)).reshape(-1, 9)
        minv = np.ctypeslib.as_array(p.to_world_inv, shape=(16,))
        o = np.array([_xf(minv, (rays["ox"][k], rays["oy"][k], rays["oz"][k]), 1.0) for k in range(n)])
        dd = np.array([_xf(minv, (rays["dx"][k], rays["dy"][k], rays["dz"][k]), 0.0) for k in range(n)])
        local = {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": dd[:, 0], "dy": dd[:, 1], "dz": dd[:, 2],
                 "tmin": rays["tmin"], "tmax": rays["tmax"]}
        t, j = brute_force(tris, local)
        t_of[:, i] = np.where(j >= 0, t, np.inf)
        better = (j >= 0) & (t < best_t)
        best_t[better] = t[better]
        best[better] = np.stack([np.full(better.sum(), i), j[better]], 1)
    hit = best[:, 0] >= 0
    assert np.array_equal(hit, z["inst"] >= 0)
    assert np.array_equal(_bits(best_t)[hit], z["t_bits"][hit])
    # where two primitives meet the ray at exactly the same t (the boxes'
    # bottom faces lie on the floor) the reference's traversal order picks
    # one: the golden's primitive must then attain that same t
    same = hit & (best[:, 0] == z["inst"]) & (best[:, 1] == z["prim"])
    tie = hit & ~same
    rows = np.nonzero(tie)[0]
    assert np.array_equal(_bits(t_of[rows, z["inst"][rows]]), z["t_bits"][rows])
    assert tie.sum() < 0.02 * hit.sum(), "ties should be rare"


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_reproduces_golden_rays(device, case):
    z, rays, hs = _load(case)
    ds = DeviceScene(device, hs)
    g, _ = device.trace_closest(ds, rays)
    short = dict(rays)
    short["tmax"] = z["any_tmax"]
    occ, _ = device.trace_any(ds, short)
    _check(z, g, occ)
    ds.free()
