"""Writes the ray-batch golden fixtures: seeded rays through two generated
scenes with the oracle's closest hits (instance, primitive, t, u, v as raw
float32 bits) and any-hit occlusion.

    python tests/golden/make_ray_goldens.py

Scenes: the Cornell box (36 triangles, `cbox`) and the scaled scene's
10,000-triangle displaced sphere (`scaled` with level=2;instances=1: one
instance of that sphere beside a 320-triangle icosphere, a floor and a
light).  The rays are stored with the hits, so the fixtures do not depend on
a random generator's stream."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

CASES = {
    "cbox": ("cbox", "", 2048, (0.0, 0.0, 0.0), 3.0),
    "sphere10k": ("scaled", "level=2;instances=1", 2048, (0.0, 0.0, 0.0), 3.5),
}


def rays_for(rng, n, center, radius):
    """Half from outside towards the scene centre region, half from inside
    in random directions; a few axis-aligned and short ones."""
    c = np.asarray(center, np.float32)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * radius
    o[n // 2:] = rng.uniform(-0.9, 0.9, size=(n - n // 2, 3))
    o = (o + c).astype(np.float32)
    tgt = (c + rng.uniform(-1.0, 1.0, size=(n, 3))).astype(np.float32)
    d = tgt - o
    d[n // 2:] = rng.normal(size=(n - n // 2, 3))
    d[:6] = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tmax = np.full(n, np.inf, np.float32)
    tmax[6:70] = rng.uniform(0.05, 2.0, 64).astype(np.float32)
    return {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(), "dx": d[:, 0].copy(),
            "dy": d[:, 1].copy(), "dz": d[:, 2].copy(), "tmin": np.full(n, 1e-4, np.float32), "tmax": tmax}


def main():
    from gianduia_amd import HostScene
    from oracle import RefScene
    for k, (gen, params, n, center, radius) in CASES.items():
        hs = HostScene.generate(gen, params)
        ref = RefScene(hs.desc())
        rays = rays_for(np.random.default_rng(2024 + len(k)), n, center, radius)
        hit = ref.trace(rays)
        short = dict(rays)
        short["tmax"] = np.where(np.isinf(rays["tmax"]), np.float32(1.5), rays["tmax"]).astype(np.float32)
        occ = ref.trace(short, any_hit=True)
        path = os.path.join(HERE, f"rays_{k}.npz")
        np.savez_compressed(path, generator=np.array(gen), params=np.array(params), **{f"ray_{a}": v for a, v in rays.items()},
                            any_tmax=short["tmax"], inst=hit["inst"].astype(np.int32), prim=hit["prim"].astype(np.int32),
                            t_bits=hit["t"].astype(np.float32).view(np.uint32), u_bits=hit["u"].astype(np.float32).view(np.uint32),
                            v_bits=hit["v"].astype(np.float32).view(np.uint32), occluded=occ.astype(np.int8))
        print(f"{path}: {n} rays, {(hit['inst'] >= 0).sum()} hits, {occ.sum()} occluded")


if __name__ == "__main__":
    main()
