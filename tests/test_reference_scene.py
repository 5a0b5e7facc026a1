"""The reference's own preview scene (docs/preview/cbox.xml with plane.obj
and bunny.obj, committed unchanged under tests/golden/reference_cbox) goes
through the drop-in: the XML loader reads it as the reference's parser would,
and the GPU render of a crop of it (the scene's own thin-lens camera at a
lower resolution) equals the oracle's path by path at 1 spp and
statistically at 16 spp with independent random streams.  The scene holds a
smooth glass bunny with a homogeneous medium inside, a rough glass, a rough
conductor and a moving plastic bunny, two analytic spheres and an area
light, under the volumetric path integrator."""
import os

import numpy as np
import pytest

from gianduia_amd import DeviceScene, HostScene
from gianduia_amd import _abi as A
from oracle import RefScene

SCENE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_cbox", "cbox.xml")


def _scene(size):
    hs = HostScene.load_xml(SCENE)
    desc = hs.desc()  # the descriptor is rebuilt on every call: change this one, hand it to the oracle
    desc.contents.camera.width = desc.contents.camera.height = size
    return hs, desc


def test_reference_scene_loads_as_written():
    hs = HostScene.load_xml(SCENE)
    d = hs.desc().contents
    assert d.integrator == A.GND_INTEGRATOR_VOLPATH and d.spp == 2056
    assert d.camera.type == A.GND_CAMERA_THINLENS and (d.camera.width, d.camera.height) == (800, 800)
    assert d.camera.filter == A.GND_FILTER_BOX and abs(d.camera.fov - 40.0) < 1e-6
    assert d.nprims == 12 and d.nmeshes == 2 and d.nmedia == 1
    assert sorted(d.meshes[i].ntris for i in range(d.nmeshes)) == [2, 4968]
    shapes = [d.prims[i].shape_type for i in range(d.nprims)]
    assert shapes.count(A.GND_SHAPE_SPHERE) == 2
    assert sum(d.prims[i].has_emitter for i in range(d.nprims)) == 1
    types = [d.materials[d.prims[i].material].type for i in range(d.nprims)]
    assert types.count(A.GND_MAT_GLASS) == 2 and types.count(A.GND_MAT_PLASTIC) == 1
    assert types.count(A.GND_MAT_CONDUCTOR) == 3  # one bunny and the two default-conductor spheres
    med = d.media[0]
    assert med.type == A.GND_MEDIUM_HOMOGENEOUS
    assert np.allclose(list(med.sigma_a), [0.5, 0.1, 0.6]) and np.allclose(list(med.sigma_s), [2.0, 2.5, 1.5])
    glass = [d.materials[d.prims[i].material] for i in range(d.nprims)
             if d.materials[d.prims[i].material].type == A.GND_MAT_GLASS]
    assert sorted(m.medium_inside for m in glass) == [-1, 0] and all(abs(m.eta - 1.6) < 1e-6 for m in glass)
    assert sum(d.prims[i].has_motion for i in range(d.nprims)) == 1


@pytest.mark.gpu
def test_reference_scene_matches_oracle(device):
    from test_gpu_render import GPU_STREAM, _check_stats, _replay_fraction
    size = 96
    hs, desc = _scene(size)
    ref = RefScene(desc)
    ds = DeviceScene(device, hs)
    job = hs.default_job(spp=1)
    job.camera.width = job.camera.height = size
    out, st = device.render(ds, [job])
    r, _ = ref.render(threads=16, spp=1)
    frac = _replay_fraction(out[0], r)
    print(f"reference cbox.xml {size}^2: {frac:.4f} of paths reproduce the oracle", flush=True)
    assert st["nan_samples"] == 0 and frac >= 0.99
    job16 = hs.default_job(spp=16, seed_offset=GPU_STREAM)
    job16.camera.width = job16.camera.height = size
    g, gsq = device.render(ds, [job16], want_sq=True)[0][0]
    r16, rsq, _ = ref.render(threads=16, spp=16, squares=True)
    _check_stats(g, gsq, r16, rsq, 16, "reference cbox.xml")
