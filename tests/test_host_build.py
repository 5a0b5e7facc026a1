"""CPU tests: the host shell's BVH build / packing / CDF equal the oracle's
restatement of the reference byte for byte; BVH invariants; the library
exports every symbol the public headers declare; XML loading."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gianduia_amd import HostScene, _abi as A
from oracle import RefScene
from scenes import SceneBuilder, random_soup, soup_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _arrays(hs):
    return hs.build().contents


def _slice(ptr, base, n, size):
    if n == 0:
        return np.zeros((0, size), np.uint8)
    addr = C.addressof(ptr.contents) + base * size
    return np.frombuffer(C.string_at(addr, n * size), np.uint8).reshape(n, size)


def _check_same(hs):
    a = _arrays(hs)
    ref = RefScene(hs.desc())
    tl = _slice(a.tlas_nodes, 0, a.n_tlas_nodes, 144)
    assert np.array_equal(tl[:, :132], ref.tlas_nodes()[:, :132]), "TLAS nodes differ"
    tp = np.frombuffer(C.string_at(a.tlas_prims, a.n_tlas_prims * 4), np.int32)
    assert np.array_equal(tp, ref.tlas_prims()), "TLAS primitive order differs"
    for m in range(a.n_meshes):
        mi = a.meshes[m]
        nodes = _slice(a.blas_nodes, mi.node_base, mi.n_nodes, 144)
        packs = _slice(a.packs, mi.pack_base, mi.n_packs, 160)
        assert np.array_equal(nodes[:, :132], ref.mesh_nodes(m)[:, :132]), f"mesh {m}: BVH4 nodes differ"
        assert np.array_equal(packs, ref.mesh_packs(m)), f"mesh {m}: triangle packs differ"
        cdf = np.frombuffer(C.string_at(C.addressof(a.cdf.contents) + 4 * mi.cdf_base, 4 * (mi.n_tris + 1)),
                            np.float32)
        rc, ta = ref.mesh_cdf(m)
        assert np.array_equal(cdf.view(np.uint32), rc.view(np.uint32)), f"mesh {m}: area CDF differs"
        assert np.float32(mi.total_area) == np.float32(ta)
    return a


def test_cbox_build_matches_oracle():
    _check_same(HostScene.generate("cbox", ""))


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 5), (2, 100), (3, 5000)])
def test_soup_build_matches_oracle(seed, n):
    b = soup_scene(seed=seed, n=n, instances=3, sphere=True)
    _check_same(HostScene.from_desc(b.desc()))


def test_degenerate_and_duplicate_triangles():
    rng = np.random.default_rng(5)
    t = random_soup(rng, 500)
    t[:50] = t[50:100]                     # exact duplicates (equal centroids)
    t[100:120, 3:6] = t[100:120, 0:3]      # degenerate (zero area)
    t[120:140] = np.tile(t[120], (20, 1))  # a stack of identical triangles
    b = SceneBuilder()
    mt = b.matte()
    b.prim(b.mesh(t), material=mt, radiance=(1, 1, 1))
    _check_same(HostScene.from_desc(b.desc()))


def test_procedural_100k_build_matches_oracle():
    # exercises the threaded builder (subtrees above its parallel threshold)
    hs = HostScene.generate("procedural", "tris=600000;seed=3")  # 327,680-triangle mesh
    _check_same(hs)


def _all_bytes(a):
    """Every array of a built scene as bytes (nodes, primitive order, mesh
    table, packs, shading data, CDF)."""
    def raw(ptr, n, size):
        return C.string_at(ptr, n * size) if n else b""
    return [raw(a.tlas_nodes, a.n_tlas_nodes, 144), raw(a.tlas_prims, a.n_tlas_prims, 4),
            raw(a.meshes, a.n_meshes, C.sizeof(A.gnd_mesh_info)), raw(a.blas_nodes, a.n_blas_nodes, 144),
            raw(a.packs, a.n_packs, 160), raw(a.shading, a.n_tris * 24, 4), raw(a.cdf, a.n_cdf, 4)]


def test_threaded_build_equals_one_thread():
    """The task-parallel build (subtree tasks, parallel packing, per-vertex
    normal gathers) gives the same bytes as one thread, shading data included."""
    params = "level=6;instances=16"
    h1, h8 = HostScene.generate("scaled", params), HostScene.generate("scaled", params)  # alive while read
    one = _all_bytes(h1.build(threads=1).contents)
    many = _all_bytes(h8.build(threads=8).contents)
    for k, (x, y) in enumerate(zip(one, many)):
        assert x == y, f"array {k} differs between 1 and 8 threads"


def test_smooth_normals_sum_faces_in_order():
    """Mesh::activate's smooth normals: per vertex the area-weighted face
    normals added in face order, then normalised.  The builder gathers each
    vertex's faces from a face-ordered list; this checks it against the
    sequential scatter (np.add.at adds in index order) in float32."""
    hs = HostScene.generate("scaled", "level=4;instances=2")
    a = hs.build(threads=8).contents
    m = hs.desc().contents.meshes[0]
    nt = m.ntris
    assert m.vertex_ids and not m.normals
    P = np.ctypeslib.as_array(m.positions, shape=(nt * 9,)).reshape(nt, 3, 3).astype(np.float32)
    vid = np.ctypeslib.as_array(m.vertex_ids, shape=(nt * 3,)).astype(np.int64)
    e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    w = np.stack([e1[:, 1] * e2[:, 2] - e2[:, 1] * e1[:, 2], e1[:, 2] * e2[:, 0] - e2[:, 2] * e1[:, 0],
                  e1[:, 0] * e2[:, 1] - e2[:, 0] * e1[:, 1]], -1).astype(np.float32)
    acc = np.zeros((vid.max() + 1, 3), np.float32)
    np.add.at(acc, vid, np.repeat(w, 3, axis=0))
    d = (acc[:, 0] * acc[:, 0] + acc[:, 1] * acc[:, 1]) + acc[:, 2] * acc[:, 2]
    ln = np.sqrt(d.astype(np.float32)).astype(np.float32)
    nrm = np.where(ln[:, None] > 0, acc / np.where(ln > 0, ln, 1)[:, None], acc).astype(np.float32)
    base = a.meshes[0].tri_base
    sh = np.ctypeslib.as_array(a.shading, shape=(a.n_tris * 24,)).reshape(-1, 24)[base:base + nt]
    got = sh[:, 9:18].reshape(nt * 3, 3)
    assert np.array_equal(got.view(np.uint32), nrm[vid].view(np.uint32))


def test_bvh_invariants():
    hs = HostScene.from_desc(soup_scene(seed=9, n=3000).desc())
    a = _arrays(hs)
    mi = a.meshes[0]
    nodes = (A.gnd_bvh4_node * mi.n_nodes).from_address(C.addressof(a.blas_nodes.contents) + 144 * mi.node_base)
    packs = (A.gnd_tri_pack4 * mi.n_packs).from_address(C.addressof(a.packs.contents) + 160 * mi.pack_base)
    seen = np.zeros(mi.n_tris, np.int32)
    for p in packs:
        for l in range(4):
            if p.prim[l] != 0xFFFFFFFF:
                seen[p.prim[l]] += 1
    assert np.all(seen == 1), "every triangle sits in exactly one pack lane"

    def walk(i, lo, hi):
        n = nodes[i]
        for c in range(4):
            if n.child_type[c] == 0:
                continue
            clo = np.array([n.min_x[c], n.min_y[c], n.min_z[c]])
            chi = np.array([n.max_x[c], n.max_y[c], n.max_z[c]])
            assert np.all(clo >= lo) and np.all(chi <= hi), "child box outside its parent"
            if n.child_type[c] == 1:
                walk(n.offset[c], clo, chi)
            else:
                for k in range(n.count[c]):
                    pk = packs[n.offset[c] + k]
                    for l in range(4):
                        if pk.prim[l] == 0xFFFFFFFF:
                            continue
                        xs = [pk.v0x[l], pk.v1x[l], pk.v2x[l]]
                        ys = [pk.v0y[l], pk.v1y[l], pk.v2y[l]]
                        zs = [pk.v0z[l], pk.v1z[l], pk.v2z[l]]
                        assert min(xs) >= clo[0] and max(xs) <= chi[0]
                        assert min(ys) >= clo[1] and max(ys) <= chi[1]
                        assert min(zs) >= clo[2] and max(zs) <= chi[2]

    walk(0, np.full(3, -np.inf), np.full(3, np.inf))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gnd_(?:hip|host)_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = A.load_library()
    names = _declared("gnd_hip.h") + _declared("gnd_host.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    # and the ctypes tables cover them
    table = set(A.HIP_API) | set(A.HOST_API)
    assert not [n for n in names if n not in table]


def test_invalid_scenes_raise():
    from gianduia_amd import GndError
    b = SceneBuilder()
    mt = b.matte()
    b.prim(b.mesh(random_soup(np.random.default_rng(0), 10)), material=mt)  # no emitter
    hs = HostScene.from_desc(b.desc())
    with pytest.raises(GndError, match="emitter"):
        hs.build()
    with pytest.raises(GndError):
        HostScene.generate("no-such-scene")


CBOX_XML = """<scene>
  <camera type="perspective">
    <transform name="toWorld"><lookat origin="0, 0, 3.9" target="0, 0, 0" up="0, 1, 0"/></transform>
    <float name="fov" value="40"/>
    <integer name="width" value="64"/><integer name="height" value="48"/>
    <filter type="box"/>
  </camera>
  <sampler type="independent"><integer name="spp" value="4"/></sampler>
  <integrator type="path"><integer name="maxDepth" value="6"/></integrator>
  <primitive>
    <transform name="toWorld"><scale x="2" y="1" z="2"/><translate x="0" y="-1" z="0"/></transform>
    <shape type="mesh" id="plane"><string name="filename" value="plane.obj"/></shape>
    <material type="matte"><color name="albedo" value="0.7 0.7 0.7"/></material>
  </primitive>
  <primitive>
    <transform name="toWorld"><rotate axis="1 0 0" angle="180"/><translate x="0" y="0.99" z="0"/></transform>
    <ref id="plane"/>
    <material type="matte"><color name="albedo" value="0 0 0"/></material>
    <emitter type="area"><color name="radiance" value="10, 10, 10"/></emitter>
  </primitive>
  <primitive>
    <shape type="sphere"><float name="radius" value="0.4"/></shape>
    <material type="conductor"><float name="roughness" value="0.3"/></material>
  </primitive>
</scene>
"""

PLANE_OBJ = """v -0.5 0 -0.5
v 0.5 0 -0.5
v 0.5 0 0.5
v -0.5 0 0.5
vn 0 1 0
f 1//1 2//1 3//1
f 1//1 3//1 4//1
"""


def test_xml_loader(tmp_path):
    (tmp_path / "plane.obj").write_text(PLANE_OBJ)
    (tmp_path / "scene.xml").write_text(CBOX_XML)
    hs = HostScene.load_xml(str(tmp_path / "scene.xml"))
    d = hs.desc().contents
    assert d.nprims == 3 and d.nmeshes == 1 and d.spp == 4 and d.max_depth == 6
    assert d.camera.width == 64 and d.camera.height == 48 and d.camera.filter == A.GND_FILTER_BOX
    assert d.prims[1].has_emitter == 1 and d.prims[2].shape_type == A.GND_SHAPE_SPHERE
    # composed transform: translate * scale (parser.cpp:162, t = step * t)
    m = np.array(d.prims[0].to_world[:]).reshape(4, 4).T
    assert np.allclose(m @ np.array([0.5, 0, 0.5, 1]), [1.0, -1.0, 1.0, 1.0])
    _check_same(hs)
    bad = tmp_path / "bad.xml"
    bad.write_text("<scene><camera type='perspective'></scene>")
    from gianduia_amd import GndError
    with pytest.raises(GndError):
        HostScene.load_xml(str(bad))


def test_xml_camera_motion_and_rolling_shutter(tmp_path):
    """<transform name="toWorldEnd"> on the camera (camera.h:20-25) and the
    thin lens's rolling-shutter properties (thinLens.cpp:29-31) reach the
    camera description."""
    (tmp_path / "plane.obj").write_text(PLANE_OBJ)
    xml = CBOX_XML.replace('<camera type="perspective">', '<camera type="thinlens">').replace(
        '<filter type="box"/>',
        '<filter type="box"/><transform name="toWorldEnd"><lookat origin="0.5, 0, 3.9" target="0, 0, 0" '
        'up="0, 1, 0"/></transform><boolean name="rolling_shutter" value="true"/>'
        '<float name="scan_duration" value="0.6"/>')
    (tmp_path / "scene.xml").write_text(xml)
    c = HostScene.load_xml(str(tmp_path / "scene.xml")).desc().contents.camera
    assert c.has_motion == 1 and c.rolling_shutter == 1
    assert abs(c.scan_duration - 0.6) < 1e-7 and abs(c.exposure_duration - 0.1) < 1e-7
    end = np.array(c.to_world_end[:]).reshape(4, 4).T
    assert np.allclose(end[:3, 3], (0.5, 0, 3.9), atol=1e-6)
    start = HostScene.load_xml(str(tmp_path / "scene.xml")).desc().contents.camera.to_world[:]
    assert not np.allclose(start, c.to_world_end[:])
