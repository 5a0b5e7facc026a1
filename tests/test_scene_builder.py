"""The object-graph entry points (gnd_host_builder_*): a binding hands over
the objects the reference's parser creates (class tag, type name, name,
PropertyList, then the parent -> child links in document order) instead of
a file name.  Here the reference parser's walk (src/core/parser.cpp:41-235:
properties of an element first, the object created from them, then its
children in order, <ref id> giving the same object to a second parent) is
replayed over scene files, and the scene built from the objects must equal
the one gnd_host_scene_load_xml reads from the same file, field for field and
array for array.

The PropertyList holds transforms already composed (Transform, matrix and
inverse); the replay takes each composed transform from the XML loader's own
result, so what is compared is the object graph -> scene step alone."""
import ctypes as C
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gianduia_amd import GndError, HostScene, SceneObjects
from gianduia_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CBOX = os.path.join(ROOT, "tests", "golden", "reference_cbox", "cbox.xml")
IDENTITY = tuple(float(i == j) for j in range(4) for i in range(4))


def _vec(s):
    return [float(x) for x in s.replace(",", " ").split()]


def _replay(path, desc):
    """Objects of the file in the order the reference's parser creates them."""
    d = desc.contents
    so = SceneObjects(os.path.dirname(path))
    ids = {}
    prim_index = [-1]

    def xf(el, name, prim):
        if el.tag == "camera":
            c = d.camera
            if name == "toWorld":
                return tuple(c.to_world), tuple(c.to_world_inv)
            return tuple(c.to_world_end), IDENTITY
        p = d.prims[prim]
        if name == "toWorld":
            return tuple(p.to_world), tuple(p.to_world_inv)
        return tuple(p.to_world_end), tuple(p.to_world_end_inv)

    def obj(el):
        prim = -1
        if el.tag == "primitive":
            prim_index[0] += 1
            prim = prim_index[0]
        props, kids = [], []
        for c in el:
            t, nm, val = c.tag, c.get("name", ""), c.get("value")
            if t == "boolean":
                props.append((nm, "boolean", val in ("true", "1")))
            elif t == "integer":
                props.append((nm, "integer", int(val)))
            elif t == "float":
                props.append((nm, "float", float(val)))
            elif t == "string":
                props.append((nm, "string", val))
            elif t in ("point", "vector", "vector3", "color"):
                if val is not None:
                    v = _vec(val)
                else:
                    v = [float(c.get(k, 0)) for k in ("rgb" if t == "color" else "xyz")]
                props.append((nm, {"vector3": "vector"}.get(t, t), v))
            elif t == "vector2":
                props.append((nm, "vector2", _vec(val) if val is not None else [float(c.get(k, 0)) for k in "xy"]))
            elif t == "transform":
                props.append((nm, "transform", xf(el, nm, prim)))
            else:
                kids.append(c)
        oid = so.add(el.tag, el.get("type") or el.tag, el.get("name"), props)
        if el.get("id"):
            ids[el.get("id")] = oid
        for k in kids:
            so.add_child(oid, ids[k.get("id")] if k.tag == "ref" else obj(k))
        return oid

    root = obj(ET.parse(path).getroot())
    hs = so.finish(root)
    so.close()
    return hs


def _bytes(ptr, n):
    return C.string_at(ptr, n) if n and ptr else b""


def _struct_bytes(s, skip=()):
    """The bytes of a ctypes struct's fields (padding and the fields in
    `skip` left out)."""
    out = b""
    for name, _ in s._fields_:
        if name not in skip:
            f = getattr(type(s), name)
            out += bytes(s)[f.offset:f.offset + f.size]
    return out


def _assert_same_desc(a, b):
    a, b = a.contents, b.contents
    for f in ("nmeshes", "nprims", "nmaterials", "nmedia", "nlights", "integrator", "max_depth", "spp"):
        assert getattr(a, f) == getattr(b, f), f
    assert np.float32(a.ao_length) == np.float32(b.ao_length)
    assert bytes(a.camera) == bytes(b.camera), "camera"
    for i in range(a.nmeshes):
        ma, mb = a.meshes[i], b.meshes[i]
        assert ma.ntris == mb.ntris, f"mesh {i}"
        n = ma.ntris
        for f, k in (("positions", 36), ("normals", 36), ("uvs", 24), ("vertex_ids", 12)):
            pa, pb = getattr(ma, f), getattr(mb, f)
            assert bool(pa) == bool(pb), f"mesh {i} {f}"
            assert _bytes(pa, n * k) == _bytes(pb, n * k), f"mesh {i} {f}"
    for f, n in (("prims", a.nprims), ("materials", a.nmaterials)):
        size = C.sizeof(getattr(a, f)._type_)
        assert _bytes(getattr(a, f), n * size) == _bytes(getattr(b, f), n * size), f
    for i in range(a.nmedia):
        ma, mb = a.media[i], b.media[i]
        assert _struct_bytes(ma, ("density",)) == _struct_bytes(mb, ("density",)), f"medium {i}"
        nv = ma.nx * ma.ny * ma.nz * 4 if ma.density else 0
        assert _bytes(ma.density, nv) == _bytes(mb.density, nv), f"medium {i} density"
    for i in range(a.nlights):
        la, lb = a.lights[i], b.lights[i]
        assert _struct_bytes(la, ("env_rgb",)) == _struct_bytes(lb, ("env_rgb",)), f"light {i}"
        nv = la.env_width * la.env_height * 12 if la.env_rgb else 0
        assert _bytes(la.env_rgb, nv) == _bytes(lb.env_rgb, nv), f"light {i} map"


def test_reference_scene_objects_equal_its_file():
    """The reference's preview scene: thin lens, shared meshes (<ref>), glass
    with a medium inside, conductor, a moving plastic primitive, spheres and
    an area light."""
    xml = HostScene.load_xml(REF_CBOX)
    obj = _replay(REF_CBOX, xml.desc())
    _assert_same_desc(xml.desc(), obj.desc())
    assert xml.desc().contents.nprims == 12
    a, b = xml.build().contents, obj.build().contents
    assert a.n_blas_nodes == b.n_blas_nodes and a.n_packs == b.n_packs
    assert _bytes(a.blas_nodes, a.n_blas_nodes * 144) == _bytes(b.blas_nodes, b.n_blas_nodes * 144)
    assert _bytes(a.tlas_nodes, a.n_tlas_nodes * 144) == _bytes(b.tlas_nodes, b.n_tlas_nodes * 144)


ALL_KINDS = """<scene>
  <camera type="thinlens">
    <transform name="toWorld"><lookat origin="0, 1, 4" target="0, 0, 0" up="0, 1, 0"/></transform>
    <transform name="toWorldEnd"><lookat origin="0.2, 1, 4" target="0, 0, 0" up="0, 1, 0"/></transform>
    <float name="fov" value="50"/><float name="lens_radius" value="0.03"/><float name="focal_distance" value="4"/>
    <float name="k1" value="0.02"/><boolean name="rolling_shutter" value="true"/>
    <integer name="width" value="40"/><integer name="height" value="30"/>
    <filter type="gaussian"><float name="alpha" value="1.5"/><vector2 name="radius" value="1.5, 1.5"/></filter>
  </camera>
  <sampler type="independent"><integer name="spp" value="3"/></sampler>
  <integrator type="ao"><integer name="maxDepth" value="5"/><float name="length" value="2.5"/></integrator>
  <emitter type="point"><point name="position" value="0.5, 1.5, 0.5"/><color name="power" value="8, 7, 6"/></emitter>
  <primitive>
    <transform name="toWorld"><scale x="6" y="1" z="6"/><translate x="0" y="-0.5" z="0"/></transform>
    <shape type="mesh"><string name="filename" value="plane.obj"/></shape>
    <material type="matte"><color name="albedo" value="0.6 0.6 0.6"/></material>
  </primitive>
  <primitive>
    <shape type="sphere" id="ball"><float name="radius" value="0.5"/></shape>
    <material type="glass"><float name="roughness" value="0.1"/>
      <medium type="homogeneous" name="inside"><color name="sigma_a" value="0.5, 0.2, 0.1"/>
        <color name="sigma_s" value="1, 1, 1"/><float name="g" value="0.3"/></medium>
    </material>
  </primitive>
  <primitive>
    <transform name="toWorld"><translate x="-1.2" y="0" z="0.3"/></transform>
    <transform name="toWorldEnd"><translate x="-1.0" y="0.1" z="0.3"/></transform>
    <ref id="ball"/>
    <material type="plastic"><float name="roughness" value="0.3"/><color name="Kd" value="0.5, 0.1, 0.1"/></material>
    <emitter type="area"><color name="radiance" value="2, 1, 0.5"/></emitter>
  </primitive>
  <emitter type="spot"><point name="position" value="-1, 2, 1"/><vector name="direction" value="0.5, -1, -0.5"/>
    <float name="spot_size" value="40"/></emitter>
  <primitive>
    <transform name="toWorld"><rotate axis="0, 1, 0" angle="30"/><translate x="1" y="0" z="0"/></transform>
    <shape type="mesh"><string name="filename" value="plane.obj"/></shape>
    <material type="conductor"><float name="roughness" value="0.2"/><color name="eta" value="0.2, 0.9, 1.1"/></material>
  </primitive>
  <emitter type="directional"><vector name="direction" value="-0.3, -1, -0.2"/><color name="irradiance" value="0.8, 0.8, 0.7"/></emitter>
  <emitter type="environment"><string name="filename" value="sky.exr"/><float name="strength" value="0.7"/></emitter>
</scene>
"""


@pytest.fixture
def all_kinds_file(tmp_path):
    import imgio
    rng = np.random.default_rng(5)
    imgio.write_exr(tmp_path / "sky.exr", {c: rng.random((8, 16)).astype(np.float32) for c in "RGB"}, ptype=2)
    (tmp_path / "plane.obj").write_text("v -0.5 0 -0.5\nv 0.5 0 -0.5\nv 0.5 0 0.5\nv -0.5 0 0.5\n"
                                        "vn 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                        "f 1/1/1 3/3/1 2/2/1\nf 1/1/1 4/4/1 3/3/1\n")
    (tmp_path / "scene.xml").write_text(ALL_KINDS)
    return str(tmp_path / "scene.xml")


def test_every_object_kind_equals_its_file(all_kinds_file):
    """Every emitter kind in document order between primitives, a Gaussian
    filter, a moving thin lens with rolling shutter, a moving primitive, a
    sphere shared through <ref>, a medium inside glass, the AO integrator."""
    xml = HostScene.load_xml(all_kinds_file)
    d = xml.desc().contents
    assert (d.nlights, d.nprims, d.nmedia) == (4, 4, 1)
    assert d.camera.has_motion == 1 and d.camera.rolling_shutter == 1 and d.prims[2].has_motion == 1
    obj = _replay(all_kinds_file, xml.desc())
    _assert_same_desc(xml.desc(), obj.desc())


def test_builder_errors_match_the_loader(tmp_path):
    so = SceneObjects(str(tmp_path))
    with pytest.raises(GndError, match="unknown object class"):
        so.add("widget", "x")
    root = so.add("scene")
    with pytest.raises(GndError, match="invalid object ids"):
        so.add_child(root, 7)
    with pytest.raises(GndError, match="no camera specified"):
        so.finish(root)
    cam = so.add("camera", "perspective", None, [("width", "integer", 8), ("height", "integer", 8)])
    so.add_child(root, cam)
    prim = so.add("primitive")
    so.add_child(root, prim)
    so.add_child(root, so.add("sampler", "independent"))
    so.add_child(root, so.add("integrator", "path"))
    with pytest.raises(GndError, match="defined without a Shape"):
        so.finish(root)
    shape = so.add("shape", "mesh", None, [("filename", "string", "missing.obj")])
    so.add_child(prim, shape)
    with pytest.raises(GndError):
        so.finish(root)
    so.close()
