"""Host side of the non-area emitters: the OpenEXR reader behind environment
maps (Image2D::loadEXR, include/gianduia/core/image2d.h:100-134), the XML
loader's scene-level <emitter> elements (Scene::addChild, scene.cpp:33-40,
and the emitter constructors: pointlight.cpp:9-12, spotlight.cpp:9-22,
directional.cpp:8-11, environment.cpp:14-27) and the generators' lights=
option.

The reference reads EXR files with OpenEXR, which this build does not have;
the files here come from tests/imgio.write_exr, a writer of the published
layout that shares no code with the library's reader."""
import numpy as np
import pytest

from gianduia_amd import GndError, HostScene, load_exr
from gianduia_amd import _abi as A

import imgio
from oracle import RefScene


@pytest.mark.parametrize("ptype", [0, 1, 2])
@pytest.mark.parametrize("compression", [0, 1, 2, 3])
def test_exr_reader_matches_writer(tmp_path, ptype, compression):
    rng = np.random.default_rng(ptype * 10 + compression)
    h, w = 37, 29  # ZIP: 16-line blocks plus a 5-line remainder
    if ptype == 0:
        ch = {c: rng.integers(0, 100000, (h, w)).astype(np.float32) for c in "RGB"}
    else:
        ch = {c: (rng.standard_normal((h, w)) * 10).astype(np.float32) for c in "RGB"}
    for c in "RGB":
        ch[c][4:20] = 2.0  # compressible rows, so the coded blocks are really coded
    path = tmp_path / "img.exr"
    imgio.write_exr(path, ch, ptype, compression, origin=(3, -2))
    raw_bytes = h * w * 3 * (2 if ptype == 1 else 4)
    if compression:
        assert path.stat().st_size < raw_bytes  # some blocks went through the codec
    img = load_exr(str(path))
    want = np.stack([ch[c] for c in "RGB"], -1)
    if ptype == 1:
        want = want.astype(np.float16).astype(np.float32)
    assert img.shape == (h, w, 3)
    assert np.array_equal(img, want)


def test_exr_every_half_value(tmp_path):
    """All 65536 binary16 patterns (zeros, subnormals, normals, inf, NaN)
    widen exactly as numpy's float16 -> float32."""
    bits = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    half = bits.view(np.float16)
    path = tmp_path / "h.exr"
    imgio.write_exr(path, {"R": half.astype(np.float32), "G": np.zeros((256, 256), np.float32),
                           "B": np.zeros((256, 256), np.float32)}, ptype=1)
    # write_exr re-rounds to half: float32(half) -> half is exact, NaN kept as NaN;
    # rewrite the R channel with the exact bit patterns
    data = bytearray(path.read_bytes())
    row = 256 * 2
    # uncompressed, one line per block: y, size, then B, G, R rows
    import struct
    hdr_end = data.index(b"screenWindowWidth") + len("screenWindowWidth\0float\0") + 4 + 4 + 1
    offs = struct.unpack("<256Q", data[hdr_end:hdr_end + 8 * 256])
    for y, o in enumerate(offs):
        r0 = o + 8 + 2 * row
        data[r0:r0 + row] = bits[y].astype("<u2").tobytes()
    path.write_bytes(bytes(data))
    img = load_exr(str(path))
    want = half.astype(np.float32)
    got = img[..., 0]
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got.view(np.uint32)[~nan], want.view(np.uint32)[~nan])


def test_exr_missing_channels_read_zero(tmp_path):
    rng = np.random.default_rng(3)
    r = rng.random((5, 7)).astype(np.float32)
    y = rng.random((5, 7)).astype(np.float32)
    imgio.write_exr(tmp_path / "ry.exr", {"R": r, "Y": y})
    img = load_exr(str(tmp_path / "ry.exr"))
    assert np.array_equal(img[..., 0], r) and not img[..., 1:].any()
    imgio.write_exr(tmp_path / "y.exr", {"Y": y})
    assert not load_exr(str(tmp_path / "y.exr")).any()  # luminance-only: black RGB


def test_exr_rejects_unsupported(tmp_path):
    ch = {c: np.ones((4, 4), np.float32) for c in "RGB"}
    p = tmp_path / "a.exr"
    imgio.write_exr(p, ch)
    data = bytearray(p.read_bytes())
    tiled = bytearray(data)
    tiled[4:8] = (2 | 0x200).to_bytes(4, "little")
    (tmp_path / "tiled.exr").write_bytes(bytes(tiled))
    with pytest.raises(GndError, match="tiled"):
        load_exr(str(tmp_path / "tiled.exr"))
    piz = bytearray(data)
    k = piz.index(b"compression\0compression\0") + len("compression\0compression\0") + 4
    piz[k] = 4
    (tmp_path / "piz.exr").write_bytes(bytes(piz))
    with pytest.raises(GndError, match="compression 4"):
        load_exr(str(tmp_path / "piz.exr"))
    (tmp_path / "trunc.exr").write_bytes(bytes(data[:len(data) - 10]))
    with pytest.raises(GndError):
        load_exr(str(tmp_path / "trunc.exr"))
    with pytest.raises(GndError):
        load_exr(str(tmp_path / "missing.exr"))


LIGHTS_XML = """<scene>
  <camera type="perspective">
    <transform name="toWorld"><lookat origin="0, 0, 3.9" target="0, 0, 0" up="0, 1, 0"/></transform>
    <integer name="width" value="16"/><integer name="height" value="12"/>
    <filter type="box"/>
  </camera>
  <sampler type="independent"><integer name="spp" value="2"/></sampler>
  <integrator type="path"><integer name="maxDepth" value="4"/></integrator>
  <emitter type="point"><point name="position" value="0, 1, 0"/><color name="power" value="5, 6, 7"/></emitter>
  <primitive>
    <shape type="sphere"><float name="radius" value="0.5"/></shape>
    <material type="matte"><color name="albedo" value="0.5 0.5 0.5"/></material>
  </primitive>
  <primitive>
    <transform name="toWorld"><translate x="0" y="2" z="0"/></transform>
    <shape type="sphere"><float name="radius" value="0.1"/></shape>
    <material type="matte"><color name="albedo" value="0 0 0"/></material>
    <emitter type="area"><color name="radiance" value="3, 3, 3"/></emitter>
  </primitive>
  <emitter type="spot"><point name="position" value="0, 2, 1"/><vector name="direction" value="0, -1, -1"/>
    <float name="spot_size" value="60"/></emitter>
  <emitter type="directional"><vector name="direction" value="1, -2, 0"/></emitter>
  <emitter type="environment"><string name="filename" value="sky.exr"/><float name="strength" value="0.5"/></emitter>
</scene>
"""


def test_xml_scene_emitters(tmp_path):
    rng = np.random.default_rng(5)
    sky = {c: rng.random((8, 16)).astype(np.float32) for c in "RGB"}
    imgio.write_exr(tmp_path / "sky.exr", sky, ptype=1, compression=3)
    (tmp_path / "scene.xml").write_text(LIGHTS_XML)
    hs = HostScene.load_xml(str(tmp_path / "scene.xml"))
    d = hs.desc().contents
    assert d.nlights == 4 and d.nprims == 2
    L = [d.lights[i] for i in range(4)]
    # the emitter list is document order: point, area sphere, spot, directional, environment
    assert [l.emitter_index for l in L] == [0, 2, 3, 4]
    assert [l.type for l in L] == [A.GND_LIGHT_POINT, A.GND_LIGHT_SPOT, A.GND_LIGHT_DIRECTIONAL,
                                   A.GND_LIGHT_ENVIRONMENT]
    assert list(L[0].position) == [0, 1, 0] and list(L[0].power) == [5, 6, 7]
    assert list(L[1].power) == [10, 10, 10] and L[1].spot_size == 60 and abs(L[1].blend - 0.15) < 1e-7
    assert list(L[1].direction) == [0, -1, -1]
    assert list(L[2].direction) == [1, -2, 0] and list(L[2].irradiance) == [1, 1, 1]
    env = L[3]
    assert (env.env_width, env.env_height, env.strength) == (16, 8, 0.5)
    px = np.ctypeslib.as_array(env.env_rgb, shape=(8 * 16 * 3,)).reshape(8, 16, 3)
    want = np.stack([sky[c] for c in "RGB"], -1).astype(np.float16).astype(np.float32)
    assert np.array_equal(px, want)
    # the oracle renders the loaded scene: all four lights reach the film
    img, _ = RefScene(hs.desc()).render(threads=4, spp=2, width=16, height=12)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0


@pytest.mark.parametrize("emitter,err", [
    ('<emitter type="environment"><string name="filename" value="sky.png"/></emitter>', "extension"),
    ('<emitter type="environment"><string name="filename" value="none.exr"/></emitter>', "path"),
    ('<emitter type="environment"><string name="filename" value="sky.exr"/></emitter>' * 2, "infinite"),
    ('<emitter type="point"/>', "position"),
    ('<emitter type="area"/>', "primitive"),
    ('<emitter type="laser"/>', "emitter type"),
])
def test_xml_emitter_errors(tmp_path, emitter, err):
    imgio.write_exr(tmp_path / "sky.exr", {c: np.ones((2, 4), np.float32) for c in "RGB"})
    (tmp_path / "sky.png").write_bytes(b"x")
    xml = LIGHTS_XML.replace(LIGHTS_XML[LIGHTS_XML.index('  <emitter type="environment">'):
                                        LIGHTS_XML.index("</scene>")], emitter + "\n")
    (tmp_path / "scene.xml").write_text(xml)
    with pytest.raises(GndError, match=err):
        HostScene.load_xml(str(tmp_path / "scene.xml"))


def test_generator_lights():
    hs = HostScene.generate("cbox", "width=8;height=8;lights=point,spot,sun,env;envres=4;envstrength=2")
    d = hs.desc().contents
    assert d.nlights == 4
    assert [d.lights[i].emitter_index for i in range(4)] == [1, 2, 3, 4]  # after the ceiling light
    assert [d.lights[i].type for i in range(4)] == [A.GND_LIGHT_POINT, A.GND_LIGHT_SPOT,
                                                    A.GND_LIGHT_DIRECTIONAL, A.GND_LIGHT_ENVIRONMENT]
    e = d.lights[3]
    assert (e.env_width, e.env_height, e.strength) == (8, 4, 2.0)
    with pytest.raises(GndError, match="light kind"):
        HostScene.generate("cbox", "lights=torch")
