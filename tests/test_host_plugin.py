"""CPU tests of the host plugin layer: integrator registry, the host
IndependentSampler (equal to the oracle's restatement, itself pinned by the
published PCG32 / SplitMix64 outputs) and the PNG / PFM / EXR film writers."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

from gianduia_amd import _abi as A, integrator_names, sampler_floats
import oracle as O
from imgio import read_exr
from oracle import fp


def test_registry_names():
    names = integrator_names()
    for n in ("path", "volpath", "path_hip", "volpath_hip", "direct", "direct_hip", "ao", "ao_hip"):
        assert n in names


@pytest.mark.parametrize("idx", [0, 1, 7, 512 * 512 * 63 + 5, 2 ** 41 + 3])
def test_host_sampler_equals_oracle(idx):
    ref = np.zeros(16, np.float32)
    O.lib().gnd_ref_sampler_floats(idx, 16, fp(ref))
    assert np.array_equal(sampler_floats(idx, 16).view(np.uint32), ref.view(np.uint32))


def _image(w=13, h=7):
    rng = np.random.default_rng(1)
    img = rng.random((h, w, 3)).astype(np.float32)
    img[0, 0] = (2.0, -1.0, 0.5)  # out of range values are clamped in PNG
    return img


def test_pfm_roundtrip(tmp_path):
    img = _image()
    p = tmp_path / "f.pfm"
    rc = A.load_library().gnd_host_image_save(str(p).encode(), img.ctypes.data_as(A.f32p), 13, 7)
    assert rc == 0
    data = p.read_bytes()
    head, rest = data.split(b"\n", 3)[:3], data.split(b"\n", 3)[3]
    assert head[0] == b"PF" and head[1] == b"13 7" and float(head[2]) < 0
    back = np.frombuffer(rest, "<f4").reshape(7, 13, 3)[::-1]   # rows stored bottom to top
    assert np.array_equal(back, img)


def _srgb8(v):
    v = np.clip(v, 0, 1)
    s = np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(v, 1 / 2.4) - 0.055)
    return np.rint(s * 255).astype(np.uint8)


def test_png_decodes(tmp_path):
    img = _image()
    p = tmp_path / "f.png"
    assert A.load_library().gnd_host_image_save(str(p).encode(), img.ctypes.data_as(A.f32p), 13, 7) == 0
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) == crc, typ
        chunks.setdefault(typ, b"")
        chunks[typ] += body
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert (w, h, depth, ctype) == (13, 7, 8, 2)
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(7, 1 + 13 * 3)
    assert np.all(rows[:, 0] == 0)
    pix = rows[:, 1:].reshape(7, 13, 3)
    assert np.abs(pix.astype(int) - _srgb8(img.astype(np.float64)).astype(int)).max() <= 1


def test_exr_roundtrip(tmp_path):
    """Film::saveEXR (R, G, B float channels, Image2D::saveEXR): every pixel
    reads back exactly through an independent reader of the file layout."""
    img = _image()
    img[0, 0] = (np.inf, -1e-40, 1e-40)   # non-finite and denormal values survive
    p = tmp_path / "f.exr"
    assert A.load_library().gnd_host_image_save(str(p).encode(), img.ctypes.data_as(A.f32p), 13, 7) == 0
    attrs, chans = read_exr(p.read_bytes())
    assert sorted(chans) == ["B", "G", "R"]
    back = np.stack([chans["R"], chans["G"], chans["B"]], axis=-1)
    assert np.array_equal(back.view(np.uint32), img.view(np.uint32))

