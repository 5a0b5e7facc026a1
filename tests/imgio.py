"""Image-file readers for the writer tests (written from the published file
layouts, independent of the library's writers)."""
import struct

import numpy as np


def read_exr(data: bytes):
    """Reader for single-part scanline OpenEXR files with uncompressed 32-bit
    float channels, written from the published file layout (magic, version,
    header attributes ending in an empty name, scanline offset table, blocks
    of y / byte count / channel rows in channel-list order).  Returns the
    header attributes and {channel: (H, W) array}."""
    magic, version = struct.unpack("<ii", data[:8])
    assert magic == 20000630 and (version & 0xFF) == 2 and (version >> 8) == 0  # scanline, single part
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        type_end = data.index(b"\0", name_end + 1)
        name, typ = data[pos:name_end].decode(), data[name_end + 1:type_end].decode()
        size = struct.unpack("<i", data[type_end + 1:type_end + 5])[0]
        attrs[name] = (typ, data[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    chans, cl, q = [], attrs["channels"][1], 0
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        ptype, _, xs, ys = struct.unpack("<iB3xii", cl[e + 1:e + 17])
        assert ptype == 2 and xs == 1 and ys == 1   # FLOAT, full resolution
        chans.append(cl[q:e].decode())
        q = e + 17
    assert chans == sorted(chans)
    assert attrs["compression"] == ("compression", b"\0")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    assert attrs["lineOrder"] == ("lineOrder", b"\0")
    assert struct.unpack("<f", attrs["pixelAspectRatio"][1])[0] == 1.0
    assert "screenWindowCenter" in attrs and "screenWindowWidth" in attrs
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack(f"<{h}Q", data[pos:pos + 8 * h])
    img = {c: np.zeros((h, w), np.float32) for c in chans}
    for row, off in enumerate(offsets):
        y, n = struct.unpack("<ii", data[off:off + 8])
        assert y == y0 + row and n == 4 * w * len(chans)
        block = np.frombuffer(data[off + 8:off + 8 + n], "<f4").reshape(len(chans), w)
        for c, v in zip(chans, block):
            img[c][row] = v
    return attrs, img


def read_pfm(data: bytes) -> np.ndarray:
    """PFM: "PF" (RGB) or "Pf" (one channel), size, negative scale =
    little-endian, rows bottom to top.  Returns (H, W, C) top row first."""
    head = data.split(b"\n", 3)
    kind, (w, h), scale = head[0], map(int, head[1].split()), float(head[2])
    assert kind in (b"PF", b"Pf") and scale < 0
    c = 3 if kind == b"PF" else 1
    return np.frombuffer(head[3], "<f4").reshape(h, w, c)[::-1]
