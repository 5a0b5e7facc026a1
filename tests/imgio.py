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


def _half_bits(a: np.ndarray) -> np.ndarray:
    return np.asarray(a, np.float16).view("<u2")


def _rle(raw: bytes) -> bytes:
    """OpenEXR RLE: runs of 3+ equal bytes as (count - 1, byte), the rest as
    (-literal count, bytes), counts up to 127 / 128."""
    out, i, n = bytearray(), 0, len(raw)
    while i < n:
        j = i
        while j + 1 < n and raw[j + 1] == raw[i] and j - i < 127:
            j += 1
        if j - i >= 2:
            out += bytes([j - i]) + raw[i:i + 1]
            i = j + 1
            continue
        k = i
        while k < n and k - i < 127 and not (k + 2 < n and raw[k] == raw[k + 1] == raw[k + 2]):
            k += 1
        out += bytes([(-(k - i)) & 0xFF]) + raw[i:k]
        i = k
    return bytes(out)


def _pre(raw: bytes) -> bytes:
    """The byte reordering ZIP and RLE apply first: even-indexed bytes then
    odd-indexed ones, then each byte replaced by its difference to the one
    before, plus 128 (mod 256)."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def write_exr(path, channels: dict, ptype: int = 2, compression: int = 0, origin=(0, 0)):
    """Scanline single-part OpenEXR writer from the published file layout:
    channels {name: (H, W) array} stored sorted by name with pixel type
    `ptype` (0 UINT, 1 HALF, 2 FLOAT) and `compression` 0 NONE, 1 RLE,
    2 ZIPS, 3 ZIP; the data window starts at `origin` (x, y).  A block that
    does not shrink is stored raw, as the format prescribes."""
    import zlib
    names = sorted(channels)
    h, w = next(iter(channels.values())).shape
    x0, y0 = origin

    def attr(name, typ, payload):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", x0, y0, x0 + w - 1, y0 + h - 1)
    head = struct.pack("<ii", 20000630, 2)
    head += attr("channels", "chlist", chl) + attr("compression", "compression", bytes([compression]))
    head += attr("dataWindow", "box2i", box) + attr("displayWindow", "box2i", box)
    head += attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    head += attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    lpb = 16 if compression == 3 else 1
    blocks = []
    for b0 in range(0, h, lpb):
        raw = bytearray()
        for y in range(b0, min(h, b0 + lpb)):
            for n in names:
                row = channels[n][y]
                if ptype == 1:
                    raw += _half_bits(row).tobytes()
                elif ptype == 0:
                    raw += np.asarray(row, "<u4").tobytes()
                else:
                    raw += np.asarray(row, "<f4").tobytes()
        raw = bytes(raw)
        if compression == 1:
            packed = _rle(_pre(raw))
        elif compression in (2, 3):
            packed = zlib.compress(_pre(raw))
        else:
            packed = raw
        if len(packed) >= len(raw):
            packed = raw
        blocks.append(struct.pack("<ii", y0 + b0, len(packed)) + packed)
    table_end = len(head) + 8 * len(blocks)
    offs, pos = [], table_end
    for b in blocks:
        offs.append(pos)
        pos += len(b)
    with open(path, "wb") as f:
        f.write(head + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(blocks))
