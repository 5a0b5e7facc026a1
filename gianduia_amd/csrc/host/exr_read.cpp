// exr_read.cpp — OpenEXR reader for environment maps.
//
// The reference loads environment images with OpenEXR (Image2D::loadEXR,
// include/gianduia/core/image2d.h:100-134): R, G and B read as 32-bit float
// slices over the data window, a channel the file lacks filled with 0.
// OpenEXR is not part of this build, so this is a reader of its own written
// from the published file layout (openexr.com "OpenEXR File Layout"):
// single-part scanline files, channels UINT / HALF / FLOAT at full
// resolution, compression NONE, RLE, ZIPS or ZIP (zlib).  Tiled, deep,
// multi-part, subsampled and PIZ / PXR24 / B44 / DWA files are rejected
// with an error naming what is unsupported.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "integrator.h"

namespace gnd {

namespace {

[[noreturn]] void bad(const std::string& path, const std::string& why) {
    throw std::runtime_error("EXR Load Error (" + path + "): " + why);
}

struct Cursor {
    const unsigned char* p;
    size_t n, i = 0;
    const std::string& path;
    template <typename T>
    T get() {
        if (n - i < sizeof(T)) bad(path, "truncated file");
        T v;
        std::memcpy(&v, p + i, sizeof(T));  // little-endian host
        i += sizeof(T);
        return v;
    }
    std::string str() {
        size_t e = i;
        while (e < n && p[e]) ++e;
        if (e == n) bad(path, "unterminated name");
        std::string s(reinterpret_cast<const char*>(p + i), e - i);
        i = e + 1;
        return s;
    }
};

// IEEE 754 binary16 -> binary32, exact for every input (subnormals, inf, NaN)
float halfToFloat(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1F, m = h & 0x3FF;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = sign;
        else {
            const float v = std::ldexp((float)m, -24);
            std::memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7F800000u | (m << 13);
    } else {
        bits = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// The reordering ZIP and RLE apply before entropy coding, undone: a byte
// delta predictor (each byte stored as the difference to the previous one,
// offset by 128), over a buffer whose first half holds the even-indexed
// bytes of the block and whose second half the odd-indexed ones.
void unpredictInterleave(std::vector<unsigned char>& t) {
    const size_t n = t.size();
    for (size_t i = 1; i < n; ++i) t[i] = (unsigned char)(t[i - 1] + t[i] - 128);
    std::vector<unsigned char> out(n);
    const size_t half = (n + 1) / 2;
    for (size_t k = 0; k < half; ++k) {
        out[2 * k] = t[k];
        if (2 * k + 1 < n) out[2 * k + 1] = t[half + k];
    }
    t.swap(out);
}

// RLE runs: a signed count byte; negative = that many literal bytes follow,
// else the next byte repeats count + 1 times
void rleDecode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t want,
               const std::string& path) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int c = (signed char)in[i++];
        if (c < 0) {
            if (i + (size_t)(-c) > n) bad(path, "corrupt RLE block");
            out.insert(out.end(), in + i, in + i + (-c));
            i += (size_t)(-c);
        } else {
            if (i >= n) bad(path, "corrupt RLE block");
            out.insert(out.end(), (size_t)c + 1, in[i++]);
        }
        if (out.size() > want) bad(path, "RLE block larger than its scanlines");
    }
    if (out.size() != want) bad(path, "RLE block size mismatch");
}

struct Channel {
    std::string name;
    int type;  // 0 UINT, 1 HALF, 2 FLOAT
};

}  // namespace

void readEXR(const std::string& path, int& width, int& height, std::vector<float>& rgb) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad(path, "cannot open");
    std::vector<unsigned char> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Cursor c{data.data(), data.size(), 0, path};
    if (c.get<int32_t>() != 20000630) bad(path, "not an OpenEXR file");
    const int32_t version = c.get<int32_t>();
    if ((version & 0xFF) != 2) bad(path, "unsupported file version");
    if (version & 0x200) bad(path, "tiled files are not supported");
    if (version & 0x800) bad(path, "deep data is not supported");
    if (version & 0x1000) bad(path, "multi-part files are not supported");

    std::vector<Channel> chans;
    int compression = -1;
    int32_t box[4] = {0, 0, -1, -1};
    bool haveBox = false;
    for (;;) {
        const std::string name = c.str();
        if (name.empty()) break;
        const std::string type = c.str();
        const int32_t size = c.get<int32_t>();
        if (size < 0 || (size_t)size > c.n - c.i) bad(path, "bad attribute size");
        Cursor a{c.p + c.i, (size_t)size, 0, path};
        if (name == "channels") {
            if (type != "chlist") bad(path, "bad channel list");
            for (;;) {
                std::string cn = a.str();
                if (cn.empty()) break;
                const int32_t pt = a.get<int32_t>();
                a.get<uint32_t>();  // pLinear + reserved
                const int32_t xs = a.get<int32_t>(), ys = a.get<int32_t>();
                if (pt < 0 || pt > 2) bad(path, "unknown pixel type");
                if (xs != 1 || ys != 1) bad(path, "subsampled channels are not supported");
                chans.push_back({cn, pt});
            }
        } else if (name == "compression") {
            compression = a.get<uint8_t>();
        } else if (name == "dataWindow") {
            for (int k = 0; k < 4; ++k) box[k] = a.get<int32_t>();
            haveBox = true;
        }
        c.i += (size_t)size;
    }
    if (chans.empty() || compression < 0 || !haveBox) bad(path, "missing a required header attribute");
    int linesPerBlock;
    switch (compression) {
        case 0: case 1: case 2: linesPerBlock = 1; break;  // NONE, RLE, ZIPS
        case 3: linesPerBlock = 16; break;                   // ZIP
        default: bad(path, "compression " + std::to_string(compression) + " is not supported");
    }
    const int64_t w = (int64_t)box[2] - box[0] + 1, h = (int64_t)box[3] - box[1] + 1;
    if (w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16)) bad(path, "bad data window");
    width = (int)w;
    height = (int)h;
    rgb.assign((size_t)w * h * 3, 0.0f);
    int target[3] = {-1, -1, -1};  // channel index of R, G, B
    size_t pixelBytes = 0;
    std::vector<size_t> chOff;  // byte offset of each channel within a scanline
    for (size_t k = 0; k < chans.size(); ++k) {
        chOff.push_back(pixelBytes * (size_t)w);
        pixelBytes += chans[k].type == 1 ? 2 : 4;
        if (chans[k].name == "R") target[0] = (int)k;
        if (chans[k].name == "G") target[1] = (int)k;
        if (chans[k].name == "B") target[2] = (int)k;
    }
    const size_t lineBytes = pixelBytes * (size_t)w;
    const int64_t nBlocks = (h + linesPerBlock - 1) / linesPerBlock;
    std::vector<uint64_t> offsets((size_t)nBlocks);
    for (auto& o : offsets) o = c.get<uint64_t>();
    std::vector<unsigned char> raw;
    for (int64_t b = 0; b < nBlocks; ++b) {
        Cursor k{data.data(), data.size(), (size_t)offsets[b], path};
        if (offsets[b] >= data.size()) bad(path, "bad block offset");
        const int32_t y0 = k.get<int32_t>();
        const int32_t packed = k.get<int32_t>();
        if (packed < 0 || (size_t)packed > k.n - k.i) bad(path, "bad block size");
        if (y0 < box[1] || y0 > box[3] || (y0 - box[1]) % linesPerBlock) bad(path, "bad block scanline");
        const int lines = (int)std::min<int64_t>(linesPerBlock, (int64_t)box[3] - y0 + 1);
        const size_t want = lineBytes * (size_t)lines;
        const unsigned char* src = data.data() + k.i;
        if (compression == 0 || (size_t)packed == want) {
            // stored as is (a compressed block no smaller than its raw bytes is kept raw)
            if ((size_t)packed != want) bad(path, "block size mismatch");
            raw.assign(src, src + want);
        } else if (compression == 1) {
            rleDecode(src, (size_t)packed, raw, want, path);
            unpredictInterleave(raw);
        } else {
            raw.resize(want);
            uLongf got = (uLongf)want;
            if (uncompress(raw.data(), &got, src, (uLong)packed) != Z_OK || got != want)
                bad(path, "corrupt zlib block");
            unpredictInterleave(raw);
        }
        for (int l = 0; l < lines; ++l) {
            const unsigned char* line = raw.data() + (size_t)l * lineBytes;
            float* dst = rgb.data() + (size_t)(y0 - box[1] + l) * w * 3;
            for (int ch = 0; ch < 3; ++ch) {
                const int s = target[ch];
                if (s < 0) continue;  // absent channel: the zero fill
                const unsigned char* q = line + chOff[s];
                for (int64_t x = 0; x < w; ++x) {
                    float v;
                    if (chans[s].type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, q + 2 * x, 2);
                        v = halfToFloat(hv);
                    } else if (chans[s].type == 2) {
                        std::memcpy(&v, q + 4 * x, 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, q + 4 * x, 4);
                        v = (float)u;
                    }
                    dst[3 * x + ch] = v;
                }
            }
        }
    }
}

}  // namespace gnd
