// image_io.cpp — PNG (8-bit, non-interlaced) and Radiance HDR (RGBE) readers
// for Texture2D::create / Scene::loadTexture (libWR/src/internal/scene.cpp:
// 246-264, which uses stb_image with 4 forced channels) and
// Scene::setEnvironmentMap (scene.cpp:289-348, stbi_loadf).
#include "image_io.h"

#include <zlib.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>

namespace wrhost {

// Largest image accepted: 2^24 texels per side (stb_image's default limit)
// and 2^28 texels in all, so sizes never overflow and a corrupt header
// cannot ask for an absurd allocation.
static bool dims_ok(uint64_t w, uint64_t h) {
    return w > 0 && h > 0 && w <= (1u << 24) && h <= (1u << 24) && w * h <= (1ull << 28);
}

static uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

static bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

static uint8_t paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

bool load_png_rgba8(const std::string& path, Image8& img) {
    std::vector<uint8_t> f;
    if (!read_file(path, f) || f.size() < 8) return false;
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (std::memcmp(f.data(), sig, 8) != 0) return false;
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (pos + 8 <= f.size()) {
        uint32_t len = be32(&f[pos]);
        const char* type = (const char*)&f[pos + 4];
        if (pos + 12 + (size_t)len > f.size()) return false;
        const uint8_t* d = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return false;
            w = be32(d); h = be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + len;
    }
    if (!dims_ok(w, h) || interlace != 0 || depth != 8) return false;
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: return false;
    }
    const size_t stride = (size_t)w * channels;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf rawLen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawLen, idat.data(), (uLong)idat.size()) != Z_OK || rawLen != raw.size()) return false;
    std::vector<uint8_t> px(stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t filter = raw[y * (stride + 1)];
        const uint8_t* src = &raw[y * (stride + 1) + 1];
        uint8_t* dst = &px[y * stride];
        const uint8_t* prev = y ? &px[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            int a = x >= (size_t)channels ? dst[x - channels] : 0;
            int b = prev ? prev[x] : 0;
            int c = (prev && x >= (size_t)channels) ? prev[x - channels] : 0;
            int v = src[x];
            switch (filter) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: v += paeth(a, b, c); break;
                default: return false;
            }
            dst[x] = (uint8_t)v;
        }
    }
    img.width = w; img.height = h;
    img.rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        uint8_t r, g, b, a = 255;
        const uint8_t* s = &px[i * channels];
        if (ctype == 0) { r = g = b = s[0]; }
        else if (ctype == 4) { r = g = b = s[0]; a = s[1]; }
        else if (ctype == 2) { r = s[0]; g = s[1]; b = s[2]; }
        else if (ctype == 6) { r = s[0]; g = s[1]; b = s[2]; a = s[3]; }
        else {
            uint32_t k = s[0];
            if ((size_t)k * 3 + 2 >= plte.size()) return false;
            r = plte[k * 3]; g = plte[k * 3 + 1]; b = plte[k * 3 + 2];
            if (k < trns.size()) a = trns[k];
        }
        img.rgba[i * 4 + 0] = r; img.rgba[i * 4 + 1] = g; img.rgba[i * 4 + 2] = b; img.rgba[i * 4 + 3] = a;
    }
    return true;
}

bool load_hdr_rgba32f(const std::string& path, ImageF& img) {
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return false;
    size_t pos = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (pos < f.size() && f[pos] != '\n') s.push_back((char)f[pos++]);
        if (pos < f.size()) ++pos;
        return pos <= f.size();
    };
    std::string s;
    line(s);
    if (s.rfind("#?", 0) != 0) return false;
    do { if (!line(s)) return false; } while (!s.empty() && pos < f.size());
    line(s);
    int w = 0, h = 0;
    char ya[3] = {0}, xa[3] = {0};
    if (std::sscanf(s.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || w <= 0 || h <= 0 ||
        !dims_ok((uint64_t)w, (uint64_t)h))
        return false;
    std::vector<uint8_t> rgbe((size_t)w * h * 4);
    for (int y = 0; y < h; ++y) {
        uint8_t* row = &rgbe[(size_t)y * w * 4];
        if (pos + 4 <= f.size() && f[pos] == 2 && f[pos + 1] == 2 && w >= 8 && w < 32768) {
            pos += 4;
            for (int c = 0; c < 4; ++c) {
                int x = 0;
                while (x < w) {
                    if (pos >= f.size()) return false;
                    int cnt = f[pos++];
                    if (cnt > 128) {
                        cnt -= 128;
                        if (pos >= f.size() || x + cnt > w) return false;
                        uint8_t v = f[pos++];
                        for (int k = 0; k < cnt; ++k) row[(x++) * 4 + c] = v;
                    } else {
                        if (cnt == 0 || pos + cnt > f.size() || x + cnt > w) return false;
                        for (int k = 0; k < cnt; ++k) row[(x++) * 4 + c] = f[pos++];
                    }
                }
            }
        } else {
            if (pos + (size_t)w * 4 > f.size()) return false;
            std::memcpy(row, &f[pos], (size_t)w * 4);
            pos += (size_t)w * 4;
        }
    }
    img.width = (uint32_t)w; img.height = (uint32_t)h;
    img.rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* e = &rgbe[i * 4];
        float scale = e[3] ? std::ldexp(1.0f, (int)e[3] - (128 + 8)) : 0.0f;
        img.rgba[i * 4 + 0] = e[0] * scale;
        img.rgba[i * 4 + 1] = e[1] * scale;
        img.rgba[i * 4 + 2] = e[2] * scale;
        img.rgba[i * 4 + 3] = 1.0f;
    }
    return true;
}

}  // namespace wrhost
