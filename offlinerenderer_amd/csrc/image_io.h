// image_io.h — minimal image readers used by the Scene texture API.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace wrhost {

struct Image8 {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba;  // row-major, 4 bytes per texel
};
struct ImageF {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;  // row-major, 4 floats per texel
};

// 8-bit PNG (gray, gray+alpha, RGB, RGBA, palette), non-interlaced,
// expanded to RGBA8. Returns false on any unsupported or corrupt input.
bool load_png_rgba8(const std::string& path, Image8& img);
// Radiance .hdr (RGBE, flat or new-style RLE scanlines) to RGBA32F.
bool load_hdr_rgba32f(const std::string& path, ImageF& img);

}  // namespace wrhost
