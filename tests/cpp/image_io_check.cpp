// image_io_check.cpp — test driver for the PNG / Radiance HDR readers
// (offlinerenderer_amd/csrc/image_io.cpp), built by tests/test_image_io.py
// with AddressSanitizer and UndefinedBehaviorSanitizer. For each path it
// prints: path, 1/0 (decoded or rejected), width, height, CRC-32 of the
// decoded texels.
#include <zlib.h>

#include <cstdio>
#include <string>

#include "../../offlinerenderer_amd/csrc/image_io.h"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const std::string p = argv[i];
        bool ok;
        uint32_t w = 0, h = 0;
        unsigned long crc = 0;
        if (p.size() > 4 && p.compare(p.size() - 4, 4, ".hdr") == 0) {
            wrhost::ImageF img;
            ok = wrhost::load_hdr_rgba32f(p, img);
            if (ok) crc = crc32(0, reinterpret_cast<const Bytef*>(img.rgba.data()), (uInt)(img.rgba.size() * 4));
            w = img.width; h = img.height;
        } else {
            wrhost::Image8 img;
            ok = wrhost::load_png_rgba8(p, img);
            if (ok) crc = crc32(0, img.rgba.data(), (uInt)img.rgba.size());
            w = img.width; h = img.height;
        }
        std::printf("%s %d %u %u %08lx\n", p.c_str(), ok ? 1 : 0, ok ? w : 0, ok ? h : 0, crc);
    }
    return 0;
}
