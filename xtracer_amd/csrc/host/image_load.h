// Texture images as the reference loads them: nimg::io::load::image
// (lib/nimg/img.cc:17-66) reads 8-bit files through stb_image, keeps only
// 3- and 4-channel images, and stores byte / 255.f per channel (alpha 1 for
// 3 channels), top row first. These decoders are written from the format
// specifications for the files stb_image reads that way: PNG (8-bit RGB /
// RGBA, not interlaced), TGA (true colour, raw or run-length, 24 / 32 bit),
// BMP (uncompressed 24 / 32 bit), binary PPM (P6, maxval 255), Radiance HDR,
// which stb_image converts to 8 bits, and baseline / progressive JPEG decoded
// with stb_image's integer arithmetic.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace xt {

struct Image {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;  // width * height * 4, row 0 at the top
};

// 0: loaded; 1: the file cannot be read; 2: a format or variant this loader
// does not decode (or one the reference rejects: 1- and 2-channel images), or
// an image with a side above 65535 pixels.
int load_image(const std::string &path, Image &out, std::string *err = nullptr);

}  // namespace xt
