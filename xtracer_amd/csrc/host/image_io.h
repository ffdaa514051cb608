// Image files from an assembled RGBA float framebuffer: the formats the
// reference saves through nimg::io::save (lib/nimg/img.cc:72-318) and the
// GUI's / CLI's "save" actions (src/frontend/gui/action.cc:95-110,
// src/frontend/cli/xtracer.cc:140-150), written without third-party code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace xt {

// The reference's 8-bit conversion (img.cc:152-163): each of r, g, b, a
// times 255, values above 255 set to 255, then truncated; row-major RGBA8.
void rgba_to_8bit(const float *rgba, uint32_t w, uint32_t h, std::vector<uint8_t> &out);

// Every writer first clamps the framebuffer the way the reference's pixels
// are clamped (negative colour to 0, alpha to [0, 1]).
// Each returns 0 on success, 1 on a bad argument or an I/O failure.
int write_png(const char *path, const float *rgba, uint32_t w, uint32_t h);   // RGBA8, deflate "stored" blocks
int write_bmp(const char *path, const float *rgba, uint32_t w, uint32_t h);   // 24-bit, alpha over (255, 0, 255)
int write_tga(const char *path, const float *rgba, uint32_t w, uint32_t h);   // 32-bit BGRA, run-length packets
int write_hdr(const char *path, const float *rgba, uint32_t w, uint32_t h);   // Radiance RGBE, run-length scanlines
int write_exr(const char *path, const float *rgba, uint32_t w, uint32_t h);   // OpenEXR scanline, FLOAT B G R, no compression
int write_pfm(const char *path, const float *rgba, uint32_t w, uint32_t h);   // float RGB, bottom row first
int write_ppm(const char *path, const float *rgba, uint32_t w, uint32_t h);   // P6, 8-bit RGB

// Picks the writer from the file extension (.png .bmp .tga .hdr .exr .pfm
// .ppm); 2 for an unknown extension.
int write_image(const char *path, const float *rgba, uint32_t w, uint32_t h);

}  // namespace xt
