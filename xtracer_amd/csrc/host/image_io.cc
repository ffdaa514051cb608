#include "image_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace xt {

namespace {

struct File {
    FILE *f = nullptr;
    bool ok = true;
    explicit File(const char *path) : f(path ? std::fopen(path, "wb") : nullptr) { ok = f != nullptr; }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void *p, size_t n) {
        if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
    }
    void u8(uint8_t v) { put(&v, 1); }
    void le16(uint16_t v) {
        uint8_t b[2] = {(uint8_t)v, (uint8_t)(v >> 8)};
        put(b, 2);
    }
    void le32(uint32_t v) {
        uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        put(b, 4);
    }
    void le64(uint64_t v) {
        le32((uint32_t)v);
        le32((uint32_t)(v >> 32));
    }
    void str(const char *s) { put(s, std::strlen(s)); }
    int close() {
        if (f && std::fclose(f) != 0) ok = false;
        f = nullptr;
        return ok ? 0 : 1;
    }
};

// ------------------------------------------------------------- checksums

uint32_t crc_table[256];
bool crc_ready = false;

// CRC-32 (ISO 3309 polynomial 0xedb88320, as PNG specifies).
uint32_t crc32(uint32_t crc, const uint8_t *p, size_t n) {
    if (!crc_ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
            crc_table[i] = c;
        }
        crc_ready = true;
    }
    crc ^= 0xffffffffu;
    for (size_t i = 0; i < n; ++i) crc = crc_table[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
    return crc ^ 0xffffffffu;
}

// Adler-32 (RFC 1950).
uint32_t adler32(const uint8_t *p, size_t n) {
    uint32_t a = 1, b = 0;
    for (size_t i = 0; i < n; ++i) {
        a = (a + p[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    return (b << 16) | a;
}

void be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void png_chunk(File &f, const char *type, const std::vector<uint8_t> &data) {
    uint8_t len[4] = {(uint8_t)(data.size() >> 24), (uint8_t)(data.size() >> 16), (uint8_t)(data.size() >> 8),
                      (uint8_t)data.size()};
    f.put(len, 4);
    std::vector<uint8_t> body(type, type + 4);
    body.insert(body.end(), data.begin(), data.end());
    f.put(body.data(), body.size());
    std::vector<uint8_t> crc;
    be32(crc, crc32(0, body.data(), body.size()));
    f.put(crc.data(), 4);
}

bool valid(const float *rgba, uint32_t w, uint32_t h) { return rgba && w && h && w < (1u << 16) && h < (1u << 16); }

// Radiance RGBE of one RGB pixel (G. Ward, "Real Pixels", Graphics Gems II,
// 1991): the shared exponent of the largest component.
void to_rgbe(const float *c, uint8_t *e) {
    const float m = std::max(c[0], std::max(c[1], c[2]));
    if (m < 1e-32f) {
        e[0] = e[1] = e[2] = e[3] = 0;
        return;
    }
    int ex = 0;
    const float scale = (float)std::frexp(m, &ex) * 256.0f / m;
    e[0] = (uint8_t)(c[0] * scale);
    e[1] = (uint8_t)(c[1] * scale);
    e[2] = (uint8_t)(c[2] * scale);
    e[3] = (uint8_t)(ex + 128);
}

// The framebuffer as the reference's Pixmap holds it: ColorRGBAf's
// constructor (lib/nimg/color.cc:35-40) sets negative r, g, b to 0 and
// clamps alpha to [0, 1].
std::vector<float> as_pixmap(const float *rgba, uint32_t w, uint32_t h) {
    std::vector<float> c(rgba, rgba + (size_t)w * h * 4);
    for (size_t i = 0; i < c.size(); i += 4) {
        for (int k = 0; k < 3; ++k) c[i + k] = c[i + k] < 0.f ? 0.f : c[i + k];
        c[i + 3] = c[i + 3] < 0.f ? 0.f : (c[i + 3] > 1.f ? 1.f : c[i + 3]);
    }
    return c;
}

}  // namespace

void rgba_to_8bit(const float *rgba, uint32_t w, uint32_t h, std::vector<uint8_t> &out) {
    out.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h * 4; ++i) {
        const float v = rgba[i] * 255.f;
        out[i] = (uint8_t)(v > 255.f ? 255.f : v);
    }
}

int write_png(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    std::vector<uint8_t> px;
    rgba_to_8bit(rgba, w, h, px);
    // filter type 0 on every row
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (1 + 4 * (size_t)w));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), px.begin() + (size_t)y * w * 4, px.begin() + (size_t)(y + 1) * w * 4);
    }
    // zlib stream of deflate "stored" blocks (RFC 1950 / 1951)
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t at = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - at);
        const bool last = at + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)n);
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)~n);
        z.push_back((uint8_t)(~n >> 8));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + n);
        at += n;
    } while (at < raw.size());
    be32(z, adler32(raw.data(), raw.size()));
    File f(path);
    const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    f.put(sig, 8);
    std::vector<uint8_t> ihdr;
    be32(ihdr, w);
    be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, no interlace
    png_chunk(f, "IHDR", ihdr);
    png_chunk(f, "IDAT", z);
    png_chunk(f, "IEND", {});
    return f.close();
}

int write_bmp(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    std::vector<uint8_t> px;
    rgba_to_8bit(rgba, w, h, px);
    const uint32_t pad = (0u - w * 3u) & 3u, row = w * 3 + pad;
    File f(path);
    f.put("BM", 2);
    f.le32(14 + 40 + row * h);
    f.le32(0);
    f.le32(14 + 40);
    f.le32(40);
    f.le32(w);
    f.le32(h);
    f.le16(1);
    f.le16(24);
    for (int k = 0; k < 6; ++k) f.le32(0);
    std::vector<uint8_t> line(row, 0);
    for (uint32_t r = 0; r < h; ++r) {  // bottom row first
        const uint8_t *src = px.data() + (size_t)(h - 1 - r) * w * 4;
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t *d = src + 4 * x;
            // no alpha in the file: composite over magenta (255, 0, 255)
            const int bg[3] = {255, 0, 255};
            uint8_t c[3];
            for (int k = 0; k < 3; ++k) c[k] = (uint8_t)(bg[k] + ((int)d[k] - bg[k]) * (int)d[3] / 255);
            line[3 * x] = c[2];
            line[3 * x + 1] = c[1];
            line[3 * x + 2] = c[0];
        }
        f.put(line.data(), row);
    }
    return f.close();
}

int write_tga(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    std::vector<uint8_t> px;
    rgba_to_8bit(rgba, w, h, px);
    File f(path);
    const uint8_t hdr[12] = {0, 0, 10, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // run-length true colour
    f.put(hdr, 12);
    f.le16((uint16_t)w);
    f.le16((uint16_t)h);
    f.u8(32);
    f.u8(8);  // 8 alpha bits, bottom-left origin
    auto bgra = [&](uint32_t y, uint32_t x, uint8_t *o) {
        const uint8_t *d = px.data() + ((size_t)y * w + x) * 4;
        o[0] = d[2];
        o[1] = d[1];
        o[2] = d[0];
        o[3] = d[3];
    };
    for (uint32_t r = 0; r < h; ++r) {
        const uint32_t y = h - 1 - r;
        uint32_t x = 0;
        while (x < w) {
            uint8_t a[4], b[4];
            bgra(y, x, a);
            // a run of equal pixels (up to 128) or a literal packet up to the next run
            uint32_t run = 1;
            while (x + run < w && run < 128) {
                bgra(y, x + run, b);
                if (std::memcmp(a, b, 4)) break;
                ++run;
            }
            if (run > 1) {
                f.u8((uint8_t)(0x80 | (run - 1)));
                f.put(a, 4);
                x += run;
                continue;
            }
            uint32_t lit = 1;
            while (x + lit < w && lit < 128) {
                uint8_t c[4];
                bgra(y, x + lit, b);
                if (x + lit + 1 < w) {
                    bgra(y, x + lit + 1, c);
                    if (!std::memcmp(b, c, 4)) break;
                }
                ++lit;
            }
            f.u8((uint8_t)(lit - 1));
            for (uint32_t k = 0; k < lit; ++k) {
                bgra(y, x + k, b);
                f.put(b, 4);
            }
            x += lit;
        }
    }
    return f.close();
}

int write_hdr(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    File f(path);
    char head[128];
    std::snprintf(head, sizeof head, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    f.str(head);
    std::vector<uint8_t> e((size_t)w * 4), comp(w);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) to_rgbe(rgba + ((size_t)y * w + x) * 4, &e[4 * x]);
        if (w < 8 || w >= 32768) {  // flat scanline
            f.put(e.data(), e.size());
            continue;
        }
        // new-style run-length scanline: each component on its own
        const uint8_t tag[4] = {2, 2, (uint8_t)(w >> 8), (uint8_t)w};
        f.put(tag, 4);
        for (int c = 0; c < 4; ++c) {
            for (uint32_t x = 0; x < w; ++x) comp[x] = e[4 * x + c];
            uint32_t x = 0;
            while (x < w) {
                uint32_t run = 1;
                while (x + run < w && run < 127 && comp[x + run] == comp[x]) ++run;
                if (run >= 4) {
                    f.u8((uint8_t)(128 + run));
                    f.u8(comp[x]);
                    x += run;
                    continue;
                }
                // literal bytes up to the next run of 4 or 128 bytes
                uint32_t lit = 0;
                while (x + lit < w && lit < 128) {
                    uint32_t r2 = 1;
                    while (x + lit + r2 < w && r2 < 4 && comp[x + lit + r2] == comp[x + lit]) ++r2;
                    if (r2 >= 4) break;
                    ++lit;
                }
                f.u8((uint8_t)lit);
                f.put(&comp[x], lit);
                x += lit;
            }
        }
    }
    return f.close();
}

int write_exr(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    std::vector<uint8_t> head;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = (const uint8_t *)p;
        head.insert(head.end(), b, b + n);
    };
    auto i32 = [&](int32_t v) { put(&v, 4); };  // the host is little-endian, like the format
    auto attr = [&](const char *name, const char *type, uint32_t size) {
        put(name, std::strlen(name) + 1);
        put(type, std::strlen(type) + 1);
        i32((int32_t)size);
    };
    const uint8_t magic[8] = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};  // version 2, single-part scanline
    put(magic, 8);
    attr("channels", "chlist", 3 * 18 + 1);
    for (const char *ch : {"B", "G", "R"}) {  // alphabetical, as the format requires
        put(ch, 2);
        i32(2);  // FLOAT
        const uint8_t lin[4] = {0, 0, 0, 0};
        put(lin, 4);
        i32(1);
        i32(1);
    }
    head.push_back(0);
    attr("compression", "compression", 1);
    head.push_back(0);  // none
    attr("dataWindow", "box2i", 16);
    i32(0);
    i32(0);
    i32((int32_t)w - 1);
    i32((int32_t)h - 1);
    attr("displayWindow", "box2i", 16);
    i32(0);
    i32(0);
    i32((int32_t)w - 1);
    i32((int32_t)h - 1);
    attr("lineOrder", "lineOrder", 1);
    head.push_back(0);  // increasing y
    attr("pixelAspectRatio", "float", 4);
    const float one = 1.f, zero[2] = {0.f, 0.f};
    put(&one, 4);
    attr("screenWindowCenter", "v2f", 8);
    put(zero, 8);
    attr("screenWindowWidth", "float", 4);
    put(&one, 4);
    head.push_back(0);  // end of header
    const uint64_t block = 8 + (uint64_t)w * 3 * 4;
    const uint64_t first = head.size() + (uint64_t)h * 8;
    File f(path);
    f.put(head.data(), head.size());
    for (uint32_t y = 0; y < h; ++y) f.le64(first + y * block);
    std::vector<float> line((size_t)w * 3);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const float *p = rgba + ((size_t)y * w + x) * 4;
            line[x] = p[2];
            line[w + x] = p[1];
            line[2 * (size_t)w + x] = p[0];
        }
        f.le32(y);
        f.le32((uint32_t)(line.size() * 4));
        f.put(line.data(), line.size() * 4);
    }
    return f.close();
}

int write_pfm(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    File f(path);
    char head[64];
    std::snprintf(head, sizeof head, "PF\n%u %u\n-1.0\n", w, h);
    f.str(head);
    std::vector<float> line((size_t)w * 3);
    for (uint32_t r = 0; r < h; ++r) {
        const float *src = rgba + (size_t)(h - 1 - r) * w * 4;
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) line[3 * x + c] = src[4 * x + c];
        f.put(line.data(), line.size() * 4);
    }
    return f.close();
}

int write_ppm(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!valid(rgba, w, h)) return 1;
    const std::vector<float> pm = as_pixmap(rgba, w, h);
    rgba = pm.data();
    std::vector<uint8_t> px;
    rgba_to_8bit(rgba, w, h, px);
    File f(path);
    char head[64];
    std::snprintf(head, sizeof head, "P6\n%u %u\n255\n", w, h);
    f.str(head);
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = px[4 * i + c];
    f.put(rgb.data(), rgb.size());
    return f.close();
}

int write_image(const char *path, const float *rgba, uint32_t w, uint32_t h) {
    if (!path) return 1;
    std::string p(path);
    const size_t dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    for (char &c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext == "png") return write_png(path, rgba, w, h);
    if (ext == "bmp") return write_bmp(path, rgba, w, h);
    if (ext == "tga") return write_tga(path, rgba, w, h);
    if (ext == "hdr") return write_hdr(path, rgba, w, h);
    if (ext == "exr") return write_exr(path, rgba, w, h);
    if (ext == "pfm") return write_pfm(path, rgba, w, h);
    if (ext == "ppm") return write_ppm(path, rgba, w, h);
    return 2;
}

}  // namespace xt
