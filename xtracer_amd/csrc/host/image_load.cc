#include "image_load.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace xt {

namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint32_t le32(const uint8_t *p) { return (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]; }
uint16_t le16(const uint8_t *p) { return (uint16_t)(p[1] << 8 | p[0]); }

int fail(std::string *err, const std::string &msg, int code) {
    if (err) *err = msg;
    return code;
}

// channels 3 or 4, bytes row-major top-down -> byte / 255.f (img.cc:30-57)
void to_float(const std::vector<uint8_t> &px, uint32_t w, uint32_t h, int ch, Image &out) {
    out.width = w;
    out.height = h;
    out.rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        for (int c = 0; c < 3; ++c) out.rgba[4 * i + c] = px[ch * i + c] / 255.f;
        out.rgba[4 * i + 3] = ch == 4 ? px[ch * i + 3] / 255.f : 1.f;
    }
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// Undoes the row filters of one (sub)image of w x h pixels: `in` holds h rows
// of a filter byte and `stride` bytes; `bpp` is the filter's byte distance.
bool unfilter(const uint8_t *in, size_t stride, uint32_t h, size_t bpp, std::vector<uint8_t> &rows) {
    rows.assign(stride * h, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t filter = in[y * (stride + 1)];
        const uint8_t *src = in + y * (stride + 1) + 1;
        uint8_t *row = &rows[y * stride];
        const uint8_t *up = y ? &rows[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0, b = up ? up[i] : 0;
            const int c = (up && i >= bpp) ? up[i - bpp] : 0;
            int pred;
            switch (filter) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: return false;
            }
            row[i] = (uint8_t)(src[i] + pred);
        }
    }
    return true;
}

// PNG as stbi_load(..., 0) returns it, for the images the reference keeps
// (3 or 4 channels): RGB, RGBA and palette images at any bit depth (16-bit
// samples keep their high byte; palette indices of 1, 2 or 4 bits), plain or
// Adam7-interlaced. A tRNS chunk adds alpha: per palette entry, or 0 where an
// RGB pixel equals the key colour (at 16 bits for 16-bit files, else the
// key's low bytes). Grey
// images come out with 1 or 2 channels, which the reference refuses.
int load_png(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, pal, trns;
    while (pos + 12 <= f.size()) {
        const uint32_t n = be32(&f[pos]);
        if (pos + 12 + (size_t)n > f.size()) return fail(err, "truncated PNG chunk", 2);
        const uint8_t *type = &f[pos + 4], *data = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (n < 13) return fail(err, "bad PNG header", 2);
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            if (data[10] || data[11] || interlace > 1) return fail(err, "bad PNG header", 2);
        } else if (!std::memcmp(type, "PLTE", 4)) {
            pal.assign(data, data + n);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(data, data + n);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + n);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)n;
    }
    if (!w || !h || w > (1u << 15) || h > (1u << 15)) return fail(err, "bad PNG size", 2);
    int file_ch;  // samples per pixel in the file
    switch (ctype) {
        case 2: file_ch = 3; break;
        case 6: file_ch = 4; break;
        case 3: file_ch = 1; break;
        default: return fail(err, "grey PNG: 1 or 2 channels, which the reference refuses", 2);
    }
    const bool ok_depth = ctype == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                     : (depth == 8 || depth == 16);
    if (!ok_depth) return fail(err, "bad PNG bit depth", 2);
    if (ctype == 3 && (pal.empty() || pal.size() % 3)) return fail(err, "PNG palette missing", 2);
    const int out_ch = ctype == 6 ? 4 : (trns.empty() ? 3 : 4);
    // the passes of the image: one, or Adam7's seven
    static const int xo[7] = {0, 4, 0, 2, 0, 1, 0}, yo[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int xs[7] = {8, 8, 4, 4, 2, 2, 1}, ys[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = interlace ? 7 : 1;
    size_t raw_len = 0;
    uint32_t pw[7], ph[7];
    size_t pstride[7];
    for (int p = 0; p < passes; ++p) {
        pw[p] = interlace ? (w - xo[p] + xs[p] - 1) / xs[p] : w;
        ph[p] = interlace ? (h - yo[p] + ys[p] - 1) / ys[p] : h;
        if (interlace && (w <= (uint32_t)xo[p] || h <= (uint32_t)yo[p])) pw[p] = ph[p] = 0;
        pstride[p] = ((size_t)pw[p] * file_ch * depth + 7) / 8;
        if (pw[p] && ph[p]) raw_len += (pstride[p] + 1) * ph[p];
    }
    std::vector<uint8_t> raw(raw_len);
    uLongf got = (uLongf)raw.size();
    if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != raw.size())
        return fail(err, "PNG data does not inflate", 2);
    std::vector<uint8_t> px((size_t)w * h * out_ch);
    std::vector<uint8_t> rows;
    const size_t bpp = std::max<size_t>(1, (size_t)file_ch * depth / 8);
    size_t at = 0;
    for (int p = 0; p < passes; ++p) {
        if (!pw[p] || !ph[p]) continue;
        if (!unfilter(&raw[at], pstride[p], ph[p], bpp, rows)) return fail(err, "bad PNG filter", 2);
        at += (pstride[p] + 1) * ph[p];
        for (uint32_t j = 0; j < ph[p]; ++j) {
            const uint8_t *row = &rows[j * pstride[p]];
            for (uint32_t i = 0; i < pw[p]; ++i) {
                const uint32_t x = interlace ? i * xs[p] + xo[p] : i, y = interlace ? j * ys[p] + yo[p] : j;
                uint8_t *o = &px[((size_t)y * w + x) * out_ch];
                if (ctype == 3) {
                    const size_t bit = (size_t)i * depth;
                    const int idx = depth == 8 ? row[i] : (row[bit / 8] >> (8 - depth - bit % 8)) & ((1 << depth) - 1);
                    if ((size_t)idx * 3 + 2 >= pal.size()) return fail(err, "PNG palette index out of range", 2);
                    o[0] = pal[3 * idx];
                    o[1] = pal[3 * idx + 1];
                    o[2] = pal[3 * idx + 2];
                    if (out_ch == 4) o[3] = (size_t)idx < trns.size() ? trns[idx] : 255;
                    continue;
                }
                bool key = out_ch == 4 && ctype == 2 && trns.size() >= 6;
                for (int c = 0; c < file_ch; ++c) {
                    if (depth == 16) {
                        const uint8_t *sv = &row[((size_t)i * file_ch + c) * 2];
                        o[c] = sv[0];  // the high byte: stb's 16 -> 8 bit reduction
                        key = key && sv[0] == trns[2 * c] && sv[1] == trns[2 * c + 1];
                    } else {
                        o[c] = row[(size_t)i * file_ch + c];
                        key = key && o[c] == trns[2 * c + 1];  // stb keeps the key's low byte
                    }
                }
                if (out_ch == 4 && ctype == 2) o[3] = key ? 0 : 255;
            }
        }
    }
    to_float(px, w, h, out_ch, out);
    return 0;
}

int load_tga(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    if (f.size() < 18) return fail(err, "truncated TGA", 2);
    const int idlen = f[0], cmap = f[1], type = f[2], bits = f[16], desc = f[17];
    const uint32_t w = le16(&f[12]), h = le16(&f[14]);
    if (cmap || (type != 2 && type != 10) || (bits != 24 && bits != 32) || !w || !h)
        return fail(err, "TGA variant not decoded (true colour, 24 / 32 bit)", 2);
    const int ch = bits / 8;
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> file_px(n * ch);
    size_t pos = 18 + idlen, k = 0;
    auto pixel = [&](size_t at) {  // BGR(A) -> RGB(A)
        file_px[at * ch] = f[pos + 2];
        file_px[at * ch + 1] = f[pos + 1];
        file_px[at * ch + 2] = f[pos];
        if (ch == 4) file_px[at * ch + 3] = f[pos + 3];
    };
    while (k < n) {
        if (type == 2) {
            if (pos + ch > f.size()) return fail(err, "truncated TGA", 2);
            pixel(k++);
            pos += ch;
            continue;
        }
        if (pos >= f.size()) return fail(err, "truncated TGA", 2);
        const int head = f[pos++], count = (head & 127) + 1;
        if (head & 128) {
            if (pos + ch > f.size()) return fail(err, "truncated TGA", 2);
            for (int r = 0; r < count && k < n; ++r) pixel(k++);
            pos += ch;
        } else {
            for (int r = 0; r < count && k < n; ++r) {
                if (pos + ch > f.size()) return fail(err, "truncated TGA", 2);
                pixel(k++);
                pos += ch;
            }
        }
    }
    // rows are stored bottom-up unless the descriptor's top-origin bit is set
    const bool top = (desc & 32) != 0;
    std::vector<uint8_t> px(n * ch);
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t src = top ? y : h - 1 - y;
        std::memcpy(&px[(size_t)y * w * ch], &file_px[(size_t)src * w * ch], (size_t)w * ch);
    }
    to_float(px, w, h, ch, out);
    return 0;
}

int load_bmp(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    if (f.size() < 54) return fail(err, "truncated BMP", 2);
    const uint32_t off = le32(&f[10]), hsz = le32(&f[14]);
    const int32_t wi = (int32_t)le32(&f[18]), hi = (int32_t)le32(&f[22]);
    const int bits = le16(&f[28]);
    const uint32_t comp = le32(&f[30]);
    if (hsz < 40 || comp != 0 || (bits != 24 && bits != 32) || wi <= 0 || hi == 0)
        return fail(err, "BMP variant not decoded (uncompressed 24 / 32 bit)", 2);
    const uint32_t w = (uint32_t)wi, h = (uint32_t)(hi < 0 ? -hi : hi);
    const int bpp = bits / 8, ch = bits == 32 ? 4 : 3;
    const size_t row = ((size_t)w * bpp + 3) & ~(size_t)3;
    if ((size_t)off + row * h > f.size()) return fail(err, "truncated BMP", 2);
    std::vector<uint8_t> px((size_t)w * h * ch);
    bool any_alpha = false;
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t src = hi < 0 ? y : h - 1 - y;  // bottom-up unless the height is negative
        const uint8_t *r = &f[off + src * row];
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t *d = &px[((size_t)y * w + x) * ch];
            d[0] = r[x * bpp + 2];
            d[1] = r[x * bpp + 1];
            d[2] = r[x * bpp];
            if (ch == 4) {
                d[3] = r[x * bpp + 3];
                any_alpha = any_alpha || d[3] != 0;
            }
        }
    }
    // a 32-bit file whose alpha bytes are all zero is opaque
    if (ch == 4 && !any_alpha)
        for (size_t i = 0; i < (size_t)w * h; ++i) px[4 * i + 3] = 255;
    to_float(px, w, h, ch, out);
    return 0;
}

int load_ppm(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    size_t pos = 2;
    long v[3];
    for (int k = 0; k < 3; ++k) {
        while (pos < f.size() && (std::isspace(f[pos]) || f[pos] == '#')) {
            if (f[pos] == '#')
                while (pos < f.size() && f[pos] != '\n') ++pos;
            else
                ++pos;
        }
        v[k] = 0;
        if (pos >= f.size() || !std::isdigit(f[pos])) return fail(err, "bad PPM header", 2);
        while (pos < f.size() && std::isdigit(f[pos])) v[k] = v[k] * 10 + (f[pos++] - '0');
    }
    ++pos;  // the single whitespace after maxval
    if (v[2] != 255 || v[0] <= 0 || v[1] <= 0 || v[0] > 32768 || v[1] > 32768)
        return fail(err, "PPM variant not decoded (maxval 255)", 2);
    const uint32_t w = (uint32_t)v[0], h = (uint32_t)v[1];
    if (pos + (size_t)w * h * 3 > f.size()) return fail(err, "truncated PPM", 2);
    std::vector<uint8_t> px(f.begin() + pos, f.begin() + pos + (size_t)w * h * 3);
    to_float(px, w, h, 3, out);
    return 0;
}

// Radiance RGBE read as stbi_load returns it (stb_image 2.x): the header's
// FORMAT=32-bit_rle_rgbe and "-Y h +X w" layout; scanlines flat for widths
// outside [8, 32767], else run-length per component (a scanline without the
// 2 2 header turns the rest of the image into flat pixels, from the second
// pixel on, as stb does); RGBE -> float by ldexp(1, e - 136), then to 8 bits
// by powf(v, 1 / 2.2f) * 255 + 0.5, clamped to [0, 255] and truncated.
// Bytes past the end read as 0.
int load_hdr(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    size_t pos = 0;
    auto get8 = [&]() -> uint8_t { return pos < f.size() ? f[pos++] : 0; };
    auto token = [&]() {
        std::string t;
        while (pos < f.size()) {
            const char c = (char)f[pos++];
            if (c == '\n') break;
            t.push_back(c);
        }
        return t;
    };
    const std::string sig = token();
    if (sig != "#?RADIANCE" && sig != "#?RGBE") return fail(err, "not a Radiance HDR file", 2);
    bool valid = false;
    for (;;) {
        const std::string t = token();
        if (t.empty()) break;
        if (t == "FORMAT=32-bit_rle_rgbe") valid = true;
    }
    if (!valid) return fail(err, "HDR format not decoded", 2);
    const std::string res = token();
    if (res.compare(0, 3, "-Y ")) return fail(err, "HDR layout not decoded", 2);
    char *e = nullptr;
    const long h = std::strtol(res.c_str() + 3, &e, 10);
    while (*e == ' ') ++e;
    if (std::strncmp(e, "+X ", 3)) return fail(err, "HDR layout not decoded", 2);
    const long w = std::strtol(e + 3, nullptr, 10);
    if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return fail(err, "bad HDR size", 2);
    std::vector<float> hdr((size_t)w * h * 3);
    auto convert = [](float *o, const uint8_t *in) {
        if (in[3]) {
            const float f1 = std::ldexp(1.0f, in[3] - (int)(128 + 8));
            o[0] = in[0] * f1;
            o[1] = in[1] * f1;
            o[2] = in[2] * f1;
        } else {
            o[0] = o[1] = o[2] = 0.f;
        }
    };
    auto flat_from = [&](size_t first) {
        for (size_t p = first; p < (size_t)w * h; ++p) {
            uint8_t rgbe[4];
            for (uint8_t &b : rgbe) b = get8();
            convert(&hdr[3 * p], rgbe);
        }
    };
    if (w < 8 || w >= 32768) {
        flat_from(0);
    } else {
        std::vector<uint8_t> line((size_t)w * 4);
        for (long j = 0; j < h; ++j) {
            const int c1 = get8(), c2 = get8();
            int len = get8();
            if (c1 != 2 || c2 != 2 || (len & 0x80)) {
                const uint8_t rgbe[4] = {(uint8_t)c1, (uint8_t)c2, (uint8_t)len, get8()};
                convert(&hdr[0], rgbe);
                flat_from(1);
                break;
            }
            len = (len << 8) | get8();
            if (len != w) return fail(err, "bad HDR scanline length", 2);
            for (int k = 0; k < 4; ++k) {
                long i = 0;
                while (i < w) {
                    int count = get8();
                    if (count > 128) {
                        const uint8_t value = get8();
                        count -= 128;
                        if (count > w - i) return fail(err, "bad HDR run", 2);
                        for (int z = 0; z < count; ++z) line[4 * i++ + k] = value;
                    } else {
                        if (count > w - i) return fail(err, "bad HDR run", 2);
                        for (int z = 0; z < count; ++z) line[4 * i++ + k] = get8();
                    }
                }
            }
            for (long i = 0; i < w; ++i) convert(&hdr[3 * ((size_t)j * w + i)], &line[4 * i]);
        }
    }
    std::vector<uint8_t> px(hdr.size());
    const float gamma = 1.0f / 2.2f;
    for (size_t i = 0; i < hdr.size(); ++i) {
        float z = std::pow(hdr[i] * 1.0f, gamma) * 255 + 0.5f;
        if (z < 0) z = 0;
        if (z > 255) z = 255;
        px[i] = (uint8_t)(int)z;
    }
    to_float(px, (uint32_t)w, (uint32_t)h, 3, out);
    return 0;
}

}  // namespace

int load_image(const std::string &path, Image &out, std::string *err) {
    std::ifstream in(path, std::ios::binary);
    if (!in.good()) return fail(err, "cannot open " + path, 1);
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    static const uint8_t png_sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    int rc;
    if (f.size() >= 8 && !std::memcmp(f.data(), png_sig, 8)) rc = load_png(f, out, err);
    else if (f.size() >= 2 && f[0] == 'B' && f[1] == 'M') rc = load_bmp(f, out, err);
    else if (f.size() >= 2 && f[0] == 'P' && f[1] == '6') rc = load_ppm(f, out, err);
    else if (f.size() >= 2 && f[0] == '#' && f[1] == '?') rc = load_hdr(f, out, err);
    else if (path.size() > 4 && (path.substr(path.size() - 4) == ".tga" || path.substr(path.size() - 4) == ".TGA"))
        rc = load_tga(f, out, err);
    else rc = fail(err, "image format not decoded: " + path, 2);
    // the device's texel addressing takes sides below 2^16
    if (!rc && (out.width > 65535 || out.height > 65535)) rc = fail(err, "image larger than 65535 pixels a side", 2);
    if (rc && err && err->find(path) == std::string::npos) *err += " (" + path + ")";
    return rc;
}

}  // namespace xt
