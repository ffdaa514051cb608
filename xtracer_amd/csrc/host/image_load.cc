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

// JPEG (baseline SOF0 / SOF1 in one scan, or progressive SOF2 over several;
// Huffman, 8-bit) as stb_image 2.14 decodes it for stbi_load: coefficients times the quantiser kept as 16-bit
// values, the integer IDCT of the IJG "islow" method with stb's 12-bit
// constants and roundings, chroma upsampled by stb's row filters (triangle
// 2x horizontal / vertical / both, replication otherwise), YCbCr -> RGB in
// stb's reduced-precision fixed point. Three components give RGB; one gives
// grey, which the reference refuses. Arithmetic-coded and lossless files are
// not decoded.
namespace jpg {

const uint8_t dezigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                              12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                              35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                              58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    // canonical code lengths -> (code, length) -> symbol
    std::vector<uint16_t> code;
    std::vector<uint8_t> len, sym;
    bool ok = false;
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int x = 0, y = 0, w2 = 0, h2 = 0, dc = 0;
    std::vector<uint8_t> plane;
    std::vector<int16_t> coef;  // progressive: 64 per block, w2 / 8 blocks a row
};

struct Bits {
    const std::vector<uint8_t> &f;
    size_t pos;
    uint32_t acc = 0;
    int n = 0;
    bool marker = false;  // a marker was met: zeros are fed from here on
    void reset(size_t p) {
        pos = p;
        acc = 0;
        n = 0;
        marker = false;
    }
    int fill() {
        while (n <= 24) {
            int b = 0;
            if (!marker && pos < f.size()) {
                b = f[pos];
                if (b == 0xff) {
                    const int c = pos + 1 < f.size() ? f[pos + 1] : 0;
                    if (c == 0) {
                        pos += 2;
                    } else {
                        marker = true;
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            acc |= (uint32_t)b << (24 - n);
            n += 8;
        }
        return 0;
    }
    int bit() {
        if (n < 1) fill();
        const int r = (int)(acc >> 31);
        acc <<= 1;
        --n;
        return r;
    }
    int get(int k) {
        int r = 0;
        for (int i = 0; i < k; ++i) r = (r << 1) | bit();
        return r;
    }
};

// the symbol of the next Huffman code, or -1
int decode(Bits &b, const Huff &h) {
    int code = 0;
    size_t k = 0;
    for (int l = 1; l <= 16; ++l) {
        code = (code << 1) | b.bit();
        while (k < h.len.size() && h.len[k] == l) {
            if (h.code[k] == code) return h.sym[k];
            ++k;
        }
    }
    return -1;
}

// JPEG's receive + extend: `k` bits as a signed magnitude category value
int extend(Bits &b, int k) {
    if (!k) return 0;
    const int v = b.get(k);
    return v < (1 << (k - 1)) ? v - (1 << k) + 1 : v;
}

inline uint8_t clamp8(int x) { return (unsigned)x > 255 ? (x < 0 ? 0 : 255) : (uint8_t)x; }
inline int f2f(double x) { return (int)(x * 4096 + 0.5); }

// The 8-point inverse DCT of the IJG "islow" method: the factorisation of
// Loeffler, Ligtenberg and Moschytz (1989) with 12-bit fixed-point
// constants. Its rotations and butterflies multiply out to one integer
// matrix per pass, and integer products are exact, so applying the matrix
// gives the factorised sums. The sums are formed in 64 bits and wrapped to
// 32, as the decoder's 32-bit int arithmetic wraps, before the rounding
// shift. Input k of a pass is frequency k; output j is sample j.
struct IdctMatrix {
    int m[8][8];
    IdctMatrix() {
        const int one = 4096, r541 = f2f(0.5411961f), r765 = f2f(0.765366865f), r1848 = f2f(-1.847759065f);
        const int r1176 = f2f(1.175875602f), r299 = f2f(0.298631336f), r2053 = f2f(2.053119869f),
                  r3073 = f2f(3.072711026f), r1501 = f2f(1.501321110f), r900 = f2f(-0.899976223f),
                  r2563 = f2f(-2.562915447f), r1962 = f2f(-1.961570560f), r390 = f2f(-0.390180644f);
        // even half (frequencies 0, 2, 4, 6), one row per output pair
        const int even[4][4] = {{one, r541 + r765, one, r541},
                                {one, r541, -one, r541 + r1848},
                                {one, -r541, -one, -(r541 + r1848)},
                                {one, -(r541 + r765), one, -r541}};
        // odd half (frequencies 1, 3, 5, 7); row i pairs with even row 3 - i
        const int odd[4][4] = {{r1176 + r900, r1176 + r1962, r1176, r299 + r1176 + r900 + r1962},
                               {r1176 + r390, r1176 + r2563, r2053 + r1176 + r2563 + r390, r1176},
                               {r1176, r3073 + r1176 + r2563 + r1962, r1176 + r2563, r1176 + r1962},
                               {r1501 + r1176 + r900 + r390, r1176, r1176 + r390, r1176 + r900}};
        for (int j = 0; j < 4; ++j) {
            for (int k = 0; k < 4; ++k) {
                m[j][2 * k] = m[7 - j][2 * k] = even[j][k];
                m[j][2 * k + 1] = odd[3 - j][k];
                m[7 - j][2 * k + 1] = -odd[3 - j][k];
            }
        }
    }
    // samples of one pass: in[k * step] -> out[j] = (sum + bias) >> shift
    void pass(const int *in, int step, int bias, int shift, int *out) const {
        for (int j = 0; j < 8; ++j) {
            int64_t acc = bias;
            for (int k = 0; k < 8; ++k) acc += (int64_t)m[j][k] * in[k * step];
            out[j] = (int)(int32_t)(uint32_t)(uint64_t)acc >> shift;
        }
    }
};

const IdctMatrix idct_matrix;

// Columns first, keeping 2 extra bits (bias 1 << 9, shift 10); then rows,
// removing those and the 12 fraction bits with rounding and the +128 level
// shift (bias 1 << 16 plus 128 << 17, shift 17), saturated to 8 bits.
void idct(uint8_t *out, int stride, const int16_t d[64]) {
    int col[64], in[8], res[8];
    for (int c = 0; c < 8; ++c) {
        for (int k = 0; k < 8; ++k) in[k] = d[8 * k + c];
        idct_matrix.pass(in, 1, 1 << 9, 10, res);
        for (int j = 0; j < 8; ++j) col[8 * j + c] = res[j];
    }
    for (int r = 0; r < 8; ++r) {
        idct_matrix.pass(col + 8 * r, 1, (1 << 16) + (128 << 17), 17, res);
        for (int j = 0; j < 8; ++j) out[r * stride + j] = clamp8(res[j]);
    }
}

// stb's row upsamplers: `near` is the nearer source row, `far` the other
void up_v2(uint8_t *o, const uint8_t *nr, const uint8_t *fr, int w) {
    for (int i = 0; i < w; ++i) o[i] = (uint8_t)((3 * nr[i] + fr[i] + 2) >> 2);
}
// 2x horizontal: output x lies between source sample x / 2 and its
// neighbour on x's side, weighted 3 : 1 toward the nearer; the first and
// last outputs copy the edge samples. The decoder this matches weights its
// last even output 3 : 1 toward the neighbour instead.
void up_h2(uint8_t *o, const uint8_t *in, int w) {
    for (int x = 0; x < 2 * w; ++x) {
        const int i = x >> 1, nb = (x & 1) ? i + 1 : i - 1;
        if (nb < 0 || nb >= w) o[x] = in[i];
        else if (x == 2 * w - 2) o[x] = (uint8_t)((3 * in[nb] + in[i] + 2) >> 2);
        else o[x] = (uint8_t)((3 * in[i] + in[nb] + 2) >> 2);
    }
}
// 2x both ways: the vertical 3 : 1 sums of the nearer and farther rows,
// then the horizontal 3 : 1 weighting of up_h2 on those sums in one
// rounding (/16), the edge outputs from their own sum alone (/4).
void up_hv2(uint8_t *o, const uint8_t *nr, const uint8_t *fr, int w) {
    auto col = [&](int i) { return 3 * nr[i] + fr[i]; };
    for (int x = 0; x < 2 * w; ++x) {
        const int i = x >> 1, nb = (x & 1) ? i + 1 : i - 1;
        o[x] = (nb < 0 || nb >= w) ? (uint8_t)((col(i) + 2) >> 2) : (uint8_t)((3 * col(i) + col(nb) + 8) >> 4);
    }
}

}  // namespace jpg

int load_jpeg(const std::vector<uint8_t> &f, Image &out, std::string *err) {
    using namespace jpg;
    int16_t q[4][64] = {};
    Huff hd[4], ha[4];
    std::vector<Comp> comps;
    uint32_t W = 0, H = 0;
    int restart = 0, hmax = 1, vmax = 1;
    size_t pos = 2;
    auto u16 = [&](size_t p) { return p + 1 < f.size() ? (int)(f[p] << 8 | f[p + 1]) : 0; };
    bool frame = false, progressive = false, laid_out = false;
    for (;;) {
        while (pos < f.size() && f[pos] != 0xff) ++pos;
        while (pos < f.size() && f[pos] == 0xff) ++pos;
        if (pos >= f.size()) return fail(err, "JPEG without image data", 2);
        const int m = f[pos++];
        if (m == 0xd8 || (m >= 0xd0 && m <= 0xd7) || m == 0x01) continue;
        if (m == 0xd9) {
            if (progressive && laid_out) break;  // the scans are all in
            return fail(err, "JPEG without scan", 2);
        }
        const int len = u16(pos);
        if (len < 2 || pos + len > f.size()) return fail(err, "truncated JPEG", 2);
        const uint8_t *d = &f[pos + 2];
        const int n = len - 2;
        if (m == 0xdb) {  // quantisation tables, stored in natural order
            for (int k = 0; k < n;) {
                const int pq = d[k] >> 4, tq = d[k] & 15;
                ++k;
                if (tq > 3 || k + 64 * (pq + 1) > n) return fail(err, "bad JPEG DQT", 2);
                for (int i = 0; i < 64; ++i, k += pq + 1)
                    q[tq][dezigzag[i]] = (int16_t)(pq ? (d[k] << 8 | d[k + 1]) : d[k]);
            }
        } else if (m == 0xc4) {  // Huffman tables
            for (int k = 0; k < n;) {
                const int tc = d[k] >> 4, th = d[k] & 15;
                if (tc > 1 || th > 3 || k + 17 > n) return fail(err, "bad JPEG DHT", 2);
                Huff h;
                int code = 0, total = 0;
                for (int l = 1; l <= 16; ++l) total += d[k + l];
                if (k + 17 + total > n) return fail(err, "bad JPEG DHT", 2);
                int at = k + 17;
                for (int l = 1; l <= 16; ++l) {
                    for (int i = 0; i < d[k + l]; ++i) {
                        h.code.push_back((uint16_t)code++);
                        h.len.push_back((uint8_t)l);
                        h.sym.push_back(d[at++]);
                    }
                    code <<= 1;
                }
                h.ok = true;
                (tc ? ha : hd)[th] = h;
                k = at;
            }
        } else if (m == 0xdd) {
            restart = u16(pos + 2);
        } else if (m == 0xc0 || m == 0xc1 || m == 0xc2) {
            // one frame per file: a second header would add components
            // without coefficient stores after the first scan's layout
            if (frame) return fail(err, "bad JPEG: second frame header", 2);
            progressive = m == 0xc2;
            if (n < 6 || d[0] != 8) return fail(err, "JPEG precision not decoded", 2);
            H = (uint32_t)(d[1] << 8 | d[2]);
            W = (uint32_t)(d[3] << 8 | d[4]);
            const int nc = d[5];
            if (!W || !H || W > 65535 || H > 65535 || (nc != 1 && nc != 3) || n < 6 + 3 * nc)
                return fail(err, "JPEG frame not decoded", 2);
            for (int c = 0; c < nc; ++c) {
                Comp k;
                k.id = d[6 + 3 * c];
                k.h = d[7 + 3 * c] >> 4;
                k.v = d[7 + 3 * c] & 15;
                k.tq = d[8 + 3 * c];
                if (k.h < 1 || k.h > 4 || k.v < 1 || k.v > 4 || k.tq > 3) return fail(err, "bad JPEG frame", 2);
                hmax = std::max(hmax, k.h);
                vmax = std::max(vmax, k.v);
                comps.push_back(k);
            }
            frame = true;
        } else if (m >= 0xc3 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc) {
            return fail(err, "arithmetic / lossless JPEG not decoded", 2);
        } else if (m == 0xda) {
            if (!frame) return fail(err, "JPEG scan before frame", 2);
            pos += len;
            const int ns = d[0];
            if (!progressive && ns != (int)comps.size()) return fail(err, "JPEG with several scans not decoded", 2);
            if (ns < 1 || ns > (int)comps.size() || n < 4 + 2 * ns) return fail(err, "bad JPEG scan", 2);
            const int ss = d[1 + 2 * ns], se = d[2 + 2 * ns], ah = d[3 + 2 * ns] >> 4, al = d[3 + 2 * ns] & 15;
            // the spectral range indexes the 64-entry zigzag table and the
            // approximation bits shift 16-bit coefficients (stb rejects the
            // same headers as "bad SOS")
            if (progressive) {
                if (ss > 63 || se > 63 || ss > se || ah > 13 || al > 13) return fail(err, "bad JPEG SOS", 2);
            } else if (ss != 0 || ah != 0 || al != 0) {
                return fail(err, "bad JPEG SOS", 2);
            }
            std::vector<int> order;
            for (int s = 0; s < ns; ++s) {
                const int id = d[1 + 2 * s];
                int c = -1;
                for (size_t k = 0; k < comps.size(); ++k)
                    if (comps[k].id == id) c = (int)k;
                if (c < 0) return fail(err, "bad JPEG scan", 2);
                comps[c].td = d[2 + 2 * s] >> 4;
                comps[c].ta = d[2 + 2 * s] & 15;
                if (comps[c].td > 3 || comps[c].ta > 3) return fail(err, "bad JPEG scan", 2);
                const bool need_dc = !progressive || (ss == 0 && ah == 0), need_ac = !progressive || ss > 0;
                if ((need_dc && !hd[comps[c].td].ok) || (need_ac && !ha[comps[c].ta].ok))
                    return fail(err, "JPEG Huffman table missing", 2);
                order.push_back(c);
            }
            // frame layout (MCU-padded planes)
            const int mcuw = hmax * 8, mcuh = vmax * 8;
            const int mx = ((int)W + mcuw - 1) / mcuw, my = ((int)H + mcuh - 1) / mcuh;
            if (!laid_out)
                for (Comp &c : comps) {
                    c.x = ((int)W * c.h + hmax - 1) / hmax;
                    c.y = ((int)H * c.v + vmax - 1) / vmax;
                    c.w2 = mx * c.h * 8;
                    c.h2 = my * c.v * 8;
                    c.plane.assign((size_t)c.w2 * c.h2, 0);
                    if (progressive) c.coef.assign((size_t)c.w2 * c.h2, 0);
                }
            laid_out = true;
            for (Comp &c : comps) c.dc = 0;
            int eob_run = 0;
            Bits b{f, pos};
            // progressive blocks (stb's decode_block_prog_dc / _ac): values
            // kept unquantised in the coefficient store until the end
            auto prog_block = [&](Comp &c, int bx, int by) -> bool {
                int16_t *data = &c.coef[((size_t)by * (c.w2 / 8) + bx) * 64];
                if (ss == 0) {
                    if (se != 0) return false;
                    if (ah == 0) {
                        for (int k = 1; k < 64; ++k) data[k] = 0;
                        const int t = decode(b, hd[c.td]);
                        if (t < 0 || t > 16) return false;
                        c.dc += t ? extend(b, t) : 0;
                        data[0] = (int16_t)(c.dc * (1 << al));
                    } else if (b.bit()) {
                        data[0] = (int16_t)(data[0] + (1 << al));
                    }
                    return true;
                }
                if (ah == 0) {
                    if (eob_run) {
                        --eob_run;
                        return true;
                    }
                    int k = ss;
                    do {
                        const int rs = decode(b, ha[c.ta]);
                        if (rs < 0) return false;
                        const int s = rs & 15, r = rs >> 4;
                        if (s == 0) {
                            if (r < 15) {
                                eob_run = (1 << r) + (r ? b.get(r) : 0) - 1;
                                break;
                            }
                            k += 16;
                        } else {
                            k += r;
                            if (k > 63) return false;
                            data[dezigzag[k++]] = (int16_t)(extend(b, s) * (1 << al));
                        }
                    } while (k <= se);
                    return true;
                }
                const int16_t bit = (int16_t)(1 << al);
                auto refine = [&](int16_t *p) {
                    if (b.bit() && (*p & bit) == 0) *p = (int16_t)(*p > 0 ? *p + bit : *p - bit);
                };
                if (eob_run) {
                    --eob_run;
                    for (int k = ss; k <= se; ++k)
                        if (data[dezigzag[k]] != 0) refine(&data[dezigzag[k]]);
                    return true;
                }
                int k = ss;
                do {
                    const int rs = decode(b, ha[c.ta]);
                    if (rs < 0) return false;
                    int s = rs & 15, r = rs >> 4;
                    if (s == 0) {
                        if (r < 15) {
                            eob_run = (1 << r) - 1 + (r ? b.get(r) : 0);
                            r = 64;
                        }
                    } else {
                        if (s != 1) return false;
                        s = b.bit() ? bit : -bit;
                    }
                    while (k <= se) {
                        int16_t *p = &data[dezigzag[k++]];
                        if (*p != 0) {
                            refine(p);
                        } else {
                            if (r == 0) {
                                *p = (int16_t)s;
                                break;
                            }
                            --r;
                        }
                    }
                } while (k <= se);
                return true;
            };
            int todo = restart ? restart : 1 << 30;
            auto block = [&](Comp &c, int bx, int by) -> bool {
                int16_t data[64] = {};
                const int t = decode(b, hd[c.td]);
                if (t < 0 || t > 16) return false;
                const int diff = t ? extend(b, t) : 0;
                c.dc += diff;
                data[0] = (int16_t)(c.dc * q[c.tq][0]);
                for (int k = 1; k < 64;) {
                    const int rs = decode(b, ha[c.ta]);
                    if (rs < 0) return false;
                    const int r = rs >> 4, s = rs & 15;
                    if (!s) {
                        if (r != 15) break;  // end of block
                        k += 16;
                        continue;
                    }
                    k += r;
                    if (k > 63) return false;
                    const int z = dezigzag[k++];
                    data[z] = (int16_t)(extend(b, s) * q[c.tq][z]);
                }
                idct(&c.plane[(size_t)by * 8 * c.w2 + (size_t)bx * 8], c.w2, data);
                return true;
            };
            // after each restart interval: the next marker must be RSTn, which
            // resets the bit reader and the DC predictions; any other marker
            // ends the scan there, as in stb
            auto restart_marker = [&]() -> bool {
                if (--todo > 0) return true;
                size_t p = b.pos;
                while (p + 1 < f.size() && !(f[p] == 0xff && f[p + 1] != 0 && f[p + 1] != 0xff)) ++p;
                if (p + 1 >= f.size() || f[p + 1] < 0xd0 || f[p + 1] > 0xd7) return false;
                b.reset(p + 2);
                for (Comp &c : comps) c.dc = 0;
                eob_run = 0;
                todo = restart;
                return true;
            };
            if (ns == 1) {
                Comp &c = comps[order[0]];
                const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
                for (int by = 0; by < bh; ++by)
                    for (int bx = 0; bx < bw; ++bx) {
                        if (!(progressive ? prog_block(c, bx, by) : block(c, bx, by)))
                            return fail(err, "corrupt JPEG data", 2);
                        if (!restart_marker()) by = bh, bx = bw;
                    }
            } else {
                for (int j = 0; j < my; ++j)
                    for (int i = 0; i < mx; ++i) {
                        for (int c : order)
                            for (int y = 0; y < comps[c].v; ++y)
                                for (int x = 0; x < comps[c].h; ++x)
                                    if (!(progressive ? prog_block(comps[c], i * comps[c].h + x, j * comps[c].v + y)
                                                      : block(comps[c], i * comps[c].h + x, j * comps[c].v + y)))
                                        return fail(err, "corrupt JPEG data", 2);
                        if (!restart_marker()) j = my, i = mx;
                    }
            }
            if (!progressive) break;
            pos = b.pos;  // the next marker follows the scan's data
            continue;
        }
        pos += len;
    }
    if (comps.size() != 3) return fail(err, "grey JPEG: 1 channel, which the reference refuses", 2);
    if (progressive)  // stb's jpeg_finish: dequantise (16-bit products) and transform every block
        for (Comp &c : comps)
            for (int by = 0; by < (c.y + 7) >> 3; ++by)
                for (int bx = 0; bx < (c.x + 7) >> 3; ++bx) {
                    int16_t *data = &c.coef[((size_t)by * (c.w2 / 8) + bx) * 64];
                    for (int k = 0; k < 64; ++k) data[k] = (int16_t)(data[k] * q[c.tq][k]);
                    idct(&c.plane[(size_t)by * 8 * c.w2 + (size_t)bx * 8], c.w2, data);
                }
    // upsample each component row by row as stb's load_jpeg_image does, then
    // convert to RGB
    struct Res {
        int hs, vs, ystep, ypos, wlo;
        size_t line0, line1;
    };
    Res r[3];
    for (int k = 0; k < 3; ++k) {
        r[k].hs = hmax / comps[k].h;
        r[k].vs = vmax / comps[k].v;
        r[k].ystep = r[k].vs >> 1;
        r[k].wlo = ((int)W + r[k].hs - 1) / r[k].hs;
        r[k].ypos = 0;
        r[k].line0 = r[k].line1 = 0;
    }
    std::vector<uint8_t> px((size_t)W * H * 3), buf[3];
    for (int k = 0; k < 3; ++k) buf[k].assign((size_t)W + 3, 0);
    const uint8_t *row[3];
    for (uint32_t j = 0; j < H; ++j) {
        for (int k = 0; k < 3; ++k) {
            Res &s = r[k];
            const Comp &c = comps[k];
            const bool bot = s.ystep >= (s.vs >> 1);
            const uint8_t *nr = &c.plane[bot ? s.line1 : s.line0], *fr = &c.plane[bot ? s.line0 : s.line1];
            uint8_t *o = buf[k].data();
            if (s.hs == 1 && s.vs == 1) {
                row[k] = nr;
            } else {
                if (s.hs == 1 && s.vs == 2) up_v2(o, nr, fr, s.wlo);
                else if (s.hs == 2 && s.vs == 1) up_h2(o, nr, s.wlo);
                else if (s.hs == 2 && s.vs == 2) up_hv2(o, nr, fr, s.wlo);
                else
                    for (int i = 0; i < s.wlo; ++i)
                        for (int t = 0; t < s.hs; ++t) o[i * s.hs + t] = nr[i];
                row[k] = o;
            }
            if (++s.ystep >= s.vs) {
                s.ystep = 0;
                s.line0 = s.line1;
                if (++s.ypos < c.y) s.line1 += (size_t)c.w2;
            }
        }
        uint8_t *o = &px[(size_t)j * W * 3];
        for (uint32_t i = 0; i < W; ++i) {
            // stb's reduced-precision YCbCr -> RGB (20-bit fixed point)
            const int yf = (row[0][i] << 20) + (1 << 19);
            const int cr = row[2][i] - 128, cb = row[1][i] - 128;
            auto fx = [](float x) { return ((int)(x * 4096.0f + 0.5f)) << 8; };
            int rr = yf + cr * fx(1.40200f);
            int gg = yf + (cr * -fx(0.71414f)) + ((cb * -fx(0.34414f)) & (int)0xffff0000);
            int bb = yf + cb * fx(1.77200f);
            o[3 * i] = clamp8(rr >> 20);
            o[3 * i + 1] = clamp8(gg >> 20);
            o[3 * i + 2] = clamp8(bb >> 20);
        }
    }
    to_float(px, W, H, 3, out);
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
    else if (f.size() >= 3 && f[0] == 0xff && f[1] == 0xd8 && f[2] == 0xff) rc = load_jpeg(f, out, err);
    else if (path.size() > 4 && (path.substr(path.size() - 4) == ".tga" || path.substr(path.size() - 4) == ".TGA"))
        rc = load_tga(f, out, err);
    else rc = fail(err, "image format not decoded: " + path, 2);
    // the device's texel addressing takes sides below 2^16
    if (!rc && (out.width > 65535 || out.height > 65535)) rc = fail(err, "image larger than 65535 pixels a side", 2);
    if (rc && err && err->find(path) == std::string::npos) *err += " (" + path + ")";
    return rc;
}

}  // namespace xt
