"""Writers of small texture image files with known bytes, for the texture
decoder tests: PNG with every row filter, TGA raw / run-length in both row
orders, BMP 24 / 32 bit bottom-up and top-down, binary PPM. Each writer takes
an (H, W, C) uint8 array whose row 0 is the image's top row."""
import struct
import zlib

import numpy as np


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def _filtered(rows, bpp, filters, start=0):
    """Filter bytes rows (list of uint8 arrays) with filters[(start + y) % len]."""
    raw = bytearray()
    prev = np.zeros(len(rows[0]), np.int64) if rows else None
    for y, r in enumerate(rows):
        row = r.astype(np.int64)
        a = np.concatenate([np.zeros(bpp, np.int64), row[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        f = filters[(start + y) % len(filters)]
        pred = [np.zeros_like(row), a, prev, (a + prev) >> 1, _paeth(a, prev, c)][f]
        raw.append(f)
        raw += ((row - pred) % 256).astype(np.uint8).tobytes()
        prev = row
    return raw


def _pack(samples, depth):
    """One row of samples (uint values) as bytes at `depth` bits each."""
    if depth == 16:
        return np.asarray(samples, ">u2").view(np.uint8)
    if depth == 8:
        return np.asarray(samples, np.uint8)
    bits = np.unpackbits(np.asarray(samples, np.uint8)[:, None], axis=1)[:, 8 - depth:].reshape(-1)
    return np.packbits(bits)


def png(path, px, filters=(0, 1, 2, 3, 4), bit_depth=8, interlace=0, palette=None, trns=None):
    """px: (H, W, C) samples. C = 3 RGB, 4 RGBA, 1 grey, or palette indices
    with `palette` ((n, 3) uint8); bit_depth 1/2/4 (palette), 8 or 16
    (samples up to 65535); Adam7 when interlace=1; `trns`: raw tRNS bytes."""
    h, w, ch = px.shape
    ctype = 3 if palette is not None else {1: 0, 3: 2, 4: 6}[ch]
    bpp = max(1, ch * bit_depth // 8)

    def rows_of(img):
        return [_pack(img[y].reshape(-1), bit_depth) for y in range(img.shape[0])]
    if not interlace:
        raw = _filtered(rows_of(px), bpp, filters)
    else:
        raw = bytearray()
        for p, (xo, yo, xs, ys) in enumerate(zip((0, 4, 0, 2, 0, 1, 0), (0, 0, 4, 0, 2, 0, 1), (8, 8, 4, 4, 2, 2, 1),
                                                  (8, 8, 8, 4, 4, 2, 2))):
            sub = px[yo::ys, xo::xs]
            if sub.shape[0] and sub.shape[1]:
                raw += _filtered(rows_of(sub), bpp, filters, start=p)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bit_depth, ctype, 0, 0, interlace))
    if palette is not None:
        data += chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        data += chunk(b"tRNS", bytes(trns))
    data += chunk(b"IDAT", zlib.compress(bytes(raw), 9)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)
    return str(path)


def tga(path, px, rle=False, top_down=False):
    h, w, ch = px.shape
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, 10 if rle else 2, 0, 0, 0, 0, 0, w, h, 8 * ch, 32 if top_down else 0)
    rows = px if top_down else px[::-1]
    bgr = rows[..., [2, 1, 0] + ([3] if ch == 4 else [])].reshape(-1, ch)
    body = bytearray()
    if not rle:
        body += bgr.tobytes()
    else:  # run-length packets of up to 128 equal pixels, literal packets otherwise
        i, n = 0, len(bgr)
        while i < n:
            j = i + 1
            while j < n and j - i < 128 and np.array_equal(bgr[j], bgr[i]):
                j += 1
            if j - i > 1:
                body.append(128 | (j - i - 1))
                body += bgr[i].tobytes()
                i = j
                continue
            j = i + 1
            while j < n and j - i < 128 and not np.array_equal(bgr[j], bgr[j - 1]):
                j += 1
            body.append(j - i - 1)
            body += bgr[i:j].tobytes()
            i = j
    with open(path, "wb") as f:
        f.write(hdr + bytes(body))
    return str(path)


def bmp(path, px, top_down=False):
    """BI_RGB, 24 bit for C=3, 32 bit for C=4 (alpha in the fourth byte)."""
    h, w, ch = px.shape
    bpp = ch
    stride = (w * bpp + 3) & ~3
    rows = px if top_down else px[::-1]
    body = bytearray()
    for r in rows:
        line = r[:, [2, 1, 0] + ([3] if ch == 4 else [])].reshape(-1).tobytes()
        body += line + b"\0" * (stride - len(line))
    info = struct.pack("<IiiHHIIiiII", 40, w, -h if top_down else h, 1, 8 * bpp, 0, len(body), 2835, 2835, 0, 0)
    head = struct.pack("<2sIHHI", b"BM", 14 + 40 + len(body), 0, 0, 14 + 40)
    with open(path, "wb") as f:
        f.write(head + info + bytes(body))
    return str(path)


def ppm(path, px):
    h, w, _ = px.shape
    with open(path, "wb") as f:
        f.write(b"P6\n# test image\n%d %d\n255\n" % (w, h) + px.astype(np.uint8).tobytes())
    return str(path)


def expected(px):
    """What the loader must return: byte / 255 as float32, alpha 1 without an alpha channel."""
    h, w, ch = px.shape
    out = np.ones((h, w, 4), np.float32)
    out[..., :ch] = px.astype(np.float32) / np.float32(255.0)
    return out
