"""Image export (xt_image_write) against the reference's writers
(nimg::io::save::{png,bmp,tga,hdr,exr}, lib/nimg/img.cc:72-318).

The fixture tests/golden/ref_images.npz holds a framebuffer and the bytes of
the files the reference wrote from it. BMP is compared byte for byte (the
format has one encoding); PNG, TGA, HDR and EXR leave the encoder choices
(deflate blocks, run-length packets, header attributes) free, so their files
are decoded and compared pixel for pixel, bit-exact: 8-bit values for PNG and
TGA, RGBE bytes for HDR, float bits for EXR. The decoders below are written
from the format specifications; PIL decodes PNG, BMP and TGA as a second
reader. Host-only: no GPU is needed."""
import io
import os
import sys

import numpy as np
import pytest

import xtracer_amd as xa
from xtracer_amd._native import XtError

XT_E_INVALID = 1

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle import oracle as O  # noqa: E402

GOLD = np.load(os.path.join(HERE, "golden", "ref_images.npz"))


def pixmap(fb):
    """The framebuffer as the reference's Pixmap holds it (color.cc:35-40)."""
    p = np.array(fb, np.float32)
    p[..., :3] = np.where(p[..., :3] < 0, np.float32(0), p[..., :3])
    p[..., 3] = np.clip(p[..., 3], 0, 1)
    return p


def to8(fb):
    v = pixmap(fb) * np.float32(255)
    return np.minimum(v, np.float32(255)).astype(np.uint8)  # truncation, as img.cc:152-163


def read_png(b):
    import zlib
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    while pos < len(b):
        n = int.from_bytes(b[pos:pos + 4], "big")
        typ = b[pos + 4:pos + 8]
        data = b[pos + 8:pos + 8 + n]
        crc = int.from_bytes(b[pos + 8 + n:pos + 12 + n], "big")
        assert zlib.crc32(typ + data) == crc, typ
        if typ == b"IHDR":
            w, h = int.from_bytes(data[0:4], "big"), int.from_bytes(data[4:8], "big")
            assert data[8:13] == bytes([8, 6, 0, 0, 0])
        elif typ == b"IDAT":
            idat += data
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = 4 * w
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - 4] if i >= 4 else 0
            c = prev[i - 4] if i >= 4 else 0
            up = prev[i]
            if f == 0:
                pred = 0
            elif f == 1:
                pred = a
            elif f == 2:
                pred = up
            elif f == 3:
                pred = (a + up) // 2
            else:
                p = a + up - c
                pa, pb, pc = abs(p - a), abs(p - up), abs(p - c)
                pred = a if pa <= pb and pa <= pc else (up if pb <= pc else c)
            cur[i] = (line[i] + pred) & 255
        out[y], prev = cur, cur
    return out.reshape(h, w, 4).astype(np.uint8)


def read_hdr(b):
    """Radiance RGBE: flat or new-style run-length scanlines."""
    assert b.startswith(b"#?RADIANCE")
    pos = b.index(b"\n\n") + 2
    end = b.index(b"\n", pos)
    f = b[pos:end].split()
    assert f[0] == b"-Y" and f[2] == b"+X"
    h, w = int(f[1]), int(f[3])
    pos = end + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and b[pos] == 2 and b[pos + 1] == 2 and (b[pos + 2] << 8 | b[pos + 3]) == w:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = b[pos]
                    pos += 1
                    if n > 128:
                        n -= 128
                        out[y, x:x + n, c] = b[pos]
                        pos += 1
                    else:
                        assert n > 0
                        out[y, x:x + n, c] = np.frombuffer(b, np.uint8, n, pos)
                        pos += n
                    x += n
                assert x == w
        else:
            out[y] = np.frombuffer(b, np.uint8, 4 * w, pos).reshape(w, 4)
            pos += 4 * w
    assert pos == len(b)
    return out


def read_exr(b):
    """OpenEXR single-part scanline file, no compression, FLOAT channels:
    returns {channel name: (H, W) float32}."""
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01])
    pos, attrs = 8, {}
    while b[pos] != 0:
        e = b.index(b"\0", pos)
        name = b[pos:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        size = int.from_bytes(b[e2 + 1:e2 + 5], "little")
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    assert attrs["compression"][1][0] == 0
    chans, cb, q = [], attrs["channels"][1], 0
    while cb[q] != 0:
        e = cb.index(b"\0", q)
        chans.append(cb[q:e].decode())
        assert int.from_bytes(cb[e + 1:e + 5], "little") == 2  # FLOAT
        q = e + 17
    x0, y0, x1, y1 = np.frombuffer(attrs["dataWindow"][1], "<i4")
    w, h = int(x1 - x0 + 1), int(y1 - y0 + 1)
    offsets = np.frombuffer(b, "<u8", h, pos)
    planes = np.full((len(chans), h, w), np.nan, np.float32)
    for o in offsets:
        o = int(o)
        y = int.from_bytes(b[o:o + 4], "little", signed=True) - int(y0)
        n = int.from_bytes(b[o + 4:o + 8], "little")
        assert n == 4 * w * len(chans)
        planes[:, y, :] = np.frombuffer(b, "<f4", w * len(chans), o + 8).reshape(len(chans), w)
    return {c: planes[i] for i, c in enumerate(chans)}


def pil_rgba(b):
    from PIL import Image
    return np.asarray(Image.open(io.BytesIO(b)).convert("RGBA"))


def written(tmp_path, fb, ext):
    p = tmp_path / ("img." + ext)
    xa.write_image(str(p), fb)
    return p.read_bytes()


def check_same(fmt, ours, ref, fb):
    if fmt == "bmp":
        assert ours == ref, "BMP bytes differ from the reference writer's"
    elif fmt in ("png", "tga"):
        if fmt == "png":
            assert np.array_equal(read_png(ours), read_png(ref))
            assert np.array_equal(read_png(ours), to8(fb))
        assert np.array_equal(pil_rgba(ours), pil_rgba(ref))
    elif fmt == "hdr":
        assert np.array_equal(read_hdr(ours), read_hdr(ref))
    elif fmt == "exr":
        a, r = read_exr(ours), read_exr(ref)
        assert list(a) == list(r) == ["B", "G", "R"]
        for c in a:
            assert np.array_equal(a[c].view(np.uint32), r[c].view(np.uint32)), c


@pytest.mark.parametrize("fmt", ["png", "bmp", "tga", "hdr", "exr"])
def test_matches_reference_writer_fixture(tmp_path, fmt):
    fb = GOLD["rgba"]
    check_same(fmt, written(tmp_path, fb, fmt), GOLD[fmt].tobytes(), fb)


def test_fixture_decodes_to_the_reference_conversion():
    """The fixture itself, read by the decoders here, is what the reference's
    conversions give: 8-bit truncation for PNG, the Pixmap clamp plus RGB
    channels for EXR, Ward's shared-exponent RGBE for HDR."""
    fb = GOLD["rgba"]
    pm = pixmap(fb)
    assert np.array_equal(read_png(GOLD["png"].tobytes()), to8(fb))
    e = read_exr(GOLD["exr"].tobytes())
    for i, c in enumerate("RGB"):
        assert np.array_equal(e[c], pm[..., i])
    rgbe = read_hdr(GOLD["hdr"].tobytes()).astype(np.float64)
    m = pm[..., :3].max(-1)
    lit = m >= 1e-32
    scale = np.ldexp(1.0, rgbe[..., 3].astype(np.int32) - 136)
    back = (rgbe[..., :3] + 0.0) * scale[..., None]
    # RGBE keeps each component to within one step of its shared exponent
    step = np.ldexp(1.0, rgbe[..., 3].astype(np.int32) - 136)
    assert np.all(np.abs(back[lit] - pm[..., :3][lit]) <= 1.5 * step[lit][:, None])
    assert np.all(rgbe[~lit] == 0)
    bmp = pil_rgba(GOLD["bmp"].tobytes())[..., :3].astype(np.int32)
    d = to8(fb).astype(np.int32)
    bg = np.array([255, 0, 255])
    num = (d[..., :3] - bg) * d[..., 3:4]
    assert np.array_equal(bmp, bg + np.sign(num) * (np.abs(num) // 255))  # C's truncating division


def test_pfm_and_ppm(tmp_path):
    fb = GOLD["rgba"]
    h, w = fb.shape[:2]
    b = written(tmp_path, fb, "pfm")
    head = b"PF\n%d %d\n-1.0\n" % (w, h)
    assert b.startswith(head)
    img = np.frombuffer(b, "<f4", offset=len(head)).reshape(h, w, 3)[::-1]
    assert np.array_equal(img, pixmap(fb)[..., :3])
    b = written(tmp_path, fb, "ppm")
    head = b"P6\n%d %d\n255\n" % (w, h)
    assert b.startswith(head)
    assert np.array_equal(np.frombuffer(b, np.uint8, offset=len(head)).reshape(h, w, 3), to8(fb)[..., :3])


def test_errors(tmp_path):
    fb = np.zeros((2, 3, 4), np.float32)
    with pytest.raises(XtError) as e:
        xa.write_image(str(tmp_path / "img.xyz"), fb)
    assert e.value.code == XT_E_INVALID
    with pytest.raises(XtError):
        xa.write_image(str(tmp_path / "no_such_dir" / "img.png"), fb)
    with pytest.raises(ValueError):
        xa.write_image(str(tmp_path / "img.png"), np.zeros((2, 3, 3), np.float32))


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker (oracle/_ref) not built")


@needs_ref
@pytest.mark.parametrize("shape", [(5, 3), (200, 300), (70, 40000)])
@pytest.mark.parametrize("fmt", ["png", "bmp", "tga", "hdr", "exr"])
def test_matches_reference_writer_live(tmp_path, fmt, shape):
    """Random framebuffers through both writers: a width below the RGBE
    run-length minimum, a PNG larger than one stored deflate block, and a
    width above the RGBE run-length maximum."""
    h, w = shape
    if fmt == "exr" and w > 1000 or fmt in ("png", "tga") and w > 1000:
        pytest.skip("pure-Python decoders at this size take too long")
    rng = np.random.default_rng(h * w)
    fb = rng.uniform(-0.3, 2.0, (h, w, 4)).astype(np.float32)
    fb[:, : w // 3] = fb[:, :1]  # runs
    p = tmp_path / ("ref." + fmt)
    O.ref_save_image(fmt, p, fb)
    check_same(fmt, written(tmp_path, fb, fmt), p.read_bytes(), fb)
