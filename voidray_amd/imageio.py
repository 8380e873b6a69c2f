"""Image files for textures and environments.

The reference reads every image with `image::open(path).to_rgb32f()` (image
0.24; core/texture.rs:36-50, voidray_common/src/environments.rs:42-55):
8-bit channels become x / 255, 16-bit ones x / 65535, float ones stay as
they are; grey is copied to R, G and B; alpha is dropped.  Here PNG, TIFF,
JPEG and the other raster formats are decoded by Pillow and OpenEXR by the
reader below; the conversion to RGB f32 is the one above.  8-bit images are
handed to the device as RGBA8 (the device applies x / 255, the same f32
division), everything else as RGB f32.

JPEG decoders differ in their inverse DCT and chroma upsampling, so a JPEG
may decode to values a level or two away from image 0.24's decoder; PNG and
TIFF are lossless and decode identically.
"""
import os
import struct
import zlib

import numpy as np

f32 = np.float32


def _u8_rgba(a):
    """(H, W[, C]) uint8 -> (H, W, 4) uint8 with R, G, B of to_rgb32f and alpha 255."""
    if a.ndim == 2:
        a = a[..., None]
    c = a.shape[2]
    rgb = np.repeat(a[..., :1], 3, axis=2) if c in (1, 2) else a[..., :3]
    out = np.empty(a.shape[:2] + (4,), np.uint8)
    out[..., :3] = rgb
    out[..., 3] = 255
    return out


def _to_rgb32f(a, scale):
    a = np.asarray(a)
    if a.ndim == 2:
        a = a[..., None]
    c = a.shape[2]
    rgb = np.repeat(a[..., :1], 3, axis=2) if c in (1, 2) else a[..., :3]
    rgb = rgb.astype(f32)
    return rgb / f32(scale) if scale != 1 else rgb


def load_image(path, want_float=False):
    """An image file as the reference's to_rgb32f sees it.

    Returns (H, W, 4) uint8 for 8-bit images (unless want_float) or (H, W, 3)
    float32.  Row 0 is the top row of the image, as in the image crate.
    """
    path = os.fspath(path)
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic == b"\x76\x2f\x31\x01":
        return read_exr(path)
    from PIL import Image
    with Image.open(path) as im:
        im.load()
        mode = im.mode
        if mode == "P":
            im = im.convert("RGBA" if "transparency" in im.info else "RGB")
            mode = im.mode
        elif mode == "1":
            im = im.convert("L")
            mode = "L"
        if mode in ("L", "LA", "RGB", "RGBA"):
            a = np.asarray(im, dtype=np.uint8)
            return _to_rgb32f(a, 255.0) if want_float else _u8_rgba(a)
        if mode.startswith("I;16"):
            return _to_rgb32f(np.asarray(im).astype(np.uint16), 65535.0)
        if mode == "F":
            return _to_rgb32f(np.asarray(im, dtype=f32), 1.0)
        if mode in ("RGBX", "CMYK", "YCbCr", "LAB", "HSV"):
            a = np.asarray(im.convert("RGB"), dtype=np.uint8)
            return _to_rgb32f(a, 255.0) if want_float else _u8_rgba(a)
    raise ValueError(f"{path}: unsupported image mode {mode}")


# ---- OpenEXR (scanline images; NONE, RLE, ZIPS and ZIP compression) ----------
# The layout follows the published OpenEXR file format: magic, version, a
# header of (name, type, size, value) attributes ending in an empty name, a
# table of chunk offsets, then chunks of (y, size, data) with each scan line
# holding every channel in name order.  RLE and ZIP data are stored with a
# byte predictor and the bytes split into even and odd halves.

_EXR_NONE, _EXR_RLE, _EXR_ZIPS, _EXR_ZIP = 0, 1, 2, 3
_EXR_LINES = {_EXR_NONE: 1, _EXR_RLE: 1, _EXR_ZIPS: 1, _EXR_ZIP: 16}
_EXR_TYPES = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}


def _exr_header(buf, pos):
    attrs = {}
    while True:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode("latin-1")
        pos = end + 1
        if not name:
            return attrs, pos
        end = buf.index(b"\0", pos)
        typ = buf[pos:end].decode("latin-1")
        pos = end + 1
        (size,) = struct.unpack_from("<i", buf, pos)
        pos += 4
        attrs[name] = (typ, buf[pos:pos + size])
        pos += size


def _exr_channels(raw):
    chans, pos = [], 0
    while raw[pos:pos + 1] != b"\0":
        end = raw.index(b"\0", pos)
        name = raw[pos:end].decode("latin-1")
        ptype, _linear, xs, ys = struct.unpack_from("<iB3xii", raw, end + 1)
        if xs != 1 or ys != 1:
            raise ValueError(f"OpenEXR channel {name}: subsampled channels are not supported")
        chans.append((name, ptype))
        pos = end + 1 + 16
    return chans


def _undo_predictor_and_split(t):
    # the predictor stored d[i] = t[i] - t[i-1] + 128 (mod 256), d[0] = t[0]:
    # t[i] = d[0] + sum_{k=1..i} (d[k] - 128) = sum_{k=0..i} (d[k] - 128) + 128
    t = ((np.cumsum(np.frombuffer(t, np.uint8).astype(np.int64) - 128) + 128) & 0xFF).astype(np.uint8)
    n = len(t)
    half = (n + 1) // 2
    out = np.empty(n, np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(data, expected):
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        c = struct.unpack_from("b", data, i)[0]
        i += 1
        if c < 0:
            out += data[i:i - c]
            i -= c
        else:
            out += bytes([data[i]]) * (c + 1)
            i += 1
    if len(out) != expected:
        raise ValueError("OpenEXR: RLE chunk decodes to the wrong size")
    return bytes(out)


def read_exr(path):
    """A scanline OpenEXR file as (H, W, 3) float32 of its R, G, B channels."""
    with open(path, "rb") as fh:
        buf = fh.read()
    if buf[:4] != b"\x76\x2f\x31\x01":
        raise ValueError(f"{path}: not an OpenEXR file")
    (version,) = struct.unpack_from("<I", buf, 4)
    if version & 0xFF != 2 or version & 0x1A00:
        raise ValueError(f"{path}: only single-part scanline OpenEXR files are supported")
    attrs, pos = _exr_header(buf, 8)
    chans = _exr_channels(attrs["channels"][1])
    comp = attrs["compression"][1][0]
    if comp not in _EXR_LINES:
        raise ValueError(f"{path}: OpenEXR compression {comp} is not supported (NONE, RLE, ZIPS, ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = _EXR_LINES[comp]
    nchunks = (h + lines - 1) // lines
    offsets = struct.unpack_from(f"<{nchunks}Q", buf, pos)
    planes = {name: np.zeros((h, w), f32) for name, _ in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8:off + 8 + size]
        row0 = y - y0
        nrows = min(lines, h - row0)
        expected = sum(_EXR_TYPES[t].itemsize for _, t in chans) * w * nrows
        if size < expected:
            if comp in (_EXR_ZIPS, _EXR_ZIP):
                data = _undo_predictor_and_split(zlib.decompress(data))
            elif comp == _EXR_RLE:
                data = _undo_predictor_and_split(_rle_decode(data, expected))
        if len(data) != expected:
            raise ValueError(f"{path}: OpenEXR chunk at y={y} has {len(data)} bytes, expected {expected}")
        p = 0
        for r in range(nrows):
            for name, t in chans:
                dt = _EXR_TYPES[t]
                n = dt.itemsize * w
                planes[name][row0 + r] = np.frombuffer(data, dt, w, p).astype(f32)
                p += n
    if all(c in planes for c in "RGB"):
        rgb = [planes["R"], planes["G"], planes["B"]]
    elif "Y" in planes:
        rgb = [planes["Y"]] * 3
    else:
        raise ValueError(f"{path}: OpenEXR file has neither R, G, B nor Y channels")
    return np.ascontiguousarray(np.stack(rgb, axis=-1))
