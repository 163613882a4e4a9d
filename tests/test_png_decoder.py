"""The scene host layer's zlib / PNG decoder (bte_png.cpp, for glTF images and
PNG albedo textures) against PNG files written here: compression by Python's
zlib (an independent DEFLATE implementation: stored, fixed-Huffman, dynamic,
Huffman-only and RLE streams), every scanline filter, Adam7 interlacing,
split IDAT chunks, every colour type and bit depth.  Expected RGBA8 values
follow the PNG specification and stb_image's 8-bit, 4-channel load (high
byte of 16-bit samples, grey below 8 bits scaled by 255 / (2^depth - 1), tRNS
colour key -> alpha 0, palette alpha from tRNS, 255 past its end)."""
import struct
import subprocess
import zlib

import numpy as np
import pytest

from test_jpeg_decoder import build_decoder

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def _pack_rows(samples, depth):
    """samples (h, w * channels) ints -> list of packed row bytes"""
    rows = []
    for r in samples:
        if depth == 8:
            rows.append(bytes(int(v) for v in r))
        elif depth == 16:
            rows.append(b"".join(struct.pack(">H", int(v)) for v in r))
        else:
            bits = "".join(format(int(v), f"0{depth}b") for v in r)
            bits += "0" * (-len(bits) % 8)
            rows.append(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
    return rows


def _filter(rows, bpp, start):
    """applies filter type (start + y) % 5 to row y; returns the filtered stream"""
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        ft = (start + y) % 5
        f = bytearray()
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                pr = 0
            elif ft == 1:
                pr = a
            elif ft == 2:
                pr = b
            elif ft == 3:
                pr = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            f.append((x - pr) & 255)
        out += bytes([ft]) + f
        prev = row
    return bytes(out)


def encode_png(samples, depth, ctype, plte=None, trns=None, interlace=False, level=6, strategy=None, split=0):
    """samples: (h, w, channels) ints in sample units."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    if interlace:
        raw = b""
        for k, (x0, y0, dx, dy) in enumerate(ADAM7):
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += _filter(_pack_rows(sub.reshape(sub.shape[0], -1), depth), bpp, k)
    else:
        raw = _filter(_pack_rows(samples.reshape(h, -1), depth), bpp, 0)
    if strategy is None:
        data = zlib.compress(raw, level)
    else:
        co = zlib.compressobj(level, zlib.DEFLATED, 15, 8, strategy)
        data = co.compress(raw) + co.flush()
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    out += _chunk(b"tEXt", b"Comment\x00test")          # ancillary chunk: ignored
    if plte is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    if split:
        for i in range(0, len(data), split):
            out += _chunk(b"IDAT", data[i:i + split])
    else:
        out += _chunk(b"IDAT", data)
    return out + _chunk(b"IEND", b"")


@pytest.fixture(scope="module")
def decoder(tmp_path_factory):
    return build_decoder(tmp_path_factory)


def _decode(decoder, data, tmp_path, ok=True):
    src, dst = tmp_path / "in.png", tmp_path / "out.rgba"
    src.write_bytes(data)
    r = subprocess.run([decoder, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    if not ok:
        return r
    assert r.returncode == 0, r.stdout + r.stderr
    raw = dst.read_bytes()
    w, h = np.frombuffer(raw[:8], np.int32)
    return np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4)


RS = np.random.default_rng(11)


def _rgba_expect(s, depth, ctype, plte=None, trns=None):
    s = s.astype(np.int64)
    h, w, _ = s.shape
    out = np.zeros((h, w, 4), np.int64)
    to8 = (lambda v: v >> 8) if depth == 16 else (lambda v: v)
    if ctype == 0:
        g = s[..., 0] * (255 // ((1 << depth) - 1)) if depth < 8 else to8(s[..., 0])
        out[..., :3] = g[..., None]
        out[..., 3] = 255
        if trns is not None:
            out[..., 3][s[..., 0] == struct.unpack(">H", trns)[0]] = 0
    elif ctype == 2:
        out[..., :3] = to8(s)
        out[..., 3] = 255
        if trns is not None:
            key = np.array(struct.unpack(">HHH", trns))
            out[..., 3][np.all(s == key, -1)] = 0
    elif ctype == 3:
        pal = np.asarray(plte, np.int64)
        out[..., :3] = pal[s[..., 0]]
        alpha = np.full(256, 255)
        if trns is not None:
            alpha[:len(trns)] = list(trns)
        out[..., 3] = alpha[s[..., 0]]
    elif ctype == 4:
        out[..., :3] = to8(s[..., :1])
        out[..., 3] = to8(s[..., 1])
    else:
        out[...] = to8(s)
    return out.astype(np.uint8)


CASES = [
    # (name, ctype, depth, h, w, extra)
    ("rgba8", 6, 8, 19, 23, {}),
    ("rgba16", 6, 16, 9, 11, {}),
    ("rgb8_key", 2, 8, 13, 17, {"key": True}),
    ("rgb16_key", 2, 16, 7, 9, {"key": True}),
    ("grey1", 0, 1, 9, 21, {}),
    ("grey2_key", 0, 2, 8, 13, {"key": True}),
    ("grey4", 0, 4, 7, 15, {}),
    ("grey8", 0, 8, 6, 10, {}),
    ("grey16_key", 0, 16, 5, 7, {"key": True}),
    ("pal1", 3, 1, 6, 19, {}),
    ("pal2", 3, 2, 7, 9, {"trns": True}),
    ("pal4", 3, 4, 8, 11, {"trns": True}),
    ("pal8", 3, 8, 10, 12, {"trns": True}),
    ("ga8", 4, 8, 9, 9, {}),
    ("ga16", 4, 16, 4, 6, {}),
]


@pytest.mark.parametrize("name,ctype,depth,h,w,extra", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_colour_types_depths_filters(decoder, tmp_path, name, ctype, depth, h, w, extra, interlace):
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    hi = (1 << depth) if ctype != 3 else min(1 << depth, 200)
    s = RS.integers(0, hi, (h, w, ch))
    plte = trns = None
    if ctype == 3:
        plte = RS.integers(0, 256, (hi, 3))
        if extra.get("trns"):
            trns = bytes(RS.integers(0, 256, max(1, hi // 2)).tolist())
    if extra.get("key"):
        key = s[h // 2, w // 3]
        s[0, 0] = key                                   # at least one pixel matches
        trns = struct.pack(">" + "H" * ch, *[int(v) for v in key])
    png = encode_png(s, depth, ctype, plte=plte, trns=trns, interlace=interlace)
    got = _decode(decoder, png, tmp_path)
    want = _rgba_expect(s, depth, ctype, plte, trns)
    assert np.array_equal(got, want), (name, np.argwhere(got != want)[:5])


@pytest.mark.parametrize("mode", ["stored", "fixed", "huffman_only", "rle", "best", "split_idat"])
def test_deflate_block_kinds(decoder, tmp_path, mode):
    # smooth data with repeats so that every block kind has matches to copy
    y, x = np.mgrid[0:64, 0:96]
    s = np.stack([(x * 3) % 256, (y * 5 + x // 7) % 256, (x ^ y) % 256, np.full_like(x, 255)], -1)
    s[10:20, 30:60] = RS.integers(0, 256, (10, 30, 4))
    kw = {"stored": dict(level=0), "fixed": dict(strategy=zlib.Z_FIXED), "huffman_only": dict(strategy=zlib.Z_HUFFMAN_ONLY),
          "rle": dict(strategy=zlib.Z_RLE), "best": dict(level=9), "split_idat": dict(split=97)}[mode]
    got = _decode(decoder, encode_png(s, 8, 6, **kw), tmp_path)
    assert np.array_equal(got, s.astype(np.uint8))


def test_png_errors(decoder, tmp_path):
    s = RS.integers(0, 256, (4, 4, 4))
    good = encode_png(s, 8, 6)
    # corrupt the compressed stream: flip bits inside IDAT's DEFLATE data
    i = good.index(b"IDAT") + 4
    bad = bytearray(good)
    for k in range(2, 12):
        bad[i + k] ^= 0x5A
    r = _decode(decoder, bytes(bad), tmp_path, ok=False)
    assert r.returncode == 1 and "error" in r.stdout
    # unknown filter type
    raw = b"".join(bytes([7]) + bytes(16) for _ in range(4))
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 4, 8, 6, 0, 0, 0))
    png += _chunk(b"IDAT", zlib.compress(raw)) + _chunk(b"IEND", b"")
    r = _decode(decoder, png, tmp_path, ok=False)
    assert r.returncode == 1 and "filter" in r.stdout
    # an unknown critical chunk and a missing header
    r = _decode(decoder, b"\x89PNG\r\n\x1a\n" + _chunk(b"ABCD", b"x") + _chunk(b"IEND", b""), tmp_path, ok=False)
    assert r.returncode == 1
