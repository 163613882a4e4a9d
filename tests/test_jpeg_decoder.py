"""The scene host layer's baseline JPEG decoder (bte_image.cpp, for texture
maps such as the reference's assets/checkerboard.jpg) against an encoder
written here from ITU-T T.81: forward DCT, quantization, zig-zag, DC
prediction, run/size coding with canonical Huffman codes, byte stuffing,
restart markers.  The decoder must return exactly the encoder's own
reconstruction (dequantized coefficients -> inverse DCT -> level shift ->
JFIF YCbCr -> RGB, all in binary64), with chroma planes upsampled as
stb_image does it (the reference's loader), allowing one unit where a value
sits on a rounding boundary.  Cases: 4:4:4 colour, 4:2:0 colour with a size
that is not a multiple of the MCU, grayscale, restart intervals, and a
hand-computed 4:2:0 seam."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bluetruthengine_amd", "csrc")
ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
                   13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52,
                   45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])


def _dct_matrix():
    c = np.zeros((8, 8))
    for x in range(8):
        for u in range(8):
            c[x, u] = (np.sqrt(0.5) if u == 0 else 1.0) * 0.5 * np.cos((2 * x + 1) * u * np.pi / 16)
    return c   # f[x] = sum_u c[x, u] F[u];  F[u] = sum_x c[x, u] f[x]


C = _dct_matrix()


class _Bits:
    def __init__(self):
        self.out, self.acc, self.n = bytearray(), 0, 0

    def put(self, value, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc, self.n = 0, 0

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)   # pad with ones


def _size(v):
    return 0 if v == 0 else int(abs(v)).bit_length()


def _bits_of(v, s):
    return v if v >= 0 else v + (1 << s) - 1


def encode(planes, sampling, quant, restart=0):
    """planes: list of (H_c, W_c) float sample planes (0..255) at their own
    resolution; sampling: (h, v) per component; returns (jpeg bytes, the
    per-component reconstruction planes)."""
    H, W = planes[0].shape
    hmax, vmax = max(s[0] for s in sampling), max(s[1] for s in sampling)
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    blocks, recon = [], []
    for (h, v), p in zip(sampling, planes):
        bw, bh = mcux * h, mcuy * v
        pad = np.pad(p, ((0, bh * 8 - p.shape[0]), (0, bw * 8 - p.shape[1])), mode="edge") - 128.0
        coef = np.zeros((bh, bw, 64), np.int64)
        rec = np.zeros((bh * 8, bw * 8))
        for by in range(bh):
            for bx in range(bw):
                blk = pad[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
                F = C.T @ blk @ C
                q = np.round(F.reshape(-1) / quant).astype(np.int64)
                coef[by, bx] = q
                deq = (q * quant).reshape(8, 8).astype(np.float64)
                rec[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = np.clip(np.floor(C @ deq @ C.T + 128.0 + 0.5), 0, 255)
        blocks.append(coef)
        recon.append(rec)
    # symbols, then canonical codes: every DC category 4 bits, every AC symbol 8 bits
    stream = []   # (kind, comp, symbol, extra value, extra length)
    preds = [0] * len(planes)
    n_mcu = 0
    for my in range(mcuy):
        for mx in range(mcux):
            if restart and n_mcu and n_mcu % restart == 0:
                stream.append(("RST", (n_mcu // restart - 1) % 8))
                preds = [0] * len(planes)
            n_mcu += 1
            for ci, (h, v) in enumerate(sampling):
                for yy in range(v):
                    for xx in range(h):
                        z = blocks[ci][my * v + yy, mx * h + xx][ZIGZAG]
                        d = int(z[0]) - preds[ci]
                        preds[ci] = int(z[0])
                        s = _size(d)
                        stream.append(("DC", ci, s, _bits_of(d, s), s))
                        run = 0
                        last = max([k for k in range(1, 64) if z[k] != 0], default=0)
                        for k in range(1, last + 1):
                            if z[k] == 0:
                                run += 1
                                continue
                            while run > 15:
                                stream.append(("AC", ci, 0xF0, 0, 0))
                                run -= 16
                            s = _size(int(z[k]))
                            stream.append(("AC", ci, (run << 4) | s, _bits_of(int(z[k]), s), s))
                            run = 0
                        if last < 63:
                            stream.append(("AC", ci, 0x00, 0, 0))
    ac_syms = sorted({t[2] for t in stream if t[0] == "AC"})
    ac_code = {sym: i for i, sym in enumerate(ac_syms)}
    bits = _Bits()
    for t in stream:
        if t[0] == "RST":
            bits.flush()
            bits.out += bytes([0xFF, 0xD0 + t[1]])
            continue
        kind, _, sym, val, length = t
        if kind == "DC":
            bits.put(sym, 4)
        else:
            bits.put(ac_code[sym], 8)
        if length:
            bits.put(val, length)
    bits.flush()

    def seg(marker, payload):
        return bytes([0xFF, marker, (len(payload) + 2) >> 8, (len(payload) + 2) & 255]) + payload
    out = bytearray(b"\xff\xd8")
    out += seg(0xDB, bytes([0]) + bytes(int(x) for x in quant[ZIGZAG]))
    nc = len(planes)
    sof = bytes([8, H >> 8, H & 255, W >> 8, W & 255, nc])
    for ci, (h, v) in enumerate(sampling):
        sof += bytes([ci + 1, (h << 4) | v, 0])
    out += seg(0xC0, sof)
    dc_counts = [0] * 16
    dc_counts[3] = 12
    out += seg(0xC4, bytes([0x00] + dc_counts + list(range(12))))
    ac_counts = [0] * 16
    ac_counts[7] = len(ac_syms)
    out += seg(0xC4, bytes([0x10] + ac_counts + ac_syms))
    if restart:
        out += seg(0xDD, bytes([restart >> 8, restart & 255]))
    sos = bytes([nc]) + b"".join(bytes([ci + 1, 0x00]) for ci in range(nc)) + bytes([0, 63, 0])
    out += seg(0xDA, sos) + bits.out + b"\xff\xd9"
    return bytes(out), recon


def _upsample(plane, hs, vs, H, W):
    """A component plane at full resolution as stb_image produces it
    (stb_image.h:3465-3527 and its row loop at :3927-3938): output row y takes
    the nearer source row y // vs and the farther one (above for the upper
    row of a pair, below for the lower, clamped); 2x factors use the 3:1
    triangle filter with stb's integer rounding, others repeat the sample."""
    wlo = -(-W // hs)
    rows = -(-H // vs)
    out = np.zeros((H, W), np.int64)
    for y in range(H):
        m = y // vs
        if vs == 2:
            far = max(m - 1, 0) if y % 2 == 0 else min(m + 1, rows - 1)
        else:
            far = m
        near_row = plane[m, :wlo].astype(np.int64)
        far_row = plane[far, :wlo].astype(np.int64)
        if hs == 2:
            t = 3 * near_row + far_row if vs == 2 else 4 * near_row
            row = np.zeros(2 * wlo, np.int64)
            if vs == 2:
                row[0], row[-1] = (t[0] + 2) >> 2, (t[-1] + 2) >> 2
            else:
                row[0], row[-1] = near_row[0], near_row[-1]
                if wlo > 1:
                    row[1] = (3 * near_row[0] + near_row[1] + 2) >> 2
                    row[-2] = (near_row[-2] * 3 + near_row[-1] + 2) >> 2
            for i in range(1, wlo):
                if vs == 2:
                    row[2 * i - 1] = (3 * t[i - 1] + t[i] + 8) >> 4
                    row[2 * i] = (3 * t[i] + t[i - 1] + 8) >> 4
                elif 0 < i < wlo - 1:
                    row[2 * i] = (3 * near_row[i] + 2 + near_row[i - 1]) >> 2
                    row[2 * i + 1] = (3 * near_row[i] + 2 + near_row[i + 1]) >> 2
            if wlo == 1:
                row[:] = (t[0] + 2) >> 2 if vs == 2 else near_row[0]
        elif vs == 2:
            row = (3 * near_row + far_row + 2) >> 2
        else:
            row = np.repeat(near_row, hs)
        out[y] = row[:W]
    return out


def _to_rgb(recon, sampling, H, W):
    hmax, vmax = max(s[0] for s in sampling), max(s[1] for s in sampling)
    s = [_upsample(r, hmax // h, vmax // v, H, W) for r, (h, v) in zip(recon, sampling)]
    if len(s) == 1:
        return np.repeat(s[0][..., None], 3, -1)
    y, cb, cr = s[0], s[1] - 128.0, s[2] - 128.0
    rgb = np.stack([y + 1.402 * cr, y - 0.344136 * cb - 0.714136 * cr, y + 1.772 * cb], -1)
    return np.clip(np.floor(rgb + 0.5), 0, 255)


@pytest.fixture(scope="module")
def decoder(tmp_path_factory):
    return build_decoder(tmp_path_factory)


def build_decoder(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("imgdec") / "image_decode")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "image_decode.cpp"),
                    os.path.join(CSRC, "bte_image.cpp"), os.path.join(CSRC, "bte_png.cpp"), "-o", exe], check=True)
    return exe


def _decode(decoder, data, tmp_path):
    src, dst = tmp_path / "in.jpg", tmp_path / "out.rgba"
    src.write_bytes(data)
    r = subprocess.run([decoder, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = dst.read_bytes()
    w, h = np.frombuffer(raw[:8], np.int32)
    return np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4)


def _image(H, W, seed):
    rs = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:W]
    rgb = np.stack([128 + 90 * np.sin(xs / 7.0), 100 + 60 * np.cos(ys / 5.0 + xs / 11.0), 40 + xs * 150.0 / W], -1)
    return np.clip(rgb + rs.normal(0, 12, rgb.shape), 0, 255)


def _ycc(rgb):
    r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
    return [0.299 * r + 0.587 * g + 0.114 * b, -0.168736 * r - 0.331264 * g + 0.5 * b + 128,
            0.5 * r - 0.418688 * g - 0.081312 * b + 128]


QUANT = np.array([16, 11, 10, 16, 24, 40, 51, 61] * 8, np.int64) // 2 + 1


@pytest.mark.parametrize("case", ["444", "420", "gray", "restart"])
def test_decoder_reproduces_the_encoders_reconstruction(decoder, tmp_path, case):
    H, W = (37, 53) if case == "420" else (32, 48)
    rgb = _image(H, W, 1)
    y, cb, cr = _ycc(rgb)
    if case == "gray":
        planes, sampling = [y], [(1, 1)]
    elif case == "420":
        sub = lambda p: np.pad(p, ((0, H % 2), (0, W % 2)), mode="edge").reshape((H + 1) // 2, 2, (W + 1) // 2, 2).mean((1, 3))
        planes, sampling = [y, sub(cb), sub(cr)], [(2, 2), (1, 1), (1, 1)]
    else:
        planes, sampling = [y, cb, cr], [(1, 1)] * 3
    data, recon = encode(planes, sampling, QUANT, restart=3 if case == "restart" else 0)
    got = _decode(decoder, data, tmp_path)
    want = _to_rgb(recon, sampling, H, W)
    diff = np.abs(got[..., :3].astype(np.int64) - want.astype(np.int64))
    assert got.shape == (H, W, 4) and np.all(got[..., 3] == 255)
    assert diff.max() <= 1 and (diff == 0).mean() > 0.999, (diff.max(), (diff == 0).mean())
    # and the picture survives the round trip (4:2:0 also loses the noise in
    # the chroma planes; a grayscale file keeps only the luma)
    src = np.repeat(y[..., None], 3, -1) if case == "gray" else rgb
    assert np.abs(got[..., :3].astype(np.float64) - src).mean() < (12.0 if case == "420" else 8.0)


def test_chroma_upsampling_hand_computed(decoder, tmp_path):
    # 32 x 32 4:2:0 picture, Y = 128 everywhere; the 16 x 16 chroma planes are
    # constant per 8 x 8 block, with values whose DC coefficient the
    # quantizer keeps exactly (8 (c - 128) a multiple of QUANT[0] = 9): Cb
    # 101 | 173 left | right, Cr 92 / 164 top / bottom.  By hand, the 3:1
    # filter gives at the seams Cb (3 * 404 + 692 + 8) >> 4 = 119 and
    # (3 * 692 + 404 + 8) >> 4 = 155 (t = 4 c), Cr (4 * 440 + 8) >> 4 = 110 and
    # (4 * 584 + 8) >> 4 = 146 (t = 3 * 92 + 164, 3 * 164 + 92); nearest
    # sampling would jump 101 -> 173 and 92 -> 164.
    H = W = 32
    y = np.full((H, W), 128.0)
    cb = np.full((16, 16), 101.0)
    cb[:, 8:] = 173.0
    cr = np.full((16, 16), 92.0)
    cr[8:, :] = 164.0
    data, recon = encode([y, cb, cr], [(2, 2), (1, 1), (1, 1)], QUANT)
    assert np.array_equal(recon[1][:16, :16], cb) and np.array_equal(recon[2][:16, :16], cr)
    cb_row = np.array([101] * 15 + [119, 155] + [173] * 15, np.float64)
    cr_col = np.array([92] * 15 + [110, 146] + [164] * 15, np.float64)
    CB, CR = np.meshgrid(cb_row - 128.0, cr_col - 128.0)
    want = np.clip(np.floor(np.stack([128 + 1.402 * CR, 128 - 0.344136 * CB - 0.714136 * CR, 128 + 1.772 * CB], -1)
                            + 0.5), 0, 255)
    got = _decode(decoder, data, tmp_path)[..., :3]
    assert np.array_equal(got, want)
    assert np.array_equal(_to_rgb(recon, [(2, 2), (1, 1), (1, 1)], H, W), want)


def test_decoder_rejects_what_it_does_not_decode(decoder, tmp_path):
    src = tmp_path / "p.jpg"
    # SOI then a progressive frame header
    src.write_bytes(b"\xff\xd8\xff\xc2\x00\x0b\x08\x00\x08\x00\x08\x01\x01\x11\x00\xff\xd9")
    r = subprocess.run([decoder, str(src), str(tmp_path / "o")], capture_output=True, text=True)
    assert r.returncode == 1 and "progressive" in r.stdout


def test_reference_checkerboard_texture(decoder, tmp_path):
    path = "/root/reference/assets/checkerboard.jpg"
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    px = _decode(decoder, open(path, "rb").read(), tmp_path)
    assert px.shape == (2048, 2048, 4)
    # a two-tone board: the channel histogram has two separated modes
    lum = px[..., :3].mean(-1)
    dark, light = (lum < 90).mean(), (lum > 110).mean()
    assert dark > 0.35 and light > 0.35
