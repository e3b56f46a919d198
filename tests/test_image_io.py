"""The scoring command's loaders on the host (no GPU): PNG in every colour
type, bit depth, row filter and Adam7, and YUV4MPEG2 headers, frames and
errors, mirroring the reference's tests/input/sequence_reader_test.cc and
input_format_chroma_test.cc:527-555. PNG fixtures are written here from the
PNG specification and the expected planes come from the source arrays, not
from the encoder. Built from tests/cpp/image_io_dump.cc and
vqm_amd/cli/image_io.cc over the library's C ABI."""
from __future__ import annotations

import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from conftest import ROOT

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


@pytest.fixture(scope="module")
def load(image_loader):
    return image_loader


# ---- a PNG writer from the specification ----------------------------------------------

def chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xffffffff)


def pack_row(samples: np.ndarray, depth: int) -> bytes:
    """One row of samples (pixels x channels flattened) at `depth` bits."""
    s = samples.astype(np.int64).ravel()
    if depth == 16:
        return s.astype(">u2").tobytes()
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    per = 8 // depth
    pad = (-len(s)) % per
    s = np.concatenate([s, np.zeros(pad, np.int64)]).reshape(-1, per)
    shifts = np.array([8 - depth * (k + 1) for k in range(per)])
    return (s << shifts).sum(axis=1).astype(np.uint8).tobytes()


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def filter_row(row: bytes, prev: bytes, kind: int, bpp: int) -> bytes:
    out = bytearray([kind])
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][kind] if kind < 5 else 0
        out.append((x - pred) & 0xff)
    return bytes(out)


def encode_png(pixels: np.ndarray, color: int, depth: int, interlace: bool = False, plte: bytes = b"",
               trns: bytes = b"", filters=(0, 1, 2, 3, 4), extra: bytes = b"") -> bytes:
    """pixels: HxWxC integer samples (palette indices for colour type 3)."""
    h, w, ch = pixels.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    k = 0
    for x0, y0, dx, dy in passes:
        sub = pixels[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = bytes(len(pack_row(sub[0], depth)))
        for r in sub:
            row = pack_row(r, depth)
            raw += filter_row(row, prev, filters[k % len(filters)], bpp)
            k += 1
            prev = row
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0)
    body = chunk(b"IHDR", ihdr) + extra
    if plte:
        body += chunk(b"PLTE", plte)
    if trns:
        body += chunk(b"tRNS", trns)
    z = zlib.compress(bytes(raw), 9)
    body += chunk(b"IDAT", z[:len(z) // 2]) + chunk(b"IDAT", z[len(z) // 2:])  # split IDAT
    return b"\x89PNG\r\n\x1a\n" + body + chunk(b"IEND", b"")


def planes_of(o, n_planes, sample_bytes):
    """Split the packed data into planes (4-byte aligned)."""
    w, h = o["width"], o["height"]
    size = w * h * sample_bytes
    stride = (size + 3) // 4 * 4
    dt = np.uint8 if sample_bytes == 1 else np.uint16
    return [np.frombuffer(o["bytes"][p * stride:p * stride + size], dt).reshape(h, w) for p in range(n_planes)]


rng = np.random.default_rng(7)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("color,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (4, 8),
                                         (4, 16), (6, 8), (6, 16)])
def test_png_direct_colour(load, color, depth, interlace):
    h, w = 13, 11  # odd sizes exercise partial Adam7 passes and sub-byte padding
    ch = {0: 1, 2: 3, 4: 2, 6: 4}[color]
    px = rng.integers(0, 1 << depth, size=(h, w, ch))
    o = load(encode_png(px, color, depth, interlace))
    assert "error" not in o, o
    assert (o["width"], o["height"]) == (w, h)
    assert o["type"] == (1 if depth == 16 else 0) and o["depth"] == (16 if depth == 16 else 0)
    assert o["family"] == (0 if color in (0, 4) else 1)
    assert o["alpha"] == (1 if color in (4, 6) else 0)
    assert o["matrix"] == (1 if color in (0, 4) else 0)  # grey keeps out of the RGB value
    assert o["transfer"] == 13 and o["primaries"] == 1 and o["range"] == 1
    colour = 1 if color in (0, 4) else 3
    planes = planes_of(o, colour + (1 if color in (4, 6) else 0), 2 if depth == 16 else 1)
    scale = 255 // ((1 << depth) - 1) if depth < 8 else 1
    for c in range(colour):
        np.testing.assert_array_equal(planes[c], px[:, :, c] * scale)
    if color in (4, 6):
        np.testing.assert_array_equal(planes[-1], px[:, :, -1])


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette_and_trns(load, depth):
    h, w = 9, 14
    n = 1 << depth
    pal = rng.integers(0, 256, size=(n, 3))
    idx = rng.integers(0, n, size=(h, w, 1))
    o = load(encode_png(idx, 3, depth, plte=pal.astype(np.uint8).tobytes()))
    planes = planes_of(o, 3, 1)
    for c in range(3):
        np.testing.assert_array_equal(planes[c], pal[idx[:, :, 0], c])
    assert o["alpha"] == 0
    # tRNS shorter than the palette: the remaining entries are opaque.
    t = rng.integers(0, 256, size=max(1, n // 2))
    o = load(encode_png(idx, 3, depth, interlace=True, plte=pal.astype(np.uint8).tobytes(),
                        trns=t.astype(np.uint8).tobytes()))
    planes = planes_of(o, 4, 1)
    alpha = np.concatenate([t, np.full(n - len(t), 255)])
    assert o["alpha"] == 1
    np.testing.assert_array_equal(planes[3], alpha[idx[:, :, 0]])
    np.testing.assert_array_equal(planes[0], pal[idx[:, :, 0], 0])


def test_png_colour_key_trns(load):
    px = rng.integers(0, 4, size=(6, 7, 3)) * 80
    key = px[2, 3]
    trns = struct.pack(">HHH", *key)
    o = load(encode_png(px, 2, 8, trns=trns))
    planes = planes_of(o, 4, 1)
    expect = np.where((px == key).all(axis=2), 0, 255)
    np.testing.assert_array_equal(planes[3], expect)
    g = rng.integers(0, 4, size=(6, 7, 1))
    o = load(encode_png(g, 0, 2, trns=struct.pack(">H", 1)))
    planes = planes_of(o, 2, 1)
    np.testing.assert_array_equal(planes[0], g[:, :, 0] * 85)
    np.testing.assert_array_equal(planes[1], np.where(g[:, :, 0] == 1, 0, 255))
    g16 = rng.integers(0, 3, size=(5, 5, 1)) * 30000
    o = load(encode_png(g16, 0, 16, trns=struct.pack(">H", 30000)))
    planes = planes_of(o, 2, 2)
    np.testing.assert_array_equal(planes[1], np.where(g16[:, :, 0] == 30000, 0, 65535))


def test_png_ancillary_chunks_and_errors(load):
    px = rng.integers(0, 256, size=(4, 4, 3))
    good = encode_png(px, 2, 8, extra=chunk(b"gAMA", struct.pack(">I", 45455)) + chunk(b"tEXt", b"k\x00v"))
    assert "error" not in load(good)
    bad_crc = bytearray(good)
    bad_crc[8 + 4 + 4 + 5] ^= 1  # a byte inside IHDR
    assert "CRC" in load(bytes(bad_crc))["error"]
    assert "error" in load(good[:len(good) - 30])
    assert "error" in load(encode_png(px, 2, 8, filters=(5,)))  # unknown row filter
    jpeg = load(b"\xff\xd8\xff\xe0" + b"\x00" * 32)
    assert "JPEG" in jpeg["error"]
    assert "unsupported image format" in load(b"GIF89a" + b"\x00" * 16)["error"]
    idx = np.full((3, 3, 1), 5)
    assert "palette index" in load(encode_png(idx, 3, 8, plte=bytes(3 * 4)))["error"]


# ---- YUV4MPEG2 (sequence_reader_test.cc, input_format_chroma_test.cc:527-555) -----------

def test_y4m_sequence(load):
    first, second = bytes([32]) * 24, bytes([192]) * 24
    data = b"YUV4MPEG2 W4 H2 F60000:2002 Ip C444\nFRAME\n" + first + b"FRAME Xfoo=bar\n" + second
    seq = load(data, sequence=True)
    assert seq[0]["rate"] == [30000, 1001]
    frames = [o for o in seq if "bytes" in o]
    assert len(frames) == 2 and seq[-1] == {"end": True}
    assert frames[0]["bytes"][0] == 32 and frames[1]["bytes"][0] == 192
    assert (frames[0]["width"], frames[0]["height"]) == (4, 2)
    still = load(data)
    assert still["bytes"] == frames[0]["bytes"]  # LoadImage keeps first-frame semantics
    assert still["range"] == 0 and still["matrix"] == 5 and still["transfer"] == 1 and still["primaries"] == 1


def test_y4m_malformed_frames(load):
    seq = load(b"YUV4MPEG2 W4 H2 F30:1 Ip C444\nFRAME\n" + bytes(24) + b"FRAME\n" + bytes(23), sequence=True)
    assert "bytes" in seq[1] and "truncated y4m frame" in seq[-1]["error"]
    seq = load(b"YUV4MPEG2 W4 H2 F30:1 Ip C444\nFRAMEoops\n", sequence=True)
    assert "error" in seq[-1]


@pytest.mark.parametrize("header", [b"YUV4MPEG2 W4 H2 F0:1 Ip C444\n", b"YUV4MPEG2 W4 H2 F30:0 Ip C444\n",
                                    b"YUV4MPEG2 W4 H2 Ip C444\n", b"YUV4MPEG2 W4 H2 F30:1 It C444\n",
                                    b"YUV4MPEG2 W4 H2 F30:1 Im C444\n"])
def test_y4m_timing_and_interlace(load, header):
    data = header + b"FRAME\n" + bytes(24)
    assert "error" in load(data, sequence=True)[0]
    assert "error" not in load(data)  # the still loader ignores temporal metadata


def test_y4m_word_samples_and_depths(load):
    o = load(b"YUV4MPEG2 W4 H2 F24:1 Ip Cmonop16\nFRAME\n" + b"\x34\x12" * 8)
    assert np.frombuffer(o["bytes"][:2], "<u2")[0] == 0x1234 and o["depth"] == 16 and o["family"] == 0
    samples = [64, 257, 512, 940, 100, 333, 601, 900, 256, 500, 700, 1000]
    o = load(b"YUV4MPEG2 W2 H2 C444p10\nFRAME\n" + np.array(samples, "<u2").tobytes())
    assert o["type"] == 1 and o["depth"] == 10
    np.testing.assert_array_equal(np.frombuffer(o["bytes"], "<u2"), samples)
    for tag in (b"C420p11", b"Cbogus", b"C420p8"):
        assert "unsupported y4m chroma tag" in load(b"YUV4MPEG2 W4 H2 F24:1 Ip " + tag + b"\nFRAME\n" + bytes(64))["error"]


@pytest.mark.parametrize("tag,location,sub", [(b"", 1, (1, 1)), (b" C420jpeg", 1, (1, 1)), (b" C420mpeg2", 0, (1, 1)),
                                              (b" C420", 0, (1, 1)), (b" C420paldv", 2, (1, 1)),
                                              (b" C422", 0, (1, 0)), (b" C411", 0, (2, 0)), (b" C444", 0, (0, 0))])
def test_y4m_chroma_tags(load, tag, location, sub):
    w, h = 8, 2
    luma = w * h
    chroma = (w >> sub[0]) * (h >> sub[1])
    o = load(b"YUV4MPEG2 W8 H2 F24:1 Ip" + tag + b"\nFRAME\n" + bytes(range(luma)) + bytes([128]) * (2 * chroma))
    assert o["location"] == location and (o["sub_w"], o["sub_h"]) == sub and o["family"] == 2
    assert o["bytes"][:luma] == bytes(range(luma))
    stride = (chroma + 3) // 4 * 4
    assert o["bytes"][luma:luma + chroma] == bytes([128]) * chroma
    assert len(o["bytes"]) == luma + 2 * stride


def test_y4m_alpha_and_geometry_errors(load):
    o = load(b"YUV4MPEG2 W4 H2 F24:1 Ip C444alpha\nFRAME\n" + bytes(range(32)))
    assert o["alpha"] == 1 and o["bytes"][24:32] == bytes(range(24, 32))
    assert "divisible" in load(b"YUV4MPEG2 W5 H2 F24:1 Ip C420\nFRAME\n" + bytes(64))["error"]
    assert "frame size" in load(b"YUV4MPEG2 H2 F24:1 Ip C444\nFRAME\n")["error"]
    assert "invalid y4m width" in load(b"YUV4MPEG2 Wx H2 C444\nFRAME\n")["error"]


# ---- the command's argument handling (no device work) ------------------------------------

def test_command_line_errors():
    """Parse errors exit 2 with the usage; unreadable inputs exit 1 before
    any device work (cli/main.cc:108-160)."""
    cli = os.path.join(ROOT, "vqm_amd", "vqm-ss2")
    if not os.path.exists(cli):
        pytest.skip("vqm-ss2 is not built")
    r = subprocess.run([cli, "--metric", "vmaf-v0", "a", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "ssimulacra2 only" in r.stderr
    r = subprocess.run([cli, "--input-matrix", "bogus", "a", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "rgb|bt709" in r.stderr
    r = subprocess.run([cli, "--blur", "iir", "a", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "sampled|recursive" in r.stderr
    r = subprocess.run([cli, "only-one"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "two inputs" in r.stderr
    r = subprocess.run([cli, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--input-transfer" in r.stdout
    r = subprocess.run([cli, "/nonexistent/a.png", "/nonexistent/b.png"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "could not open" in r.stderr
