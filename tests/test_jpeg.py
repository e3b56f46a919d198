"""The scoring command's JPEG decoder (vqm_amd/cli/jpeg.cc) against
libjpeg-turbo, through Pillow: the reference reads JPEG with libjpeg's
default set-up (tools/image_io/io.cc:277-355), so every file must decode to
the same 8-bit samples. Files come from Pillow's encoder (baseline,
optimised and progressive, 4:4:4 / 4:2:2 / 4:2:0, grey, restart markers,
qualities down to 1) and from tests/jpeg_writer.py for what Pillow does not
write (4:4:0 and 4:1:1 sampling, mixed factors, RGB by component ids or an
Adobe marker, one scan per component). No GPU."""
from __future__ import annotations

import io

import numpy as np
import pytest

from jpeg_writer import encode

PIL = pytest.importorskip("PIL.Image")


def decoded_planes(d: dict) -> np.ndarray:
    """(H, W, C) samples of the loader's packed planes."""
    w, h = d["width"], d["height"]
    n = w * h
    stride = (n + 3) // 4 * 4
    channels = 1 if d["family"] == 0 else 3
    raw = np.frombuffer(d["bytes"], np.uint8)
    return np.stack([raw[c * stride:c * stride + n].reshape(h, w) for c in range(channels)], axis=-1)


def reference(data: bytes) -> np.ndarray:
    img = PIL.open(io.BytesIO(data))
    arr = np.asarray(img)
    return arr[..., None] if arr.ndim == 2 else arr


def check(load, data: bytes, what: str):
    got = load(data)
    assert "error" not in got, f"{what}: {got.get('error')}"
    want = reference(data)
    planes = decoded_planes(got)
    assert planes.shape == want.shape, what
    bad = np.argwhere(planes != want)
    assert bad.size == 0, (f"{what}: {len(bad)} samples differ, first at {bad[0].tolist()}: "
                           f"{int(planes[tuple(bad[0])])} vs {int(want[tuple(bad[0])])}")
    if planes.shape[-1] == 1:
        assert got["family"] == 0 and got["matrix"] == 1
    else:
        assert got["family"] == 1 and got["matrix"] == 0
    assert got["type"] == 0 and got["transfer"] == 13 and got["primaries"] == 1 and got["range"] == 1
    return planes


def content(w: int, h: int, kind: str, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if kind == "noise":
        return rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    y, x = np.mgrid[0:h, 0:w]
    base = np.stack([x * 255 // max(w - 1, 1), y * 255 // max(h - 1, 1), (x * 7 + y * 13) % 256], -1)
    return np.clip(base + rng.normal(0, 6, base.shape), 0, 255).astype(np.uint8)


def pillow_jpeg(arr: np.ndarray, **kw) -> bytes:
    b = io.BytesIO()
    PIL.fromarray(arr).save(b, "JPEG", **kw)
    return b.getvalue()


SIZES = [(1, 1), (2, 3), (7, 9), (16, 16), (17, 8), (37, 53), (64, 48), (101, 77), (250, 3), (3, 250)]


@pytest.mark.parametrize("size", SIZES)
@pytest.mark.parametrize("subsampling", [0, 1, 2])
def test_baseline_matches_libjpeg(image_loader, size, subsampling):
    w, h = size
    for kind, quality in (("smooth", 90), ("noise", 40)):
        arr = content(w, h, kind, seed=w * 31 + h)
        check(image_loader, pillow_jpeg(arr, quality=quality, subsampling=subsampling), f"{size} {kind}")


@pytest.mark.parametrize("size", [(7, 9), (37, 53), (101, 77), (250, 3)])
@pytest.mark.parametrize("subsampling", [0, 1, 2])
def test_progressive_and_optimised_match_libjpeg(image_loader, size, subsampling):
    w, h = size
    arr = content(w, h, "smooth", seed=7)
    noisy = content(w, h, "noise", seed=8)
    check(image_loader, pillow_jpeg(arr, quality=85, subsampling=subsampling, progressive=True), "progressive")
    check(image_loader, pillow_jpeg(noisy, quality=60, subsampling=subsampling, progressive=True,
                                    optimize=True), "progressive optimised")
    check(image_loader, pillow_jpeg(arr, quality=75, subsampling=subsampling, optimize=True), "optimised")


@pytest.mark.parametrize("quality", [1, 5, 100])
def test_extreme_quality_exercises_range_limits(image_loader, quality):
    """Coarse quantisation overshoots the 0..255 range (range table);
    quality 100 makes large coefficients."""
    for kind in ("noise", "smooth"):
        arr = content(45, 37, kind, seed=quality)
        check(image_loader, pillow_jpeg(arr, quality=quality, subsampling=2), f"q{quality} {kind}")
        check(image_loader, pillow_jpeg(arr, quality=quality, subsampling=0, progressive=True), f"q{quality} prog")


def test_restart_markers(image_loader):
    arr = content(77, 45, "smooth", seed=3)
    for kw in (dict(restart_marker_blocks=1), dict(restart_marker_blocks=3), dict(restart_marker_rows=1)):
        check(image_loader, pillow_jpeg(arr, quality=80, subsampling=2, **kw), str(kw))
        check(image_loader, pillow_jpeg(arr, quality=80, subsampling=0, progressive=True, **kw), f"prog {kw}")


def test_grey(image_loader):
    for w, h in ((1, 1), (33, 17), (64, 64)):
        g = content(w, h, "smooth", seed=w)[..., 0]
        check(image_loader, pillow_jpeg(g, quality=90), "grey")
        check(image_loader, pillow_jpeg(g, quality=50, progressive=True), "grey progressive")


def down(plane: np.ndarray, fx: int, fy: int, w: int, h: int, hmax: int, vmax: int, hc: int, vc: int):
    """A component at its downsampled size: box means of fx x fy cells."""
    dw, dh = -(-w * hc // hmax), -(-h * vc // vmax)
    pad = np.pad(plane.astype(np.float64), ((0, dh * fy - h), (0, dw * fx - w)), mode="edge")
    return np.round(pad.reshape(dh, fy, dw, fx).mean(axis=(1, 3))).astype(np.uint8)


@pytest.mark.parametrize("sampling", [
    [(1, 2), (1, 1), (1, 1)],   # 4:4:0: vertical triangle filter
    [(4, 1), (1, 1), (1, 1)],   # 4:1:1: box replication
    [(2, 2), (1, 2), (1, 1)],   # mixed: 2:1 vertical for Cb, 2x2 for Cr
    [(4, 1), (2, 1), (1, 1)],   # 2:1 horizontal (fancy) for Cb, 4:1 box for Cr
    [(2, 2), (2, 1), (1, 1)],   # 2:1 vertical for Cb, 2x2 for Cr
    [(1, 1), (1, 1), (1, 1)],
])
@pytest.mark.parametrize("size", [(3, 5), (19, 13), (40, 24), (67, 35)])
def test_sampling_factors(image_loader, sampling, size):
    w, h = size
    arr = content(w, h, "smooth", seed=w + h)
    hmax = max(s[0] for s in sampling)
    vmax = max(s[1] for s in sampling)
    planes = [down(arr[..., c], hmax // hc, vmax // vc, w, h, hmax, vmax, hc, vc)
              for c, (hc, vc) in enumerate(sampling)]
    check(image_loader, encode(planes, w, h, sampling), f"{sampling} {size}")
    check(image_loader, encode(planes, w, h, sampling, restart=2), f"{sampling} {size} restart")
    check(image_loader, encode(planes, w, h, sampling, separate_scans=True, restart=3),
          f"{sampling} {size} one scan per component")


def test_colour_space_selection(image_loader):
    """jdapimin.c's choice for three components: JFIF means YCbCr; Adobe
    transform 0 means RGB (no conversion), 1 YCbCr; without either, ids
    'R','G','B' mean RGB and any other ids YCbCr."""
    w, h = 24, 16
    arr = content(w, h, "smooth", seed=5)
    planes = [arr[..., c] for c in range(3)]
    s = [(1, 1)] * 3
    rgb_ids = [ord("R"), ord("G"), ord("B")]
    cases = {
        "jfif": encode(planes, w, h, s),
        "adobe rgb": encode(planes, w, h, s, jfif=False, adobe_transform=0),
        "adobe ycc": encode(planes, w, h, s, jfif=False, adobe_transform=1),
        "rgb ids": encode(planes, w, h, s, jfif=False, ids=rgb_ids),
        "rgb ids under jfif": encode(planes, w, h, s, ids=rgb_ids),
        "other ids": encode(planes, w, h, s, jfif=False, ids=[7, 8, 9]),
    }
    out = {name: check(image_loader, data, name) for name, data in cases.items()}
    # Stored as RGB: the samples come back within the coding error; read as
    # YCbCr the same planes come back far from them.
    assert np.abs(out["adobe rgb"].astype(int) - arr).mean() < 4
    assert np.abs(out["jfif"].astype(int) - arr).mean() > 20
    assert np.array_equal(out["adobe rgb"], out["rgb ids"])
    assert not np.array_equal(out["adobe rgb"], out["jfif"])


def test_grey_with_sampling_factors(image_loader):
    """One component: its MCU is one block whatever its factors."""
    g = content(30, 21, "smooth", seed=2)[..., 0]
    check(image_loader, encode([g], 30, 21, [(2, 2)]), "grey 2x2 factors")


def test_errors(image_loader):
    arr = content(40, 32, "noise", seed=1)
    data = pillow_jpeg(arr, quality=90)
    # Entropy data cut short: the reference's libjpeg set-up fails on
    # JWRN_HIT_MARKER (io.cc:264-273).
    sos = data.index(b"\xff\xda")
    cut = image_loader(data[:sos + (len(data) - sos) // 2])
    assert "premature end of data segment" in cut["error"]
    # 12-bit precision, arithmetic coding, four components.
    sof = data.index(b"\xff\xc0")
    twelve = bytearray(data)
    twelve[sof + 4] = 12
    assert "precision 12" in image_loader(bytes(twelve))["error"]
    arith = bytearray(data)
    arith[sof + 1] = 0xC9
    assert "Unsupported JPEG process" in image_loader(bytes(arith))["error"]
    b = io.BytesIO()
    PIL.fromarray(arr).convert("CMYK").save(b, "JPEG")
    assert "number of components" in image_loader(b.getvalue())["error"]
    assert "error" in image_loader(b"\xff\xd8\xff\xd9")


def dht_segment(tc_th: int, counts, values) -> bytes:
    body = bytes([tc_th]) + bytes(counts) + bytes(values)
    return b"\xff\xc4" + (len(body) + 2).to_bytes(2, "big") + body


@pytest.mark.parametrize("counts", [
    [200] + [0] * 15,                 # 200 one-bit codes: far more than 2 fit
    [2] + [0] * 15,                   # both one-bit codes: the second is all ones
    [0, 0, 8] + [0] * 13,             # eight three-bit codes: 111 is all ones
    [0, 0, 0, 20] + [0] * 12,         # twenty four-bit codes where sixteen exist
    [1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2],  # the last sixteen-bit code is all ones
])
def test_malformed_huffman_tables_fail(image_loader, counts):
    """Per-length code counts that do not fit their lengths are rejected
    before any lookahead entry is written (a DHT segment alone, after SOI)."""
    n = sum(counts)
    data = b"\xff\xd8" + dht_segment(0x10, counts, [i & 0xFF for i in range(n)]) + b"\xff\xd9"
    assert "Bogus Huffman table definition" in image_loader(data)["error"]


def test_huffman_counts_past_256_fail(image_loader):
    counts = [0] * 8 + [255, 2] + [0] * 6
    data = b"\xff\xd8" + dht_segment(0x10, counts, [0] * 257) + b"\xff\xd9"
    assert "Bogus Huffman table definition" in image_loader(data)["error"]
