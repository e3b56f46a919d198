"""The vqm-ss2 command on the HIP path: its printed score and --raw norms
against the library's own results for the same inputs (cli/main.cc:42-53,
report.cc:221-248, video.cc:276-340), the 32x32 goldens through PNG files, and
a 4:2:0 Y4M frame against an FP64 float reference
(input_format_chroma_test.cc:557-640)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import format_ref as R
from conftest import GOLDEN_B, GOLDEN_TOL, ROOT, gradient_image
from test_image_io import encode_png
import vqm_amd
from vqm_amd import (AlphaType, ChromaFilter, ChromaLocation, ColorFamily, ColorPrimaries as CP, Image,
                     ImageFormat, MatrixCoefficients as MC, PixelRange, PixelType, Size,
                     Ssimulacra2Config, Ssimulacra2Pipeline, TransferCharacteristics as TC)

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "vqm_amd", "vqm-ss2")


def run(*args) -> list:
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def channel_major(result) -> np.ndarray:
    """report.cc:221-248's order from the 108-entry level-major layout."""
    out = []
    for c in range(3):
        for s in result.scales:
            for n in range(2):
                out += [s.avg_ssim[c * 2 + n], s.avg_edgediff[c * 4 + n], s.avg_edgediff[c * 4 + 2 + n]]
    return np.array(out)


def rgb_png(flat: np.ndarray, side: int = 32) -> bytes:
    px = flat.reshape(3, side, side).transpose(1, 2, 0)
    return encode_png(px, 2, 8)


def test_png_pair_scores_like_the_library(tmp_path):
    a, b = gradient_image(0), gradient_image(1)
    pa, pb = tmp_path / "a.png", tmp_path / "b.png"
    pa.write_bytes(rgb_png(a))
    pb.write_bytes(rgb_png(b))
    lines = run(pa, pb)
    assert len(lines) == 1
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32))) as p:
        expect = p.Compute(Image(Size(32, 32), ImageFormat.srgb8(), a), Image(Size(32, 32), ImageFormat.srgb8(), b))
    assert float(lines[0]) == pytest.approx(expect.Score(), abs=1e-8)
    assert float(lines[0]) == pytest.approx(GOLDEN_B, abs=GOLDEN_TOL)
    raw = run("--raw", pa, pb)
    assert len(raw) == 2 and raw[1] == lines[0]
    values = np.array([float(v) for v in raw[0].rstrip(",").split(",")])
    assert values.size == 3 * len(expect.scales) * 6
    np.testing.assert_allclose(values, channel_major(expect), atol=1e-12)
    assert float(run(pa, pa)[0]) == 100.0
    for planes in ("f16", "u16"):
        # The reference's half-plane bound (pipeline_test.cc:785-834).
        assert float(run("--planes", planes, pa, pb)[0]) == pytest.approx(expect.Score(), abs=0.1)


def test_png_with_alpha_takes_the_worse_background(tmp_path):
    """16-bit grey + alpha: ComputeWorstBackground over 0.1 and 0.9."""
    rng = np.random.default_rng(3)
    y, x = np.mgrid[0:32, 0:32]
    grey = (x * 2000 + y * 50).astype(np.int64)
    alpha = rng.integers(20000, 65536, size=(32, 32))
    dist = np.clip(grey + rng.integers(-900, 900, size=(32, 32)), 0, 65535)
    pa, pb = tmp_path / "a.png", tmp_path / "b.png"
    pa.write_bytes(encode_png(np.stack([grey, alpha], -1), 4, 16))
    pb.write_bytes(encode_png(np.stack([dist, alpha], -1), 4, 16))
    score = float(run(pa, pb)[0])
    fmt = ImageFormat(type=PixelType.Word, depth=16, family=ColorFamily.Grey, alpha=AlphaType.Straight,
                      matrix=MC.Bt709, transfer=TC.Srgb, primaries=CP.Bt709)

    def image(g):
        data = np.concatenate([g.astype(np.uint16).ravel(), alpha.astype(np.uint16).ravel()])
        return Image(Size(32, 32), fmt, data.view(np.uint8))
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32))) as p:
        dark = p.Compute(image(grey), image(dist), 0.1).Score()
        bright = p.Compute(image(grey), image(dist), 0.9).Score()
    assert dark != bright
    assert score == pytest.approx(min(dark, bright), abs=1e-8)


def y4m_420(luma: np.ndarray, u: np.ndarray, v: np.ndarray, tag=b"C420jpeg") -> bytes:
    h, w = luma.shape
    return (b"YUV4MPEG2 W%d H%d F25:1 Ip A1:1 %s\nFRAME\n" % (w, h, tag) + luma.astype(np.uint8).tobytes()
            + u.astype(np.uint8).tobytes() + v.astype(np.uint8).tobytes())


def test_y4m_frame_matches_the_float_reference(tmp_path):
    """A 4:2:0 8-bit frame with centred chroma decodes like its FP64
    float reference: limited-range codes normalised, chroma upsampled
    bilinearly, BT.470BG matrix, BT.709 transfer."""
    side, half = 16, 8
    y, x = np.mgrid[0:side, 0:side]
    luma = (20 + x * 13 + y * 7) & 0xff
    i = np.arange(half * half).reshape(half, half)
    u = np.where(i % half < half // 2, 60, 200)
    v = np.where(i // half < half // 2, 180, 90)
    path = tmp_path / "f.y4m"
    path.write_bytes(y4m_420(luma, u, v))
    fmt = ImageFormat(type=PixelType.Byte, family=ColorFamily.Yuv, range=PixelRange.Limited, matrix=MC.Bt470Bg,
                      transfer=TC.Bt709, primaries=CP.Bt709, chroma_location=ChromaLocation.Center,
                      subsample_w=1, subsample_h=1)
    coded = Image(Size(side, side), fmt, np.concatenate([luma.ravel(), u.ravel(), v.ravel()]).astype(np.uint8))
    vals = {0: (luma - 16) / 219.0}
    for k, plane in ((1, u), (2, v)):
        vals[k] = R.upsample((plane - 128) / 224.0, 1, 1, ChromaFilter.Bilinear, ChromaLocation.Center, side, side)
    ref_fmt = ImageFormat(type=PixelType.Float, family=ColorFamily.Yuv, matrix=MC.Bt470Bg, transfer=TC.Bt709,
                          primaries=CP.Bt709)
    reference = R.pack(ref_fmt, Size(side, side), vals)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(side, side))) as p:
        pair, identity = p.Compute(coded, reference), p.Compute(reference, reference)
        delta = float(np.max(np.abs(pair.norms() - identity.norms())))
        assert delta <= 2e-4, delta
        # The command reads the same frame from the file: identical norms.
        dist_path = tmp_path / "g.y4m"
        dist_path.write_bytes(y4m_420(luma, np.roll(u, 1, axis=1), v))
        dist = Image(Size(side, side), fmt,
                     np.concatenate([luma.ravel(), np.roll(u, 1, axis=1).ravel(), v.ravel()]).astype(np.uint8))
        expect = p.Compute(coded, dist)
    raw = run("--raw", path, dist_path)
    values = np.array([float(t) for t in raw[0].rstrip(",").split(",")])
    np.testing.assert_allclose(values, channel_major(expect), atol=1e-12)
    assert float(raw[1]) == pytest.approx(expect.Score(), abs=1e-8)


def test_y4m_sequence_scores_the_mean_of_its_frames(tmp_path):
    side, half = 16, 8
    rng = np.random.default_rng(11)
    frames_a, frames_b = [], []
    for _ in range(3):
        luma = rng.integers(16, 236, size=(side, side))
        u, v = rng.integers(16, 241, size=(2, half, half))
        frames_a.append((luma, u, v))
        frames_b.append((np.clip(luma + rng.integers(-6, 7, size=luma.shape), 16, 235), u, v))

    def write(path, frames, rate=b"F25:1"):
        body = b"YUV4MPEG2 W16 H16 %s Ip C420mpeg2\n" % rate
        for l, u, v in frames:
            body += b"FRAME\n" + l.astype(np.uint8).tobytes() + u.astype(np.uint8).tobytes() + v.astype(np.uint8).tobytes()
        path.write_bytes(body)
    pa, pb = tmp_path / "a.y4m", tmp_path / "b.y4m"
    write(pa, frames_a)
    write(pb, frames_b)
    fmt = ImageFormat(type=PixelType.Byte, family=ColorFamily.Yuv, range=PixelRange.Limited, matrix=MC.Bt470Bg,
                      transfer=TC.Bt709, primaries=CP.Bt709, chroma_location=ChromaLocation.Left,
                      subsample_w=1, subsample_h=1)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(side, side), chroma_filter=ChromaFilter.Bicubic)) as p:
        scores = [p.Compute(Image(Size(side, side), fmt, np.concatenate([x.ravel() for x in fa]).astype(np.uint8)),
                            Image(Size(side, side), fmt, np.concatenate([x.ravel() for x in fb]).astype(np.uint8))
                            ).Score() for fa, fb in zip(frames_a, frames_b)]
    out = run("--chroma-filter", "bicubic", pa, pb)
    assert float(out[0]) == pytest.approx(sum(scores) / 3, abs=1e-8)
    raw = run("--raw", "--chroma-filter", "bicubic", pa, pb)
    assert len(raw) == 4
    # Overrides apply to both inputs: full range changes the score.
    full = float(run("--chroma-filter", "bicubic", "--input-range", "full", pa, pb)[0])
    assert abs(full - float(out[0])) > 1e-6
    # Mismatched frame counts and rates are refused.
    write(pb, frames_b[:2])
    r = subprocess.run([CLI, str(pa), str(pb)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Frame count mismatch" in r.stderr
    write(pb, frames_b, rate=b"F30:1")
    r = subprocess.run([CLI, str(pa), str(pb)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Frame rate mismatch" in r.stderr


def test_jpeg_pair_scores_like_its_decoded_samples(tmp_path):
    """JPEG inputs (io.cc:277-355): the score of two JPEG files is the
    library's score of the samples libjpeg-turbo decodes from them (read
    through Pillow), 4:2:0 colour and progressive grey."""
    PIL = pytest.importorskip("PIL.Image")
    import io as _io
    rng = np.random.default_rng(4)
    y, x = np.mgrid[0:72, 0:96]
    base = np.stack([x * 2, y * 3, (x + y) % 256], -1).astype(np.float64)
    orig = np.clip(base + rng.normal(0, 4, base.shape), 0, 255).astype(np.uint8)
    for mode, kw in (("RGB", dict(quality=92, subsampling=2)), ("L", dict(quality=90, progressive=True))):
        src = orig if mode == "RGB" else orig[..., 0]
        paths, samples = [], []
        for i, q in enumerate((kw["quality"], 35)):
            b = _io.BytesIO()
            PIL.fromarray(src).save(b, "JPEG", **{**kw, "quality": q})
            path = tmp_path / f"{mode}{i}.jpg"
            path.write_bytes(b.getvalue())
            paths.append(path)
            samples.append(np.asarray(PIL.open(_io.BytesIO(b.getvalue()))))
        fmt = ImageFormat.srgb8()
        if mode == "L":
            fmt = ImageFormat(type=PixelType.Byte, family=ColorFamily.Grey, matrix=MC.Bt709)
        def planes(arr):
            arr = arr[..., None] if arr.ndim == 2 else arr
            n = 96 * 72
            out = np.zeros(arr.shape[-1] * ((n + 3) // 4 * 4), np.uint8)
            for c in range(arr.shape[-1]):
                out[c * ((n + 3) // 4 * 4):c * ((n + 3) // 4 * 4) + n] = arr[..., c].ravel()
            return out
        with Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(96, 72))) as p:
            want = p.Compute(Image(Size(96, 72), fmt, planes(samples[0])),
                             Image(Size(96, 72), fmt, planes(samples[1]))).Score()
        got = float(run(paths[0], paths[1])[0])
        assert got == pytest.approx(want, abs=1e-8), mode
        assert got < 100.0
