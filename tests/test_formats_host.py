"""The general input surface on the host (no GPU): format validation, plane
layout sizes and the decode scale/offset, against the reference's own
expectations (tests/input/input_format_test.cc:77-205,
tests/input/input_format_chroma_test.cc:430-524)."""
from __future__ import annotations

import numpy as np
import pytest

import vqm_amd
from vqm_amd import (AlphaType, ColorFamily, ColorPrimaries, ImageFormat, InvalidArgument,
                     MatrixCoefficients, PixelRange, PixelType, Size, TransferCharacteristics)

S = Size(32, 32)


def pinned(alpha=AlphaType.NoAlpha, type_=PixelType.Float) -> ImageFormat:
    """tests/support.h:106-117."""
    return ImageFormat(type=type_, family=ColorFamily.Rgb, alpha=alpha, matrix=MatrixCoefficients.Rgb,
                       transfer=TransferCharacteristics.Srgb, primaries=ColorPrimaries.Bt709)


def rejects(fmt, match, size=S):
    with pytest.raises(InvalidArgument, match=match):
        vqm_amd.validate_image_format(fmt, size)


def test_rejected_formats():
    """input_format_chroma_test.cc:430-510."""
    rgb = pinned()
    f = pinned(); f.family, f.matrix = ColorFamily.Yuv, MatrixCoefficients.Rgb
    rejects(f, "YUV images cannot declare RGB")
    f = pinned(); f.subsample_w = 1
    rejects(f, "subsampled RGB images are not supported")
    f = pinned(); f.family, f.matrix, f.subsample_h = ColorFamily.Grey, MatrixCoefficients.Bt709, 1
    rejects(f, "subsampled greyscale images are not supported")
    f = pinned(); f.matrix = MatrixCoefficients.Bt709
    rejects(f, "RGB images cannot declare YUV")
    f = pinned(); f.family, f.matrix, f.primaries = ColorFamily.Yuv, MatrixCoefficients.Bt2020Cl, ColorPrimaries.Bt2020
    rejects(f, "constant-luminance matrix coefficients")
    f = pinned(); f.family, f.matrix, f.primaries = ColorFamily.Yuv, 14, ColorPrimaries.Bt2020
    rejects(f, "ICtCp requires")
    f = pinned(); f.transfer = TransferCharacteristics.Unspecified
    rejects(f, "transfer characteristics are required")
    f = pinned(); f.primaries = ColorPrimaries.Unspecified
    rejects(f, "color primaries are required")
    f = pinned(); f.family, f.matrix = ColorFamily.Yuv, 11  # ST 2085 YDzDx
    rejects(f, "unsupported matrix coefficients")
    f = pinned(); f.type, f.depth = PixelType.Byte, 12
    rejects(f, "bit depth exceeds the pixel type")
    f = pinned(); f.type, f.depth, f.range = PixelType.Byte, 7, PixelRange.Limited
    rejects(f, "limited range requires a bit depth of at least 8")
    f = pinned(); f.family, f.matrix, f.subsample_w = ColorFamily.Yuv, MatrixCoefficients.Bt709, 1
    rejects(f, "divisible by the subsampling", Size(33, 32))
    f = pinned(); f.family, f.matrix = ColorFamily.Yuv, MatrixCoefficients.Unspecified
    rejects(f, "matrix coefficients are required")
    f = pinned(); f.family, f.matrix, f.primaries = ColorFamily.Yuv, 12, 10  # derived NCL, XYZ
    rejects(f, "XYZ color primaries")
    vqm_amd.validate_image_format(rgb, S)  # and the base format passes


def test_enum_validation():
    """input_format_test.cc:77-120: unknown enum values are refused before
    anything is packed."""
    for field in ("type", "family", "alpha", "range", "chroma_location", "transfer", "primaries",
                  "matrix"):
        f = ImageFormat()
        setattr(f, field, 255)
        with pytest.raises(InvalidArgument):
            vqm_amd.validate_image_format(f, S)
    f = ImageFormat(family=ColorFamily.Grey, matrix=255)
    with pytest.raises(InvalidArgument):
        vqm_amd.validate_image_format(f, S)


def test_accepted_formats():
    ok = [
        ImageFormat(type=PixelType.Word, depth=10, range=PixelRange.Limited, family=ColorFamily.Yuv,
                    matrix=MatrixCoefficients.Bt2020Ncl, transfer=TransferCharacteristics.St2084,
                    primaries=ColorPrimaries.Bt2020, subsample_w=1, subsample_h=1),
        ImageFormat(type=PixelType.Byte, family=ColorFamily.Grey, matrix=MatrixCoefficients.Unspecified,
                    primaries=ColorPrimaries.Unspecified),
        ImageFormat(type=PixelType.Half, alpha=AlphaType.Premultiplied),
        ImageFormat(family=ColorFamily.Yuv, matrix=14, transfer=TransferCharacteristics.AribB67,
                    primaries=ColorPrimaries.Bt2020),
        ImageFormat(family=ColorFamily.Yuv, matrix=MatrixCoefficients.Bt2020Cl,
                    transfer=TransferCharacteristics.Bt709, primaries=ColorPrimaries.Bt2020),
    ]
    for f in ok:
        vqm_amd.validate_image_format(f, S)


def test_required_bytes_boundaries():
    """input_format_test.cc:122-177."""
    f = pinned(type_=PixelType.Byte)
    assert vqm_amd.required_bytes(f, Size(3, 1)) == 12
    f.type = PixelType.Word
    assert vqm_amd.required_bytes(f, Size(3, 1)) == 24
    f.type, f.family, f.subsample_w, f.subsample_h = PixelType.Byte, ColorFamily.Yuv, 1, 1
    assert vqm_amd.required_bytes(f, Size(8, 4)) == 48
    f.alpha = AlphaType.Straight
    assert vqm_amd.required_bytes(f, Size(8, 4)) == 80
    for shift in (3, 64, 255):
        g = pinned(type_=PixelType.Byte)
        g.subsample_w = shift
        with pytest.raises(InvalidArgument):
            vqm_amd.required_bytes(g, Size(8, 8))


def test_scale_offset():
    """input_format_test.cc:179-205 and the depth/range/chroma rules of
    input.cc:714-757: limited luma 16..235 and chroma 16..240 around 128,
    full-range chroma centred at the midpoint, alpha always full."""
    f = pinned(type_=PixelType.Byte)
    f.depth, f.range = 8, PixelRange.Limited
    assert vqm_amd.plane_scale_offset(f) == (np.float32(1 / 219), np.float32(-16 / 219))
    assert vqm_amd.plane_scale_offset(f, chroma=True) == (np.float32(1 / 224), np.float32(-128 / 224))
    assert vqm_amd.plane_scale_offset(f, alpha=True) == (np.float32(1 / 255), 0.0)
    f.range = PixelRange.Full
    assert vqm_amd.plane_scale_offset(f, chroma=True) == (np.float32(1 / 255), np.float32(-128 / 255))
    f.type, f.depth = PixelType.Word, 10
    f.range = PixelRange.Limited
    assert vqm_amd.plane_scale_offset(f) == (np.float32(1 / 876), np.float32(-64 / 876))
    f.matrix = MatrixCoefficients.YCgCo  # YCgCo chroma keeps the luma range
    assert vqm_amd.plane_scale_offset(f, chroma=True) == (np.float32(1 / 876), np.float32(-512 / 876))
    f = pinned()
    assert vqm_amd.plane_scale_offset(f) == (1.0, 0.0)  # float samples are normalised
    f = pinned(type_=PixelType.Byte)
    f.depth, f.range = 7, PixelRange.Limited
    with pytest.raises(InvalidArgument):
        vqm_amd.plane_scale_offset(f)


# ---- pins of the FP64 format reference the GPU A/B tests use -------------------

def test_format_reference_closed_forms():
    """format_ref against published constants it does not contain: the
    luminance rows of BT.709 and BT.2020 derived from their primaries, the
    BT.709 decode matrix, PQ and HLG anchor points, and the ICtCp and
    YCgCo matrices on white."""
    import format_ref as R
    from vqm_amd import ChromaFilter, ChromaLocation, ColorPrimaries as CP, MatrixCoefficients as MC
    from vqm_amd import TransferCharacteristics as TC
    np.testing.assert_allclose(R.rgb_to_xyz(CP.Bt709)[1], [0.2126, 0.7152, 0.0722], atol=1e-4)
    np.testing.assert_allclose(R.rgb_to_xyz(CP.Bt2020)[1], [0.2627, 0.6780, 0.0593], atol=1e-4)
    np.testing.assert_allclose(R.rgb_to_xyz(CP.Bt709)[:, 0], [0.4124, 0.2126, 0.0193], atol=1e-4)
    np.testing.assert_allclose(R.gamut(CP.Bt709), np.eye(3), atol=1e-12)
    np.testing.assert_allclose(R.gamut(CP.Bt2020)[0], [1.6605, -0.5876, -0.0728], atol=1e-4)
    np.testing.assert_allclose(R.ncl_to_rgb(MC.Bt709), [[1, 0, 1.5748], [1, -0.18733, -0.46813],
                                                        [1, 1.8556, 0]], atol=1e-4)
    kr, kb = R.derived_krkb(CP.Bt470Bg)
    assert (kr, kb) == pytest.approx((0.2220, 0.0713), abs=2e-4)
    assert R.decode(TC.St2084, 1.0) == pytest.approx(100.0, rel=1e-12)
    assert R.decode(TC.St2084, 0.5080784) == pytest.approx(1.0, rel=1e-4)  # 100 cd/m2
    assert R.hlg_inverse_oetf(0.5) == pytest.approx(1.0 / 12.0, rel=1e-12)
    assert R.hlg_inverse_oetf(0.5 + 1e-12) == pytest.approx(1.0 / 12.0, rel=1e-6)  # continuous
    assert R.hlg_inverse_oetf(1.0) == pytest.approx(1.0, rel=1e-6)
    assert R.decode(TC.Srgb, 0.04045) == pytest.approx(R.decode(TC.Srgb, 0.04045 + 1e-12), rel=1e-6)
    for v in (0.01, 0.3, 0.9):
        assert R.bt709_inverse_oetf(R.bt709_oetf(v)) == pytest.approx(v, rel=1e-12)
    white = R.np.linalg.inv(R.lms_to_rgb2020()) @ np.ones(3)
    np.testing.assert_allclose(white, np.ones(3), atol=1e-12)
    for t in (TC.St2084, TC.AribB67):
        np.testing.assert_allclose(R.lms_to_ictcp(t) @ np.ones(3), [1, 0, 0], atol=1e-12)
    np.testing.assert_allclose(R.ncl_to_rgb(MC.YCgCo) @ [1, 0, 0], np.ones(3), atol=1e-12)
    # A BT.2020 constant-luminance neutral decodes to itself through the curve.
    assert R.cl_to_rgb(0.2627, 0.0593, (0.5, 0.0, 0.0)) == pytest.approx(
        (R.bt709_inverse_oetf(0.5),) * 3, rel=1e-12)
    # Resampler: textbook 2x bilinear phases for centre and left siting.
    c = R.chroma_shift(ChromaLocation.Center, True, 1)
    assert R.axis_taps(ChromaFilter.Bilinear, 0, 2.0, c, 8) == [(0, 0.25), (0, 0.75)]
    assert R.axis_taps(ChromaFilter.Bilinear, 1, 2.0, c, 8) == [(0, 0.75), (1, 0.25)]
    left = R.chroma_shift(ChromaLocation.Left, True, 1)
    assert R.axis_taps(ChromaFilter.Bilinear, 1, 2.0, left, 8) == [(0, 0.5), (1, 0.5)]
    assert R.axis_taps(ChromaFilter.Bilinear, 15, 2.0, c, 8)[-1][0] == 7  # mirrored edge
    for flt in ChromaFilter:
        for out in range(16):
            assert sum(w for _, w in R.axis_taps(flt, out, 2.0, c, 8)) == pytest.approx(1.0)
    # Co-sited samples reproduce the source exactly under every filter.
    for flt in ChromaFilter:
        taps = R.axis_taps(flt, 2, 2.0, left, 8)
        assert sum(w for i, w in taps if i == 1) == pytest.approx(1.0, abs=1e-12)


def test_format_reference_pack_layout():
    """format_ref.pack's plane layout agrees with the library's RequiredBytes
    for every family, subsampling and alpha combination it packs."""
    import format_ref as R
    from vqm_amd import ColorPrimaries as CP, MatrixCoefficients as MC
    for fam, mat, sub, alpha in [(ColorFamily.Rgb, MC.Rgb, (0, 0), AlphaType.NoAlpha),
                                 (ColorFamily.Grey, MC.Unspecified, (0, 0), AlphaType.Straight),
                                 (ColorFamily.Yuv, MC.Bt709, (1, 1), AlphaType.Premultiplied),
                                 (ColorFamily.Yuv, MC.Bt709, (2, 0), AlphaType.NoAlpha)]:
        for t in PixelType:
            f = ImageFormat(type=t, family=fam, matrix=mat, alpha=alpha, subsample_w=sub[0],
                            subsample_h=sub[1], primaries=CP.Bt709)
            size = Size(12, 6)
            vals = {p: np.full((h, w), 0.25) for p, w, h in R.plane_shapes(f, size)}
            img = R.pack(f, size, vals)
            assert len(img.data) == vqm_amd.required_bytes(f, size)
