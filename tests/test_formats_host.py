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
