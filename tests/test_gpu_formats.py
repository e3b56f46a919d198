"""The general input surface on the HIP path: every sample type, range, transfer,
matrix, primaries set, chroma siting and filter, alpha type and family the
reference accepts, checked the reference's way
(tests/input/input_format_test_support.h:4-15): frame A in the format under
test, frame B as linear BT.709 floats computed in FP64 by format_ref, and

    max |norms(A, B) - norms(B, B)| <= tolerance

Cases and tolerances follow tests/input/input_format_test.cc:205-462 and
tests/input/input_format_chroma_test.cc:100-430 (2e-4 for rounding, 4e-4 for
the PQ and HLG curves, 1e-2 for limited against full range)."""
from __future__ import annotations

import numpy as np
import pytest

import format_ref as R
from conftest import gradient_image
from vqm_amd import (AlphaType, ChromaFilter, ChromaLocation, ColorFamily, ColorPrimaries as CP,
                     Image, ImageFormat, MatrixCoefficients as MC, PixelRange, PixelType, Size,
                     Ssimulacra2Config, Ssimulacra2Pipeline, TransferCharacteristics as TC)

pytestmark = pytest.mark.gpu

SIDE = 32
S = Size(SIDE, SIDE)
ROUNDING = 2e-4
PQ_CURVE = 4e-4
HLG_CURVE = 4e-4

_PIPELINES = {}


@pytest.fixture(autouse=True, scope="module")
def _close_pipelines():
    yield
    for p in _PIPELINES.values():
        p.close()
    _PIPELINES.clear()


def pipeline(flt: int = ChromaFilter.Bilinear) -> Ssimulacra2Pipeline:
    if flt not in _PIPELINES:
        _PIPELINES[flt] = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, chroma_filter=int(flt)))
    return _PIPELINES[flt]


def max_norm_delta(a, b) -> float:
    assert len(a.scales) == len(b.scales)
    return float(np.max(np.abs(a.norms() - b.norms())))


def check_agrees(p, a: Image, b: Image, tol: float, what: str, background: float = 0.5):
    pair = p.Compute(a, b, background)
    identity = p.Compute(b, b, background)
    delta = max_norm_delta(pair, identity)
    assert delta <= tol, f"{what}: conversion differs by {delta:.3e} (tolerance {tol:.1e})"
    return delta


def planes3(f) -> dict:
    """Per-pixel content from f(x, y) -> 3 values."""
    out = np.zeros((3, SIDE, SIDE))
    for y in range(SIDE):
        for x in range(SIDE):
            out[:, y, x] = f(x, y)
    return {0: out[0], 1: out[1], 2: out[2]}


def uniform(v, alpha: float = 1.0) -> dict:
    return {i: np.full((SIDE, SIDE), float(v[i])) for i in range(3)} | {3: np.full((SIDE, SIDE), alpha)}


def ramp(x: int, y: int) -> float:
    """The diagonal ramp every A/B case sweeps, in [0.02, 0.98]
    (input_format_test.cc:54-62): structured content, since uniform frames
    saturate the norms at ~1e-4 whatever the conversion's accuracy."""
    return 0.02 + 0.96 * (x + y) / (2.0 * (SIDE - 1))


def ladder(grid: int) -> dict:
    """input_format_test.cc:10-32: samples on a `grid`-level ladder, exact in
    any type with log2(grid)+1 bits or more."""
    y, x = np.mgrid[0:SIDE, 0:SIDE]
    return {0: (x * 7 % (grid + 1)) / grid, 1: (y * 13 % (grid + 1)) / grid,
            2: ((x * 11 + y * 5) % (grid + 1)) / grid}


def linear(values: dict) -> Image:
    return R.pack(R.linear_rgb_format(), S, values)


def yuv_format(sub_w=0, sub_h=0, location=ChromaLocation.Left, matrix=MC.Bt709,
               transfer=TC.Linear, primaries=CP.Bt709) -> ImageFormat:
    return ImageFormat(type=PixelType.Float, family=ColorFamily.Yuv, matrix=matrix, transfer=transfer,
                       primaries=primaries, subsample_w=sub_w, subsample_h=sub_h,
                       chroma_location=location)


# ---- sample types, depth and range (input_format_test.cc:205-286) ---------------

def test_srgb8_table_path_matches_the_general_decode():
    """The u8 sRGB table path and the per-pixel decode of the same content as
    Float sRGB see the same linear image."""
    orig = gradient_image(0)
    vals = {i: orig[i * 1024:(i + 1) * 1024].reshape(SIDE, SIDE) / 255.0 for i in range(3)}
    table = Image(S, ImageFormat.srgb8(), orig)
    general = R.pack(R.pinned_format(), S, vals)
    check_agrees(pipeline(), table, general, ROUNDING, "u8 table vs float sRGB")
    ref = linear({i: np.vectorize(lambda v: R.decode(TC.Srgb, v))(vals[i]) for i in range(3)})
    check_agrees(pipeline(), table, ref, ROUNDING, "u8 table vs FP64 linear")


def test_float_half_word_agree():
    base = R.pinned_format()
    half = R.pinned_format(type_=PixelType.Half)
    v = ladder(1024)
    check_agrees(pipeline(), R.pack(half, S, v), R.pack(base, S, v), ROUNDING, "half vs float")
    word = R.pinned_format(type_=PixelType.Word)
    word.depth = 16
    v = ladder(65535)
    check_agrees(pipeline(), R.pack(word, S, v), R.pack(base, S, v), ROUNDING, "word16 vs float")


def test_integer_quantisation_agrees():
    base = R.pinned_format()
    byte = R.pinned_format(type_=PixelType.Byte)
    v = ladder(255)
    check_agrees(pipeline(), R.pack(byte, S, v), R.pack(base, S, v), ROUNDING, "byte vs float")
    word10 = R.pinned_format(type_=PixelType.Word)
    word10.depth = 10
    v10 = ladder(1023)
    check_agrees(pipeline(), R.pack(word10, S, v10), R.pack(base, S, v10), ROUNDING, "word10 vs float")
    limited = R.pinned_format(type_=PixelType.Byte)
    limited.range = PixelRange.Limited
    full_frame, limited_frame = R.pack(byte, S, v), R.pack(limited, S, v)
    p = pipeline()
    delta = max_norm_delta(p.Compute(limited_frame, full_frame), p.Compute(full_frame, full_frame))
    # The two quantisations differ by up to half a limited code step (1/438);
    # a wrong limited mapping is off by orders of magnitude more.
    assert delta <= 1e-2, f"limited vs full: {delta:.3e}"
    assert delta > 0.0


def test_full_range_byte_chroma_absolute_codes():
    """input_format_chroma_test.cc:104-139: codes written directly, so the
    absolute mapping is checked independently of the packer: 8-bit full-range
    chroma is centred on code 128."""
    fmt = yuv_format()
    fmt.type = PixelType.Byte
    data = np.zeros(3 * SIDE * SIDE, np.uint8)
    data[:1024], data[1024:2048], data[2048:] = 128, 128, 255
    coded = Image(S, fmt, data)
    reference = R.pack(yuv_format(), S, uniform((128 / 255, 0.0, 127 / 255)))
    check_agrees(pipeline(), coded, reference, ROUNDING, "full-range byte chroma")


def test_limited_range_word_codes():
    """10-bit limited range: luma 64..940, chroma 64..960 about 512, written
    as codes and compared with their decoded FP64 values."""
    fmt = yuv_format()
    fmt.type, fmt.depth, fmt.range = PixelType.Word, 10, PixelRange.Limited
    y, x = np.mgrid[0:SIDE, 0:SIDE]
    luma = 64 + (x * 27 + y * 3) % 877
    cb = 64 + (x * 13 + y * 17) % 897
    cr = 64 + (x * 5 + y * 29) % 897
    data = np.concatenate([luma.ravel(), cb.ravel(), cr.ravel()]).astype(np.uint16)
    coded = Image(S, fmt, data.view(np.uint8))
    vals = {0: (luma - 64) / 876.0, 1: (cb - 512) / 896.0, 2: (cr - 512) / 896.0}
    m = R.ncl_to_rgb(MC.Bt709)
    yuv = np.stack([vals[0], vals[1], vals[2]])
    rgb = np.einsum("ij,jyx->iyx", m, yuv)
    check_agrees(pipeline(), coded, linear({i: rgb[i] for i in range(3)}), ROUNDING, "limited word10")


# ---- transfer functions (input_format_test.cc:288-378) ----------------------------

TRANSFERS = [
    (TC.Bt709, False), (TC.Bt601, False), (TC.Bt2020_10, False), (TC.Bt2020_12, False),
    (TC.St240M, False), (TC.Bt470M, False), (TC.Bt470Bg, False), (TC.Srgb, False),
    (TC.Log100, True), (TC.Log316, True), (TC.Iec61966_2_4, True), (TC.Bt1361, True),
    (TC.St428, False), (TC.Linear, False), (TC.St2084, True), (TC.St2084, False),
    (TC.AribB67, True), (TC.AribB67, False),
]


@pytest.mark.parametrize("transfer,upper", TRANSFERS, ids=lambda v: str(int(v)))
def test_transfer_functions(transfer, upper):
    fmt = R.pinned_format()
    fmt.transfer = transfer

    def sample(x, y):
        v = ramp(x, y)
        return 0.5 + 0.5 * v if upper else 0.5 * v

    def decoded(x, y):
        v = sample(x, y)
        if transfer == TC.AribB67:
            lin = R.hlg_to_linear((v, v, v), 0.2126, 0.7152, 0.0722)[0]
        else:
            lin = R.decode(transfer, v)
        return (lin, lin, lin)

    a = R.pack(fmt, S, planes3(lambda x, y: (sample(x, y),) * 3))
    b = linear(planes3(decoded))
    tol = PQ_CURVE if transfer == TC.St2084 else HLG_CURVE if transfer == TC.AribB67 else ROUNDING
    check_agrees(pipeline(), a, b, tol, f"transfer {int(transfer)}")


def test_hlg_luminance_ootf_on_chromatic_content():
    """The OOTF scales the triple by Y^0.2, which differs from a per-channel
    1.2 power for every non-neutral pixel (input_format_test.cc:364-376)."""
    fmt = R.pinned_format()
    fmt.transfer = TC.AribB67
    hlg = (0.4, 0.55, 0.7)
    a = R.pack(fmt, S, uniform(hlg))
    b = linear(uniform(R.hlg_to_linear(hlg, 0.2126, 0.7152, 0.0722)))
    check_agrees(pipeline(), a, b, HLG_CURVE, "transfer HLG chromatic")
    per_channel = linear(uniform([10.0 * R.hlg_inverse_oetf(v) ** 1.2 for v in hlg]))
    p = pipeline()
    assert max_norm_delta(p.Compute(a, per_channel), p.Compute(per_channel, per_channel)) > HLG_CURVE


# ---- matrices and primaries (input_format_test.cc:380-462) ------------------------

MATRICES = [
    (MC.Bt709, CP.Bt709, False), (MC.Bt470Bg, CP.Bt709, False), (MC.St170M, CP.Bt709, False),
    (MC.St240M, CP.Bt709, False), (MC.Fcc, CP.Bt709, False), (MC.YCgCo, CP.Bt709, False),
    (MC.Bt2020Ncl, CP.Bt2020, False), (MC.Bt2020Ncl, CP.Bt470Bg, False),
    (MC.ChromaticityDerivedNcl, CP.Bt2020, True), (MC.ChromaticityDerivedNcl, CP.Bt470Bg, True),
    (MC.ChromaticityDerivedNcl, CP.Ebu3213E, True),
]


@pytest.mark.parametrize("matrix,primaries,derived", MATRICES, ids=lambda v: str(int(v)))
def test_matrix_coefficients(matrix, primaries, derived):
    fmt = yuv_format(matrix=matrix, primaries=primaries)

    def plane(x, y):
        t = ramp(x, y)
        return (0.1 + 0.8 * t, 0.25 - 0.4 * t, -0.25 + 0.5 * t)

    ncl = R.derived_ncl_to_rgb(primaries) if derived else R.ncl_to_rgb(matrix)
    to_709 = R.gamut(primaries)
    b = linear(planes3(lambda x, y: to_709 @ (ncl @ np.array(plane(x, y)))))
    check_agrees(pipeline(), R.pack(fmt, S, planes3(plane)), b, ROUNDING,
                 f"matrix {int(matrix)} primaries {int(primaries)}")


PRIMARIES = [CP.Bt470M, CP.Bt470Bg, CP.St170M, CP.St240M, CP.Film, CP.Bt2020, CP.St428,
             CP.St431_2, CP.St432_1, CP.Ebu3213E]


@pytest.mark.parametrize("primaries", PRIMARIES, ids=lambda v: str(int(v)))
def test_primaries(primaries):
    fmt = R.linear_rgb_format()
    fmt.primaries = primaries

    def sample(x, y):
        t = ramp(x, y)
        return (0.1 + 0.8 * t, 0.9 - 0.8 * t, 0.05 + 0.5 * t)

    m = R.gamut(primaries)
    b = linear(planes3(lambda x, y: m @ np.array(sample(x, y))))
    check_agrees(pipeline(), R.pack(fmt, S, planes3(sample)), b, ROUNDING, f"primaries {int(primaries)}")


def test_greyscale():
    """Grey: one plane, R = G = B = Y, through the transfer curve."""
    fmt = ImageFormat(type=PixelType.Word, depth=12, family=ColorFamily.Grey, matrix=MC.Unspecified,
                      transfer=TC.Bt709, primaries=CP.Bt709)
    v = np.array([[ramp(x, y) for x in range(SIDE)] for y in range(SIDE)])
    codes = np.floor(v * 4095 + 0.5)
    a = Image(S, fmt, codes.astype(np.uint16).view(np.uint8))
    lin = (codes / 4095.0) ** 2.4
    check_agrees(pipeline(), a, linear({0: lin, 1: lin, 2: lin}), ROUNDING, "grey word12")


# ---- constant luminance and ICtCp (input_format_chroma_test.cc:268-350) -----------

@pytest.mark.parametrize("matrix,primaries", [(MC.Bt2020Cl, CP.Bt2020),
                                              (MC.ChromaticityDerivedCl, CP.Bt709),
                                              (MC.ChromaticityDerivedCl, CP.Bt470Bg)],
                         ids=lambda v: str(int(v)))
def test_constant_luminance(matrix, primaries):
    yuv = (0.5, 0.1, -0.2)
    fmt = yuv_format(matrix=matrix, transfer=TC.Bt709, primaries=primaries)
    kr, kb = R.KRKB[MC.Bt2020Cl] if matrix == MC.Bt2020Cl else R.derived_krkb(primaries)
    rgb = R.gamut(primaries) @ np.array(R.cl_to_rgb(kr, kb, yuv))
    check_agrees(pipeline(), R.pack(fmt, S, uniform(yuv)), linear(uniform(rgb)), ROUNDING,
                 f"CL matrix {int(matrix)} primaries {int(primaries)}")


@pytest.mark.parametrize("transfer", [TC.St2084, TC.AribB67], ids=lambda v: str(int(v)))
def test_ictcp(transfer):
    ictcp = (0.5, 0.15, -0.12)
    fmt = yuv_format(matrix=MC.Bt2100ICtCp, transfer=transfer, primaries=CP.Bt2020)
    lms = np.linalg.solve(R.lms_to_ictcp(transfer), np.array(ictcp))
    if transfer == TC.AribB67:
        dec = np.array(R.hlg_to_linear(lms, 0.2627, 0.6780, 0.0593))
    else:
        dec = np.array([R.decode(transfer, v) for v in lms])
    rgb = R.gamut(CP.Bt2020) @ (R.lms_to_rgb2020() @ dec)
    tol = HLG_CURVE if transfer == TC.AribB67 else ROUNDING
    check_agrees(pipeline(), R.pack(fmt, S, uniform(ictcp)), linear(uniform(rgb)), tol,
                 f"ictcp transfer {int(transfer)}")


# ---- chroma subsampling (input_format_chroma_test.cc:12-99,141-266) ----------------

def subsampled_content(sub_w: int, sub_h: int) -> dict:
    """A luma ramp with a sharp chroma edge, so the upsampling kernel shows."""
    cw, ch = SIDE >> sub_w, SIDE >> sub_h
    y, x = np.mgrid[0:SIDE, 0:SIDE]
    cy, cx = np.mgrid[0:ch, 0:cw]
    return {0: (x / (SIDE - 1)).astype(np.float32).astype(np.float64),
            1: np.where(cx < cw // 2, 0.2, 0.8).astype(np.float32).astype(np.float64),
            2: np.where(cy < ch // 2, 0.7, 0.3).astype(np.float32).astype(np.float64)}


def upsampled_content(sub_w, sub_h, flt, location) -> dict:
    v = subsampled_content(sub_w, sub_h)
    return {0: v[0], 1: R.upsample(v[1], sub_w, sub_h, flt, location, SIDE, SIDE),
            2: R.upsample(v[2], sub_w, sub_h, flt, location, SIDE, SIDE)}


SUBSAMPLING = [
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.Left),
    (1, 1, ChromaFilter.Bicubic, ChromaLocation.Left),
    (1, 1, ChromaFilter.Lanczos, ChromaLocation.Left),
    (1, 0, ChromaFilter.Bilinear, ChromaLocation.Left),
    (2, 0, ChromaFilter.Bilinear, ChromaLocation.Left),
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.Center),
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.TopLeft),
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.Top),
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.BottomLeft),
    (1, 1, ChromaFilter.Bilinear, ChromaLocation.Bottom),
    (1, 1, ChromaFilter.Bicubic, ChromaLocation.Bottom),
    (2, 2, ChromaFilter.Lanczos, ChromaLocation.Center),
    (0, 1, ChromaFilter.Bicubic, ChromaLocation.Top),
]


@pytest.mark.parametrize("sub_w,sub_h,flt,location", SUBSAMPLING, ids=lambda v: str(int(v)))
def test_subsampling_matches_reference(sub_w, sub_h, flt, location):
    a = R.pack(yuv_format(sub_w, sub_h, location), S, subsampled_content(sub_w, sub_h))
    b = R.pack(yuv_format(0, 0, location), S, upsampled_content(sub_w, sub_h, flt, location))
    check_agrees(pipeline(flt), a, b, ROUNDING,
                 f"subsampling {sub_w}x{sub_h} filter {int(flt)} siting {int(location)}")


# 4x on one axis with 1x or 2x on the other, beyond the reference's cases.
MIXED_RATIOS = [
    (0, 2, ChromaFilter.Lanczos, ChromaLocation.Top),
    (1, 2, ChromaFilter.Bicubic, ChromaLocation.BottomLeft),
    (2, 1, ChromaFilter.Bilinear, ChromaLocation.TopLeft),
]


@pytest.mark.parametrize("sub_w,sub_h,flt,location", SUBSAMPLING + MIXED_RATIOS, ids=lambda v: str(int(v)))
def test_subsampling_across_tiles(sub_w, sub_h, flt, location):
    """The same A/B check on a 136 x 72 frame: 3 x 3 prepare tiles of 64 x 32
    pixels, the last row and column ragged, each tile upsampling from its own
    window of chroma taps. The chroma samples are independent random values,
    so a tap read from the wrong sample moves the norms far past the
    tolerance."""
    w, h = 136, 72
    size = Size(w, h)
    rng = np.random.default_rng(100 * sub_w + 10 * sub_h + int(flt) + 1000 * int(location))
    cw, ch = w >> sub_w, h >> sub_h
    y, x = np.mgrid[0:h, 0:w]
    planes = {0: (0.2 + 0.6 * (x + y) / (w + h)).astype(np.float32).astype(np.float64),
              1: rng.uniform(-0.2, 0.2, (ch, cw)).astype(np.float32).astype(np.float64),
              2: rng.uniform(-0.2, 0.2, (ch, cw)).astype(np.float32).astype(np.float64)}
    up = {0: planes[0], 1: R.upsample(planes[1], sub_w, sub_h, flt, location, w, h),
          2: R.upsample(planes[2], sub_w, sub_h, flt, location, w, h)}
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size, chroma_filter=int(flt))) as p:
        a = R.pack(yuv_format(sub_w, sub_h, location), size, planes)
        b = R.pack(yuv_format(0, 0, location), size, up)
        pair, identity = p.Compute(a, b), p.Compute(b, b)
        delta = max_norm_delta(pair, identity)
        assert delta <= ROUNDING, (f"subsampling {sub_w}x{sub_h} filter {int(flt)} siting {int(location)}: "
                                   f"{delta:.3e}")


def test_chroma_filter_and_siting_change_the_score():
    """input_format_chroma_test.cc:185-266: the three filters give different
    scores on a chroma-shifted pair; an edge scores differently left- and
    centre-sited; identical pairs stay identical."""
    fmt = yuv_format(1, 1)
    content = R.pack(fmt, S, subsampled_content(1, 1))
    shifted = subsampled_content(1, 1)
    shifted[1] = np.roll(shifted[1].ravel(), -1).reshape(shifted[1].shape)
    distorted = R.pack(fmt, S, shifted)
    scores = [pipeline(f).Compute(content, distorted).Score() for f in ChromaFilter]
    assert abs(scores[0] - scores[1]) > 1e-3
    assert abs(scores[0] - scores[2]) > 1e-3
    assert abs(scores[1] - scores[2]) > 1e-3
    edge = subsampled_content(1, 1)
    edge[1] = np.where(np.arange(edge[1].size).reshape(edge[1].shape) == 0, 0.0, 1.0)
    left = R.pack(yuv_format(1, 1, ChromaLocation.Left), S, edge)
    center = R.pack(yuv_format(1, 1, ChromaLocation.Center), S, edge)
    p = pipeline()
    left_score = p.Compute(left, left).Score()
    assert left_score == pytest.approx(p.Compute(center, center).Score(), abs=1e-9)
    assert abs(p.Compute(left, center).Score() - left_score) > 1e-3


# ---- alpha (input_format_chroma_test.cc:352-397) ---------------------------------

def test_alpha_types():
    """Straight colour S with alpha A equals premultiplied S*A, blended over
    the background in the encoded domain; a zero premultiplied alpha does
    not divide by zero."""
    y, x = np.mgrid[0:SIDE, 0:SIDE]
    a = ((x + 3 * y) % 7) / 6.0
    s = [x / (SIDE - 1), y / (SIDE - 1), ((x * 7 + y * 11) % 64) / 63.0]
    straight = {0: s[0], 1: s[1], 2: s[2], 3: a}
    premul = {0: s[0] * a, 1: s[1] * a, 2: s[2] * a, 3: a}
    fa, fb = R.pinned_format(AlphaType.Straight), R.pinned_format(AlphaType.Premultiplied)
    ia, ib = R.pack(fa, S, straight), R.pack(fb, S, premul)
    p = pipeline()
    for background in (0.1, 0.9):
        check_agrees(p, ia, ib, ROUNDING, "alpha types", background)
        blended = {c: np.vectorize(lambda v: R.decode(TC.Srgb, v))(s[c] * a + background * (1 - a))
                   for c in range(3)}
        check_agrees(p, ia, linear(blended), ROUNDING, f"alpha blend over {background}", background)
    zero = R.pack(fb, S, {i: np.zeros((SIDE, SIDE)) for i in range(4)})
    res = p.Compute(zero, zero)
    assert np.all(np.isfinite(res.norms()))
    score = res.Score()
    assert np.isfinite(score) and score <= 100.0


def test_mixed_table_and_general_pair_and_back():
    """A pair of one u8 sRGB and one general image, then the u8 pair again:
    the per-pair decode choice does not leak between submissions."""
    orig, dist = gradient_image(0), gradient_image(1)
    p = pipeline()
    u8 = p.Compute(Image(S, ImageFormat.srgb8(), orig), Image(S, ImageFormat.srgb8(), dist))
    vals = {i: dist[i * 1024:(i + 1) * 1024].reshape(SIDE, SIDE) / 255.0 for i in range(3)}
    word = R.pinned_format(type_=PixelType.Word)  # k/255 is the exact code 257k
    mixed = p.Compute(Image(S, ImageFormat.srgb8(), orig), R.pack(word, S, vals))
    assert max_norm_delta(mixed, u8) <= ROUNDING
    again = p.Compute(Image(S, ImageFormat.srgb8(), orig), Image(S, ImageFormat.srgb8(), dist))
    assert np.array_equal(again.norms(), u8.norms())


def test_hdr_yuv420_at_360p():
    """A 640x360 10-bit limited-range BT.2020 PQ 4:2:0 frame (Lanczos chroma,
    top-left siting) against its FP64 decode in linear BT.709: the A/B bar of
    the PQ curve at a real frame size, on structured content."""
    w, h = 640, 360
    size = Size(w, h)
    y, x = np.mgrid[0:h, 0:w]
    luma = 0.3 + 0.15 * np.sin(x / 23.0) * np.cos(y / 17.0) + 0.08 * ((x // 40 + y // 30) % 2)
    cy, cx = np.mgrid[0:h // 2, 0:w // 2]
    cb = 0.08 * np.sin(cx / 11.0 + cy / 29.0)
    cr = -0.06 * np.cos(cx / 13.0 - cy / 7.0)
    fmt = ImageFormat(type=PixelType.Word, depth=10, range=PixelRange.Limited, family=ColorFamily.Yuv,
                      matrix=MC.Bt2020Ncl, transfer=TC.St2084, primaries=CP.Bt2020, subsample_w=1,
                      subsample_h=1, chroma_location=ChromaLocation.TopLeft)
    a = R.pack(fmt, size, {0: luma, 1: cb, 2: cr})
    # The FP64 decode of the packed codes: codes -> normalised, upsample,
    # matrix, PQ, gamut.
    raw = np.frombuffer(a.data.tobytes(), np.uint16).astype(np.float64)
    n_l, n_c = w * h, (w // 2) * (h // 2)
    stride_l = (n_l * 2 + 3) // 4 * 4 // 2
    stride_c = (n_c * 2 + 3) // 4 * 4 // 2
    yl = (raw[:n_l].reshape(h, w) - 64) / 876.0
    u = (raw[stride_l:stride_l + n_c].reshape(h // 2, w // 2) - 512) / 896.0
    v = (raw[stride_l + stride_c:stride_l + stride_c + n_c].reshape(h // 2, w // 2) - 512) / 896.0
    uu = R.upsample(u, 1, 1, ChromaFilter.Lanczos, ChromaLocation.TopLeft, w, h)
    vv = R.upsample(v, 1, 1, ChromaFilter.Lanczos, ChromaLocation.TopLeft, w, h)
    rgb = np.einsum("ij,jyx->iyx", R.ncl_to_rgb(MC.Bt2020Ncl), np.stack([yl, uu, vv]))
    lin = np.vectorize(lambda t: R.decode(TC.St2084, t))(rgb)
    out = np.einsum("ij,jyx->iyx", R.gamut(CP.Bt2020), lin)
    b = R.pack(R.linear_rgb_format(), size, {0: out[0], 1: out[1], 2: out[2]})
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size, chroma_filter=ChromaFilter.Lanczos)) as p:
        check_agrees(p, a, b, PQ_CURVE, "360p HDR 4:2:0")
