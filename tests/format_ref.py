"""An FP64 restatement of the input conversions, written from the standards, so
the A/B tests of the general input surface check the GPU decode against
something other than the library's own host tables.

Every A/B case builds frame A in the format under test and frame B as linear
BT.709 RGB floats computed here; (A, B) must then score like (B, B). The
reference does the same (tests/input/input_format_test_support.h:4-15) with its
own double reference (tests/input/input_reference.h); the formulas are the
published ones:

* BT.709 / BT.1886 OETF pair, IEC 61966-2-1 sRGB, SMPTE ST 2084 (PQ), ARIB
  STD-B67 (HLG) with the BT.2100 luminance OOTF, log 100 / log 316, ST 428,
  with zimg's display-referred scaling (PQ 10000 cd/m2 and HLG 1000 cd/m2 over
  a 100 cd/m2 reference white);
* YUV matrices from (Kr, Kb), YCgCo, chromaticity-derived Kr/Kb
  (ITU-T H.273 eq. 39-44), the BT.2020 constant-luminance decode, BT.2100
  ICtCp;
* RGB->XYZ from primaries and white point (SMPTE RP 177);
* zimg's resampler taps for the chroma upsampling (filter support centred on
  the output sample, mirrored edges, weights normalised per output sample).
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

import vqm_amd
from vqm_amd import (AlphaType, ChromaFilter, ChromaLocation, ColorFamily, ColorPrimaries, Image,
                     ImageFormat, MatrixCoefficients as MC, PixelType, Size,
                     TransferCharacteristics as TC)

# ---- transfer curves ---------------------------------------------------------

_BT709_A = 1.09929682680944
_BT709_B = 0.018053968510807


def bt709_oetf(x: float) -> float:
    x = max(x, 0.0)
    return 4.5 * x if x < _BT709_B else _BT709_A * x ** 0.45 - (_BT709_A - 1.0)


def bt709_inverse_oetf(x: float) -> float:
    x = max(x, 0.0)
    if x < 4.5 * _BT709_B:
        return x / 4.5
    return ((x + _BT709_A - 1.0) / _BT709_A) ** (1.0 / 0.45)


def hlg_inverse_oetf(x: float) -> float:
    a, b, c = 0.17883277, 0.28466892, 0.55991073
    x = max(x, 0.0)
    return x * x / 3.0 if x <= 0.5 else (math.exp((x - c) / a) + b) / 12.0


def decode(transfer: int, v: float) -> float:
    """Display-referred linear light of one encoded sample (HLG: the
    per-channel inverse OETF only; hlg_to_linear applies the OOTF)."""
    t = int(transfer)
    if t == TC.Linear or t == TC.Unspecified:
        return v
    if t in (TC.Bt709, TC.Bt601, TC.Bt2020_10, TC.Bt2020_12, TC.St240M):
        return 0.0 if v < 0.0 else v ** 2.4  # BT.1886 with a zero black level
    if t == TC.Bt470M:
        return 0.0 if v < 0.0 else v ** 2.2
    if t == TC.Bt470Bg:
        return 0.0 if v < 0.0 else v ** 2.8
    if t == TC.Srgb:
        return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4
    if t == TC.Log100:
        return 0.01 if v <= 0.0 else 10.0 ** (2.0 * (v - 1.0))
    if t == TC.Log316:
        return math.sqrt(10.0) / 1000.0 if v <= 0.0 else 10.0 ** (2.5 * (v - 1.0))
    if t == TC.Iec61966_2_4:  # xvYCC: BT.1886 inside [0, 1], odd-symmetric OETF outside
        if 0.0 <= v <= 1.0:
            return v ** 2.4
        return math.copysign(bt709_inverse_oetf(abs(v)), v)
    if t == TC.Bt1361:  # extended gamut: negative part on a quarter scale
        if 0.0 <= v <= 1.0:
            return v ** 2.4
        return bt709_inverse_oetf(v) if v >= 0.0 else -bt709_inverse_oetf(-4.0 * v) / 4.0
    if t == TC.St2084:
        if v <= 0.0:
            return 0.0
        m1, m2 = 2610.0 / 16384.0, 2523.0 / 4096.0 * 128.0
        c1, c2, c3 = 3424.0 / 4096.0, 2413.0 / 4096.0 * 32.0, 2392.0 / 4096.0 * 32.0
        p = v ** (1.0 / m2)
        return 100.0 * (max(p - c1, 0.0) / max(c2 - c3 * p, 1.17549435e-38)) ** (1.0 / m1)
    if t == TC.St428:
        return 0.0 if v < 0.0 else 52.37 / 48.0 * v ** 2.6
    if t == TC.AribB67:
        return hlg_inverse_oetf(v)
    raise ValueError(f"transfer {t}")


def hlg_to_linear(v: Sequence[float], kr: float, kg: float, kb: float,
                  scale: float = 10.0) -> Tuple[float, float, float]:
    """BT.2100 HLG display decode: per-channel inverse OETF, then the luminance
    OOTF Y^(gamma-1) with gamma 1.2 at 1000 cd/m2, times 1000/100."""
    c = [hlg_inverse_oetf(x) for x in v]
    y = max(kr * c[0] + kg * c[1] + kb * c[2], 1.17549435e-38)
    g = y ** 0.2
    return (c[0] * g * scale, c[1] * g * scale, c[2] * g * scale)


# ---- matrices ------------------------------------------------------------------

KRKB: Dict[int, Tuple[float, float]] = {
    MC.Fcc: (0.30, 0.11),
    MC.St240M: (0.212, 0.087),
    MC.Bt470Bg: (0.299, 0.114),
    MC.St170M: (0.299, 0.114),
    MC.Bt709: (0.2126, 0.0722),
    MC.Bt2020Ncl: (0.2627, 0.0593),
    MC.Bt2020Cl: (0.2627, 0.0593),
}


def rgb_to_yuv(kr: float, kb: float) -> np.ndarray:
    kg = 1.0 - kr - kb
    return np.array([[kr, kg, kb],
                     [-kr / (2 - 2 * kb), -kg / (2 - 2 * kb), 0.5],
                     [0.5, -kg / (2 - 2 * kr), -kb / (2 - 2 * kr)]])


def ncl_to_rgb(matrix: int) -> np.ndarray:
    if int(matrix) == MC.YCgCo:
        # Y = R/4 + G/2 + B/4, Cg = -R/4 + G/2 - B/4, Co = (R - B)/2.
        fwd = np.array([[0.25, 0.5, 0.25], [-0.25, 0.5, -0.25], [0.5, 0.0, -0.5]])
        return np.linalg.inv(fwd)
    kr, kb = KRKB[int(matrix)]
    return np.linalg.inv(rgb_to_yuv(kr, kb))


PRIMARIES: Dict[int, Tuple[Tuple[float, float], ...]] = {
    ColorPrimaries.Bt470M: ((0.670, 0.330), (0.210, 0.710), (0.140, 0.080)),
    ColorPrimaries.Bt470Bg: ((0.640, 0.330), (0.290, 0.600), (0.150, 0.060)),
    ColorPrimaries.St170M: ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)),
    ColorPrimaries.St240M: ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)),
    ColorPrimaries.Bt709: ((0.640, 0.330), (0.300, 0.600), (0.150, 0.060)),
    ColorPrimaries.Film: ((0.681, 0.319), (0.243, 0.692), (0.145, 0.049)),
    ColorPrimaries.Bt2020: ((0.708, 0.292), (0.170, 0.797), (0.131, 0.046)),
    ColorPrimaries.St431_2: ((0.680, 0.320), (0.265, 0.690), (0.150, 0.060)),
    ColorPrimaries.St432_1: ((0.680, 0.320), (0.265, 0.690), (0.150, 0.060)),
    ColorPrimaries.Ebu3213E: ((0.630, 0.340), (0.295, 0.605), (0.155, 0.077)),
}


def white_point(primaries: int) -> Tuple[float, float]:
    p = int(primaries)
    if p in (ColorPrimaries.Bt470M, ColorPrimaries.Film):
        return (0.310, 0.316)  # illuminant C
    if p == ColorPrimaries.St428:
        return (1.0 / 3.0, 1.0 / 3.0)  # equal-energy E
    if p == ColorPrimaries.St431_2:
        return (0.314, 0.351)  # DCI white
    return (0.3127, 0.3290)  # D65


def _xyz(xy: Tuple[float, float]) -> np.ndarray:
    x, y = xy
    return np.array([x / y, 1.0, (1.0 - x - y) / y])


def rgb_to_xyz(primaries: int) -> np.ndarray:
    """SMPTE RP 177: columns are the primaries' XYZ scaled so RGB=1 is white.
    ST 428 is XYZ itself."""
    if int(primaries) == ColorPrimaries.St428:
        return np.eye(3)
    m = np.stack([_xyz(xy) for xy in PRIMARIES[int(primaries)]], axis=1)
    s = np.linalg.solve(m, _xyz(white_point(primaries)))
    return m * s[None, :]


def gamut(src: int, dst: int = ColorPrimaries.Bt709) -> np.ndarray:
    return np.linalg.solve(rgb_to_xyz(dst), rgb_to_xyz(src))


def derived_krkb(primaries: int) -> Tuple[float, float]:
    """ITU-T H.273 eq. 39-44; zimg reads BT.709 and BT.2020 off their
    published constants."""
    p = int(primaries)
    if p == ColorPrimaries.Bt709:
        return KRKB[MC.Bt709]
    if p == ColorPrimaries.Bt2020:
        return KRKB[MC.Bt2020Ncl]
    m = rgb_to_xyz(p)
    return float(m[1, 0]), float(m[1, 2])  # the Y row: luminance of R and B


def derived_ncl_to_rgb(primaries: int) -> np.ndarray:
    kr, kb = derived_krkb(primaries)
    return np.linalg.inv(rgb_to_yuv(kr, kb))


def cl_to_rgb(kr: float, kb: float, yuv: Sequence[float]) -> Tuple[float, float, float]:
    """BT.2020 constant luminance decode: Y', B'-Y' and R'-Y' with the
    sign-dependent scale factors, then linear G from the luminance equation."""
    kg = 1.0 - kr - kb
    nb, pb = bt709_oetf(1.0 - kb), 1.0 - bt709_oetf(kb)
    nr, pr = bt709_oetf(1.0 - kr), 1.0 - bt709_oetf(kr)
    b = bt709_inverse_oetf(yuv[0] + yuv[1] * 2.0 * (nb if yuv[1] < 0.0 else pb))
    r = bt709_inverse_oetf(yuv[0] + yuv[2] * 2.0 * (nr if yuv[2] < 0.0 else pr))
    y = bt709_inverse_oetf(yuv[0])
    return (r, (y - kr * r - kb * b) / kg, b)


def lms_to_ictcp(transfer: int) -> np.ndarray:
    """BT.2100 table 7: the PQ and HLG forms of the I'Ct'Cp matrix."""
    if int(transfer) == TC.AribB67:
        return np.array([[2048, 2048, 0], [3625, -7465, 3840], [9500, -9212, -288]]) / 4096.0
    return np.array([[2048, 2048, 0], [6610, -13613, 7003], [17933, -17390, -543]]) / 4096.0


def lms_to_rgb2020() -> np.ndarray:
    return np.linalg.inv(np.array([[1688, 2146, 262], [683, 2951, 462], [99, 309, 3688]]) / 4096.0)


# ---- chroma upsampling -----------------------------------------------------------

def _sinc(x: float) -> float:
    return 1.0 if x == 0.0 else math.sin(math.pi * x) / (math.pi * x)


def filter_kernel(flt: int, x: float) -> float:
    x = abs(x)
    if flt == ChromaFilter.Bilinear:
        return max(1.0 - x, 0.0)
    if flt == ChromaFilter.Bicubic:  # Catmull-Rom (B=0, C=1/2)
        if x < 1.0:
            return 1.0 - 2.5 * x * x + 1.5 * x ** 3
        return 2.0 - 4.0 * x + 2.5 * x * x - 0.5 * x ** 3 if x < 2.0 else 0.0
    return _sinc(x) * _sinc(x / 3.0) if x < 3.0 else 0.0  # Lanczos-3


def filter_support(flt: int) -> int:
    return {ChromaFilter.Bilinear: 2, ChromaFilter.Bicubic: 4, ChromaFilter.Lanczos: 6}[int(flt)]


def axis_taps(flt: int, out: int, scale: float, shift: float, n: int) -> List[Tuple[int, float]]:
    """(source index, weight) for output sample `out` of an axis upsampled by
    `scale` from `n` samples: the filter's support placed around the output
    centre mapped into source coordinates, the first tap rounded half up so
    that integer steps of the position move every tap by one, indices mirrored
    at the edges and weights normalised to one."""
    taps = filter_support(flt)
    pos = (out + 0.5) / scale + shift
    first = math.floor(pos - taps / 2.0 + 0.5) + 0.5
    weights = [filter_kernel(flt, first + j - pos) for j in range(taps)]
    total = sum(weights)
    result = []
    for j in range(taps):
        x = first + j
        if x < 0.0:
            x = -x
        elif x >= n:
            x = 2.0 * n - x
        result.append((int(math.floor(min(max(x, 0.0), n - 1.0))), weights[j] / total))
    return result


def chroma_shift(location: int, horizontal: bool, sub: int) -> float:
    """Chroma sample position relative to the luma grid, in chroma samples:
    left-sited chroma sits on the first luma column, top/bottom-sited on the
    first/last luma row of its block."""
    off = 0.5 - 0.5 / (1 << sub)
    loc = int(location)
    if horizontal:
        return off if loc in (ChromaLocation.Left, ChromaLocation.TopLeft, ChromaLocation.BottomLeft) else 0.0
    if loc in (ChromaLocation.TopLeft, ChromaLocation.Top):
        return off
    if loc in (ChromaLocation.BottomLeft, ChromaLocation.Bottom):
        return -off
    return 0.0


def upsample(plane: np.ndarray, sub_w: int, sub_h: int, flt: int, location: int,
             out_w: int, out_h: int) -> np.ndarray:
    ch, cw = plane.shape
    sx, sy = chroma_shift(location, True, sub_w), chroma_shift(location, False, sub_h)
    wx = np.zeros((out_w, cw))
    wy = np.zeros((out_h, ch))
    for x in range(out_w):
        for i, w in axis_taps(flt, x, float(1 << sub_w), sx, cw):
            wx[x, i] += w
    for y in range(out_h):
        for i, w in axis_taps(flt, y, float(1 << sub_h), sy, ch):
            wy[y, i] += w
    return wy @ plane.astype(np.float64) @ wx.T


# ---- packing ------------------------------------------------------------------------

def effective_depth(fmt: ImageFormat) -> int:
    if fmt.depth:
        return int(fmt.depth)
    return {PixelType.Byte: 8, PixelType.Word: 16, PixelType.Half: 16, PixelType.Float: 32}[int(fmt.type)]


def plane_shapes(fmt: ImageFormat, size: Size) -> List[Tuple[int, int, int]]:
    """(plane index, width, height) of each present plane, in storage order."""
    planes = [0] if int(fmt.family) == ColorFamily.Grey else [0, 1, 2]
    if int(fmt.alpha) != AlphaType.NoAlpha:
        planes.append(3)
    out = []
    for p in planes:
        w, h = size.width, size.height
        if p in (1, 2) and int(fmt.family) == ColorFamily.Yuv:
            w, h = w >> int(fmt.subsample_w), h >> int(fmt.subsample_h)
        out.append((p, w, h))
    return out


def pack(fmt: ImageFormat, size: Size, values: Dict[int, np.ndarray]) -> Image:
    """An Image of `fmt` whose planes decode to `values` (plane index -> HxW
    array of decoded, normalised samples): codes are (v - offset) / scale,
    clamped to the code range and rounded half up for integer types, RNE
    binary16 for Half. Planes are packed on 4-byte boundaries."""
    elem = {PixelType.Byte: 1, PixelType.Word: 2, PixelType.Half: 2, PixelType.Float: 4}[int(fmt.type)]
    integer = int(fmt.type) in (PixelType.Byte, PixelType.Word)
    total = vqm_amd.required_bytes(fmt, size)
    buf = np.zeros(total, np.uint8)
    off = 0
    for p, w, h in plane_shapes(fmt, size):
        chroma = p in (1, 2) and int(fmt.family) == ColorFamily.Yuv
        scale, offset = vqm_amd.plane_scale_offset(fmt, chroma=chroma, alpha=(p == 3))
        v = np.asarray(values[p], np.float64).reshape(h, w)
        code = (v - float(offset)) / float(scale)
        if integer:
            code = np.floor(np.clip(code, 0.0, float((1 << effective_depth(fmt)) - 1)) + 0.5)
            raw = code.astype({1: np.uint8, 2: np.uint16}[elem])
        elif int(fmt.type) == PixelType.Half:
            raw = code.astype(np.float32).astype(np.float16)
        else:
            raw = code.astype(np.float32)
        nbytes = w * h * elem
        buf[off:off + nbytes] = np.frombuffer(raw.tobytes(), np.uint8)
        off += (nbytes + 3) // 4 * 4
    assert off == total, (off, total)
    return Image(size, fmt, buf)


def linear_rgb_format() -> ImageFormat:
    return ImageFormat(type=PixelType.Float, family=ColorFamily.Rgb, matrix=MC.Rgb,
                       transfer=TC.Linear, primaries=ColorPrimaries.Bt709)


def pinned_format(alpha: int = AlphaType.NoAlpha, type_: int = PixelType.Float) -> ImageFormat:
    return ImageFormat(type=type_, family=ColorFamily.Rgb, alpha=alpha, matrix=MC.Rgb,
                       transfer=TC.Srgb, primaries=ColorPrimaries.Bt709)
