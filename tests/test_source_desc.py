"""The decode descriptor the prepare kernel receives (host only): matrices,
curve ids, luma weights, constant-luminance constants, chroma siting and
plane layout for every matrix, primaries set and transfer the reference
accepts, against format_ref's FP64 restatement of the standards. Built from
tests/cpp/source_desc_dump.cc with the library's own host sources."""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

import format_ref as R
from conftest import ROOT
from vqm_amd import (AlphaType, ChromaLocation, ColorFamily, ColorPrimaries as CP, ImageFormat,
                     MatrixCoefficients as MC, PixelRange, PixelType, Size,
                     TransferCharacteristics as TC)

CURVE = {  # the kernel's curve ids (TransferId) and peak factors
    TC.Linear: (0, 1.0), TC.Bt709: (1, 1.0), TC.Bt601: (1, 1.0), TC.Bt2020_10: (1, 1.0),
    TC.Bt2020_12: (1, 1.0), TC.St240M: (1, 1.0), TC.Bt470M: (2, 1.0), TC.Bt470Bg: (3, 1.0),
    TC.Srgb: (4, 1.0), TC.Log100: (5, 1.0), TC.Log316: (6, 1.0), TC.Iec61966_2_4: (7, 1.0),
    TC.Bt1361: (8, 1.0), TC.St2084: (9, 100.0), TC.St428: (10, 1.0), TC.AribB67: (11, 10.0),
}


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("desc") / "source_desc_dump")
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "vqm_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "source_desc_dump.cc"),
                    os.path.join(ROOT, "vqm_amd", "csrc", "ss2_host.cc"),
                    os.path.join(ROOT, "vqm_amd", "csrc", "ss2_source.cc"), "-o", exe], check=True)

    def run(f: ImageFormat, size: Size = Size(32, 32), flt: int = 0) -> dict:
        args = [f.type, f.family, f.alpha, f.range, f.matrix, f.transfer, f.primaries,
                f.chroma_location, f.subsample_w, f.subsample_h, f.depth, size.width, size.height, flt]
        out = subprocess.run([exe] + [str(int(a)) for a in args], capture_output=True, text=True,
                             check=True, timeout=60).stdout
        d = json.loads(out)
        assert "error" not in d, d
        return d
    return run


def close(a, b, tol=2e-6):
    np.testing.assert_allclose(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel(),
                               rtol=tol, atol=tol)


def yuv(matrix, primaries=CP.Bt709, transfer=TC.Linear, **kw) -> ImageFormat:
    return ImageFormat(family=ColorFamily.Yuv, matrix=matrix, primaries=primaries, transfer=transfer,
                       **kw)


@pytest.mark.parametrize("matrix,primaries", [
    (MC.Bt709, CP.Bt709), (MC.Bt470Bg, CP.Bt709), (MC.St170M, CP.Bt709), (MC.St240M, CP.Bt709),
    (MC.Fcc, CP.Bt709), (MC.YCgCo, CP.Bt709), (MC.Bt2020Ncl, CP.Bt2020), (MC.Bt2020Ncl, CP.Bt470Bg),
    (MC.ChromaticityDerivedNcl, CP.Bt2020), (MC.ChromaticityDerivedNcl, CP.Bt470Bg),
    (MC.ChromaticityDerivedNcl, CP.Ebu3213E), (MC.ChromaticityDerivedNcl, CP.Bt709),
    (MC.ChromaticityDerivedNcl, CP.Film)], ids=lambda v: str(int(v)))
def test_ncl_matrices(dump, matrix, primaries):
    d = dump(yuv(matrix, primaries))
    derived = matrix == MC.ChromaticityDerivedNcl
    close(d["m1"], R.derived_ncl_to_rgb(primaries) if derived else R.ncl_to_rgb(matrix))
    assert d["mode"] == 0 and d["identity1"] == 0
    close(d["m2"], R.gamut(primaries))
    assert d["identity2"] == (1 if primaries == CP.Bt709 else 0)


@pytest.mark.parametrize("primaries", [CP.Bt470M, CP.Bt470Bg, CP.St170M, CP.St240M, CP.Film,
                                       CP.Bt2020, CP.St428, CP.St431_2, CP.St432_1, CP.Ebu3213E],
                         ids=lambda v: str(int(v)))
def test_gamut(dump, primaries):
    d = dump(ImageFormat(transfer=TC.Linear, primaries=primaries))
    assert d["identity1"] == 1 and d["identity2"] == 0
    close(d["m2"], R.gamut(primaries))


@pytest.mark.parametrize("transfer", list(CURVE), ids=lambda v: str(int(v)))
def test_transfer_curves(dump, transfer):
    d = dump(ImageFormat(transfer=transfer))
    curve, peak = CURVE[transfer]
    assert d["transfer"] == curve
    assert d["k"][3] == pytest.approx(peak)
    if transfer == TC.AribB67:  # OOTF weights of the BT.709 primaries
        close(d["k"][:3], [0.2126, 0.7152, 0.0722], 1e-4)
        close(d["k"][:3], R.rgb_to_xyz(CP.Bt709)[1], 1e-4)


def test_hlg_weights_follow_the_primaries(dump):
    d = dump(ImageFormat(transfer=TC.AribB67, primaries=CP.Bt2020))
    close(d["k"][:3], [0.2627, 0.6780, 0.0593])
    d = dump(yuv(MC.Bt709, CP.Bt470Bg, TC.AribB67))
    kr, kb = R.derived_krkb(CP.Bt470Bg)
    close(d["k"][:3], [kr, 1 - kr - kb, kb])


@pytest.mark.parametrize("matrix,primaries", [(MC.Bt2020Cl, CP.Bt2020), (MC.ChromaticityDerivedCl, CP.Bt709),
                                              (MC.ChromaticityDerivedCl, CP.Bt470Bg)],
                         ids=lambda v: str(int(v)))
def test_constant_luminance(dump, matrix, primaries):
    d = dump(yuv(matrix, primaries, TC.Bt709))
    assert d["mode"] == 1
    kr, kb = R.KRKB[MC.Bt2020Cl] if matrix == MC.Bt2020Cl else R.derived_krkb(primaries)
    close(d["k"][:3], [kr, 1 - kr - kb, kb])
    close(d["cl"], [R.bt709_oetf(1 - kb), 1 - R.bt709_oetf(kb), R.bt709_oetf(1 - kr),
                    1 - R.bt709_oetf(kr)])
    close(d["m2"], R.gamut(primaries))
    # The descriptor's CL decode restated in FP64 reproduces format_ref's.
    for e in [(0.5, 0.1, -0.2), (0.3, -0.25, 0.3), (0.9, 0.0, 0.0)]:
        nb, pb, nr, pr = d["cl"]
        b = R.bt709_inverse_oetf(e[0] + e[1] * 2 * (nb if e[1] < 0 else pb))
        r = R.bt709_inverse_oetf(e[0] + e[2] * 2 * (nr if e[2] < 0 else pr))
        g = (R.bt709_inverse_oetf(e[0]) - d["k"][0] * r - d["k"][2] * b) / d["k"][1]
        close([r, g, b], R.cl_to_rgb(kr, kb, e), 1e-5)


@pytest.mark.parametrize("transfer", [TC.St2084, TC.AribB67], ids=lambda v: str(int(v)))
def test_ictcp(dump, transfer):
    d = dump(yuv(MC.Bt2100ICtCp, CP.Bt2020, transfer))
    assert d["mode"] == 2
    close(d["m1"], np.linalg.inv(R.lms_to_ictcp(transfer)))
    close(d["m2"], R.gamut(CP.Bt2020) @ R.lms_to_rgb2020())
    if transfer == TC.AribB67:
        close(d["k"][:3], [0.2627, 0.6780, 0.0593])


@pytest.mark.parametrize("location", list(ChromaLocation), ids=lambda v: str(int(v)))
@pytest.mark.parametrize("sub", [(1, 1), (1, 0), (2, 0), (0, 1), (2, 2)], ids=str)
def test_chroma_siting(dump, location, sub):
    d = dump(yuv(MC.Bt709, chroma_location=location, subsample_w=sub[0], subsample_h=sub[1]))
    expect = [R.chroma_shift(location, True, sub[0]) if sub[0] else 0.0,
              R.chroma_shift(location, False, sub[1]) if sub[1] else 0.0]
    close(d["shift"], expect)


def test_plane_layout_and_classes(dump):
    f = ImageFormat(type=PixelType.Word, depth=10, range=PixelRange.Limited, family=ColorFamily.Yuv,
                    matrix=MC.Bt709, alpha=AlphaType.Straight, subsample_w=1, subsample_h=1,
                    transfer=TC.Bt709)
    d = dump(f, Size(24, 10))
    shapes = R.plane_shapes(f, Size(24, 10))
    off = 0
    for p, w, h in shapes:
        assert d["plane_off"][p] == off and d["plane_w"][p] == w and d["plane_h"][p] == h
        off += (w * h * 2 + 3) // 4 * 4
    close(d["scale"], [1 / 876, 1 / 896, 1 / 1023])
    close(d["offset"], [-64 / 876, -512 / 896, 0.0])
    grey = dump(ImageFormat(family=ColorFamily.Grey, matrix=MC.Unspecified, transfer=TC.Bt709,
                            primaries=CP.Bt2020))
    assert grey["identity1"] == 1 and grey["identity2"] == 1  # grey carries no gamut
    assert dump(ImageFormat.srgb8())["srgb8"] == 1
    assert dump(ImageFormat())["srgb8"] == 0


def _first_taps(n: int, sub: int, shift: float, taps: int) -> np.ndarray:
    """The prepare kernel's first chroma tap of output samples 0 .. n-1 in
    FP32, operation by operation: pos = (out + 0.5) * 2^-sub + shift, then
    round-half-up of pos - taps / 2."""
    f = np.float32
    pos = (np.arange(n, dtype=f) + f(0.5)) * f(2.0 ** -sub) + f(shift)
    v = pos - f(0.5 * taps)
    return np.where(v < 0, np.floor(v + f(0.5)), np.floor(v + f(0.49999997))).astype(np.int64)


def test_chroma_window_fits_every_tile(dump):
    """Each 64 x 32 prepare tile stages the chroma taps its pixels read in an
    LDS window of kChromaWindow entries. For every subsampling, siting and
    filter, on 3840 x 2160 and a ragged 136 x 72, the window fits, and the
    first taps are non-decreasing, so every pixel's taps lie in its tile's
    window."""
    import re
    src = open(os.path.join(ROOT, "vqm_amd", "csrc", "ss2_kernels.hip")).read()
    capacity = int(re.search(r"kChromaWindow = (\d+);", src).group(1))
    worst = 0
    for sw, sh in [(1, 1), (1, 0), (2, 0), (0, 1), (2, 2), (0, 2), (1, 2), (2, 1)]:
        for location in ChromaLocation:
            d = dump(yuv(MC.Bt709, chroma_location=location, subsample_w=sw, subsample_h=sh))
            sx, sy = d["shift"]
            for taps in (2, 4, 6):
                for w, h in ((3840, 2160), (136, 72)):
                    fx = _first_taps(w, sw, sx, taps)
                    fy = _first_taps(h, sh, sy, taps)
                    assert np.all(np.diff(fx) >= 0) and np.all(np.diff(fy) >= 0)
                    x0 = np.arange(0, w, 64)
                    y0 = np.arange(0, h, 32)
                    cols = fx[np.minimum(x0 + 63, w - 1)] - fx[x0] + taps
                    rows = fy[np.minimum(y0 + 31, h - 1)] - fy[y0] + taps
                    worst = max(worst, int(cols.max() * rows.max()))
    assert worst <= capacity
    assert worst == 1518  # Lanczos-3 with 2x vertical subsampling only: 69 x 22
