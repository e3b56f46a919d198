"""Pins of the CPU oracle (oracle/): each part checked against the reference's
own goldens and contracts or against what the mathematics fixes -- never
against a retyping of the oracle's own formula."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import (GOLDEN_B, GOLDEN_C, GOLDEN_TOL, SCORE_ABS, assert_features_close,
                      gradient_image, ramp_image)


# ---- end to end: the reference's goldens and identity ---------------------

@pytest.mark.parametrize("form", ["F64", "F32_REF", "F32_DEV"])
def test_reference_goldens(oracle_lib, form):
    """pipeline_test.cc:146-174: gradient vs gradient_B / gradient_C, +-5e-3."""
    mode = getattr(oracle_lib, form)
    orig = gradient_image(0)
    n, norms_b, _ = oracle_lib.compute(mode, 32, 32, orig, gradient_image(1))
    _, norms_c, _ = oracle_lib.compute(mode, 32, 32, orig, gradient_image(2))
    assert n == 4
    assert abs(oracle_lib.score(n, norms_b) - GOLDEN_B) <= GOLDEN_TOL
    assert abs(oracle_lib.score(n, norms_c) - GOLDEN_C) <= GOLDEN_TOL


def test_reference_order_tracks_the_reference_goldens(oracle_lib):
    """The FP32 form in the shaders' operation order lands within 1e-3 of the
    reference's GPU goldens (its remaining gap is driver transcendentals)."""
    orig = gradient_image(0)
    for dist, golden in ((gradient_image(1), GOLDEN_B), (gradient_image(2), GOLDEN_C)):
        n, norms, _ = oracle_lib.compute(oracle_lib.F32_REF, 32, 32, orig, dist)
        assert abs(oracle_lib.score(n, norms) - golden) <= SCORE_ABS


@pytest.mark.parametrize("form", ["F64", "F32_REF", "F32_DEV"])
@pytest.mark.parametrize("size", [(32, 32), (8, 8), (13, 9), (64, 64), (65, 65), (129, 97)])
def test_identity_is_exactly_100(oracle_lib, form, size):
    """pipeline_test.cc:120-144 and :760-783: identical planes cancel before
    any division rounds, so every map is exactly zero."""
    w, h = size
    img = ramp_image(w, h)
    n, norms, sums = oracle_lib.compute(getattr(oracle_lib, form), w, h, img, img)
    assert np.all(sums == 0.0)
    assert oracle_lib.score(n, norms) == 100.0


# ---- scale plan (pipeline.cc:28-48) ---------------------------------------

@pytest.mark.parametrize("size,count", [((32, 32), 4), ((8, 8), 2), ((13, 9), 2), ((64, 64), 5),
                                        ((65, 65), 5), ((129, 97), 5)])
def test_scale_counts(oracle_lib, size, count):
    """pipeline_test.cc:126-128 and :767-776."""
    assert len(oracle_lib.plan_scales(*size)) == count


def test_scale_sizes_of_the_workloads(oracle_lib):
    """Level sizes halve rounding up; the previous level gates the next."""
    assert oracle_lib.plan_scales(1920, 1080) == [(1920, 1080), (960, 540), (480, 270), (240, 135),
                                                  (120, 68), (60, 34)]
    assert oracle_lib.plan_scales(3840, 2160)[-1] == (120, 68)
    assert oracle_lib.plan_scales(640, 360)[-1] == (20, 12)
    assert oracle_lib.plan_scales(1280, 720)[-1] == (40, 23)
    assert oracle_lib.plan_scales(2560, 1440)[-1] == (80, 45)
    assert oracle_lib.plan_scales(512, 512)[-1] == (16, 16)
    with pytest.raises(ValueError):
        oracle_lib.plan_scales(4, 8)
    with pytest.raises(ValueError):
        oracle_lib.plan_scales(8, 7)


# ---- pruning (score.cc:91-114) --------------------------------------------

def active(oracle_lib, scales):
    return ["".join("XYB"[c] for c in range(3) if not oracle_lib.channel_pruned(c, s, scales))
            for s in range(scales)]


def test_pruning_mask_six_scales(oracle_lib):
    """docs/ssimulacra2.md:96-104: at six scales 1:1 keeps Y, 1:2 and 1:4 X,Y,
    1:8 and 1:16 all three, 1:32 none -- 7 of 18 pairs pruned."""
    assert active(oracle_lib, 6) == ["Y", "XY", "XY", "XYB", "XYB", ""]


def test_pruning_mask_four_scales(oracle_lib):
    """pipeline_test.cc:163-169: the 32x32 pair runs four scales and three maps
    are pruned (the flat weight run puts X at 1:1, Y at 1:2, B at the 4th)."""
    assert active(oracle_lib, 4) == ["YB", "XB", "XYB", "XY"]


# ---- score mapping (score_test.cc) -----------------------------------------

def test_score_rejects_nonfinite_norms(oracle_lib):
    """score_test.cc:11-32: every one of the 108 entries is checked."""
    for bad in (math.inf, -math.inf, math.nan):
        for i in range(108):
            norms = np.zeros(108)
            norms[i] = bad
            with pytest.raises(ValueError):
                oracle_lib.score(6, norms)


def test_score_overflow_and_range(oracle_lib):
    """score_test.cc:34-46."""
    assert oracle_lib.score(0, np.zeros(108)) == 100.0
    assert oracle_lib.score(6, np.zeros(108)) == 100.0
    norms = np.zeros(108)
    norms[3] = np.finfo(np.float64).max
    with pytest.raises(ValueError):
        oracle_lib.score(6, norms)
    norms[3] = 1e200
    with pytest.raises(ValueError):
        oracle_lib.score(6, norms)
    norms[3] = 100.0
    s = oracle_lib.score(6, norms)
    assert math.isfinite(s) and s < 0.0
    with pytest.raises(ValueError):
        oracle_lib.score(7, np.zeros(108))


def test_score_mapping_closed_form(oracle_lib):
    """One weighted norm through the published polynomial and power: the Y
    1:1 4-norm weight 7.466890328078848 (score.cc:59) at six scales."""
    norms = np.zeros(108)
    norms[3] = 0.01  # level 0, avg_ssim[Y * 2 + 1]
    e = 7.466890328078848 * 0.01 * 0.9562382616834844
    m = 2.326765642916932 * e - 0.020884521182843837 * e * e + 6.248496625763138e-05 * e ** 3
    assert oracle_lib.score(6, norms) == pytest.approx(100.0 - 10.0 * m ** 0.6276336467831387, abs=1e-12)


# ---- taps, decode, XYB ------------------------------------------------------

def test_gaussian_taps(oracle_lib):
    """pipeline.cc:274-291 sampled at sigma 1.5, radius ceil(3.5 sigma) = 6."""
    t = oracle_lib.gaussian_taps()
    k = np.arange(-6, 7) / 1.5
    w = np.exp(-0.5 * k * k)
    assert np.array_equal(t, (w / w.sum()).astype(np.float32))
    assert np.array_equal(t, t[::-1])
    np.testing.assert_allclose(t[:7], [8.92210686e-05, 1.02819658e-03, 7.5974022e-03,
                                       3.59943481e-02, 0.109341175, 0.212967529, 0.265964257],
                               rtol=1e-7)


def test_srgb_decode(oracle_lib):
    """IEC 61966-2-1 EOTF of code/255 (color.glsl:135-137)."""
    lut64 = oracle_lib.srgb_lut_f64()
    lut32 = oracle_lib.srgb_lut_f32()
    v = np.arange(256) / 255.0
    want = np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(lut64, want, rtol=1e-14, atol=0)
    # FP32: the decode product and the two FP32 operations before the power
    # each round, and the exponent 2.4 multiplies their relative error.
    np.testing.assert_allclose(lut32, want, rtol=1e-6, atol=1e-12)
    assert lut32[0] == 0.0 and lut32[255] == 1.0
    assert np.all(np.diff(lut32) > 0)


def test_xyb_closed_forms(oracle_lib):
    """Black: all three opsin mixes equal the bias, so X and B sit exactly at
    their offsets and Y at 0.01 plus cbrt(bias) - kCbrtBias (the constant is
    the FP32-rounded cube root, prepare.comp:50). Grey keeps X at its offset
    (opsin rows 0 and 1 both sum to 1). FP32 forms agree with FP64 to FP32
    precision."""
    bias, cbias = 0.0037930732552754493, 0.15595419704914093
    black = oracle_lib.xyb_f64(0.0, 0.0, 0.0)
    assert black[0] == 0.42 and black[2] == 0.55
    assert black[1] == pytest.approx(0.01 + (np.cbrt(bias) - cbias), abs=1e-16)
    for v in (0.05, 0.2, 0.5, 1.0):
        g = oracle_lib.xyb_f64(v, v, v)
        assert abs(g[0] - 0.42) < 1e-15
        assert g[1] == pytest.approx(np.cbrt(v + bias) - cbias + 0.01, abs=1e-15)
    # The device form (bias cancelled, fused): black and greys keep X and B
    # exactly at their offsets, as the reference order does.
    for v in (0.0, 0.05, 0.5, 1.0):
        dev = oracle_lib.xyb_f32(v, v, v, 1)
        assert dev[0] == np.float32(0.42) and dev[2] == np.float32(0.55)
    rng = np.random.default_rng(1)
    for _ in range(200):
        r, g, b = rng.random(3).astype(np.float32)
        ref = oracle_lib.xyb_f64(float(r), float(g), float(b))
        for kind in (0, 1):
            got = oracle_lib.xyb_f32(float(r), float(g), float(b), kind)
            np.testing.assert_allclose(got, ref, atol=3e-6)


def test_device_cube_root(oracle_lib):
    """The kernels' cube root is faithful (<= 0.75 ulp) on the XYB
    domain and exact at perfect cubes of powers of two; a zero or negative
    mix is clamped to the smallest normal, whose root 2^-42 vanishes against
    the bias (c = -bias exactly, as for a zero mix)."""
    assert oracle_lib.cbrt_dev(0.0) == 2.0 ** -42
    assert oracle_lib.cbrt_dev(-3.0) == 2.0 ** -42
    cbias = np.float32(0.15595419704914093)  # prepare.comp:49
    assert np.float32(2.0 ** -42) - cbias == -cbias
    assert oracle_lib.cbrt_dev(0.125) == 0.5
    assert oracle_lib.cbrt_dev(1.0) == 1.0
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.float32(0.0037930732) * (1 + rng.random(2000)),
                         rng.random(5000).astype(np.float32) + np.float32(0.004)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        y = oracle_lib.cbrt_dev(float(x))
        exact = np.cbrt(np.float64(x))
        ulp = np.spacing(np.float32(exact))
        worst = max(worst, abs(y - exact) / ulp)
    assert worst <= 0.75
    # Every 4099th float of [2^-126, 64) (HDR mixes reach past 1): the seed's
    # error, largest where the bits sit mid-binade, is left within 0.84 ulp,
    # and the power-of-two cubes 2^(3k) have roots exactly 2^k.
    lo = int(np.float32(2.0 ** -126).view(np.uint32))
    hi = int(np.float32(64.0).view(np.uint32))
    worst = 0.0
    for bits in range(lo, hi, 4099):
        x = float(np.uint32(bits).view(np.float32))
        exact = np.cbrt(np.float64(x))
        worst = max(worst, abs(oracle_lib.cbrt_dev(x) - exact) / np.spacing(np.float32(exact)))
    assert worst <= 0.84
    for k in range(-40, 3):
        assert oracle_lib.cbrt_dev(2.0 ** (3 * k)) == 2.0 ** k


def test_device_reciprocal(oracle_lib):
    """The kernels' reciprocal (shared by the difference form's SSIM' and
    edge quotients) is the correctly rounded 1/x over the map's denominators
    (den x det_orig in [c2, ~10]): within half an ulp of the FP64 quotient,
    and numpy's FP32 division."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([np.float32(2.0) ** rng.uniform(-11, 4, 20000),
                         np.float32(0.0009) * (1 + rng.random(2000))]).astype(np.float32)
    for x in xs:
        got = oracle_lib.rcp_dev(float(x))
        exact = 1.0 / np.float64(x)
        assert abs(got - exact) <= 0.5 * np.spacing(np.float32(exact))
        assert got == float(np.float32(1.0) / x)
    for e in range(-11, 5):
        assert oracle_lib.rcp_dev(2.0 ** e) == 2.0 ** -e


# ---- blur and pyramid --------------------------------------------------------

def brute_force_blur(plane: np.ndarray) -> np.ndarray:
    """Direct 2D convolution with the separable Gaussian, zeros outside."""
    k = np.arange(-6, 7) / 1.5
    w = np.exp(-0.5 * k * k)
    w = (w / w.sum()).astype(np.float32).astype(np.float64)
    h, wd = plane.shape
    pad = np.pad(plane.astype(np.float64), 6)
    out = np.zeros((h, wd))
    for i in range(13):
        for j in range(13):
            out += w[i] * w[j] * pad[i:i + h, j:j + wd]
    return out


@pytest.mark.parametrize("shape", [(8, 8), (9, 13), (20, 31), (1, 17)])
def test_blur_against_brute_force(oracle_lib, shape):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    plane = rng.random(shape)
    want = brute_force_blur(plane)
    np.testing.assert_allclose(oracle_lib.blur_f64(plane), want, rtol=1e-13, atol=1e-15)
    p32 = plane.astype(np.float32)
    want32 = brute_force_blur(p32)
    for order in (0, 1):
        np.testing.assert_allclose(oracle_lib.blur_f32(p32, order), want32, rtol=2e-6, atol=1e-7)


def box_pyramid(levels0: np.ndarray, sizes):
    """Clamped 2x2 box pyramid via edge padding and block means (numpy)."""
    out = [levels0]
    cur = levels0
    for (w, h) in sizes[1:]:
        c, ph, pw = cur.shape
        padded = np.pad(cur, ((0, 0), (0, 2 * h - ph), (0, 2 * w - pw)), mode="edge")
        cur = padded.reshape(c, h, 2, w, 2).mean(axis=(2, 4))
        out.append(cur)
    return out


@pytest.mark.parametrize("size", [(32, 32), (13, 9), (65, 65), (129, 97), (40, 23)])
def test_pyramid_is_the_clamped_box(oracle_lib, size):
    """prepare.comp:100-111: each level is the 2x2 mean of the previous
    level's linear RGB, the last row and column duplicated on odd sizes."""
    w, h = size
    img = ramp_image(w, h, seed=3)
    sizes = oracle_lib.plan_scales(w, h)
    lut = oracle_lib.srgb_lut_f64()
    stride = (w * h + 3) // 4 * 4
    base = np.stack([lut[img[c * stride:c * stride + w * h]].reshape(h, w) for c in range(3)])
    want = box_pyramid(base, sizes)
    lut32 = oracle_lib.srgb_lut_f32().astype(np.float64)
    base32 = np.stack([lut32[img[c * stride:c * stride + w * h]].reshape(h, w) for c in range(3)])
    want32 = box_pyramid(base32, sizes)
    got64 = oracle_lib.pyramid(w, h, img, f64=True)
    got32 = oracle_lib.pyramid(w, h, img, f64=False)
    for l in range(len(sizes)):
        np.testing.assert_allclose(got64[l], want[l], rtol=1e-14, atol=1e-16)
        # Three FP32 additions per level, compounded over the levels.
        np.testing.assert_allclose(got32[l], want32[l], rtol=2.5e-7 * (l + 1), atol=1e-12)


# ---- the FP32 forms against FP64 on the synthetic workload -------------------

@pytest.mark.parametrize("strength", [(4, 1.0), (12, 3.0), (24, 8.0)])
def test_fp32_forms_within_contract_of_fp64(oracle_lib, strength):
    """512x512 pairs: both FP32 forms stay within the parity contract of the
    FP64 form (|dscore| <= 1e-3, features 1e-4 relative + 2e-6)."""
    from vqm_amd import synth
    a, b = synth.pair(0, 0, 512, 512, strength)
    n, f64, _ = oracle_lib.compute(oracle_lib.F64, 512, 512, a, b)
    for form in (oracle_lib.F32_REF, oracle_lib.F32_DEV):
        _, f32, _ = oracle_lib.compute(form, 512, 512, a, b)
        assert abs(oracle_lib.score(n, f32) - oracle_lib.score(n, f64)) <= SCORE_ABS
        assert_features_close(f32, f64, what="fp32 vs fp64")


@pytest.mark.parametrize("form", ["F32_REF", "F32_DEV"])
def test_block_sums_within_the_fp32_bound(oracle_lib, form):
    """The FP32 block sums (one column by eight rows) against the same form
    summed in FP64 per pixel: every map value is >= 0, so an FP32 sum of at
    most 8 of them is within 7 u (u = 2^-24) of its exact sum, and so is
    every level sum; the 4-norms (fourth roots) within a quarter of that.
    The blocks must also change the result (the sums really are FP32). The
    block shape itself is a choice of this path, not of the reference; the
    GPU parity tests hold the kernels to it bit for bit."""
    from vqm_amd import synth
    a, b = synth.pair(0, 4, 300, 200, (12, 3.0))
    mode = getattr(oracle_lib, form)
    per_pixel = oracle_lib.Mode.from_buffer_copy(mode)
    per_pixel.sum_blocks = 0
    n, blocks, _ = oracle_lib.compute(mode, 300, 200, a, b)
    _, exact, _ = oracle_lib.compute(per_pixel, 300, 200, a, b)
    assert not np.array_equal(blocks, exact)
    bound = 7 * 2.0 ** -24 * 1.001
    for k in range(18 * n):
        rel = bound / 4 if k % 18 in (1, 3, 5, 7, 9, 11, 13, 15, 17) else bound
        assert abs(blocks[k] - exact[k]) <= rel * abs(exact[k]) + 1e-300, (k, blocks[k], exact[k])


def _bf16_reference(v: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even) through
    float64: keep 8 significant bits of the float32 significand."""
    m, e = np.frexp(v.astype(np.float64))        # v = m * 2^e, 0.5 <= |m| < 1
    scaled = np.ldexp(m, 8)                      # 8 significant bits above the point
    return np.ldexp(np.rint(scaled), e - 8).astype(np.float32)


def test_plane_roundings(oracle_lib):
    """The narrow plane formats against independent references: numpy's
    IEEE binary16 conversion (round to nearest even), a float64 restatement
    of bfloat16 rounding, and numpy's rint for unsigned 16-bit fixed point
    q / 65536 with clamping; ties and edges included."""
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.uniform(0.0, 1.0, 4000), rng.uniform(-2.0, 2.0, 1000),
        (np.arange(0, 2000) + 0.5) / 65536.0,             # unorm16 ties
        np.array([0.0, -0.0, 1e-9, 0.93, 0.9999, 65535.5 / 65536, 1.0, 3.0, -1e-3]),
        (np.arange(1, 400) + 0.5) * 2.0 ** -11,           # binary16 ties near 0.1-0.2
        np.float32(1.0) + (np.arange(0, 200) * 2 + 1) * np.float32(2.0 ** -8) / 2,  # bf16 ties
    ]).astype(np.float32)
    f16 = vals.astype(np.float16).astype(np.float32)
    bf16 = _bf16_reference(vals)
    u16 = (np.clip(np.rint(np.maximum(vals.astype(np.float64), 0.0) * 65536.0), 0, 65535) / 65536.0).astype(np.float32)
    for v, a, b, c in zip(vals, f16, bf16, u16):
        assert oracle_lib.round_plane(1, float(v)) == a, v
        assert oracle_lib.round_plane(2, float(v)) == b, v
        assert oracle_lib.round_plane(3, float(v)) == c, v
    assert oracle_lib.round_plane(3, float("nan")) == 0.0
    assert oracle_lib.round_plane(0, 0.123) == np.float32(0.123)


def test_xyb_planes_and_the_half_plane_build(oracle_lib):
    """The oracle's stored planes against a numpy restatement of the pyramid:
    clamped 2x2 boxes summed ((((a + b) + c) + d) * 0.25 in float32, numpy's
    IEEE binary16 rounding of the linear averages for the reference's
    half-plane build (prepare.comp:100-111, ssimu2_common.glsl:84-96) and of
    the stored XYB values, XYB per sample through the oracle's own xyb_f32
    (pinned by test_xyb_closed_form)."""
    w, h = 37, 22
    rng = np.random.default_rng(21)
    stride = (w * h + 3) // 4 * 4
    img = np.zeros(3 * stride, np.uint8)
    for c in range(3):
        img[c * stride:c * stride + w * h] = rng.integers(0, 256, w * h)
    lut = oracle_lib.srgb_lut_f32()
    sizes = oracle_lib.plan_scales(w, h)

    def restated(round_linear: bool, round_planes: bool):
        lin = np.stack([lut[img[c * stride:c * stride + w * h]].reshape(h, w) for c in range(3)]).astype(np.float32)
        out = []
        for l, (lw, lh) in enumerate(sizes):
            if l > 0:
                ph, pw = lin.shape[1:]
                ys = np.minimum(np.arange(lh)[:, None] * 2 + np.array([0, 1])[None, :], ph - 1)
                xs = np.minimum(np.arange(lw)[:, None] * 2 + np.array([0, 1])[None, :], pw - 1)
                a = lin[:, ys[:, 0][:, None], xs[:, 0][None, :]]
                b = lin[:, ys[:, 0][:, None], xs[:, 1][None, :]]
                c = lin[:, ys[:, 1][:, None], xs[:, 0][None, :]]
                d = lin[:, ys[:, 1][:, None], xs[:, 1][None, :]]
                lin = ((((a + b).astype(np.float32) + c).astype(np.float32) + d).astype(np.float32)
                       * np.float32(0.25)).astype(np.float32)
                if round_linear:
                    lin = lin.astype(np.float16).astype(np.float32)
            stored = any(not oracle_lib.channel_pruned(c, l, len(sizes)) for c in range(3))
            x = np.zeros((3, lh, lw), np.float32)
            if stored:
                for yy in range(lh):
                    for xx in range(lw):
                        x[:, yy, xx] = oracle_lib.xyb_f32(*lin[:, yy, xx])
                if round_planes:
                    x = x.astype(np.float16).astype(np.float32)
            out.append(x)
        return out

    for mode, rl, rp in [(oracle_lib.F32_DEV, False, False),
                         (oracle_lib.f32_dev_planes(1), False, True),
                         (oracle_lib.f32_dev_for(1), True, True)]:
        got = oracle_lib.xyb_planes(mode, w, h, img)
        for g, e in zip(got, restated(rl, rp)):
            np.testing.assert_array_equal(g, e)
    assert oracle_lib.f32_dev_for(1).round_linear == 1
    assert oracle_lib.f32_dev_for(3).round_linear == 0 and oracle_lib.f32_dev_for(2).round_linear == 0
    # Rounding the linear samples moves levels >= 1 only.
    a = oracle_lib.xyb_planes(oracle_lib.f32_dev_planes(1), w, h, img)
    b = oracle_lib.xyb_planes(oracle_lib.f32_dev_for(1), w, h, img)
    np.testing.assert_array_equal(a[0], b[0])
    assert any(not np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))


# ---- the recursive Gaussian mode ---------------------------------------------------------

CLOUDINARY_B = 93.09970849   # pipeline_test.cc:152-153: Cloudinary ssimulacra2 2.1
CLOUDINARY_C = 92.11674825


def _recursive_gaussian_row(x: np.ndarray, sigma: float = 1.5) -> np.ndarray:
    """The three second-order recursions of Charalampidis (2016) on one row,
    FP64, zero state and zero input outside the row, written from the paper
    (eq. 33-35, 37, 44, 50, 52-57) with its 3x3 solve for beta."""
    n_r = int(round(3.2795 * sigma + 0.2546))
    w = np.array([1, 3, 5]) * np.pi / (2 * n_r)
    p = np.array([1 / np.tan(w[0] / 2), -1 / np.tan(w[1] / 2), 1 / np.tan(w[2] / 2)])
    r = np.array([p[0] ** 2 / np.sin(w[0]), -p[1] ** 2 / np.sin(w[1]), p[2] ** 2 / np.sin(w[2])])
    rho = np.exp(-0.5 * sigma ** 2 * w ** 2) / n_r
    d13, d35, d51 = p[0] * r[1] - r[0] * p[1], p[1] * r[2] - r[1] * p[2], p[2] * r[0] - r[2] * p[0]
    z15, z35 = d35 / d13, d51 / d13
    a = np.array([p, r, [z15, z35, 1.0]])
    beta = np.linalg.solve(a, [1.0, n_r * n_r - sigma ** 2, z15 * rho[0] + z35 * rho[1] + rho[2]])
    n2 = -beta * np.cos(w * (n_r + 1))
    d1 = -2 * np.cos(w)
    size = len(x)
    out = np.zeros(size)
    prev, prev2 = np.zeros(3), np.zeros(3)
    for n in range(-n_r + 1, size):
        left, right = n - n_r - 1, n + n_r - 1
        s = (x[left] if left >= 0 else 0.0) + (x[right] if right < size else 0.0)
        y = n2 * s - d1 * prev - prev2
        prev2, prev = prev, y
        if n >= 0:
            out[n] = y.sum()
    return out


def test_cosine_taps_are_the_recursive_gaussian(oracle_lib):
    """The FIR kernel the recursive mode applies equals what the recursive
    Gaussian's recursions compute on zero-padded rows, to FP64 rounding; it
    has unit gain, variance sigma^2 = 2.25 and nine non-zero taps."""
    k = oracle_lib.cosine_taps()
    m = np.arange(-6, 7)
    assert k.sum() == pytest.approx(1.0, abs=1e-14)
    assert (k * m * m).sum() == pytest.approx(2.25, abs=1e-13)
    assert np.all(k[[0, 1, 11, 12]] == 0.0) and np.all(k[2:11] > 0)
    np.testing.assert_array_equal(k, k[::-1])
    rng = np.random.default_rng(9)
    for size in (1, 2, 5, 9, 10, 11, 17, 40):
        x = rng.uniform(-1, 1, size)
        fir = np.convolve(np.pad(x, 6), k, mode="valid")
        np.testing.assert_allclose(_recursive_gaussian_row(x), fir, rtol=0, atol=1e-12)
    # The library's kernel table is this kernel rounded to FP32.
    import vqm_amd
    np.testing.assert_array_equal(vqm_amd.blur_taps(vqm_amd.Blur.Recursive), k.astype(np.float32))
    np.testing.assert_array_equal(vqm_amd.blur_taps(vqm_amd.Blur.Sampled), oracle_lib.gaussian_taps())
    with pytest.raises(vqm_amd.InvalidArgument):
        vqm_amd.blur_taps(2)


def test_recursive_mode_against_cloudinary_scores(oracle_lib):
    """The recursive mode on the reference's 32x32 gradient pairs against the
    scores the reference records for Cloudinary's ssimulacra2 binary
    (pipeline_test.cc:152-153). The binary also uses an sRGB polynomial and an
    approximate cube root, which the reference credits with the rest of the
    gap (docs/divergence.md:29-35); 0.05 bounds it here. The FP32 device form
    keeps the parity contract against FP64."""
    orig = gradient_image(0)
    for variant, cloud in ((1, CLOUDINARY_B), (2, CLOUDINARY_C)):
        dist = gradient_image(variant)
        n, f64, _ = oracle_lib.compute(oracle_lib.with_blur(oracle_lib.F64, 1), 32, 32, orig, dist)
        s64 = oracle_lib.score(n, f64)
        assert abs(s64 - cloud) <= 0.05, (variant, s64)
        _, dev, _ = oracle_lib.compute(oracle_lib.with_blur(oracle_lib.F32_DEV, 1), 32, 32, orig, dist)
        assert abs(oracle_lib.score(n, dev) - s64) <= SCORE_ABS
        assert_features_close(dev, f64, what="recursive mode, device form vs FP64")
        _, sampled, _ = oracle_lib.compute(oracle_lib.F64, 32, 32, orig, dist)
        assert abs(oracle_lib.score(n, sampled) - s64) > 0.02  # the kernel matters


def test_device_map_form_is_the_reference_map(oracle_lib):
    """The kernels' SSIM' form (blur of (a - b)^2 in place of a^2 and b^2,
    den = num_s + v) is the reference's map rearranged: in FP64 arithmetic it
    equals the direct form to rounding on random moments, and on the oracle's
    pairs the device form stays closer to FP64 than the reference's FP32
    order does wherever that order strays past 1e-4."""
    rng = np.random.default_rng(4)
    a, b = rng.uniform(0, 1, (2, 10000))
    w = rng.uniform(0, 1, (10000, 13))
    w /= w.sum(axis=1, keepdims=True)
    # Moments of 13-sample windows: any weights give the same identity.
    xa, xb = rng.uniform(0, 1, (2, 10000, 13))
    mu1, mu2 = (w * xa).sum(1), (w * xb).sum(1)
    s11, s22, s12 = (w * xa * xa).sum(1), (w * xb * xb).sum(1), (w * xa * xb).sum(1)
    e = (w * (xa - xb) ** 2).sum(1)
    c2 = 0.0009
    num_m = 1 - (mu1 - mu2) ** 2
    num_s = 2 * (s12 - mu1 * mu2) + c2
    den = (s11 - mu1 ** 2) + (s22 - mu2 ** 2) + c2
    direct = 1 - num_m * num_s / den
    t = mu1 - mu2
    v = e - t * t
    rearranged = (v + t * t * num_s) / (num_s + v)
    np.testing.assert_allclose(rearranged, direct, rtol=1e-9, atol=1e-12)
    orig = gradient_image(0)
    for variant in (1, 2):
        dist = gradient_image(variant)
        n, f64, _ = oracle_lib.compute(oracle_lib.F64, 32, 32, orig, dist)
        _, dev, _ = oracle_lib.compute(oracle_lib.F32_DEV, 32, 32, orig, dist)
        _, ref, _ = oracle_lib.compute(oracle_lib.F32_REF, 32, 32, orig, dist)
        s64 = oracle_lib.score(n, f64)
        assert abs(oracle_lib.score(n, dev) - s64) <= max(1e-4, abs(oracle_lib.score(n, ref) - s64))
