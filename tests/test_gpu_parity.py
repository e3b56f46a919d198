"""Parity of the HIP path with the oracle, through the C ABI.

Both map forms run bit for bit per pixel with their oracle form
(``oracle.form_for``): the reference form with the FP32 ref form (the
reference shaders' operation order, correctly rounded cube roots), the
difference form with the FP32 dev form, each pixel's map values and each
FP32 block sum (one column by eight rows) alike. Their 108 norms therefore
agree to the order of the FP64 sums (relative 1e-10).

Bars of the reference form (the default):
  * the reference's 32x32 goldens within 1e-3
    (tests/metrics/ssimulacra2/pipeline_test.cc:146-174);
  * features within 1e-4 relative of the FP32 ref form with no absolute
    floor, and |dscore| <= 1e-3, at 512x512, 1080p and 4K.
Bars of the difference form: |dscore| <= 1e-3 and features within 1e-4
relative + 2e-6 of the FP32 ref form and of FP64 on the workloads; the
goldens within the reference's own 5e-3 (on gradient pair C the reference's
FP32 order is 2.5e-3 from FP64 and the difference form follows FP64).
Identity scores exactly 100; results repeat bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import (GOLDEN_B, GOLDEN_C, GOLDEN_TOL, SCORE_ABS, assert_features_close,
                      gradient_image, ramp_image, random_pair)
import vqm_amd
from vqm_amd import Image, ImageFormat, MapForm, Size, Ssimulacra2Config, Ssimulacra2Pipeline, synth

pytestmark = pytest.mark.gpu

DEV_REL = 1e-10
MAPS = [MapForm.Reference, MapForm.Difference]


def image(w, h, data) -> Image:
    return Image(Size(w, h), ImageFormat.srgb8(), data)


def gpu(w, h, a, b, **cfg):
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), **cfg))
    r = p.Compute(image(w, h, a), image(w, h, b))
    p.close()
    return len(r.scales), r.norms(), r


def oracle_form(oracle_lib, **cfg):
    return oracle_lib.form_for(int(cfg.get("plane_format") or 0), int(cfg.get("map_form", 0)),
                               int(cfg.get("blur", 0)))


def check_against_oracle(oracle_lib, w, h, a, b, **cfg):
    """The GPU under `cfg` against its oracle form, bit for bit per pixel."""
    n, got, res = gpu(w, h, a, b, **cfg)
    n_o, want, _ = oracle_lib.compute(oracle_form(oracle_lib, **cfg), w, h, a, b)
    assert n == n_o
    bad = np.nonzero(np.abs(got - want) > DEV_REL * np.abs(want))[0]
    assert bad.size == 0, (f"{w}x{h} {cfg}: GPU vs its oracle form: features "
                           f"{[(int(i) // 18, int(i) % 18) for i in bad]} (level, index) differ; "
                           f"got {got[bad]} want {want[bad]}")
    return n, got, res


def test_reference_goldens_on_gpu(oracle_lib):
    """pipeline_test.cc:146-174 with the default (reference) form: both
    goldens within 1e-3 (the FP32 ref form lands -1.1e-4 and +4.6e-4 from
    them), bit-level with the FP32 ref form."""
    orig = gradient_image(0)
    for variant, golden in ((1, GOLDEN_B), (2, GOLDEN_C)):
        dist = gradient_image(variant)
        n, norms, res = check_against_oracle(oracle_lib, 32, 32, orig, dist)
        assert n == 4
        assert abs(res.Score() - golden) <= SCORE_ABS
        n2, explicit, _ = gpu(32, 32, orig, dist, map_form=MapForm.Reference)
        assert np.array_equal(norms, explicit)


def test_difference_form_goldens(oracle_lib):
    """The difference form on the goldens: the reference's own 5e-3, and
    FP64 within 1e-3 (on gradient C the reference's FP32 order is itself
    2.5e-3 from FP64; this form is 6e-5 from it)."""
    orig = gradient_image(0)
    for variant, golden in ((1, GOLDEN_B), (2, GOLDEN_C)):
        dist = gradient_image(variant)
        n, norms, res = check_against_oracle(oracle_lib, 32, 32, orig, dist, map_form=MapForm.Difference)
        assert abs(res.Score() - golden) <= GOLDEN_TOL
        _, f64, _ = oracle_lib.compute(oracle_lib.F64, 32, 32, orig, dist)
        assert abs(res.Score() - oracle_lib.score(n, f64)) <= SCORE_ABS
        assert_features_close(norms, f64, what="32x32 golden pair, difference form vs FP64")


def test_default_map_form_follows_the_plane_format(oracle_lib):
    """VQM_MAP_DEFAULT is the reference form, and the difference form for
    16-bit fixed-point planes."""
    a, b = synth.pair(0, 1, 96, 64, synth.STRENGTHS[1])
    _, default, _ = gpu(96, 64, a, b)
    _, ref, _ = gpu(96, 64, a, b, map_form=MapForm.Reference)
    _, diff, _ = gpu(96, 64, a, b, map_form=MapForm.Difference)
    assert np.array_equal(default, ref) and not np.array_equal(default, diff)
    _, u16, _ = gpu(96, 64, a, b, plane_format=vqm_amd.PLANES_U16)
    _, u16_diff, _ = gpu(96, 64, a, b, plane_format=vqm_amd.PLANES_U16, map_form=MapForm.Difference)
    assert np.array_equal(u16, u16_diff)


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("size,scales", [((8, 8), 2), ((13, 9), 2), ((32, 32), 4), ((64, 64), 5),
                                         ((65, 65), 5), ((129, 97), 5), ((300, 41), 4)])
def test_identity_and_small_geometries(size, scales, map_form):
    """pipeline_test.cc:120-144, 760-783: identity is exactly 100 and repeats."""
    w, h = size
    img = ramp_image(w, h)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), map_form=map_form))
    first = p.Compute(image(w, h, img), image(w, h, img))
    second = p.Compute(image(w, h, img), image(w, h, img))
    assert len(first.scales) == scales
    assert first.Score() == 100.0
    assert np.array_equal(first.norms(), second.norms())


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("size", [(32, 32), (97, 61), (200, 130), (257, 129), (1000, 37), (37, 600),
                                  (128, 64), (129, 65)])
def test_random_pairs_match_oracle(oracle_lib, size, map_form):
    """Pure-noise pairs, the FP32 worst case: bit-level with the form's
    oracle. The reference's own FP32 operation order drifts up to 1.6e-3 from
    FP64 on these, so the difference form is held to twice that drift (or
    1e-3, whichever is larger) against FP64."""
    w, h = size
    a, b = random_pair(w, h, seed=w * 7 + h)
    n, got, _ = check_against_oracle(oracle_lib, w, h, a, b, map_form=map_form)
    if map_form == MapForm.Difference:
        _, f64, _ = oracle_lib.compute(oracle_lib.F64, w, h, a, b)
        _, ref, _ = oracle_lib.compute(oracle_lib.F32_REF, w, h, a, b)
        s64 = oracle_lib.score(n, f64)
        bound = max(SCORE_ABS, 2.0 * abs(oracle_lib.score(n, ref) - s64))
        assert abs(oracle_lib.score(n, got) - s64) <= bound


def test_seeded_random_geometries_match_oracle(oracle_lib):
    """24 geometries drawn from a fixed seed (8 to 400 on each side), each
    pair through a pipeline at its default strip height and all 24 in one
    batch launch at the batch's default heights: every feature within
    DEV_REL of the reference form's oracle (bit-level per pixel; only the
    FP64 grouping of strip sums differs)."""
    rng = np.random.default_rng(2026)
    sizes = [(int(rng.integers(8, 401)), int(rng.integers(8, 401))) for _ in range(24)]
    want, dev = [], []
    for i, (w, h) in enumerate(sizes):
        print(f"random geometry {i}: {w}x{h}", flush=True)
        a, b = random_pair(w, h, seed=1000 + i)
        n, got, _ = check_against_oracle(oracle_lib, w, h, a, b)
        want.append((n, got))
        da, db = vqm_amd.DeviceBuffer(a.nbytes), vqm_amd.DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        dev.append((da, db))
    batch = vqm_amd.Ssimulacra2Batch(max_pairs=len(sizes))
    scales, norms = batch.Finish(batch.Submit([(da.ptr, db.ptr, w, h) for (da, db), (w, h) in zip(dev, sizes)]),
                                 raw=True)
    batch.close()
    for i, ((n, got), (w, h)) in enumerate(zip(want, sizes)):
        assert scales[i] == n
        bad = np.nonzero(np.abs(norms[i] - got) > DEV_REL * np.abs(got))[0]
        assert bad.size == 0, f"{w}x{h}: batch vs pipeline features {bad.tolist()}"
    for da, db in dev:
        da.free()
        db.free()


def check_workload(oracle_lib, w, h, a, b, map_form):
    """The workload bars of each form against the FP32 ref form (and FP64)."""
    n, got, res = check_against_oracle(oracle_lib, w, h, a, b, map_form=map_form)
    _, ref, _ = oracle_lib.compute(oracle_lib.F32_REF, w, h, a, b)
    assert abs(res.Score() - oracle_lib.score(n, ref)) <= SCORE_ABS
    if map_form == MapForm.Reference:
        assert_features_close(got, ref, abs_=0.0, what=f"{w}x{h} reference form vs FP32 ref, no floor")
    else:
        assert_features_close(got, ref, what=f"{w}x{h} difference form vs FP32 ref")
        _, f64, _ = oracle_lib.compute(oracle_lib.F64, w, h, a, b)
        assert abs(res.Score() - oracle_lib.score(n, f64)) <= SCORE_ABS
        assert_features_close(got, f64, what=f"{w}x{h} difference form vs FP64")
    return n


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("strength", synth.STRENGTHS)
def test_synthetic_512_matches_oracle(oracle_lib, strength, map_form):
    a, b = synth.pair(0, 0, 512, 512, strength)
    check_workload(oracle_lib, 512, 512, a, b, map_form)


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("strength", synth.STRENGTHS)
def test_1080p_matches_oracle(oracle_lib, strength, map_form):
    a, b = synth.pair(1, 0, 1920, 1080, strength)
    assert check_workload(oracle_lib, 1920, 1080, a, b, map_form) == 6


@pytest.mark.parametrize("map_form", MAPS)
def test_4k_matches_oracle(oracle_lib, map_form):
    a, b = synth.pair(2, 0, 3840, 2160, synth.STRENGTHS[1])
    check_workload(oracle_lib, 3840, 2160, a, b, map_form)


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F16, vqm_amd.PLANES_BF16, vqm_amd.PLANES_U16])
def test_narrow_planes_match_oracle(oracle_lib, fmt, map_form):
    """FP16/BF16/fixed-point stored planes: bit-level parity with the
    oracle form rounding the same planes; FP16 stays within the reference's
    half-plane bounds of the FP32 result (pipeline_test.cc:785-834: score 0.1,
    norm 2e-4), and the fixed-point planes within them by far."""
    orig = gradient_image(0)
    for variant in (1, 2):
        dist = gradient_image(variant)
        n, got, res = check_against_oracle(oracle_lib, 32, 32, orig, dist, plane_format=fmt, map_form=map_form)
        _, full, _ = oracle_lib.compute(oracle_lib.form_for(0, int(map_form)), 32, 32, orig, dist)
        if fmt != vqm_amd.PLANES_BF16:
            assert abs(res.Score() - oracle_lib.score(n, full)) <= 0.1
            assert np.max(np.abs(got - full)) <= 2e-4
    a, b = synth.pair(0, 0, 512, 512, synth.STRENGTHS[1])
    check_against_oracle(oracle_lib, 512, 512, a, b, plane_format=fmt, map_form=map_form)


@pytest.mark.parametrize("strength", synth.STRENGTHS)
def test_mixed_form_keeps_the_fp32_contract_at_1080p(oracle_lib, strength):
    """The mixed form (unsigned 16-bit fixed-point planes, FP32 arithmetic,
    FP64 sums; the difference form by default) is bit-exact with its oracle
    form and meets the same score and feature bounds against FP64 as the
    FP32 form."""
    a, b = synth.pair(1, 7, 1920, 1080, strength)
    n, got, res = check_against_oracle(oracle_lib, 1920, 1080, a, b, plane_format=vqm_amd.PLANES_U16)
    _, f64, _ = oracle_lib.compute(oracle_lib.F64, 1920, 1080, a, b)
    assert abs(res.Score() - oracle_lib.score(n, f64)) <= SCORE_ABS
    assert_features_close(got, f64, what="1080p mixed form vs FP64")


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("strip_cols", [64, 128])
@pytest.mark.parametrize("strip_rows", [8, 16, 128])
def test_strip_shape_does_not_change_results(oracle_lib, strip_rows, strip_cols, map_form):
    """Every strip shape (height, and width: one or two output columns per
    lane) is bit-level with the oracle form."""
    a, b = random_pair(300, 200, seed=9)
    check_against_oracle(oracle_lib, 300, 200, a, b, strip_rows=strip_rows, strip_cols=strip_cols,
                         map_form=map_form)


@pytest.mark.parametrize("map_form", MAPS)
def test_narrow_strips_at_1080p(oracle_lib, map_form):
    """64-column strips (one output column per lane) at 1920x1080, FP32 and
    fixed-point planes: bit-level with the oracle form."""
    a, b = synth.pair(0, 21, 1920, 1080, synth.STRENGTHS[1])
    check_against_oracle(oracle_lib, 1920, 1080, a, b, strip_cols=64, map_form=map_form)
    check_against_oracle(oracle_lib, 1920, 1080, a, b, strip_cols=64, map_form=map_form,
                         plane_format=vqm_amd.PLANES_U16)


def test_repeatable_and_interleaved():
    """pipeline_test.cc:335-357: results repeat bit for bit, also after
    another pair ran in between."""
    orig, da, db = gradient_image(0), gradient_image(1), gradient_image(2)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32)))
    first = p.Compute(image(32, 32, orig), image(32, 32, da)).norms()
    p.Compute(image(32, 32, orig), image(32, 32, db))
    again = p.Compute(image(32, 32, orig), image(32, 32, da)).norms()
    assert np.array_equal(first, again)


@pytest.mark.parametrize("map_form", MAPS)
@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F32, vqm_amd.PLANES_U16])
def test_recursive_blur_matches_oracle(oracle_lib, fmt, map_form):
    """The recursive-Gaussian mode (the truncated-cosine kernel) is bit-level
    equal to its oracle form with that kernel, on the goldens and a 512x512
    synthetic pair, in FP32 and fixed-point planes; the batch path agrees
    with the pipeline."""
    orig = gradient_image(0)
    rec = vqm_amd.Blur.Recursive
    for variant in (1, 2):
        dist = gradient_image(variant)
        check_against_oracle(oracle_lib, 32, 32, orig, dist, plane_format=fmt, blur=rec, map_form=map_form)
        _, got, _ = gpu(32, 32, orig, dist, plane_format=fmt, blur=rec, map_form=map_form)
        _, sampled, _ = gpu(32, 32, orig, dist, plane_format=fmt, map_form=map_form)
        assert not np.array_equal(got, sampled)
    a, b = synth.pair(0, 3, 512, 512, synth.STRENGTHS[1])
    n, got, _ = check_against_oracle(oracle_lib, 512, 512, a, b, plane_format=fmt, blur=rec, map_form=map_form,
                                     strip_rows=64)
    dev = [vqm_amd.DeviceBuffer(a.nbytes), vqm_amd.DeviceBuffer(b.nbytes)]
    dev[0].upload(a)
    dev[1].upload(b)
    batch = vqm_amd.Ssimulacra2Batch(max_pairs=2, plane_format=fmt, blur=rec, strip_rows=64, map_form=map_form)
    scales, norms = batch.Finish(batch.Submit([(dev[0].ptr, dev[1].ptr, 512, 512)] * 2), raw=True)
    batch.close()
    for k in range(2):
        np.testing.assert_array_equal(norms[k][:18 * n], got[:18 * n])
    for d in dev:
        d.free()
