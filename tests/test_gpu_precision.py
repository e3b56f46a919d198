"""Plane precision at 3840x2160: the FP32 form and the mixed form (unsigned
16-bit fixed-point planes) against the FP64 oracle
on 50 synthetic pairs (three strengths), with FP16 and BF16 planes reported
beside them. Every pair of every form is scored in one batched launch.

Bars: FP32 and mixed forms |dscore| <= 1e-3 and features within 1e-4
relative + 2e-6 absolute of FP64; FP16 within the reference's half-plane
bound of 0.1 in score (pipeline_test.cc:785-834). BF16 has no bar: its
error is printed (it is points of score, which is why the mixed form exists).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import FEATURE_ABS, FEATURE_REL, SCORE_ABS
import vqm_amd
from vqm_amd import DeviceBuffer, Ssimulacra2Batch, synth

pytestmark = pytest.mark.gpu

W, H = 3840, 2160
PAIRS = int(os.environ.get("VQM_PRECISION_PAIRS", "50"))


def test_4k_plane_precision_against_fp64(oracle_lib):
    pairs = [synth.pair(2, 100 + i, W, H, synth.STRENGTHS[i % 3]) for i in range(PAIRS)]
    keep, refs = [], []
    for a, b in pairs:
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        keep += [da, db]
        refs.append((da.ptr, db.ptr, W, H))
    threads = max(1, min(16, os.cpu_count() or 1))
    f64 = oracle_lib.compute_many(oracle_lib.F64, W, H, pairs, threads)
    s64 = np.array([oracle_lib.score(6, f64[i]) for i in range(PAIRS)])
    report = {}
    for name, fmt in (("f32", vqm_amd.PLANES_F32), ("mixed", vqm_amd.PLANES_U16),
                      ("f16", vqm_amd.PLANES_F16), ("bf16", vqm_amd.PLANES_BF16)):
        batch = Ssimulacra2Batch(max_pairs=PAIRS, plane_format=fmt)
        scales, norms = batch.Finish(batch.Submit(refs), raw=True)
        batch.close()
        assert np.all(scales == 6) and np.all(np.isfinite(norms))
        scores = np.array([oracle_lib.score(6, norms[i]) for i in range(PAIRS)])
        dscore = np.abs(scores - s64)
        excess = np.abs(norms - f64) - (FEATURE_REL * np.abs(f64) + FEATURE_ABS)
        report[name] = (dscore.max(), dscore.mean(), int((excess > 0).sum()))
        print(f"{name:6s} max |dscore| {dscore.max():.3e} mean {dscore.mean():.3e} "
              f"features over the bound: {report[name][2]} of {norms.size}")
    for name in ("f32", "mixed"):
        assert report[name][0] <= SCORE_ABS, name
        assert report[name][2] == 0, name
    assert report["f16"][0] <= 0.1
