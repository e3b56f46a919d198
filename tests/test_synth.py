"""The synthetic workload generator: deterministic, thread-count independent,
and shaped as specified (sinusoid field plus uniform noise; block-mean
quantisation plus Gaussian noise)."""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from conftest import ROOT
from vqm_amd import _lib as L
from vqm_amd import synth

FIXTURE = os.path.join(ROOT, "tests", "golden", "synth_hashes.json")


def test_generator_is_deterministic_and_thread_independent():
    a1 = synth.reference_image(0, 7, 256, 128)
    out = np.empty_like(a1)
    L.synth().vqm_synth_reference(0, 7, 256, 128, 512.0, out.ctypes.data_as(ctypes.c_void_p), 1)
    assert np.array_equal(a1, out)
    b1 = synth.distorted_image(0, 7, 256, 128, a1, 12, 3.0)
    out2 = np.empty_like(a1)
    L.synth().vqm_synth_distort(0, 7, 256, 128, a1.ctypes.data_as(ctypes.c_void_p), 12, 3.0,
                                out2.ctypes.data_as(ctypes.c_void_p), 1)
    assert np.array_equal(b1, out2)
    assert not np.array_equal(a1, synth.reference_image(0, 8, 256, 128))


def test_generator_statistics():
    a, b = synth.pair(1, 0, 640, 360, (12, 3.0))
    n = 640 * 360
    ra = a[:n].astype(np.float64)
    assert 100 < ra.mean() < 156  # centred on 128
    assert ra.std() > 10          # the sinusoids are there
    diff = b[:n].astype(np.float64) - ra
    # Block quantisation (step 12, error up to 6) plus noise sigma 3.
    assert 2.0 < diff.std() < 6.0
    assert abs(diff.mean()) < 1.0


def test_generator_matches_committed_hashes():
    """The hashes recorded when the workloads were fixed."""
    with open(FIXTURE) as f:
        want = json.load(f)
    for key, digest in want.items():
        cfg, idx, w, h, step, sigma = key.split(":")
        a, b = synth.pair(int(cfg), int(idx), int(w), int(h), (int(step), float(sigma)))
        assert f"{synth.digest(a):016x}:{synth.digest(b):016x}" == digest, key
