"""The oracle against the committed fixtures (tests/golden/): the reference's
goldens plus regression vectors of small seeded pairs."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_B, GOLDEN_C, GOLDEN_TOL, ROOT, gradient_image
from vqm_amd import synth

VECTORS = os.path.join(ROOT, "tests", "golden", "oracle_vectors.json")


def load():
    with open(VECTORS) as f:
        return json.load(f)


def inputs(entry):
    src = entry["source"]
    if src["generator"].startswith("tests/support.h"):
        return gradient_image(0), gradient_image(1 if src["variant"] == "gradient_b" else 2)
    return synth.pair(src["cfg"], src["pair"], entry["width"], entry["height"],
                      (src["step"], src["sigma"]))


def test_fixture_goldens_are_the_references():
    v = load()
    assert abs(v["gradient_b"]["F32_REF"]["score"] - GOLDEN_B) <= GOLDEN_TOL
    assert abs(v["gradient_c"]["F32_REF"]["score"] - GOLDEN_C) <= GOLDEN_TOL


@pytest.mark.parametrize("name", sorted(load().keys()))
def test_oracle_reproduces_fixture(oracle_lib, name):
    entry = load()[name]
    a, b = inputs(entry)
    for form in ("F64", "F32_DEV", "F32_REF"):
        n, norms, _ = oracle_lib.compute(getattr(oracle_lib, form), entry["width"], entry["height"], a, b)
        want = entry[form]
        assert n == want["scales"]
        np.testing.assert_allclose(norms, want["norms"], rtol=1e-12, atol=0)
        assert oracle_lib.score(n, norms) == pytest.approx(want["score"], abs=1e-10)
