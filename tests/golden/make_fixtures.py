"""Writes the committed fixtures of tests/golden/:

  synth_hashes.json    FNV-1a hashes of generated pairs (workload generator)
  oracle_vectors.json  108 norms and scores of small seeded pairs from the
                       oracle's FP64 and FP32 device forms (regression
                       vectors; the oracle itself is pinned by
                       tests/test_oracle.py against the reference's goldens
                       and closed forms)

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from conftest import gradient_image  # noqa: E402
from vqm_amd import synth  # noqa: E402

HASH_CASES = [(0, 0, 512, 512, 12, 3.0), (3, 1, 640, 360, 4, 1.0), (3, 2, 1280, 720, 24, 8.0)]
VECTOR_CASES = [  # name, cfg, pair, width, height, step, sigma
    ("synth_64x64_s12", 0, 11, 64, 64, 12, 3.0),
    ("synth_97x61_s4", 0, 12, 97, 61, 4, 1.0),
    ("synth_200x130_s24", 0, 13, 200, 130, 24, 8.0),
]


def main() -> None:
    hashes = {}
    for cfg, idx, w, h, step, sigma in HASH_CASES:
        a, b = synth.pair(cfg, idx, w, h, (step, sigma))
        hashes[f"{cfg}:{idx}:{w}:{h}:{step}:{sigma}"] = f"{synth.digest(a):016x}:{synth.digest(b):016x}"
    with open(os.path.join(HERE, "synth_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)

    vectors = {}
    cases = [("gradient_b", None), ("gradient_c", None)] + [(c[0], c) for c in VECTOR_CASES]
    for name, spec in cases:
        if spec is None:
            w = h = 32
            a = gradient_image(0)
            b = gradient_image(1 if name == "gradient_b" else 2)
            src = {"generator": "tests/support.h MakeGradient", "variant": name}
        else:
            _, cfg, idx, w, h, step, sigma = spec
            a, b = synth.pair(cfg, idx, w, h, (step, sigma))
            src = {"generator": "vqm_amd.synth", "cfg": cfg, "pair": idx, "step": step,
                   "sigma": sigma}
        entry = {"width": w, "height": h, "source": src}
        for form in ("F64", "F32_DEV", "F32_REF"):
            n, norms, _ = oracle.compute(getattr(oracle, form), w, h, a, b)
            entry[form] = {"scales": n, "score": oracle.score(n, norms),
                           "norms": [float(v) for v in norms]}
        vectors[name] = entry
    with open(os.path.join(HERE, "oracle_vectors.json"), "w") as f:
        json.dump(vectors, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
