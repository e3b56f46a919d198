"""Deterministic synthetic 8-bit sRGB pairs for the SSIMULACRA 2 workloads
(generator in csrc/synth.cc; the reference ships none).

Workload ids (the `cfg` of every seed):
  0  one 512x512 pair, sinusoid periods 8-128 px, strength {step 12, sigma 3}
  1  1920x1080 pairs, periods 8-512 px, three strengths
  2  3840x2160 pairs, same generator
  3  batches of uneven sizes from the five-size ladder
  4  serving requests from the ladder with weights 0.1/0.2/0.35/0.2/0.15
"""
from __future__ import annotations

import ctypes
import os
from typing import Tuple

import numpy as np

from . import _lib as L

STRENGTHS = ((4, 1.0), (12, 3.0), (24, 8.0))  # (block step, noise sigma)
LADDER = ((640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))
SERVE_WEIGHTS = (0.1, 0.2, 0.35, 0.2, 0.15)


def image_bytes(width: int, height: int) -> int:
    return int(L.synth().vqm_synth_image_bytes(width, height))


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def reference_image(cfg: int, pair: int, width: int, height: int,
                    pmax: float | None = None) -> np.ndarray:
    if pmax is None:
        pmax = 128.0 if (width, height) == (512, 512) else 512.0
    out = np.empty(image_bytes(width, height), np.uint8)
    L.synth().vqm_synth_reference(cfg, pair, width, height, float(pmax),
                                  out.ctypes.data_as(ctypes.c_void_p), _threads())
    return out


def distorted_image(cfg: int, pair: int, width: int, height: int, ref: np.ndarray,
                    step: int, sigma: float) -> np.ndarray:
    out = np.empty_like(ref)
    L.synth().vqm_synth_distort(cfg, pair, width, height, ref.ctypes.data_as(ctypes.c_void_p),
                                int(step), float(sigma), out.ctypes.data_as(ctypes.c_void_p),
                                _threads())
    return out


def pair(cfg: int, index: int, width: int, height: int, strength: Tuple[int, float],
         pmax: float | None = None) -> Tuple[np.ndarray, np.ndarray]:
    ref = reference_image(cfg, index, width, height, pmax)
    return ref, distorted_image(cfg, index, width, height, ref, *strength)


def draw(cfg: int, index: int, k: int) -> float:
    return float(L.synth().vqm_synth_draw(cfg, index, k))


def ladder_pair_shape(cfg: int, index: int, weights=None) -> Tuple[Tuple[int, int], Tuple[int, float]]:
    """Size and strength of pair `index` of the batch (uniform sizes) or
    serving (weighted sizes) workloads."""
    u = draw(cfg, index, 0)
    if weights is None:
        size = LADDER[min(int(u * len(LADDER)), len(LADDER) - 1)]
    else:
        acc, size = 0.0, LADDER[-1]
        for s, wgt in zip(LADDER, weights):
            acc += wgt
            if u < acc:
                size = s
                break
    strength = STRENGTHS[min(int(draw(cfg, index, 1) * 3), 2)]
    return size, strength


def digest(data: np.ndarray) -> int:
    arr = np.ascontiguousarray(data, np.uint8)
    return int(L.synth().vqm_synth_hash(arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))
