"""The XYB cube roots, checked for every float of their domain on the CPU.

tests/cpp/cbrt_exhaustive.hip compiles the library's own host-and-device
cube roots (vqm_amd/csrc/ss2_cbrt.h) for the host and checks, for every float
in [2^-126, 2^24) (1.26e9 values; the opsin mixes of linear RGB in [0, 1]
lie in [0.0038, 1.01], HDR decodes reach ~100):

* ``cbrt_rn`` is the correctly rounded cube root: the true root lies strictly
  between the cubes of the midpoints on either side of the result, decided
  in exact 128-bit integer arithmetic;
* the oracle's ``ss2o_cbrt_rn`` (the FP32 ref form's root, found by a
  different method: libm's FP64 root rounded to FP32, then moved across
  midpoints by exact integer comparisons) returns the same float;
* ``cbrt_fast`` stays within 0.84 ulp.

Every step of the device functions is an IEEE conversion, add, multiply or
fused multiply-add, so the host build computes what the device computes.
"""
from __future__ import annotations

import os
import re
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def cbrt_exe(tmp_path_factory, oracle_lib):
    exe = str(tmp_path_factory.mktemp("cbrt") / "cbrt_exhaustive")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++20",
                    "-ffp-contract=off", "-Xarch_host", "-mfma",
                    "-I", os.path.join(ROOT, "vqm_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "cbrt_exhaustive.hip"),
                    "-o", exe, "-ldl", "-lpthread"], check=True, capture_output=True)
    return exe


def run_range(exe: str, first: int, end: int):
    lib = os.path.join(ROOT, "oracle", "libss2_oracle.so")
    r = subprocess.run([exe, lib, hex(first), hex(end)], capture_output=True, text=True, timeout=600)
    m = re.search(r"checked (\d+) floats .*wrong (\d+), oracle differs (\d+), "
                  r"cbrt_fast max error ([0-9.]+) ulp", r.stdout)
    assert m, r.stdout + r.stderr
    return r.returncode, int(m.group(1)), int(m.group(2)), int(m.group(3)), float(m.group(4))


def test_cube_roots_exhaustive(cbrt_exe):
    first, end = 0x00800000, 0x4B800000  # [2^-126, 2^24)
    rc, checked, wrong, differs, fast_err = run_range(cbrt_exe, first, end)
    assert checked == end - first
    assert wrong == 0 and differs == 0
    assert fast_err <= 0.84
    assert rc == 0


def test_checker_catches_a_wrong_root(oracle_lib):
    """The exact midpoint comparison the checks rest on rejects the
    neighbours of the correctly rounded root (and accepts the root) on a
    sample that includes exact cubes, powers of two and binade edges."""
    import ctypes
    import numpy as np
    lib = oracle_lib.lib()
    lib.ss2o_cbrt_rn.restype = ctypes.c_float
    lib.ss2o_cbrt_rn.argtypes = [ctypes.c_float]
    xs = [2.0 ** -126, 2.0 ** -125, 0.125, 1.0, 8.0, 27.0, 1000.0, 2.0 ** 23 - 1, 0.0037930732552754493]
    rng = np.random.default_rng(5)
    xs += list(rng.uniform(0.004, 1.2, 200).astype(np.float32))
    for x in xs:
        x32 = float(np.float32(x))
        got = np.float32(lib.ss2o_cbrt_rn(x32))
        # Against exact rational arithmetic: |got^3 - x| is smallest among
        # got and its neighbours (the rounded root minimises the distance of
        # the root, and the cube is monotone; compare distances of roots via
        # exact Fractions of the midpoints instead).
        from fractions import Fraction
        fx = Fraction(x32)
        up = np.nextafter(got, np.float32(np.inf))
        down = np.nextafter(got, np.float32(0))
        mid_up = (Fraction(float(got)) + Fraction(float(up))) / 2
        mid_down = (Fraction(float(got)) + Fraction(float(down))) / 2
        assert mid_down ** 3 < fx < mid_up ** 3, x32
    # Exact cubes come back exact.
    for k in range(1, 200):
        v = float(k) ** 3
        assert lib.ss2o_cbrt_rn(v) == float(k)
