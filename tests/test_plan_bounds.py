"""Builds and runs tests/cpp/plan_bounds_test.cc (host only): every address
the kernels compute from a work plan stays inside its buffer."""
from __future__ import annotations

import os
import subprocess
import tempfile

from conftest import ROOT


def test_plan_bounds():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "plan_bounds_test")
        subprocess.run(["g++", "-std=c++20", "-O1", "-Wall",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "vqm_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "plan_bounds_test.cc"),
                        os.path.join(ROOT, "vqm_amd", "csrc", "ss2_host.cc"),
                        os.path.join(ROOT, "vqm_amd", "csrc", "ss2_source.cc"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr


def test_range_allocator():
    """The workspace arena's first-fit bookkeeping against a byte-map model."""
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "range_allocator_test")
        subprocess.run(["g++", "-std=c++20", "-O2", "-Wall",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "vqm_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "range_allocator_test.cc"),
                        os.path.join(ROOT, "vqm_amd", "csrc", "ss2_host.cc"),
                        os.path.join(ROOT, "vqm_amd", "csrc", "ss2_source.cc"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr


def test_host_copy():
    """The threaded host copy of streamed images equals std::memcpy at the
    chunking thresholds, with concurrent callers and in a forked child."""
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "hostcopy_test")
        subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-pthread",
                        "-I", os.path.join(ROOT, "vqm_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "hostcopy_test.cc"),
                        os.path.join(ROOT, "vqm_amd", "csrc", "ss2_hostcopy.cc"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "failures: 0" in r.stdout
