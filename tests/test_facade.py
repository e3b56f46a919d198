"""The header-only C++ facade (include/ssimulacra2.h) compiled against
libvqm_hip.so: host contract on the CPU, the pipeline contract on the GPU."""
from __future__ import annotations

import os
import subprocess

import pytest

from conftest import ROOT

LIBDIR = os.path.join(ROOT, "vqm_amd")


@pytest.fixture(scope="module")
def facade_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_test")
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "facade_test.cc"), "-L", LIBDIR, "-lvqm_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def test_facade_host_contract(facade_exe):
    r = subprocess.run([facade_exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_facade_on_device(facade_exe):
    r = subprocess.run([facade_exe, "--gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
