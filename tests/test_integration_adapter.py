"""The adapter of INTEGRATION.md section 1, compiled as written.

The C++ block of that section is extracted, compiled against a declaration
of the reference's backend seam (tests/cpp/seam/: the types its signatures
name and the pure virtual functions of internal::Ssimulacra2GpuPipeline,
ssimulacra2_internal.h:102-152) and this repository's C ABI, and linked with
libvqm_hip.so into a program that constructs the adapter through its
factory. The program is not run (that needs a device): it proves the class
implements every pure virtual and every C call it makes exists."""
from __future__ import annotations

import os
import re
import subprocess

from conftest import ROOT

MAIN = r'''
#include <memory>
#include "metrics/ssimulacra2/ssimulacra2_internal.h"
namespace vqm::internal {
std::unique_ptr<Ssimulacra2GpuPipeline> CreateHipSsimulacra2Pipeline(const Ssimulacra2Config& c);
}
int main(int argc, char**) {
    if (argc > 5) {  // never taken: the link is the test
        auto p = vqm::internal::CreateHipSsimulacra2Pipeline(vqm::Ssimulacra2Config{{64, 64}});
        p->RequireIdle();
    }
    return 0;
}
'''


def adapter_source() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text[text.index("## 1. "):text.index("## 2. ")]
    blocks = re.findall(r"```cpp\n(.*?)```", section, re.S)
    assert len(blocks) == 1, "section 1 holds one C++ block"
    return blocks[0]


def test_adapter_compiles_and_links(tmp_path):
    src = tmp_path / "hip.cc"
    src.write_text(adapter_source())
    main = tmp_path / "main.cc"
    main.write_text(MAIN)
    exe = tmp_path / "adapter"
    lib = os.path.join(ROOT, "vqm_amd")
    r = subprocess.run(["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "tests", "cpp", "seam"),
                        "-I", os.path.join(ROOT, "include"), str(src), str(main),
                        "-L", lib, "-lvqm_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
