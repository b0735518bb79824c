"""The C++ adapter header (include/phx_phare.hpp) compiles against the reference's own
ParticleArray and Field types and links against libphx.so (a compile-and-link check: running it
needs a GPU). Skipped where the PHARE source tree or the built library is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
LIB = os.path.join(ROOT, "phare_amd", "libphx.so")


@pytest.mark.skipif(not os.path.isdir(REF) or not os.path.exists(LIB) or not shutil.which("g++"),
                    reason="needs the PHARE source tree, libphx.so and g++")
def test_adapter_compiles_against_reference_types(tmp_path):
    out = str(tmp_path / "adapter_check")
    cmd = ["g++", "-std=c++20", "-O1", "-I" + REF, "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "adapter_check.cpp"), "-o", out,
           os.path.join(ROOT, "phare_amd", "libphx.so"), "-Wl,-rpath," + os.path.join(ROOT, "phare_amd"),
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.exists(out)
