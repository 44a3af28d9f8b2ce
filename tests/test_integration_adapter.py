"""The reference-side adapter (anacapa_amd/adapter/HipPathIntegrator.cpp) compiles against the
reference's own headers, i.e. it is a drop-in IIntegrator for RenderSession."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "anacapa")), reason="reference checkout absent")
def test_adapter_compiles_against_reference_headers():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    subprocess.check_call([cxx, "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(REF, "include"),
                           "-I", os.path.join(REF, "src"), "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "anacapa_amd", "adapter", "HipPathIntegrator.cpp")])
