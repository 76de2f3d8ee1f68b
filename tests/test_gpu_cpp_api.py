"""The C++ drop-in header (include/cufhe_gpu.hpp) exercised the way the
reference's own gate test is (reference test/test_gate_gpu.cc): a compiled
program, built in-tree by cufhe_amd/build.py, run on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_gate_api_truth_tables():
    exe = os.path.join(ROOT, "tests", "cpp", "test_gate_gpu.bin")
    assert os.path.exists(exe), "build with `python cufhe_amd/build.py` first"
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(p.stdout)
    assert p.returncode == 0 and "PASS" in p.stdout
